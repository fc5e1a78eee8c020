// SPDX-License-Identifier: MIT
// gfx950 kernels of the generalized bulk-flow spectrum.
//
// Stage 1 (geometry_kernel): for every (realisation, k-hat, reference bubble
//   a) and every surface direction x = (1, sin th cos ph, sin th sin ph,
//   cos th), find the first bubble that collides with a along x and the
//   collision delay t_c - t_a, with the reference's exact decision rule
//   (rust/bubble-gw/src/many_bubbles/generalized_bulk_flow.rs:127-206,
//   1003-1018, 421-471).  Each lane owns one direction; the wavefront walks the
//   candidate list of a sorted by a lower bound of the delay and stops as soon
//   as that bound proves no later candidate can matter for any of its lanes
//   (a ballot), first to find the minimum-delay candidate, then to verify it.
//   A lane whose candidate fails verification, or that meets non-finite data,
//   re-runs the reference's own full scan, so decisions are bit-exact.  Each
//   ring's delays are then sorted in LDS (bitonic) for stage 2.
//
// Stage 2 (spectrum_kernel): per item, ring by ring, the phi integral
//   B_pm(t) of compute_b_integral (:481-586) is read off prefix sums over the
//   sorted delays -- the directions already collided at time t are exactly a
//   prefix of the sorted list -- and the quadrature weights of the A and C
//   integrals (:588-664, 775-873) are folded in.  Lanes own omegas and
//   accumulate C_pm(omega) = sum_t V_pm(t) exp(i omega psi(t)), with
//   psi = (t - z_a) - (t - t_a) cos th, in registers:
//     FP64 form:  one double sincos per (t, omega) term;
//     mixed form: psi arguments reduced in double; the phase of a block of
//                 R consecutive steps factorised as anchor * E_r with
//                 E_r = exp(i omega r dt (1 - cos th)); float sincos and
//                 float block sums, Kahan-compensated accumulation.
//
// Stage 3 (reduce_kernel): sum the per-bubble partial C in bubble order
//   (deterministic) and form S = 8 pi w^3 (|C+|^2 + |Cx|^2).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bgw_common.h"
#include "bgw_device.h"
#include "bgw_kernels.h"

namespace bgw {

namespace {

constexpr int kGeoThreads = 256;
constexpr int kGeoLdsCand = 128;
constexpr int kSortCap = 1024;
constexpr double kInf = __builtin_inf();

// A candidate in the k-hat frame, as staged in LDS.
struct alignas(16) RotCand {
  double d0, d1, d2, d3;  // rotated delta
  double dsq;             // stored delta_squared (unrotated)
  double m;               // scalar(rotated delta, rotated delta)
  double key;
  int32_t idx, flags;
};
static_assert(sizeof(RotCand) == 64, "");

__device__ __forceinline__ RotCand rotate_entry(const CandEntry& e, const double* R, int identity) {
  RotCand o;
  o.d0 = e.d[0];
  if (identity) {
    o.d1 = e.d[1];
    o.d2 = e.d[2];
    o.d3 = e.d[3];
  } else {
    o.d1 = (R[0] * e.d[1] + R[1] * e.d[2]) + R[2] * e.d[3];
    o.d2 = (R[3] * e.d[1] + R[4] * e.d[2]) + R[5] * e.d[3];
    o.d3 = (R[6] * e.d[1] + R[7] * e.d[2]) + R[8] * e.d[3];
  }
  o.dsq = e.dsq;
  o.m = lscalar(o.d0, o.d1, o.d2, o.d3, o.d0, o.d1, o.d2, o.d3);
  o.key = e.key;
  o.idx = e.idx;
  o.flags = e.flags;
  return o;
}

// check_collision_point (:1003-1018): true iff b reaches the point on a's
// wall along x before c does.  b.m / c.m are the recomputed norms.
__device__ __forceinline__ bool collision_point_ok(const RotCand& b, const RotCand& c, double x1,
                                                   double x2, double x3) {
  double dbx = lscalar(b.d0, b.d1, b.d2, b.d3, 1.0, x1, x2, x3);
  if (dbx <= 0.0) return false;
  double nba = b.m, nca = c.m;
  double v0 = b.d0 * nca - c.d0 * nba;
  double v1 = b.d1 * nca - c.d1 * nba;
  double v2 = b.d2 * nca - c.d2 * nba;
  double v3 = b.d3 * nca - c.d3 * nba;
  return lscalar(v0, v1, v2, v3, 1.0, x1, x2, x3) > 0.0;
}

__device__ __forceinline__ RotCand fetch_cand(int s, const RotCand* lds, int n_lds,
                                              const CandEntry* list, const double* R,
                                              int identity) {
  if (s < n_lds) return lds[s];
  return rotate_entry(list[s], R, identity);
}

// The reference's scan in original bubble order (:162-199).  Used for lanes
// whose fast path cannot certify its answer.
__device__ int reference_scan(const CandEntry* list, const int32_t* inv, int a, int n_int,
                              int n_tot, const double* R, int identity, double x1, double x2,
                              double x3) {
  double best = kInf;
  int bidx = -1;
  for (int b = 0; b < n_tot; ++b) {
    if (b < n_int && b == a) continue;
    RotCand eb = rotate_entry(list[inv[b]], R, identity);
    double dot = lscalar(eb.d0, eb.d1, eb.d2, eb.d3, 1.0, x1, x2, x3);
    if (dot <= 1e-10) continue;
    double dtv = eb.dsq / (2.0 * dot);
    if (dtv <= 0.0 || dtv >= best) continue;
    bool first = true;
    for (int c = 0; c < n_tot; ++c) {
      if ((c < n_int && c == a) || c == b) continue;
      RotCand ec = rotate_entry(list[inv[c]], R, identity);
      if (!collision_point_ok(eb, ec, x1, x2, x3)) {
        first = false;
        break;
      }
    }
    if (first) {
      bidx = b;
      best = dtv;
    }
  }
  return bidx;
}

__device__ __forceinline__ bool wave_any(bool p) { return __any(p ? 1 : 0) != 0; }

}  // namespace

// ---------------------------------------------------------------------------
// Stage 1: first collisions
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kGeoThreads) void geometry_kernel(GeoArgs g) {
  __shared__ RotCand cand_lds[kGeoLdsCand];
  __shared__ double sort_key[kSortCap];
  __shared__ int32_t sort_val[kSortCap];

  const int item_id = blockIdx.x;
  const Item it = g.items[item_id];
  const RealInfo ri = g.real[it.r];
  const int n_cand = ri.n_tot - 1;
  const CandEntry* list = g.cand + ri.cand_off + static_cast<int64_t>(it.a) * n_cand;
  const int32_t* inv = g.inv + ri.inv_off + static_cast<int64_t>(it.a) * ri.n_tot;
  double R[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) R[q] = g.identity ? 0.0 : g.rot[9 * it.k + q];
  const int identity = g.identity;

  const int n_lds = n_cand < kGeoLdsCand ? n_cand : kGeoLdsCand;
  for (int s = threadIdx.x; s < n_lds; s += blockDim.x) cand_lds[s] = rotate_entry(list[s], R, identity);
  __syncthreads();

  const int n_phi = g.n_phi;
  const int P = g.n_phi_pow2;
  const int G = kSortCap / P;  // rings per sort group
  const int64_t n_dir = static_cast<int64_t>(g.n_c) * n_phi;
  double* tau_out = g.tau + static_cast<int64_t>(item_id) * n_dir;
  uint16_t* perm_out = g.perm + static_cast<int64_t>(item_id) * n_dir;

  // Each wavefront takes a compact patch of ph rings x pw azimuths, so its
  // lanes share first colliders and stop their culled scans together.
  const int pw = P < 8 ? P : 8;
  const int ph = 64 / pw;
  const int patches_per_row = P / pw;
  for (int i0 = 0; i0 < g.n_c; i0 += G) {
    const int gcount = min(G, g.n_c - i0);
    const int nslot = gcount * P;
    const int n_patch = ((gcount + ph - 1) / ph) * patches_per_row;
    const int nwork_round = ((n_patch * 64 + blockDim.x - 1) / blockDim.x) * blockDim.x;
    for (int w = threadIdx.x; w < nwork_round; w += blockDim.x) {
      const int patch = w >> 6, lane = w & 63;
      const int prow = patch / patches_per_row, pcol = patch % patches_per_row;
      const int rring = prow * ph + lane / pw;  // ring within the group
      const int j = pcol * pw + lane % pw;
      const int ring = i0 + rring;
      const int slot = rring * P + j;           // sort-buffer position
      const bool in_group = patch < n_patch && rring < gcount;
      const bool active = in_group && j < n_phi;
      double x1 = 0.0, x2 = 0.0, x3 = 0.0;
      if (active) {
        const double st = g.sin_t[ring];
        x1 = st * g.cos_p[j];
        x2 = st * g.sin_p[j];
        x3 = g.cos_t[ring];
      }
      bool fb = active && (ri.nonfinite || !(isfinite(x1) && isfinite(x2) && isfinite(x3)));
      const bool act = active && !fb;

      // Pass 1: minimum collision delay over the candidates (dot > 1e-10,
      // delay > 0), ties to the lowest bubble index.  The list is sorted by a
      // lower bound `key` of the delay: once key * (1 - 1e-7) > best for every
      // lane, no later candidate can reach or tie a lane's best.
      double best = kInf;
      int bidx = 0x7fffffff, bpos = -1;
      for (int s = 0; s < n_cand; ++s) {
        const RotCand e = fetch_cand(s, cand_lds, n_lds, list, R, identity);
        const bool need = act && !(e.key * (1.0 - 1e-7) > best);
        if (!wave_any(need)) break;
        if (need) {
          const double dot = lscalar(e.d0, e.d1, e.d2, e.d3, 1.0, x1, x2, x3);
          if (!(dot <= 1e-10)) {
            const double dtv = e.dsq / (2.0 * dot);
            if (dtv != dtv) {
              fb = true;
            } else if (dtv > 0.0 && (dtv < best || (dtv == best && e.idx < bidx))) {
              best = dtv;
              bidx = e.idx;
              bpos = s;
            }
          }
        }
      }

      // Pass 2: verify the minimum against every other bubble c with the
      // reference's cross-multiplied test.  For a regular, non-grazing
      // candidate, bubbles with key > best * (1 + 1e-4) provably pass, so the
      // wave again walks a prefix of the list.
      const bool p2 = act && !fb && bpos >= 0;
      RotCand b0{};
      double thr = kInf;
      if (p2) {
        b0 = fetch_cand(bpos, cand_lds, n_lds, list, R, identity);
        const double scale = fabs(b0.d0) + sqrt(b0.d1 * b0.d1 + b0.d2 * b0.d2 + b0.d3 * b0.d3);
        const bool regular = (b0.flags & kCandRegular) != 0;
        const bool grazing = !(best <= 1e3 * scale);
        if (regular && !grazing) thr = best * (1.0 + 1e-4);
      }
      bool ok = true;
      for (int s = 0; s < n_cand; ++s) {
        const RotCand e = fetch_cand(s, cand_lds, n_lds, list, R, identity);
        const bool need = p2 && ok && !(e.key > thr);
        if (!wave_any(need)) break;
        if (need && s != bpos && !collision_point_ok(b0, e, x1, x2, x3)) ok = false;
      }
      if (p2 && !ok) fb = true;

      int result = -1;
      if (fb) {
        result = reference_scan(list, inv, it.a, ri.n_int, ri.n_tot, R, identity, x1, x2, x3);
      } else if (p2) {
        result = bidx;
      }
      double tau = kInf;
      if (active) {
        double dtab = 0.0;  // compute_delta_tab (:446-468)
        if (result >= 0) {
          const RotCand eb = fetch_cand(inv[result], cand_lds, n_lds, list, R, identity);
          const double dot = lscalar(eb.d0, eb.d1, eb.d2, eb.d3, 1.0, x1, x2, x3);
          dtab = fabs(dot) < 1e-10 ? 0.0 : eb.dsq / (2.0 * dot);
          if (dtab > 0.0) tau = dtab;  // AlreadyCollided needs delta_tab > 0 (:405)
        }
        const int64_t dir = static_cast<int64_t>(ring) * n_phi + j;
        if (g.first_out) {
          int code = -1;
          if (result >= 0) code = result < ri.n_int ? result : -(result - ri.n_int) - 2;
          g.first_out[static_cast<int64_t>(item_id) * n_dir + dir] = code;
        }
        if (g.dtab_out) g.dtab_out[static_cast<int64_t>(item_id) * n_dir + dir] = dtab;
      }
      if (in_group) {
        sort_key[slot] = tau;
        sort_val[slot] = j;
      }
    }
    __syncthreads();
    // Bitonic sort of every P-slot segment (one ring each), ascending in
    // (delay, phi index).  Padding slots carry phi index >= n_phi, so they
    // sort after the directions that never collide (delay +inf as well).
    for (int k = 2; k <= P; k <<= 1) {
      for (int jj = k >> 1; jj > 0; jj >>= 1) {
        for (int e = threadIdx.x; e < nslot; e += blockDim.x) {
          const int p = e ^ jj;
          if (p > e) {
            // direction from the index within the ring's segment, so every
            // segment ends ascending (k == P merges the whole segment)
            const bool up = ((e & (P - 1)) & k) == 0;
            const double ke = sort_key[e], kp = sort_key[p];
            const bool greater = ke > kp || (ke == kp && sort_val[e] > sort_val[p]);
            if (greater == up) {
              sort_key[e] = kp;
              sort_key[p] = ke;
              const int32_t t = sort_val[e];
              sort_val[e] = sort_val[p];
              sort_val[p] = t;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int e = threadIdx.x; e < nslot; e += blockDim.x) {
      const int ring = i0 + e / P;
      const int r = e % P;
      if (r < n_phi) {
        tau_out[static_cast<int64_t>(ring) * n_phi + r] = sort_key[e];
        perm_out[static_cast<int64_t>(ring) * n_phi + r] = static_cast<uint16_t>(sort_val[e]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage 2 helpers
// ---------------------------------------------------------------------------
namespace {

// Inclusive scan of arr[0..n) by one whole wavefront (reverse: suffix sums).
__device__ void wave_scan(double* arr, int n, bool reverse) {
  const int lane = threadIdx.x & 63;
  double carry = 0.0;
  for (int base = 0; base < n; base += 64) {
    const int e = base + lane;
    const int idx = reverse ? (n - 1 - e) : e;
    double v = e < n ? arr[idx] : 0.0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double u = __shfl_up(v, off, 64);
      if (lane >= off) v += u;
    }
    v += carry;
    if (e < n) arr[idx] = v;
    carry = __shfl(v, 63, 64);
  }
}

__device__ __forceinline__ int count_le(const double* sorted, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sorted[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double cutoff_weight(const SpecArgs& g, double t) {
  if (g.cutoff_kind == 0 || t < g.t_cut) return 1.0;
  const double d = t - g.t_cut;
  return exp(-(d * d) * g.inv_t0_sq);
}

constexpr int kTimeChunk = 256;

// Per-ring tables in LDS, laid out after each other, then the per-item time
// tables (when they fit) and the per-chunk phase inputs.
struct RingLds {
  double* tau;  // [n_phi] sorted delays
  double* gc;   // [n_phi] w_j cos 2phi_j in sorted order
  double* gs;   // [n_phi]
  double* uc;   // [n_phi + 1] suffix sums of gc (uncollided part)
  double* us;   // [n_phi + 1]
  double* cc;   // [NT][n_phi + 1] prefix sums of gc * tau^p (collided part)
  double* cs;   // [NT][n_phi + 1]
  double* cl;   // [NT][n_phi + 1] prefix sums of gc * tau^p ln tau (adjoint only)
  double* cls;  // [NT][n_phi + 1]
  double* tb;   // [n_t] K dt w_t cutoff(t) (t - t_a)^3, or null
  double* tf;   // [NT][n_t] coef_q (t - t_a)^-p_q, or null
  double* psi;  // [kTimeChunk] phase argument per time step
  double* v;    // [kTimeChunk][2 * n_sets] weights V+ / Vx per time step
  float* vf;    // the same weights in float (mixed form), aliases v
  int stride;   // n_phi + 1
};

__host__ __device__ inline size_t ring_lds_doubles(int n_phi, int nt_terms) {
  return 3 * static_cast<size_t>(n_phi) + 2 * static_cast<size_t>(n_phi + 1) +
         2 * static_cast<size_t>(nt_terms) * (n_phi + 1);
}

__host__ __device__ inline size_t time_table_doubles(int n_t, int nt_terms) {
  return static_cast<size_t>(n_t) * (1 + nt_terms);
}

__device__ RingLds carve(double* smem, int n_phi, int nt_terms, int n_t, bool time_tables) {
  RingLds L;
  L.stride = n_phi + 1;
  double* p = smem;
  L.tau = p; p += n_phi;
  L.gc = p; p += n_phi;
  L.gs = p; p += n_phi;
  L.uc = p; p += L.stride;
  L.us = p; p += L.stride;
  L.cc = p; p += static_cast<size_t>(nt_terms) * L.stride;
  L.cs = p; p += static_cast<size_t>(nt_terms) * L.stride;
  L.cl = nullptr;
  L.cls = nullptr;
  L.tb = nullptr;
  L.tf = nullptr;
  if (time_tables) {
    L.tb = p; p += n_t;
    L.tf = p; p += static_cast<size_t>(nt_terms) * n_t;
  }
  L.psi = p; p += kTimeChunk;
  L.v = p;
  L.vf = reinterpret_cast<float*>(p);
  return L;
}

// Loads ring i of an item and builds its prefix/suffix tables.  Returns the
// first time index with a contribution and whether the ring is skipped.
__device__ int prepare_ring(const SpecArgs& g, const RingLds& L, int64_t item_id, int i,
                            double t_a, const RealInfo& ri, bool* skip) {
  const int n_phi = g.n_phi;
  const int nterms = g.n_sets * g.n_terms;
  const int64_t base = (item_id * g.n_c + i) * static_cast<int64_t>(n_phi);
  for (int r = threadIdx.x; r < n_phi; r += blockDim.x) {
    const double tau = g.tau[base + r];
    const int j = g.perm[base + r];
    const double gcv = g.gcos[j], gsv = g.gsin[j];
    L.tau[r] = tau;
    L.gc[r] = gcv;
    L.gs[r] = gsv;
    L.uc[r] = gcv;
    L.us[r] = gsv;
    const bool fin = tau < kInf;
    for (int q = 0; q < nterms; ++q) {
      const double pw = fin ? pow(tau, g.pows[q]) : 0.0;
      L.cc[q * L.stride + r + 1] = gcv * pw;
      L.cs[q * L.stride + r + 1] = gsv * pw;
      if (L.cl) {
        const double lw = fin ? pw * log(tau) : 0.0;
        L.cl[q * L.stride + r + 1] = gcv * lw;
        L.cls[q * L.stride + r + 1] = gsv * lw;
      }
    }
  }
  if (threadIdx.x == 0) {
    L.uc[n_phi] = 0.0;
    L.us[n_phi] = 0.0;
    for (int q = 0; q < nterms; ++q) {
      L.cc[q * L.stride] = 0.0;
      L.cs[q * L.stride] = 0.0;
      if (L.cl) {
        L.cl[q * L.stride] = 0.0;
        L.cls[q * L.stride] = 0.0;
      }
    }
  }
  __syncthreads();
  const int n_waves = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6;
  const int n_prefix = 2 * nterms * (L.cl ? 2 : 1);
  const int n_arrays = 2 + n_prefix;
  for (int arr = wave; arr < n_arrays; arr += n_waves) {
    if (arr == 0) wave_scan(L.uc, n_phi, true);
    else if (arr == 1) wave_scan(L.us, n_phi, true);
    else {
      const int e = arr - 2;
      const int q = (e >> 1) % nterms;
      const bool logs = e >= 2 * nterms;
      double* base_arr = (e & 1) ? (logs ? L.cls : L.cs) : (logs ? L.cl : L.cc);
      wave_scan(base_arr + q * L.stride + 1, n_phi, false);
    }
  }
  __syncthreads();
  // A ring that never collides contributes B = 0.5 sin^2 dphi dt^3 sum_j w_j
  // cos(2 phi_j) (resp. sin), a full trapezoid period that vanishes up to
  // rounding for n_phi >= 4.  Such rings, and the time steps before a ring's
  // first collision, are skipped when their sum is at the rounding level.
  const double tau_min = L.tau[0];
  const bool noise = fabs(L.uc[0]) <= 1e-12 * n_phi && fabs(L.us[0]) <= 1e-12 * n_phi;
  const bool drop_early = g.skip_uncollided && noise && !g.has_damping;
  *skip = drop_early && !(tau_min < kInf);
  // first k with t_k - t_a > 0 (and >= tau_min when dropping early steps)
  int lo = 0, hi = g.n_t;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const double t = ri.t_begin + ri.t_step * static_cast<double>(mid);
    const double dta = t - t_a;
    const bool pred = dta > 0.0 && (!drop_early || dta >= tau_min);
    if (pred) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// K dt w_t cutoff(t) (t - t_a)^3 and coef_q (t - t_a)^-p_q for every time step of
// an item (ring-independent parts of the weights).  Zero where t <= t_a.
__device__ void prepare_item_tables(const SpecArgs& g, const RingLds& L, const Item& it,
                                    const RealInfo& ri) {
  const int nterms = g.n_sets * g.n_terms;
  for (int k = threadIdx.x; k < g.n_t; k += blockDim.x) {
    const double t = ri.t_begin + ri.t_step * static_cast<double>(k);
    const double dta = t - it.t_a;
    double base = 0.0;
    if (dta > 0.0) {
      const double wt = (k == 0 || k == g.n_t - 1) ? 0.5 : 1.0;
      base = g.kfac * ri.dt * wt * cutoff_weight(g, t) * ((dta * dta) * dta);
    }
    L.tb[k] = base;
    for (int q = 0; q < nterms; ++q)
      L.tf[static_cast<size_t>(q) * g.n_t + k] = dta > 0.0 ? g.coef[q] * pow(dta, -g.pows[q]) : 0.0;
  }
}

// V_pm(t) and psi(t) for the time steps [t0, t0 + len) of ring i.  FLOAT_V:
// the weights are stored in float (mixed form).
template <bool FLOAT_V>
__device__ void prepare_times(const SpecArgs& g, const RingLds& L, int i, int t0, int len,
                              const Item& it, const RealInfo& ri) {
  const int n_phi = g.n_phi;
  const int ns = g.n_sets;
  const double c_i = g.ring_c[i];
  const double rf = g.ring_fac[i];
  for (int l = threadIdx.x; l < kTimeChunk; l += blockDim.x) {
    double vv[2 * 8];
    for (int s = 0; s < 2 * ns; ++s) vv[s] = 0.0;
    double psi = 0.0;
    if (l < len) {
      const int k = t0 + l;
      const double t = ri.t_begin + ri.t_step * static_cast<double>(k);
      const double dta = t - it.t_a;
      psi = (t - it.z_a) - dta * c_i;
      if (dta > 0.0) {  // compute_b_integral returns zeros for dta <= 0 (:541-543)
        if (!g.has_damping) {
          const int cnt = count_le(L.tau, n_phi, dta);
          double base;
          if (L.tb) {
            base = L.tb[k] * rf;
          } else {
            const double wt = (k == 0 || k == g.n_t - 1) ? 0.5 : 1.0;
            base = g.kfac * ri.dt * wt * cutoff_weight(g, t) * ((dta * dta) * dta) * rf;
          }
          for (int s = 0; s < ns; ++s) {
            double bp = L.uc[cnt], bx = L.us[cnt];
            for (int kk = 0; kk < g.n_terms; ++kk) {
              const int q = s * g.n_terms + kk;
              const double f = L.tf ? L.tf[static_cast<size_t>(q) * g.n_t + k]
                                    : g.coef[q] * pow(dta, -g.pows[q]);
              bp = __builtin_fma(f, L.cc[q * L.stride + cnt], bp);
              bx = __builtin_fma(f, L.cs[q * L.stride + cnt], bx);
            }
            vv[2 * s + 0] = base * bp;
            vv[2 * s + 1] = base * bx;
          }
        } else {
          // Direct phi sum with the damping factor exp(-dta (1 - ratio) / width).
          const double wt = (k == 0 || k == g.n_t - 1) ? 0.5 : 1.0;
          const double base = g.kfac * ri.dt * wt * cutoff_weight(g, t) * ((dta * dta) * dta) * rf;
          for (int s = 0; s < ns; ++s) {
            double bp = 0.0, bx = 0.0;
            for (int r = 0; r < n_phi; ++r) {
              const double tau = L.tau[r];
              double f = 1.0;
              if (tau <= dta) {
                const double ratio = tau / dta;
                f = 0.0;
                for (int kk = 0; kk < g.n_terms; ++kk) {
                  const int q = s * g.n_terms + kk;
                  f += g.coef[q] * pow(ratio, g.pows[q]);
                }
                f *= exp((-dta) * (1.0 - ratio) / g.damping);
              }
              bp = __builtin_fma(L.gc[r], f, bp);
              bx = __builtin_fma(L.gs[r], f, bx);
            }
            vv[2 * s + 0] = base * bp;
            vv[2 * s + 1] = base * bx;
          }
        }
      }
    }
    L.psi[l] = psi;
    if constexpr (FLOAT_V) {
      for (int s = 0; s < 2 * ns; ++s) L.vf[l * 2 * ns + s] = static_cast<float>(vv[s]);
    } else {
      for (int s = 0; s < 2 * ns; ++s) L.v[l * 2 * ns + s] = vv[s];
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Stage 2: phase sums
// ---------------------------------------------------------------------------
// MODE 0: FP64, one double sincos per (t, omega) term (the parity form).
// MODE 1: FP64, block-factorised phases: exp(i w psi(t0 + r)) = A(t0) E_r with
//         E_r = exp(i w r dt (1 - cos th)); double E_r (recurrence), double
//         anchors A, double block sums.
// MODE 2: mixed, the same factorisation with the phase arguments reduced in
//         double and the sincos, E_r table, weights and block sums in float;
//         block results accumulated in float with Kahan compensation.
template <int NS, int OPL, int MODE, int RB>
__global__ __launch_bounds__(256) void spectrum_kernel(SpecArgs g) {
  using Acc = typename std::conditional<MODE == 2, float, double>::type;
  extern __shared__ double smem[];
  const int item_id = g.order ? g.order[blockIdx.x] : static_cast<int>(blockIdx.x);
  const Item it = g.items[item_id];
  const RealInfo ri = g.real[it.r];
  const RingLds L = carve(smem, g.n_phi, g.n_sets * g.n_terms, g.n_t, g.time_tables != 0);
  if (L.tb) prepare_item_tables(g, L, it, ri);
  __syncthreads();

  double om[OPL];
#pragma unroll
  for (int q = 0; q < OPL; ++q) {
    const int wi = threadIdx.x + q * blockDim.x;
    om[q] = wi < g.n_w ? g.omega[ri.w_off + wi] : 0.0;
  }
  // running sums (double), or float sums + Kahan compensations in MODE 2
  Acc acc[OPL][NS][4];
  float comp[MODE == 2 ? OPL : 1][MODE == 2 ? NS : 1][4];
#pragma unroll
  for (int q = 0; q < OPL; ++q)
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[q][s][c] = Acc(0);
#pragma unroll
  for (int q = 0; q < (MODE == 2 ? OPL : 1); ++q)
#pragma unroll
    for (int s = 0; s < (MODE == 2 ? NS : 1); ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) comp[q][s][c] = 0.0f;

  for (int i = 0; i < g.n_c; ++i) {
    if (g.ring_fac[i] == 0.0) continue;  // sin^2 theta = 0: B = 0 exactly
    bool skip;
    const int t_lo = prepare_ring(g, L, item_id, i, it.t_a, ri, &skip);
    if (skip || t_lo >= g.n_t) {
      __syncthreads();
      continue;
    }
    // E_r = exp(i omega r step (1 - cos th)), r < RB, by a double recurrence
    constexpr int kE = MODE == 0 ? 1 : RB;
    Acc er[OPL][kE], ei[OPL][kE];
    if constexpr (MODE != 0) {
      const double delta = ri.t_step * (1.0 - g.ring_c[i]);
#pragma unroll
      for (int q = 0; q < OPL; ++q) {
        double s1, c1;
        sincos_f64(om[q] * delta, &s1, &c1);
        double cr = 1.0, ci = 0.0;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          er[q][r] = static_cast<Acc>(cr);
          ei[q][r] = static_cast<Acc>(ci);
          const double nr = cr * c1 - ci * s1;
          const double ni = cr * s1 + ci * c1;
          cr = nr;
          ci = ni;
        }
      }
    }
    for (int t0 = t_lo; t0 < g.n_t; t0 += kTimeChunk) {
      const int len = min(kTimeChunk, g.n_t - t0);
      prepare_times<MODE == 2>(g, L, i, t0, len, it, ri);
      __syncthreads();
      if constexpr (MODE == 0) {
        for (int l = 0; l < len; ++l) {
          const double psi = L.psi[l];
          const double* v = L.v + l * 2 * NS;
#pragma unroll
          for (int q = 0; q < OPL; ++q) {
            double sn, cs;
            sincos_f64(om[q] * psi, &sn, &cs);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
              const double vp = v[2 * s], vx = v[2 * s + 1];
              acc[q][s][0] = __builtin_fma(vp, cs, acc[q][s][0]);
              acc[q][s][1] = __builtin_fma(vp, sn, acc[q][s][1]);
              acc[q][s][2] = __builtin_fma(vx, cs, acc[q][s][2]);
              acc[q][s][3] = __builtin_fma(vx, sn, acc[q][s][3]);
            }
          }
        }
      } else {
        for (int l0 = 0; l0 < len; l0 += RB) {
          const double psi0 = L.psi[l0];
          Acc bs[OPL][NS][4];
#pragma unroll
          for (int q = 0; q < OPL; ++q)
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
              for (int c = 0; c < 4; ++c) bs[q][s][c] = Acc(0);
#pragma unroll
          for (int r = 0; r < RB; ++r) {
            const Acc* v;  // zero-padded past len
            if constexpr (MODE == 2) v = L.vf + (l0 + r) * 2 * NS;
            else v = L.v + (l0 + r) * 2 * NS;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
              const Acc vp = v[2 * s];
              const Acc vx = v[2 * s + 1];
#pragma unroll
              for (int q = 0; q < OPL; ++q) {
                bs[q][s][0] = fma(vp, er[q][r], bs[q][s][0]);
                bs[q][s][1] = fma(vp, ei[q][r], bs[q][s][1]);
                bs[q][s][2] = fma(vx, er[q][r], bs[q][s][2]);
                bs[q][s][3] = fma(vx, ei[q][r], bs[q][s][3]);
              }
            }
          }
#pragma unroll
          for (int q = 0; q < OPL; ++q) {
            Acc as, ac;
            if constexpr (MODE == 2) {
              sincos_f64_to_f32(om[q] * psi0, &as, &ac);
            } else {
              sincos_f64(om[q] * psi0, &as, &ac);
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
              // (ac + i as) * (b0 + i b1), likewise for the cross polarisation
              const Acc p0 = fma(ac, bs[q][s][0], -as * bs[q][s][1]);
              const Acc p1 = fma(ac, bs[q][s][1], as * bs[q][s][0]);
              const Acc p2 = fma(ac, bs[q][s][2], -as * bs[q][s][3]);
              const Acc p3 = fma(ac, bs[q][s][3], as * bs[q][s][2]);
              if constexpr (MODE == 2) {
                kahan_add(acc[q][s][0], comp[q][s][0], p0);
                kahan_add(acc[q][s][1], comp[q][s][1], p1);
                kahan_add(acc[q][s][2], comp[q][s][2], p2);
                kahan_add(acc[q][s][3], comp[q][s][3], p3);
              } else {
                acc[q][s][0] += p0;
                acc[q][s][1] += p1;
                acc[q][s][2] += p2;
                acc[q][s][3] += p3;
              }
            }
          }
        }
      }
      __syncthreads();
    }
  }
  // partial[item][pm][s][w] complex
  const int64_t nq = static_cast<int64_t>(g.n_sets_out) * g.n_w;
  double* out = g.partial + static_cast<int64_t>(item_id) * 2 * nq * 2;
#pragma unroll
  for (int q = 0; q < OPL; ++q) {
    const int wi = threadIdx.x + q * blockDim.x;
    if (wi >= g.n_w) continue;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (MODE == 2) {
          c[j] = static_cast<double>(acc[q][s][j]) - static_cast<double>(comp[q][s][j]);
        } else {
          c[j] = static_cast<double>(acc[q][s][j]);
        }
      }
      const int64_t ip = (static_cast<int64_t>(g.set0 + s) * g.n_w + wi) * 2;
      const int64_t ix = ((nq + static_cast<int64_t>(g.set0 + s) * g.n_w) + wi) * 2;
      out[ip] = c[0];
      out[ip + 1] = c[1];
      out[ix] = c[2];
      out[ix + 1] = c[3];
    }
  }
}

// ---------------------------------------------------------------------------
// Stage 3: deterministic sum over reference bubbles, spectrum
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_kernel(RedArgs g) {
  const int grp = blockIdx.x;
  const int i0 = g.group_off[grp], i1 = g.group_off[grp + 1];
  const int64_t nq = static_cast<int64_t>(g.n_sets) * g.n_w;
  const RealInfo ri = g.real[g.group_r[grp]];
  const double eight_pi = 8.0 * 3.14159265358979323846;
  for (int64_t q = threadIdx.x; q < nq; q += blockDim.x) {
    double pr = 0.0, pi = 0.0, xr = 0.0, xi = 0.0;
    for (int item = i0; item < i1; ++item) {
      const double* p = g.partial + static_cast<int64_t>(item) * 2 * nq * 2;
      pr += p[q * 2];
      pi += p[q * 2 + 1];
      xr += p[(nq + q) * 2];
      xi += p[(nq + q) * 2 + 1];
    }
    if (g.c_out) {
      double* c = g.c_out + static_cast<int64_t>(grp) * 2 * nq * 2;
      c[q * 2] = pr;
      c[q * 2 + 1] = pi;
      c[(nq + q) * 2] = xr;
      c[(nq + q) * 2 + 1] = xi;
    }
    if (g.s_out) {
      const double w = g.omega[ri.w_off + (q % g.n_w)];
      const double ap = hypot(pr, pi), ax = hypot(xr, xi);
      g.s_out[static_cast<int64_t>(grp) * nq + q] = eight_pi * pow(w, 3.0) * (ap * ap + ax * ax);
    }
  }
}

// ---------------------------------------------------------------------------
// Stage 2, time-resolved: the integrand of the time integral
// ---------------------------------------------------------------------------
// compute_c_integrand_fixed_bubble (generalized_bulk_flow.rs:666-767): per time
// step t the sum over rings of V_pm(t, i) exp(i w psi(t, i)) (no time cutoff,
// zero for t <= t_a).  Lanes own omegas; rings run in order and each lane adds
// its ring contribution into its own entries of the item's slab
//   partial[item][pm][s][t][w]   (complex, sets of the whole request)
// so no two lanes touch the same entry.  FP64 sincos, as the FP64 form.
template <int NS, int OPL>
__global__ __launch_bounds__(256) void integrand_kernel(SpecArgs g) {
  extern __shared__ double smem[];
  const int item_id = g.order ? g.order[blockIdx.x] : static_cast<int>(blockIdx.x);
  const Item it = g.items[item_id];
  const RealInfo ri = g.real[it.r];
  const RingLds L = carve(smem, g.n_phi, g.n_sets * g.n_terms, g.n_t, false);
  const int64_t slab = static_cast<int64_t>(g.n_t) * g.n_w;  // one (pm, s)
  double* out = g.partial + static_cast<int64_t>(blockIdx.x) * 2 * g.n_sets_out * slab * 2;
  double om[OPL];
#pragma unroll
  for (int q = 0; q < OPL; ++q) {
    const int wi = threadIdx.x + q * blockDim.x;
    om[q] = wi < g.n_w ? g.omega[ri.w_off + wi] : 0.0;
    if (wi >= g.n_w) continue;
    for (int pm = 0; pm < 2; ++pm)
      for (int s = 0; s < NS; ++s) {
        double* o = out + ((pm * g.n_sets_out + g.set0 + s) * slab + wi) * 2;
        for (int k = 0; k < g.n_t; ++k) {
          o[static_cast<int64_t>(k) * g.n_w * 2] = 0.0;
          o[static_cast<int64_t>(k) * g.n_w * 2 + 1] = 0.0;
        }
      }
  }
  for (int i = 0; i < g.n_c; ++i) {
    if (g.ring_fac[i] == 0.0) continue;
    bool skip;
    const int t_lo = prepare_ring(g, L, item_id, i, it.t_a, ri, &skip);
    if (skip || t_lo >= g.n_t) {
      __syncthreads();
      continue;
    }
    for (int t0 = t_lo; t0 < g.n_t; t0 += kTimeChunk) {
      const int len = min(kTimeChunk, g.n_t - t0);
      prepare_times<false>(g, L, i, t0, len, it, ri);
      __syncthreads();
      for (int l = 0; l < len; ++l) {
        const double psi = L.psi[l];
        const double* v = L.v + l * 2 * NS;
        const int64_t kofs = static_cast<int64_t>(t0 + l) * g.n_w;
#pragma unroll
        for (int q = 0; q < OPL; ++q) {
          const int wi = threadIdx.x + q * blockDim.x;
          if (wi >= g.n_w) continue;
          double sn, cs;
          sincos_f64(om[q] * psi, &sn, &cs);
#pragma unroll
          for (int s = 0; s < NS; ++s) {
            double* op = out + ((g.set0 + s) * slab + kofs + wi) * 2;
            double* ox = out + ((g.n_sets_out + g.set0 + s) * slab + kofs + wi) * 2;
            const double vp = v[2 * s], vx = v[2 * s + 1];
            op[0] = __builtin_fma(vp, cs, op[0]);
            op[1] = __builtin_fma(vp, sn, op[1]);
            ox[0] = __builtin_fma(vx, cs, ox[0]);
            ox[1] = __builtin_fma(vx, sn, ox[1]);
          }
        }
      }
      __syncthreads();
    }
  }
}

// acc[e] = (first ? 0 : acc[e]) + sum over the chunk's items, in item order.
__global__ __launch_bounds__(256) void accumulate_kernel(const double* __restrict__ partial,
                                                         int n_chunk_items, int64_t n_elem,
                                                         int first, double* __restrict__ acc) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n_elem;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double sum = first ? 0.0 : acc[e];
    for (int it = 0; it < n_chunk_items; ++it) sum += partial[static_cast<int64_t>(it) * n_elem + e];
    acc[e] = sum;
  }
}

// ---------------------------------------------------------------------------
// Stage 2 adjoint: gradient of a real loss with respect to the scaling
// parameters (coefficients and powers)
// ---------------------------------------------------------------------------
// C_pm,s(w) = sum_{a,i,t} V_pm,s(t, i) exp(i w psi(t, i)) is linear in the
// weights V, so for a loss L with cotangent G = dL/dRe C + i dL/dIm C
//   dL/dtheta = sum_{a,i,t} sum_{pm,s} W_pm,s(t, i) dV_pm,s(t, i)/dtheta,
//   W_pm,s(t, i) = sum_w Re(conj(G_pm,s(w)) exp(i w psi(t, i))),
// i.e. the phase sum transposed (over omega instead of over t).  The masks
// are recomputed by the geometry kernel of the same call.  V depends on the
// parameters only through the directions already collided at t (tau_j <= dta,
// a prefix of the sorted ring):
//   V       = base (U[cnt] + sum_q coef_q dta^-p_q P_q[cnt]),  P_q: prefix of g tau^p_q
//   dV/dc_q = base dta^-p_q P_q[cnt]
//   dV/dp_q = base coef_q dta^-p_q (Lg_q[cnt] - ln(dta) P_q[cnt]),  Lg_q: prefix of g tau^p_q ln tau
// so only time steps with dta >= min tau contribute.  With a damping width the
// same derivatives are summed directly over the collided directions.
//
// Lanes own (block of RB consecutive steps, omega class m mod M); the phase of
// step t0 + r is exp(i w psi(t0)) E^r with E = exp(i w step (1 - cos th)), by
// an FP64 recurrence.  The M partial sums of a block are combined by
// cross-lane shuffles in a fixed order and every lane keeps its own gradient
// partials, summed in thread order at the end: the result is deterministic.
namespace {

constexpr int kAdjChunk = 1024;  // time steps of W held in LDS at once

struct AdjLds {
  RingLds ring;  // tau, gc, gs, uc, us, cc, cs, cl, cls
  double* om;    // [n_w]
  double* G;     // [n_w][2 * NS][2] cotangent, index (pm * NS + s)
  double* E;     // [n_w][2] exp(i w step (1 - cos th)) of the current ring
  double* W;     // [kAdjChunk][2 * NS]
  double* acc;   // [2 * nterms][blockDim] per-lane gradient partial sums
};

__device__ AdjLds carve_adj(double* smem, int n_phi, int nterms, int n_w, int ns) {
  AdjLds A;
  RingLds& L = A.ring;
  L.stride = n_phi + 1;
  double* p = smem;
  L.tau = p; p += n_phi;
  L.gc = p; p += n_phi;
  L.gs = p; p += n_phi;
  L.uc = p; p += L.stride;
  L.us = p; p += L.stride;
  L.cc = p; p += static_cast<size_t>(nterms) * L.stride;
  L.cs = p; p += static_cast<size_t>(nterms) * L.stride;
  L.cl = p; p += static_cast<size_t>(nterms) * L.stride;
  L.cls = p; p += static_cast<size_t>(nterms) * L.stride;
  L.tb = nullptr;
  L.tf = nullptr;
  L.psi = nullptr;
  L.v = nullptr;
  L.vf = nullptr;
  A.om = p; p += n_w;
  A.G = p; p += static_cast<size_t>(n_w) * 4 * ns;
  A.E = p; p += static_cast<size_t>(n_w) * 2;
  A.W = p; p += static_cast<size_t>(kAdjChunk) * 2 * ns;
  A.acc = p;
  return A;
}

}  // namespace

size_t adjoint_lds_bytes(const SpecArgs& a, int threads) {
  const size_t nterms = static_cast<size_t>(a.n_sets) * a.n_terms;
  const size_t stride = static_cast<size_t>(a.n_phi) + 1;
  const size_t doubles = 3 * static_cast<size_t>(a.n_phi) + 2 * stride + 4 * nterms * stride +
                         static_cast<size_t>(a.n_w) * (1 + 4 * a.n_sets + 2) +
                         static_cast<size_t>(kAdjChunk) * 2 * a.n_sets +
                         2 * nterms * static_cast<size_t>(threads);
  return doubles * sizeof(double);
}

template <int NS, int RB>
__global__ __launch_bounds__(256) void adjoint_kernel(SpecArgs g) {
  extern __shared__ double smem[];
  const int item_id = g.order ? g.order[blockIdx.x] : static_cast<int>(blockIdx.x);
  const Item it = g.items[item_id];
  const RealInfo ri = g.real[it.r];
  const int nterms = NS * g.n_terms;
  const int bd = blockDim.x;
  const int tid = threadIdx.x;
  const AdjLds A = carve_adj(smem, g.n_phi, nterms, g.n_w, NS);
  const RingLds& L = A.ring;

  // cotangent of this item's (realisation, k-hat) group, sets [set0, set0 + NS)
  const int64_t nq = static_cast<int64_t>(g.n_sets_out) * g.n_w;
  const double* gsrc = g.grad_c + static_cast<int64_t>(it.group) * 2 * nq * 2;
  for (int j = tid; j < g.n_w; j += bd) {
    A.om[j] = g.omega[ri.w_off + j];
    for (int pm = 0; pm < 2; ++pm)
      for (int s = 0; s < NS; ++s) {
        const int64_t src = (pm * nq + static_cast<int64_t>(g.set0 + s) * g.n_w + j) * 2;
        double* dst = A.G + (static_cast<int64_t>(j) * 2 * NS + pm * NS + s) * 2;
        dst[0] = gsrc[src];
        dst[1] = gsrc[src + 1];
      }
  }
  for (int e = 0; e < 2 * nterms; ++e) A.acc[e * bd + tid] = 0.0;
  __syncthreads();

  for (int i = 0; i < g.n_c; ++i) {
    const double rf = g.ring_fac[i];
    if (rf == 0.0) continue;  // sin^2 theta = 0: B = 0 exactly
    bool unused;
    (void)prepare_ring(g, L, item_id, i, it.t_a, ri, &unused);
    const double tau_min = L.tau[0];
    if (!(tau_min < kInf)) {  // ring never collides: no parameter dependence
      __syncthreads();
      continue;
    }
    int lo = 0, hi = g.n_t;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const double t = ri.t_begin + ri.t_step * static_cast<double>(mid);
      const double dta = t - it.t_a;
      if (dta > 0.0 && dta >= tau_min) hi = mid;
      else lo = mid + 1;
    }
    const int t_lo = lo;
    if (t_lo >= g.n_t) {
      __syncthreads();
      continue;
    }
    const double c_i = g.ring_c[i];
    const double delta = ri.t_step * (1.0 - c_i);
    for (int j = tid; j < g.n_w; j += bd) {
      double sn, cs;
      sincos_f64(A.om[j] * delta, &sn, &cs);
      A.E[2 * j] = cs;
      A.E[2 * j + 1] = sn;
    }
    __syncthreads();

    for (int c0 = t_lo; c0 < g.n_t; c0 += kAdjChunk) {
      const int len = min(kAdjChunk, g.n_t - c0);
      const int n_blk = (len + RB - 1) / RB;
      int mshift = 6;  // M = 2^mshift omega classes per block, M <= 64
      while (mshift > 0 && (n_blk << mshift) > bd) --mshift;
      const int M = 1 << mshift;
      const int n_units = n_blk << mshift;
      const int n_iter = (n_units + bd - 1) / bd;
      for (int itn = 0; itn < n_iter; ++itn) {
        const int u = itn * bd + tid;
        const int blk = u >> mshift;
        const int m = u & (M - 1);
        const bool live = u < n_units;
        double w[RB][2 * NS];
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
          for (int c = 0; c < 2 * NS; ++c) w[r][c] = 0.0;
        if (live) {
          const int k0 = c0 + blk * RB;
          const double t0 = ri.t_begin + ri.t_step * static_cast<double>(k0);
          const double psi0 = (t0 - it.z_a) - (t0 - it.t_a) * c_i;
          for (int j = m; j < g.n_w; j += M) {
            double zi, zr;
            sincos_f64(A.om[j] * psi0, &zi, &zr);
            const double er = A.E[2 * j], ei = A.E[2 * j + 1];
            double gr[2 * NS], gi[2 * NS];
#pragma unroll
            for (int c = 0; c < 2 * NS; ++c) {
              gr[c] = A.G[(j * 2 * NS + c) * 2];
              gi[c] = A.G[(j * 2 * NS + c) * 2 + 1];
            }
#pragma unroll
            for (int r = 0; r < RB; ++r) {
#pragma unroll
              for (int c = 0; c < 2 * NS; ++c)
                w[r][c] = __builtin_fma(gr[c], zr, __builtin_fma(gi[c], zi, w[r][c]));
              const double nr = zr * er - zi * ei;
              const double ni = zr * ei + zi * er;
              zr = nr;
              zi = ni;
            }
          }
        }
        // sum the M omega classes of each block (adjacent lanes of one wave)
        for (int off = M >> 1; off > 0; off >>= 1) {
#pragma unroll
          for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < 2 * NS; ++c) w[r][c] += __shfl_xor(w[r][c], off, 64);
        }
        if (live && m == 0) {
          const int l0 = blk * RB;
#pragma unroll
          for (int r = 0; r < RB; ++r) {
            if (l0 + r < len) {
#pragma unroll
              for (int c = 0; c < 2 * NS; ++c) A.W[(l0 + r) * 2 * NS + c] = w[r][c];
            }
          }
        }
      }
      __syncthreads();
      // contraction with dV/dtheta
      for (int l = tid; l < len; l += bd) {
        const int k = c0 + l;
        const double t = ri.t_begin + ri.t_step * static_cast<double>(k);
        const double dta = t - it.t_a;  // >= tau_min > 0
        const double wt = (k == 0 || k == g.n_t - 1) ? 0.5 : 1.0;
        const double base = g.kfac * ri.dt * wt * cutoff_weight(g, t) * ((dta * dta) * dta) * rf;
        const double* Wt = A.W + l * 2 * NS;
        const int cnt = count_le(L.tau, g.n_phi, dta);
        if (!g.has_damping) {
          const double ldta = log(dta);
          for (int s = 0; s < NS; ++s) {
            const double wp = Wt[s], wx = Wt[NS + s];
            for (int kk = 0; kk < g.n_terms; ++kk) {
              const int q = s * g.n_terms + kk;
              const double dmp = pow(dta, -g.pows[q]);
              const double pp = L.cc[q * L.stride + cnt], px = L.cs[q * L.stride + cnt];
              const double lp = L.cl[q * L.stride + cnt], lx = L.cls[q * L.stride + cnt];
              A.acc[(2 * q) * bd + tid] += base * dmp * (wp * pp + wx * px);
              A.acc[(2 * q + 1) * bd + tid] +=
                  base * g.coef[q] * dmp * (wp * (lp - ldta * pp) + wx * (lx - ldta * px));
            }
          }
        } else {
          for (int s = 0; s < NS; ++s) {
            const double wp = Wt[s], wx = Wt[NS + s];
            for (int kk = 0; kk < g.n_terms; ++kk) {
              const int q = s * g.n_terms + kk;
              double dcp = 0.0, dcx = 0.0, dpp = 0.0, dpx = 0.0;
              for (int r = 0; r < cnt; ++r) {
                const double ratio = L.tau[r] / dta;
                const double d = exp((-dta) * (1.0 - ratio) / g.damping);
                const double rp = pow(ratio, g.pows[q]) * d;
                const double rl = rp * log(ratio);
                dcp = __builtin_fma(L.gc[r], rp, dcp);
                dcx = __builtin_fma(L.gs[r], rp, dcx);
                dpp = __builtin_fma(L.gc[r], rl, dpp);
                dpx = __builtin_fma(L.gs[r], rl, dpx);
              }
              A.acc[(2 * q) * bd + tid] += base * (wp * dcp + wx * dcx);
              A.acc[(2 * q + 1) * bd + tid] += base * g.coef[q] * (wp * dpp + wx * dpx);
            }
          }
        }
      }
      __syncthreads();
    }
  }
  // grad_out[item][which][set][term], which 0: d/dcoef, 1: d/dpow
  for (int e = tid; e < 2 * nterms; e += bd) {
    const int q = e >> 1, which = e & 1;
    const int s = q / g.n_terms, kk = q % g.n_terms;
    double sum = 0.0;
    for (int l = 0; l < bd; ++l) sum += A.acc[e * bd + l];
    const int64_t o = ((static_cast<int64_t>(item_id) * 2 + which) * g.n_sets_out + g.set0 + s) *
                          g.n_terms + kk;
    g.grad_out[o] = sum;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_geometry(const GeoArgs& a, int n_items, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(geometry_kernel, dim3(n_items), dim3(kGeoThreads), 0, s, a);
  return hipGetLastError();
}

bool spectrum_time_tables(const SpecArgs& a) {
  return time_table_doubles(a.n_t, a.n_sets * a.n_terms) * sizeof(double) <= 48 * 1024;
}

size_t spectrum_lds_bytes(const SpecArgs& a, int /*mode*/) {
  const int nterms = a.n_sets * a.n_terms;
  size_t doubles = ring_lds_doubles(a.n_phi, nterms) + kTimeChunk +
                   static_cast<size_t>(kTimeChunk) * 2 * a.n_sets;
  if (a.time_tables) doubles += time_table_doubles(a.n_t, nterms);
  return doubles * sizeof(double);
}

namespace {
template <int NS, int OPL, int MODE, int RB>
hipError_t launch_spec_t(const SpecArgs& a, int n_items, int threads, size_t lds, hipStream_t s) {
  auto k = spectrum_kernel<NS, OPL, MODE, RB>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_items), dim3(threads), lds, s, a);
  return hipGetLastError();
}

template <int NS, int MODE, int RB>
hipError_t launch_spec_opl(const SpecArgs& a, int n_items, int opl, int threads, size_t lds,
                           hipStream_t s) {
  switch (opl) {
    case 1: return launch_spec_t<NS, 1, MODE, RB>(a, n_items, threads, lds, s);
    case 2: return launch_spec_t<NS, 2, MODE, RB>(a, n_items, threads, lds, s);
    default: return launch_spec_t<NS, 4, MODE, RB>(a, n_items, threads, lds, s);
  }
}

template <int NS>
hipError_t launch_spec_ns(const SpecArgs& a, int n_items, int mode, int opl, int threads,
                          size_t lds, hipStream_t s) {
  if (mode == 0) return launch_spec_opl<NS, 0, 16>(a, n_items, opl, threads, lds, s);
  const bool rb8 = a.mixed_block == 8;
  if (mode == 1)
    return rb8 ? launch_spec_opl<NS, 1, 8>(a, n_items, opl, threads, lds, s)
               : launch_spec_opl<NS, 1, 16>(a, n_items, opl, threads, lds, s);
  return rb8 ? launch_spec_opl<NS, 2, 8>(a, n_items, opl, threads, lds, s)
             : launch_spec_opl<NS, 2, 16>(a, n_items, opl, threads, lds, s);
}
}  // namespace

hipError_t launch_spectrum(const SpecArgs& a, int n_items, int mode, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  if (mode < 0 || mode > 2) return hipErrorInvalidValue;
  // Lanes own omegas: threads = 64..256 (a multiple of 64), up to 4 omegas
  // per lane.  a.threads (0 = automatic) is the tunable choice.
  int threads = a.threads;
  if (threads <= 0) {
    threads = ((a.n_w + 63) / 64) * 64;
    if (threads > 256) threads = 256;
  }
  if (threads < 64 || threads > 256 || threads % 64) return hipErrorInvalidValue;
  const int opl_needed = (a.n_w + threads - 1) / threads;
  if (opl_needed > 4) return hipErrorInvalidValue;
  const int opl = opl_needed <= 1 ? 1 : (opl_needed <= 2 ? 2 : 4);
  const size_t lds = spectrum_lds_bytes(a, mode);
  switch (a.n_sets) {
    case 1: return launch_spec_ns<1>(a, n_items, mode, opl, threads, lds, s);
    case 2: return launch_spec_ns<2>(a, n_items, mode, opl, threads, lds, s);
    case 3: return launch_spec_ns<3>(a, n_items, mode, opl, threads, lds, s);
    case 4: return launch_spec_ns<4>(a, n_items, mode, opl, threads, lds, s);
    case 5: return launch_spec_ns<5>(a, n_items, mode, opl, threads, lds, s);
    case 6: return launch_spec_ns<6>(a, n_items, mode, opl, threads, lds, s);
    default: return hipErrorInvalidValue;
  }
}

namespace {
template <int NS, int RB>
hipError_t launch_adj_t(const SpecArgs& a, int n_items, int threads, size_t lds, hipStream_t s) {
  auto k = adjoint_kernel<NS, RB>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_items), dim3(threads), lds, s, a);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_adjoint(const SpecArgs& a, int n_items, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  const int threads = kAdjointThreads;
  const size_t lds = adjoint_lds_bytes(a, threads);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  switch (a.n_sets) {
    case 1: return launch_adj_t<1, 16>(a, n_items, threads, lds, s);
    case 2: return launch_adj_t<2, 8>(a, n_items, threads, lds, s);
    default: return hipErrorInvalidValue;
  }
}

namespace {
template <int NS, int OPL>
hipError_t launch_int_t(const SpecArgs& a, int n_items, int threads, size_t lds, hipStream_t s) {
  auto k = integrand_kernel<NS, OPL>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_items), dim3(threads), lds, s, a);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_int_ns(const SpecArgs& a, int n_items, int opl, int threads, size_t lds,
                         hipStream_t s) {
  switch (opl) {
    case 1: return launch_int_t<NS, 1>(a, n_items, threads, lds, s);
    case 2: return launch_int_t<NS, 2>(a, n_items, threads, lds, s);
    default: return launch_int_t<NS, 4>(a, n_items, threads, lds, s);
  }
}
}  // namespace

hipError_t launch_integrand(const SpecArgs& a, int n_items, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  int threads = ((a.n_w + 63) / 64) * 64;
  if (threads > 256) threads = 256;
  const int opl_needed = (a.n_w + threads - 1) / threads;
  if (opl_needed > 4) return hipErrorInvalidValue;
  const int opl = opl_needed <= 1 ? 1 : (opl_needed <= 2 ? 2 : 4);
  SpecArgs b = a;
  b.time_tables = 0;
  const size_t lds = spectrum_lds_bytes(b, 0);
  switch (a.n_sets) {
    case 1: return launch_int_ns<1>(b, n_items, opl, threads, lds, s);
    case 2: return launch_int_ns<2>(b, n_items, opl, threads, lds, s);
    case 3: return launch_int_ns<3>(b, n_items, opl, threads, lds, s);
    case 4: return launch_int_ns<4>(b, n_items, opl, threads, lds, s);
    case 5: return launch_int_ns<5>(b, n_items, opl, threads, lds, s);
    case 6: return launch_int_ns<6>(b, n_items, opl, threads, lds, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_accumulate(const double* partial, int n_chunk_items, int64_t n_elem, int first,
                             double* acc, hipStream_t s) {
  if (n_elem == 0) return hipSuccess;
  int64_t blocks = (n_elem + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(accumulate_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                     partial, n_chunk_items, n_elem, first, acc);
  return hipGetLastError();
}

hipError_t launch_reduce(const RedArgs& a, int n_groups, hipStream_t s) {
  if (n_groups == 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_kernel, dim3(n_groups), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace bgw
