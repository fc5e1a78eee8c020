// SPDX-License-Identifier: MIT
// Device-side stages shared by the kernel translation units: the first-
// collision geometry of one item and the phase sums of one item.
#pragma once

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
constexpr int kDirPerLane = 1;  // surface directions per lane in the first-collider scan
constexpr double kInf = __builtin_inf();

// A candidate in the k-hat frame, as staged in LDS.
struct alignas(16) RotCand {
  double d0, d1, d2, d3;  // rotated delta
  double hdsq;            // half the stored delta_squared (unrotated): the delay
                          //   dsq / (2 dot) is hdsq / dot, the same double
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
  o.hdsq = 0.5 * e.dsq;
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
    double dtv = eb.hdsq / dot;
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

// The predicate's ballot stays a scalar mask (no per-lane materialisation).
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

}  // namespace


namespace {

// Exchange with lane ^ m for m = 1, 2, 4, 8 (partners in the same 16-lane
// row) by DPP moves, which stay in the VALU instead of going through the LDS
// crossbar as ds_bpermute does: xor 1 and 2 are quad permutations
// (quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E), xor 8 rotates the row by 8
// (row_ror:8 = 0x128), and xor 4 fills banks 0 and 2 from 4 lanes up
// (row_shl:4 = 0x104, bank mask 0101) and banks 1 and 3 from 4 lanes down
// (row_shr:4 = 0x114, bank mask 1010).  Every lane of the row must be active.
__device__ __forceinline__ int xor_lanes(int v, int m) {
  if (m == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
  if (m == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
  if (m == 4) {
    const int t = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xF, 0x5, false);
    return __builtin_amdgcn_update_dpp(t, v, 0x114, 0xF, 0xA, false);
  }
  return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);
}

__device__ __forceinline__ double xor_lanes(double v, int m) {
  const long long b = __double_as_longlong(v);
  const int lo = xor_lanes(static_cast<int>(b), m);
  const int hi = xor_lanes(static_cast<int>(b >> 32), m);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double dta_at(const RealInfo& ri, int k, double t_a) {
  return (ri.t_begin + ri.t_step * static_cast<double>(k)) - t_a;
}

// First k in [0, n_t] with t_k - t_a >= tau (n_t if none), t_k = t_begin +
// t_step k as the other forms compute it.  t_k - t_a is non-decreasing in k,
// so #{tau_r <= t_k - t_a} = #{first_step(tau_r) <= k}.  An estimate from the
// grid (inv_step = 1 / t_step, or 0: the estimate only picks where the exact
// comparisons start), checked and corrected step by step; a binary search
// where the estimate is far off.
__device__ __forceinline__ int first_step_reaching(const RealInfo& ri, int n_t, double t_a,
                                                   double tau, double inv_step) {
  if (!(tau < kInf)) return n_t;
  int k = n_t;
  if (ri.t_step > 0.0) {
    const double est = (t_a + tau - ri.t_begin) * inv_step;
    if (est < static_cast<double>(n_t)) k = est > 0.0 ? static_cast<int>(est) : 0;
  }
  for (int it = 0; it < 4; ++it) {
    if (k < n_t && dta_at(ri, k, t_a) < tau) ++k;
    else if (k > 0 && dta_at(ri, k - 1, t_a) >= tau) --k;
    else return k;
  }
  int lo = 0, hi = n_t;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (dta_at(ri, mid, t_a) >= tau) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}


// Ring slots are sorted by a 31-bit code: the first time step at which the
// direction has collided (first_step_reaching; n_t when it never does within
// the grid, 0 for every direction when there is no grid) above the phi index.
// The directions collided at step k are then exactly a prefix of the ring for
// every k, which is all the phase sums read from the order; padding slots
// (phi index >= n_phi) come after every direction.
constexpr int kSortPhiBits = 10;  // phi index < kSortCap
__device__ __forceinline__ int32_t ring_sort_code(int n_t, const RealInfo& ri, double t_a,
                                                  double tau, double inv_step, int jd,
                                                  int n_phi) {
  if (jd >= n_phi) return (kSortStepMax << kSortPhiBits) | jd;
  const int step = n_t > 0 ? first_step_reaching(ri, n_t, t_a, tau, inv_step) : 0;
  return (step << kSortPhiBits) | jd;
}

// Bitonic sort of the P slots of one ring held by one wavefront: lane l
// holds ring-local elements l + 64 r (r < P / 64) when P >= 64, or element
// l % P of ring l / P when P < 64.  Partners less than 64 apart are exchanged
// across lanes (up to 8 apart by DPP moves, 16 and 32 apart by shuffles),
// farther ones inside a lane; no LDS traffic and no workgroup barrier.  Keys
// are ring_sort_code values (one 32-bit compare per exchange), all distinct.
template <int P>
__device__ void sort_rings_p(const double* sort_key, const int32_t* sort_val, int gcount, int i0,
                             int n_phi, double* tau_out, uint16_t* perm_out) {
  constexpr int EPL = P >= 64 ? P / 64 : 1;   // elements per lane
  constexpr int RPW = P >= 64 ? 1 : 64 / P;   // rings per wavefront
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int n_waves = blockDim.x >> 6;
  const int e0 = P >= 64 ? lane : (lane % P);
  const int rsub = P >= 64 ? 0 : lane / P;
  for (int rb = wave * RPW; rb < gcount; rb += n_waves * RPW) {
    const int rr = rb + rsub;
    const bool valid = rr < gcount;
    int32_t key[EPL];
#pragma unroll
    for (int r = 0; r < EPL; ++r) {
      const int e = e0 + 64 * r;
      key[r] = valid ? sort_val[rr * P + e] : 0x7fffffff - e;
    }
#pragma unroll
    for (int k = 2; k <= P; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        if (j >= 64) {
#pragma unroll
          for (int r = 0; r < EPL; ++r) {
            const int rp = r ^ (j >> 6);
            if (rp > r) {
              const int e = e0 + 64 * r;
              const bool up = (e & k) == 0;
              if ((key[r] > key[rp]) == up) {
                const int32_t t = key[r];
                key[r] = key[rp];
                key[rp] = t;
              }
            }
          }
        } else {
#pragma unroll
          for (int r = 0; r < EPL; ++r) {
            const int e = e0 + 64 * r;
            const int32_t pk = j <= 8 ? xor_lanes(key[r], j) : __shfl_xor(key[r], j, 64);
            const bool lower = (e & j) == 0;
            const bool up = (e & k) == 0;
            const bool mine_greater = key[r] > pk;
            // the lower element of the pair keeps the minimum when ascending
            if ((lower == up) ? mine_greater : !mine_greater) key[r] = pk;
          }
        }
      }
    }
    if (valid) {
#pragma unroll
      for (int r = 0; r < EPL; ++r) {
        const int e = e0 + 64 * r;
        if (e < n_phi) {
          const int jd = key[r] & ((1 << kSortPhiBits) - 1);
          tau_out[static_cast<int64_t>(i0 + rr) * n_phi + e] = sort_key[rr * P + jd];
          perm_out[static_cast<int64_t>(i0 + rr) * n_phi + e] = static_cast<uint16_t>(jd);
        }
      }
    }
  }
}

// Rings of more than 128 slots: the same network over the workgroup in LDS,
// one barrier per stage.
__device__ void sort_rings_lds(int P, double* sort_key, int32_t* sort_val, int gcount, int i0,
                               int n_phi, double* tau_out, uint16_t* perm_out) {
  const int nslot = gcount * P;
  for (int k = 2; k <= P; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int e = threadIdx.x; e < nslot; e += blockDim.x) {
        const int p = e ^ jj;
        if (p > e) {
          // direction from the index within the ring's segment, so every
          // segment ends ascending (k == P merges the whole segment)
          const bool up = ((e & (P - 1)) & k) == 0;
          const int32_t ke = sort_val[e], kp = sort_val[p];
          if ((ke > kp) == up) {
            sort_val[e] = kp;
            sort_val[p] = ke;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < nslot; e += blockDim.x) {
    const int rr = e / P;
    const int r = e % P;
    if (r < n_phi) {
      const int jd = sort_val[e] & ((1 << kSortPhiBits) - 1);
      tau_out[static_cast<int64_t>(i0 + rr) * n_phi + r] = sort_key[rr * P + jd];
      perm_out[static_cast<int64_t>(i0 + rr) * n_phi + r] = static_cast<uint16_t>(jd);
    }
  }
}

// SP > 0: rings of exactly SP slots; SP < 0: the LDS network; SP == 0: picked
// from P at run time.  A kernel that fixes SP carries only one network, which
// keeps its register budget to that network's (the run-time switch over all of
// them costs the first-collider scan a third of its occupancy).
template <int SP>
__device__ void sort_rings(int P, double* sort_key, int32_t* sort_val, int gcount, int i0,
                           int n_phi, double* tau_out, uint16_t* perm_out) {
  if constexpr (SP > 0) {
    sort_rings_p<SP>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out);
    return;
  } else if constexpr (SP < 0) {
    sort_rings_lds(P, sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out);
    return;
  }
  switch (P) {
    case 2: sort_rings_p<2>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    case 4: sort_rings_p<4>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    case 8: sort_rings_p<8>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    case 16: sort_rings_p<16>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    case 32: sort_rings_p<32>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    case 64: sort_rings_p<64>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    case 128: sort_rings_p<128>(sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
    default: sort_rings_lds(P, sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out); break;
  }
}

// One item's first colliders and ordered per-ring delays for rings
// [ring_begin, ring_end) (all of them by default), written to tau_out /
// perm_out (the item's n_c * n_phi row).  Called by the whole workgroup; the
// LDS arrays hold kGeoLdsCand candidates and kSortCap sort slots.  Ends
// with a barrier.  SP: the ring sort, as for sort_rings (SP > 0 requires
// g.n_phi_pow2 == SP).
template <int SP = 0>
__device__ void geometry_item(const GeoArgs& g, int item_id, double* tau_out,
                              uint16_t* perm_out, RotCand* cand_lds, double* sort_key,
                              int32_t* sort_val, int ring_begin = 0, int ring_end = -1) {
  if (ring_end < 0) ring_end = g.n_c;
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
  const double inv_step = ri.t_step > 0.0 ? 1.0 / ri.t_step : 0.0;
  const int64_t n_dir = static_cast<int64_t>(g.n_c) * n_phi;

  // Each wavefront takes kDirPerLane neighbouring patches of ph rings x pw
  // azimuths, one direction of each per lane: its lanes share first colliders
  // and stop their culled scans together, and every candidate fetched and
  // every loop step serves kDirPerLane directions.
  constexpr int D = kDirPerLane;
  const int pw = P < 8 ? P : 8;
  const int ph = 64 / pw;
  const int patches_per_row = P / pw;
  for (int i0 = ring_begin; i0 < ring_end; i0 += G) {
    const int gcount = min(G, ring_end - i0);
    const int n_patch = ((gcount + ph - 1) / ph) * patches_per_row;
    const int n_multi = (n_patch + D - 1) / D;
    const int nwork_round = ((n_multi * 64 + blockDim.x - 1) / blockDim.x) * blockDim.x;
    for (int w = threadIdx.x; w < nwork_round; w += blockDim.x) {
      const int lane = w & 63;
      int ring[D], jd[D], slot[D];
      bool in_group[D], active[D], fb[D], act[D];
      double x1[D], x2[D], x3[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int patch = (w >> 6) * D + d;
        const int prow = patch / patches_per_row, pcol = patch % patches_per_row;
        const int rring = prow * ph + lane / pw;  // ring within the group
        jd[d] = pcol * pw + lane % pw;
        ring[d] = i0 + rring;
        slot[d] = rring * P + jd[d];              // sort-buffer position
        in_group[d] = patch < n_patch && rring < gcount;
        active[d] = in_group[d] && jd[d] < n_phi;
        x1[d] = 0.0;
        x2[d] = 0.0;
        x3[d] = 0.0;
        if (active[d]) {
          const double st = g.sin_t[ring[d]];
          x1[d] = st * g.cos_p[jd[d]];
          x2[d] = st * g.sin_p[jd[d]];
          x3[d] = g.cos_t[ring[d]];
        }
        fb[d] = active[d] &&
                (ri.nonfinite || !(isfinite(x1[d]) && isfinite(x2[d]) && isfinite(x3[d])));
        act[d] = active[d] && !fb[d];
      }

      // Pass 1: minimum collision delay over the candidates (dot > 1e-10,
      // delay > 0), ties to the lowest bubble index, and the smallest delay of
      // any other candidate with dot > 0 (`second`).  The list is sorted by a
      // lower bound `key` of the delay: the walk goes on while key <= best *
      // (1 + 1e-4) (1 + 1e-6) for some direction of the wave; later candidates
      // can neither reach nor tie a best (key (1 - 1e-7) > best) and are beyond
      // what the verification below has to see (key > best (1 + 1e-4)).
      // hdsq > lim (1 + 1e-12) dot proves hdsq / dot > lim (the relative slack
      // exceeds the rounding of both sides), so the division runs only when a
      // candidate can lower the best or the second delay.
      double best[D], second[D], stop_lim[D], div_lim[D], bscale[D];
      int bidx[D], bpos[D], bflags[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        best[d] = kInf;
        second[d] = kInf;
        // NaN for inactive lanes: key <= stop_lim is then false, so that the
        // exit ballot is one comparison
        stop_lim[d] = act[d] ? kInf : __builtin_nan("");
        div_lim[d] = kInf;
        bscale[d] = 0.0;
        bidx[d] = 0x7fffffff;
        bpos[d] = -1;
        bflags[d] = 0;
      }
      // irregular candidates (key -inf) sort first
      const bool irregular = n_cand > 0 && !(cand_lds[0].flags & kCandRegular);
      // Candidates go in pairs (two per loop step: one exit ballot and one
      // candidate ballot per pair).  The second of a pair is tested with the
      // limits from before the first's update, which are never tighter: it may
      // be divided where a single step would have skipped it, but a candidate
      // that passes only the looser test can lower neither the best delay nor
      // the second (its delay exceeds both), so the results are the same.
      // Two copies of the walk: staged candidates from LDS, the rest from HBM
      // (one loop over both would read the fields through generic pointers).
      auto update = [&](const RotCand& e, int s, double dot, int d) {
        const double dtv = e.hdsq / dot;
        if (dtv != dtv) {
          fb[d] = true;
        } else if (dot > 1e-10 && dtv > 0.0 &&
                   (dtv < best[d] || (dtv == best[d] && e.idx < bidx[d]))) {
          second[d] = fmin(second[d], best[d]);
          best[d] = dtv;
          bidx[d] = e.idx;
          bpos[d] = s;
          // the minimum's scale and flags, for the verification rule below
          bscale[d] = fabs(e.d0) + sqrt(e.d1 * e.d1 + e.d2 * e.d2 + e.d3 * e.d3);
          bflags[d] = e.flags;
        } else {
          second[d] = fmin(second[d], dtv);
        }
        stop_lim[d] = best[d] * ((1.0 + 1e-4) * (1.0 + 1e-6));
        div_lim[d] = fmax(best[d], second[d]) * (1.0 + 1e-12);
      };
      auto step1 = [&](const RotCand& ea, const RotCand& eb, int s, bool has_b) -> bool {
        bool need_a[D], cand_a[D], cand_b[D], any_cand = false;
        double dot_a[D], dot_b[D];
        uint64_t need_mask = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          need_a[d] = ea.key <= stop_lim[d];
          need_mask |= __builtin_amdgcn_ballot_w64(need_a[d]);
        }
        if (!need_mask) return false;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const bool need_b = has_b && eb.key <= stop_lim[d];
          dot_a[d] = lscalar(ea.d0, ea.d1, ea.d2, ea.d3, 1.0, x1[d], x2[d], x3[d]);
          dot_b[d] = lscalar(eb.d0, eb.d1, eb.d2, eb.d3, 1.0, x1[d], x2[d], x3[d]);
          cand_a[d] = need_a[d] && dot_a[d] > 0.0 && !(ea.hdsq > div_lim[d] * dot_a[d]);
          cand_b[d] = need_b && dot_b[d] > 0.0 && !(eb.hdsq > div_lim[d] * dot_b[d]);
          any_cand = any_cand || cand_a[d] || cand_b[d];
        }
        if (wave_any(any_cand)) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            if (cand_a[d]) update(ea, s, dot_a[d], d);
            if (cand_b[d]) update(eb, s + 1, dot_b[d], d);
          }
        }
        return true;
      };
      {
        int s = 0;
        bool more = true;
        // the staged list through a pointer held in a VGPR (an opaque zero
        // offset), so that both candidates of a pair are read at immediate
        // offsets from one address that advances by one add per pair; the
        // second read stays inside the array (s + 1 < kGeoLdsCand) and is
        // unused past the staged candidates
        int vzero;
        asm volatile("v_mov_b32 %0, 0" : "=v"(vzero));
        const RotCand* cl = cand_lds + vzero;
        for (; more && s < n_lds; s += 2) {
          const RotCand ea = cl[s];
          const RotCand eb = cl[s + 1];
          more = step1(ea, eb, s, s + 1 < n_lds);
        }
        for (; more && s < n_cand; s += 2) {
          const bool has_b = s + 1 < n_cand;
          const RotCand ea = rotate_entry(list[s], R, identity);
          const RotCand eb = rotate_entry(list[has_b ? s + 1 : s], R, identity);
          more = step1(ea, eb, s, has_b);
        }
      }

      // Pass 2: verify each minimum against every other bubble c with the
      // reference's cross-multiplied test.  For a regular, non-grazing
      // minimum, a regular bubble c whose delay along the direction exceeds
      // best * (1 + 1e-4), or that does not reach the direction (dot <= 0),
      // provably passes: so bubbles with key > best * (1 + 1e-4) are never
      // tested, and when pass 1 saw no other candidate within that margin
      // (second > best * (1 + 1e-4)) and the list holds no irregular bubble,
      // nothing needs testing at all.  The wave walks a prefix of the list
      // only when one of its directions needs it.
      bool p2[D], ok[D];
      double thr[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const bool minimum = act[d] && !fb[d] && bpos[d] >= 0;
        ok[d] = true;
        thr[d] = kInf;
        if (minimum) {
          const bool regular = (bflags[d] & kCandRegular) != 0;
          const bool grazing = !(best[d] <= 1e3 * bscale[d]);
          if (regular && !grazing) thr[d] = best[d] * (1.0 + 1e-4);
        }
        p2[d] = minimum && (irregular || !(second[d] > thr[d]));
      }
      bool any_p2 = false;
#pragma unroll
      for (int d = 0; d < D; ++d) any_p2 = any_p2 || p2[d];
      if (wave_any(any_p2)) {
        RotCand b0[D];
#pragma unroll
        for (int d = 0; d < D; ++d)
          b0[d] = p2[d] ? fetch_cand(bpos[d], cand_lds, n_lds, list, R, identity) : RotCand{};
        // the verification limit of each lane, NaN where there is nothing
        // (more) to verify, so that the exit ballot is one comparison
        double lim2[D];
#pragma unroll
        for (int d = 0; d < D; ++d) lim2[d] = p2[d] ? thr[d] : __builtin_nan("");
        auto step2 = [&](const RotCand& e, int s) -> bool {
          bool need[D];
          uint64_t need_mask = 0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            need[d] = e.key <= lim2[d];
            need_mask |= __builtin_amdgcn_ballot_w64(need[d]);
          }
          if (!need_mask) return false;
#pragma unroll
          for (int d = 0; d < D; ++d)
            if (need[d] && s != bpos[d] && !collision_point_ok(b0[d], e, x1[d], x2[d], x3[d])) {
              ok[d] = false;
              lim2[d] = __builtin_nan("");
            }
          return true;
        };
        int s = 0;
        bool more = true;
        int vzero;
        asm volatile("v_mov_b32 %0, 0" : "=v"(vzero));
        const RotCand* cl = cand_lds + vzero;  // one VGPR address, as in pass 1
        for (; more && s < n_lds; ++s) {
          const RotCand e = cl[s];
          more = step2(e, s);
        }
        for (; more && s < n_cand; ++s) more = step2(rotate_entry(list[s], R, identity), s);
      }

#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (p2[d] && !ok[d]) fb[d] = true;
        int result = -1;
        if (fb[d]) {
          result = reference_scan(list, inv, it.a, ri.n_int, ri.n_tot, R, identity, x1[d], x2[d],
                                  x3[d]);
        } else if (act[d] && bpos[d] >= 0) {
          result = bidx[d];
        }
        double tau = kInf;
        if (active[d]) {
          double dtab = 0.0;  // compute_delta_tab (:446-468)
          if (result >= 0) {
            if (fb[d]) {
              const RotCand eb = fetch_cand(inv[result], cand_lds, n_lds, list, R, identity);
              const double dot = lscalar(eb.d0, eb.d1, eb.d2, eb.d3, 1.0, x1[d], x2[d], x3[d]);
              dtab = fabs(dot) < 1e-10 ? 0.0 : eb.hdsq / dot;
            } else {
              // the verified minimum's delay: pass 1 computed it with the same operands
              // and operations (its dot > 1e-10)
              dtab = best[d];
            }
            if (dtab > 0.0) tau = dtab;  // AlreadyCollided needs delta_tab > 0 (:405)
          }
          const int64_t dir = static_cast<int64_t>(ring[d]) * n_phi + jd[d];
          if (g.first_out) {
            int code = -1;
            if (result >= 0) code = result < ri.n_int ? result : -(result - ri.n_int) - 2;
            g.first_out[static_cast<int64_t>(item_id) * n_dir + dir] = code;
          }
          if (g.dtab_out) g.dtab_out[static_cast<int64_t>(item_id) * n_dir + dir] = dtab;
        }
        if (in_group[d]) {
          sort_key[slot[d]] = tau;
          sort_val[slot[d]] = ring_sort_code(g.n_t, ri, it.t_a, tau, inv_step, jd[d], n_phi);
        }
      }
    }
    __syncthreads();
    // Sort every ring's P slots by ring_sort_code, one ring per wavefront in
    // registers (P <= 128), the delays following their phi indices.
    sort_rings<SP>(P, sort_key, sort_val, gcount, i0, n_phi, tau_out, perm_out);
    __syncthreads();
  }
}

}  // namespace

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

// x^p for a scaling power p: exact-integer exponents |p| <= 32 (the usual
// bulk-flow powers 3, 4, 5, ...) by repeated squaring -- at most 2 log2|p|
// roundings, a few ulp, against ~100 instructions of the general pow -- and
// pow otherwise.  The branch is uniform: p is a parameter of the call.
__device__ __forceinline__ double pow_param(double x, double p) {
  const double ap = fabs(p);
  if (ap <= 32.0 && ap == __builtin_rint(ap)) {
    int n = static_cast<int>(ap);
    double r = 1.0, b = x;
    while (n) {
      if (n & 1) r *= b;
      n >>= 1;
      if (n) b *= b;
    }
    return p < 0.0 ? 1.0 / r : r;
  }
  return pow(x, p);
}

// Number of a ring's delays tau[0..n) <= x, for x = t_k - t_a at a grid step
// k: the ring is ordered by collision step, so the entries <= x are exactly
// a prefix (those of steps <= k) and a binary search finds its end.
__device__ __forceinline__ int count_le(const double* tau, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (tau[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double cutoff_weight(const SpecArgs& g, double t) {
  if (g.cutoff_kind == 0 || t < g.t_cut) return 1.0;
  const double d = t - g.t_cut;
  return exp(-(d * d) * g.inv_t0_sq);
}

// Time steps per chunk of per-step weights in LDS (SpecArgs::time_chunk): one
// exact anchor per chunk and omega in the blocked and mixed forms.  Both
// lengths are multiples of every block length (8, 16).
constexpr int kTimeChunkMin = 256;
constexpr int kTimeChunkMax = 512;
__host__ __device__ inline int time_chunk_of(const SpecArgs& a) {
  return a.time_chunk == kTimeChunkMax ? kTimeChunkMax : kTimeChunkMin;
}

// Per-ring tables in LDS, laid out after each other, then the per-item time
// tables (when they fit) and the per-chunk phase inputs.
struct RingLds {
  double* tau;  // [n_phi] delays in ring order (by collision step)
  double* gc;   // [n_phi] w_j cos 2phi_j in ring order
  double* gs;   // [n_phi]
  double* uc;   // [n_phi + 1] suffix sums of gc (uncollided part)
  double* us;   // [n_phi + 1]
  double* cc;   // [NT][n_phi + 1] prefix sums of gc * tau^p (collided part)
  double* cs;   // [NT][n_phi + 1]
  double* cl;   // [NT][n_phi + 1] prefix sums of gc * tau^p ln tau (adjoint only)
  double* cls;  // [NT][n_phi + 1]
  double* tb;   // [n_t] K dt w_t cutoff(t) (t - t_a)^3, or null
  double* tf;   // [NT][n_t] coef_q (t - t_a)^-p_q, or null
  double* psi;  // [time chunk] phase argument per time step
  double* v;    // [time chunk][2 * n_sets] weights V+ / Vx per time step
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

__device__ RingLds carve(double* smem, int n_phi, int nt_terms, int n_t, bool time_tables,
                        int chunk) {
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
  L.psi = p; p += chunk;
  L.v = p;
  L.vf = reinterpret_cast<float*>(p);
  return L;
}

// Loads ring i of an item (row tau_item of g.tau / g.perm) and builds its
// prefix/suffix tables.  Returns the
// first time index with a contribution and whether the ring is skipped.
// Surface-weight set of an item: the k-hat slot when several sets are stacked.
__device__ __forceinline__ int surface_of(const SpecArgs& g, const Item& it) {
  return g.surfaces > 1 ? it.k : 0;
}

__device__ int prepare_ring(const SpecArgs& g, const RingLds& L, int64_t tau_item, int i,
                            double t_a, const RealInfo& ri, bool* skip, int surf = 0) {
  const int n_phi = g.n_phi;
  const int nterms = g.n_sets * g.n_terms;
  const int64_t base = (tau_item * g.n_c + i) * static_cast<int64_t>(n_phi);
  const double* gcos = g.gcos + static_cast<int64_t>(surf) * n_phi;
  const double* gsin = g.gsin + static_cast<int64_t>(surf) * n_phi;
  for (int r = threadIdx.x; r < n_phi; r += blockDim.x) {
    const double tau = g.tau[base + r];
    const int j = g.perm[base + r];
    const double gcv = gcos[j], gsv = gsin[j];
    L.tau[r] = tau;
    L.gc[r] = gcv;
    L.gs[r] = gsv;
    L.uc[r] = gcv;
    L.us[r] = gsv;
    const bool fin = tau < kInf;
    for (int q = 0; q < nterms; ++q) {
      const double pw = fin ? pow_param(tau, g.pows[q]) : 0.0;
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
      L.tf[static_cast<size_t>(q) * g.n_t + k] = dta > 0.0 ? g.coef[q] * pow_param(dta, -g.pows[q]) : 0.0;
  }
}

// V_pm(t) and psi(t) for the time steps [t0, t0 + len) of ring i.  FLOAT_V:
// the weights are stored in float (mixed form).
template <bool FLOAT_V, int NS>
__device__ void prepare_times(const SpecArgs& g, const RingLds& L, int i, int t0, int len,
                              const Item& it, const RealInfo& ri) {
  const int n_phi = g.n_phi;
  constexpr int ns = NS;
  const double c_i = g.ring_c[i];
  const double rf = g.ring_fac[static_cast<int64_t>(surface_of(g, it)) * g.n_c + i];
  const int chunk = time_chunk_of(g);
  for (int l = threadIdx.x; l < chunk; l += blockDim.x) {
    double vv[2 * NS];
#pragma unroll
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
#pragma unroll
          for (int s = 0; s < ns; ++s) {
            double bp = L.uc[cnt], bx = L.us[cnt];
            for (int kk = 0; kk < g.n_terms; ++kk) {
              const int q = s * g.n_terms + kk;
              const double f = L.tf ? L.tf[static_cast<size_t>(q) * g.n_t + k]
                                    : g.coef[q] * pow_param(dta, -g.pows[q]);
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
#pragma unroll
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
                  f += g.coef[q] * pow_param(ratio, g.pows[q]);
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
#pragma unroll
      for (int s = 0; s < 2 * ns; ++s) L.vf[l * 2 * ns + s] = static_cast<float>(vv[s]);
    } else {
#pragma unroll
      for (int s = 0; s < 2 * ns; ++s) L.v[l * 2 * ns + s] = vv[s];
    }
  }
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Stage 2: phase sums
// ---------------------------------------------------------------------------
// MODE 0: FP64, one double sincos per (t, omega) term (the parity form).
// MODE 1: FP64, block-factorised phases: exp(i w psi(t0 + r)) = A(t0) E_r with
//         E_r = exp(i w r dt (1 - cos th)), taken about each block's centre so
//         that the steps pair up with conjugate phases E_{+-(j+1/2)} (the
//         block's pair sums and differences formed once in LDS); double
//         half-step table (recurrence), double block sums combined by Horner's
//         rule in E_RB over a chunk's blocks, one exact anchor sincos per chunk.
// MODE 2: mixed, the same factorisation and FP64 anchors, with the half-step
//         table, weights, pair sums and block sums in float; block results
//         rotated by the FP64 anchors and accumulated in FP64.
//
// The block-centre pair sums of a chunk's weights, in place: for each block of
// RB steps from l0 and j < RB / 2, the weights at steps l0 + RB/2 + j (a) and
// l0 + RB/2 - 1 - j (b), +-(j + 1/2) steps from the block's centre, become
// v_a + v_b (at b) and v_a - v_b (at a).  Each pair belongs to one lane; the
// barrier after it publishes the sums.
template <int RB, int NS, typename V>
__device__ void pair_block_weights(V* v, int len) {
  constexpr int RH = RB / 2;
  const int n_pair = (len + RB - 1) / RB * RH * 2 * NS;
  for (int u = threadIdx.x; u < n_pair; u += blockDim.x) {
    const int c = u % (2 * NS), bj = u / (2 * NS);
    const int l0 = bj / RH * RB, j = bj % RH;
    V* pa = v + (l0 + RH + j) * 2 * NS + c;
    V* pb = v + (l0 + RH - 1 - j) * 2 * NS + c;
    const V va = *pa, vb = *pb;
    *pb = va + vb;
    *pa = va - vb;
  }
  __syncthreads();
}

// spectrum_item: one item's C contribution; lane l holds omegas l + q * blockDim.x,
// res[q][s] = (Re C+, Im C+, Re Cx, Im Cx) of set s.  tau_item indexes g.tau /
// g.perm.  Called by the whole workgroup.
template <int NS, int OPL, int MODE, int RB>
__device__ void spectrum_item(const SpecArgs& g, int item_id, int64_t tau_item, double* smem,
                              double (&res)[OPL][NS][4]) {
  using Acc = typename std::conditional<MODE == 2, float, double>::type;
  const Item it = g.items[item_id];
  const RealInfo ri = g.real[it.r];
  const int chunk = time_chunk_of(g);
  const int surf = surface_of(g, it);
  const RingLds L = carve(smem, g.n_phi, g.n_sets * g.n_terms, g.n_t, g.time_tables != 0, chunk);
  if (L.tb) prepare_item_tables(g, L, it, ri);
  __syncthreads();

  double om[OPL];
#pragma unroll
  for (int q = 0; q < OPL; ++q) {
    const int wi = threadIdx.x + q * blockDim.x;
    om[q] = wi < g.n_w ? g.omega[ri.w_off + wi] : 0.0;
  }
  // running sums, double in every form
  double acc[OPL][NS][4];
#pragma unroll
  for (int q = 0; q < OPL; ++q)
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[q][s][c] = 0.0;

  for (int i = 0; i < g.n_c; ++i) {
    if (g.ring_fac[static_cast<int64_t>(surf) * g.n_c + i] == 0.0) continue;  // B = 0 exactly
    bool skip;
    const int t_lo = prepare_ring(g, L, tau_item, i, it.t_a, ri, &skip, surf);
    if (skip || t_lo >= g.n_t) {
      __syncthreads();
      continue;
    }
    // The blocks of RB steps are anchored at their centre, (RB - 1) / 2 steps
    // in: their steps lie at +-(j + 1/2) steps from it, j < RB / 2, with the
    // half-step phases F_j = E_{j + 1/2} = E_j E_{1/2} and their conjugates,
    // E_r = exp(i omega r step (1 - cos th)) by a double recurrence; E_RB
    // advances a block anchor to the next block.
    constexpr int kE = MODE == 0 ? 1 : RB / 2;
    Acc er[OPL][kE], ei[OPL][kE];
    double ebr[OPL], ebi[OPL];
    const double delta = ri.t_step * (1.0 - g.ring_c[i]);
    if constexpr (MODE != 0) {
#pragma unroll
      for (int q = 0; q < OPL; ++q) {
        // E_{1/2} by a sincos; E_1 by its own in the FP64 form (its square
        // doubles the rounding of E_1, which the Horner steps carry through
        // the chunk), as the square of E_{1/2} in the mixed form (whose float
        // tables round far above it)
        double sh, ch, s1, c1;
        sincos_f64(om[q] * (0.5 * delta), &sh, &ch);
        if constexpr (MODE == 1) {
          sincos_f64(om[q] * delta, &s1, &c1);
        } else {
          c1 = __builtin_fma(ch, ch, -(sh * sh));
          s1 = 2.0 * (sh * ch);
        }
        double cr = 1.0, ci = 0.0;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          if (r < kE) {
            er[q][r] = static_cast<Acc>(cr * ch - ci * sh);
            ei[q][r] = static_cast<Acc>(cr * sh + ci * ch);
          }
          const double nr = cr * c1 - ci * s1;
          const double ni = cr * s1 + ci * c1;
          cr = nr;
          ci = ni;
        }
        ebr[q] = cr;
        ebi[q] = ci;
      }
    }
    for (int t0 = t_lo; t0 < g.n_t; t0 += chunk) {
      const int len = min(chunk, g.n_t - t0);
      prepare_times<MODE == 2, NS>(g, L, i, t0, len, it, ri);
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
        constexpr int RH = RB / 2;
        if constexpr (MODE == 1) {
          // FP64: Horner over the chunk's blocks, last to first,
          //   H <- H E_RB + B_b,   C += exp(i w psi(t0)) E_{(RB-1)/2} H,
          // with one exact FP64 sincos per chunk and omega.  B_b is the block
          // sum about the block's centre: its steps RB/2 + j and RB/2 - 1 - j
          // sit at +-(j + 1/2) steps from it, with conjugate phases, so
          //   v_a F_j + v_b conj(F_j) = (v_a + v_b) Re F_j + i (v_a - v_b) Im F_j
          // and the sums and differences, formed once per chunk in place of the
          // weights (the sum at step RB/2 - 1 - j, the difference at RB/2 + j),
          // serve every lane's omegas: per block and omega, 2 RB multiply-adds
          // per polarisation instead of 4 RB.  A block's sums are loaded into
          // registers before its FMA chain so the LDS reads overlap.
          pair_block_weights<RB, NS>(L.v, len);
          double hh[OPL][NS][4];
#pragma unroll
          for (int q = 0; q < OPL; ++q)
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
              for (int c = 0; c < 4; ++c) hh[q][s][c] = 0.0;
          const int nb = (len + RB - 1) / RB;
          for (int b = nb - 1; b >= 0; --b) {
            const int l0 = b * RB;
            double sv[RH][2 * NS], dv[RH][2 * NS];  // weights zero-padded past len
#pragma unroll
            for (int j = 0; j < RH; ++j)
#pragma unroll
              for (int c = 0; c < 2 * NS; ++c) {
                sv[j][c] = L.v[(l0 + RH - 1 - j) * 2 * NS + c];
                dv[j][c] = L.v[(l0 + RH + j) * 2 * NS + c];
              }
#pragma unroll
            for (int q = 0; q < OPL; ++q) {
#pragma unroll
              for (int s = 0; s < NS; ++s) {
                // H E_RB, then the block sum on top
                double b0 = __builtin_fma(hh[q][s][0], ebr[q], -(hh[q][s][1] * ebi[q]));
                double b1 = __builtin_fma(hh[q][s][0], ebi[q], hh[q][s][1] * ebr[q]);
                double b2 = __builtin_fma(hh[q][s][2], ebr[q], -(hh[q][s][3] * ebi[q]));
                double b3 = __builtin_fma(hh[q][s][2], ebi[q], hh[q][s][3] * ebr[q]);
#pragma unroll
                for (int j = 0; j < RH; ++j) {
                  b0 = __builtin_fma(sv[j][2 * s], static_cast<double>(er[q][j]), b0);
                  b1 = __builtin_fma(dv[j][2 * s], static_cast<double>(ei[q][j]), b1);
                  b2 = __builtin_fma(sv[j][2 * s + 1], static_cast<double>(er[q][j]), b2);
                  b3 = __builtin_fma(dv[j][2 * s + 1], static_cast<double>(ei[q][j]), b3);
                }
                hh[q][s][0] = b0;
                hh[q][s][1] = b1;
                hh[q][s][2] = b2;
                hh[q][s][3] = b3;
              }
            }
          }
          const double psi0 = L.psi[0];
#pragma unroll
          for (int q = 0; q < OPL; ++q) {
            // the anchor of the first block's centre: exp(i w psi(t0)) F_{RH-1}
            double as0, ac0;
            sincos_f64(om[q] * psi0, &as0, &ac0);
            const double fr = static_cast<double>(er[q][RH - 1]);
            const double fi = static_cast<double>(ei[q][RH - 1]);
            const double ac = __builtin_fma(ac0, fr, -(as0 * fi));
            const double as = __builtin_fma(ac0, fi, as0 * fr);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
              acc[q][s][0] += __builtin_fma(ac, hh[q][s][0], -(as * hh[q][s][1]));
              acc[q][s][1] += __builtin_fma(ac, hh[q][s][1], as * hh[q][s][0]);
              acc[q][s][2] += __builtin_fma(ac, hh[q][s][2], -(as * hh[q][s][3]));
              acc[q][s][3] += __builtin_fma(ac, hh[q][s][3], as * hh[q][s][2]);
            }
          }
        } else {
        // Mixed: an exact FP64 sincos gives the anchor exp(i w psi) of the
        // chunk's first block centre, each following block's anchor is the
        // previous one times E_RB (FP64, as the FP64 form's Horner steps);
        // block sums about the centre, from the float pair sums and
        // differences as in the FP64 form, rotated by the anchor and
        // accumulated in FP64.
        pair_block_weights<RB, NS>(L.vf, len);
        {
          double anc_r[OPL], anc_i[OPL];
          const double psi0 = L.psi[0] + 0.5 * static_cast<double>(RB - 1) * delta;
#pragma unroll
          for (int q = 0; q < OPL; ++q) sincos_f64(om[q] * psi0, &anc_i[q], &anc_r[q]);
          for (int l0 = 0; l0 < len; l0 += RB) {
            Acc sv[RH][2 * NS], dv[RH][2 * NS];  // weights zero-padded past len
#pragma unroll
            for (int j = 0; j < RH; ++j)
#pragma unroll
              for (int c = 0; c < 2 * NS; ++c) {
                sv[j][c] = L.vf[(l0 + RH - 1 - j) * 2 * NS + c];
                dv[j][c] = L.vf[(l0 + RH + j) * 2 * NS + c];
              }
            Acc bs[OPL][NS][4];
#pragma unroll
            for (int q = 0; q < OPL; ++q)
#pragma unroll
              for (int s = 0; s < NS; ++s) {
                bs[q][s][0] = sv[0][2 * s] * er[q][0];
                bs[q][s][1] = dv[0][2 * s] * ei[q][0];
                bs[q][s][2] = sv[0][2 * s + 1] * er[q][0];
                bs[q][s][3] = dv[0][2 * s + 1] * ei[q][0];
              }
#pragma unroll
            for (int j = 1; j < RH; ++j) {
#pragma unroll
              for (int s = 0; s < NS; ++s) {
#pragma unroll
                for (int q = 0; q < OPL; ++q) {
                  bs[q][s][0] = fma(sv[j][2 * s], er[q][j], bs[q][s][0]);
                  bs[q][s][1] = fma(dv[j][2 * s], ei[q][j], bs[q][s][1]);
                  bs[q][s][2] = fma(sv[j][2 * s + 1], er[q][j], bs[q][s][2]);
                  bs[q][s][3] = fma(dv[j][2 * s + 1], ei[q][j], bs[q][s][3]);
                }
              }
            }
#pragma unroll
            for (int q = 0; q < OPL; ++q) {
              const double as = anc_i[q], ac = anc_r[q];
#pragma unroll
              for (int s = 0; s < NS; ++s) {
                // C += (ac + i as) (b0 + i b1), likewise for the cross
                // polarisation: the float block sums widened, rotated by the
                // FP64 anchor and accumulated in FP64 (two multiply-adds per
                // component, at the rate of the float ones)
                const double b0 = bs[q][s][0], b1 = bs[q][s][1];
                const double b2 = bs[q][s][2], b3 = bs[q][s][3];
                acc[q][s][0] = __builtin_fma(-as, b1, __builtin_fma(ac, b0, acc[q][s][0]));
                acc[q][s][1] = __builtin_fma(as, b0, __builtin_fma(ac, b1, acc[q][s][1]));
                acc[q][s][2] = __builtin_fma(-as, b3, __builtin_fma(ac, b2, acc[q][s][2]));
                acc[q][s][3] = __builtin_fma(as, b2, __builtin_fma(ac, b3, acc[q][s][3]));
              }
              const double nr = __builtin_fma(anc_r[q], ebr[q], -(anc_i[q] * ebi[q]));
              const double ni = __builtin_fma(anc_r[q], ebi[q], anc_i[q] * ebr[q]);
              anc_r[q] = nr;
              anc_i[q] = ni;
            }
          }
        }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < OPL; ++q)
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) res[q][s][j] = acc[q][s][j];
}

}  // namespace

}  // namespace bgw
