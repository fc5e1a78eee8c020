// SPDX-License-Identifier: MIT
// Phase sums on the matrix cores: the FP64 matrix form.
//
// A tile holds items of one realisation and one reference bubble a seen from
// several k-hat frames (or several surface-weight sets), so they share t_a and
// the time grid.  With psi = (t - z_a) - (t - t_a) c_i = t (1 - c_i) + t_a c_i
// - z_a (c_i = cos theta_i), only z_a depends on the item, and for ring i the
// phase sums of every item of the tile are one matrix product
//
//   Y[w, col] += sum_k A_i[w, k] V_i[k, col],
//   A_i[w, k]  = exp(i w (t_k (1 - c_i) + t_a c_i)),
//   V_i[k, col] = the weight of time step k in ring i for col = (item, set, +/x),
//
// summed over the rings into the same accumulators; then C_item(w) =
// exp(-i w z_a) Y[w, col].  Re A and Im A are the A operands of two
// v_mfma_f64_16x16x4_f64 chains that share the V operand, so each term costs
// the four FP64 multiply-adds of a complex-by-real product, issued by the
// matrix cores.
//
//  * A operand: a lane holds one (omega, step) element, advanced four steps at
//    a time by an FP64 complex product with exp(i w 4 dt (1 - c_i)), re-anchored
//    by an exact FP64 sincos every 64 steps (at most 15 products between
//    anchors).
//  * V operand: computed into LDS 16 steps at a time (double-buffered) from
//    per-ring tables of each item -- for every sorted direction the first
//    time index at which it has collided, the suffix sums of the uncollided
//    weights and the prefix sums of the collided ones -- with the same
//    arithmetic as the other forms' weights (compute_b_integral, :481-586,
//    folded with the quadrature weights of :588-664, 775-873).
//  * Operand layouts of v_mfma_f64_16x16x4_f64 (measured on the device):
//    A[m][k] in lane m + 16 k, B[k][n] in lane n + 16 k, D[m][n] in element
//    m / 4 of lane n + 16 (m % 4).
// The accumulators stay in VGPRs (this file is compiled with
// -mllvm -amdgpu-mfma-vgpr-form=1; in AGPRs every accumulator would be copied
// out and back around each matrix step).
#include "bgw_stages.h"

namespace bgw {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// time steps of weights per LDS buffer: 16 matrix steps of 4, with an exact
// anchor of the phase factors at the first (at most 15 recurrence products)
constexpr int kMxChunk = 64;

__device__ __forceinline__ double dta_at(const RealInfo& ri, int k, double t_a) {
  return (ri.t_begin + ri.t_step * static_cast<double>(k)) - t_a;
}

// First k in [0, n_t] with t_k - t_a >= tau (n_t if none), t_k = t_begin +
// t_step k as the other forms compute it.  t_k - t_a is non-decreasing in k,
// so #{tau_r <= t_k - t_a} = #{first_step(tau_r) <= k}.  An estimate from the
// grid, checked and corrected step by step; a binary search where the
// estimate is far off.
__device__ __forceinline__ int first_step_reaching(const RealInfo& ri, int n_t, double t_a,
                                                   double tau) {
  if (!(tau < kInf)) return n_t;
  int k = n_t;
  if (ri.t_step > 0.0) {
    const double est = (t_a + tau - ri.t_begin) / ri.t_step;
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

__device__ __forceinline__ int first_live_step(const RealInfo& ri, int n_t, double t_a) {
  int lo = 0, hi = n_t;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (dta_at(ri, mid, t_a) > 0.0) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// number of entries of the ascending row[0..n) that are <= k
__device__ __forceinline__ int count_steps_le(const uint16_t* row, int n, int k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (static_cast<int>(row[mid]) <= k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Inclusive scan by one wavefront of the n elements arr[idx * stride]
// (reverse: suffix sums), in the order of wave_scan.
__device__ void wave_scan_strided(double* arr, int n, int stride, bool reverse) {
  const int lane = threadIdx.x & 63;
  double carry = 0.0;
  for (int base = 0; base < n; base += 64) {
    const int e = base + lane;
    const int idx = reverse ? (n - 1 - e) : e;
    double v = e < n ? arr[idx * stride] : 0.0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double u = __shfl_up(v, off, 64);
      if (lane >= off) v += u;
    }
    v += carry;
    if (e < n) arr[idx * stride] = v;
    carry = __shfl(v, 63, 64);
  }
}

// Per-ring tables of a tile in LDS.  Item j's record c (c = 0..n_phi, the
// number of directions already collided) holds W = 2 + 2 n_sets n_terms
// doubles: U+, Ux (the uncollided weights w_j cos 2 phi_j, w_j sin 2 phi_j
// summed over the sorted directions >= c) and, per term q, P+_q, Px_q (the
// collided weights times tau^p_q summed over the directions < c).  Records of
// one item are consecutive; items are TS doubles apart (odd, to spread the
// items over the LDS banks).
struct MxLayout {
  int W, TS, KS;  // record width, item stride (doubles), collided-step row stride (uint16)
};

__host__ __device__ inline MxLayout mx_layout(int n_phi, int n_sets, int n_terms) {
  MxLayout L;
  L.W = 2 + 2 * n_sets * n_terms;
  L.TS = (n_phi + 1) * L.W;
  L.TS += (L.TS & 1) ? 0 : 1;
  const int dw = (n_phi + 1) / 2;  // 32-bit words of one row
  L.KS = 2 * (dw | 1);             // an odd number of words per row
  return L;
}

// Row stride of the weight buffers in doubles: 16 columns per column tile,
// padded so that the four rows a matrix step reads (lanes 0-15, 16-31, 32-47,
// 48-63) fall on alternating halves of the 64 LDS banks.
template <int CT>
constexpr int vb_stride() { return CT == 1 ? 16 : 48; }

// Weights of time steps [kc, kc + kMxChunk) of the current ring into vb, by
// the producer lanes: lane p takes item p % 16 and the four consecutive steps
// kc + 4 (p / 16) + u (consecutive lanes write consecutive columns), the first
// collided count found by a binary search over the item's sorted first
// collided steps and carried over the next three.  tt: the tile's time tables
// (per step K dt w_t cutoff(t) (t - t_a)^3, then coef_q (t - t_a)^-p_q per
// term), so the weight is the one the other forms compute, with the same
// roundings.  FAST: one set of one term (the loads of a step are independent).
template <int CT, bool FAST>
__device__ __forceinline__ void produce_chunk(const SpecArgs& g, const MxLayout& L, int p, int kc,
                                              int ni, const double* tab, const uint16_t* kap,
                                              const double* tt, const int* tlo_s,
                                              const double* rf_s, double* vb) {
  constexpr int VS = vb_stride<CT>();
  const int j = p & 15;
  if (j >= ni) return;
  const int nsets = FAST ? 1 : g.n_sets, n_terms = FAST ? 1 : g.n_terms;
  const int n_phi = g.n_phi, n_t = g.n_t;
  const int W = FAST ? 4 : L.W;
  const int TW = 1 + nsets * n_terms;
  const double* T = tab + static_cast<size_t>(j) * L.TS;
  const uint16_t* kr = kap + j * L.KS;
  const int kfirst = kc + 4 * (p >> 4);
  const int tlo = tlo_s[j];
  const double rf = rf_s[j];
  int cnt = -1;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = kfirst + u;
    double* dst = vb + (k - kc) * VS + j * 2 * nsets;
    if (k < n_t && k >= tlo) {
      if (cnt < 0) {
        cnt = count_steps_le(kr, n_phi, k);
      } else {
        while (cnt < n_phi && static_cast<int>(kr[cnt]) <= k) ++cnt;
      }
      const double* rec = T + cnt * W;
      const double* tk = tt + static_cast<size_t>(k) * TW;
      const double base = tk[0] * rf;
      if (FAST) {
        const double up = rec[0], ux = rec[1], pp = rec[2], px = rec[3], f = tk[1];
        dst[0] = base * __builtin_fma(f, pp, up);
        dst[1] = base * __builtin_fma(f, px, ux);
      } else {
        for (int s = 0; s < nsets; ++s) {
          double bp = rec[0], bx = rec[1];
          for (int kt = 0; kt < n_terms; ++kt) {
            const int q = s * n_terms + kt;
            const double f = tk[1 + q];
            bp = __builtin_fma(f, rec[2 + 2 * q], bp);
            bx = __builtin_fma(f, rec[3 + 2 * q], bx);
          }
          dst[2 * s] = base * bp;
          dst[2 * s + 1] = base * bx;
        }
      }
    } else {
      for (int s = 0; s < 2 * nsets; ++s) dst[s] = 0.0;
    }
  }
}

// The matrix steps of one chunk: 16 steps of 4 time steps, from an exact
// anchor of the phase factors at the chunk's first step.  parg: t_k (1 - c_i)
// + t_a c_i at the lane's step kc + lane / 16.  ct_lo[c]: first step with a
// weight in column tile c (uniform).
template <int OT, int CT>
__device__ __forceinline__ void consume_chunk(const double* vb, int kc, int n_t, int lane,
                                              const double (&om)[OT], double parg,
                                              const double (&sr)[OT], const double (&si)[OT],
                                              const int (&ct_lo)[CT], v4d (&acc_re)[OT][CT],
                                              v4d (&acc_im)[OT][CT]) {
  constexpr int VS = vb_stride<CT>();
  const int kk = lane >> 4;
  double er[OT], ei[OT];
#pragma unroll
  for (int o = 0; o < OT; ++o) sincos_f64(om[o] * parg, &ei[o], &er[o]);
  const double* src = vb + kk * VS + (lane & 15);
  double bv[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) bv[c] = src[c * 16];
  const int n_ms = min(kMxChunk / 4, (n_t - kc + 3) >> 2);
  for (int ms = 0; ms < n_ms; ++ms) {
    const int kb = kc + 4 * ms;
    double bn[CT];
    const double* nsrc = src + (ms + 1 < kMxChunk / 4 ? 4 * (ms + 1) * VS : 0);
#pragma unroll
    for (int c = 0; c < CT; ++c) bn[c] = nsrc[c * 16];  // the next step's operand, early
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      if (kb + 4 <= ct_lo[c]) continue;  // the tile's columns have no weight yet
#pragma unroll
      for (int o = 0; o < OT; ++o) {
        acc_re[o][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(er[o], bv[c], acc_re[o][c], 0, 0, 0);
        acc_im[o][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei[o], bv[c], acc_im[o][c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int o = 0; o < OT; ++o) {
      const double nr = __builtin_fma(er[o], sr[o], -(ei[o] * si[o]));
      const double nim = __builtin_fma(er[o], si[o], ei[o] * sr[o]);
      er[o] = nr;
      ei[o] = nim;
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) bv[c] = bn[c];
  }
}

// Workgroups of kMxMfmaWaves matrix wavefronts (OT omega tiles of 16 each) and
// kMxProdWaves producer wavefronts; CT column tiles of 16.  Per ring all lanes
// build the items' tables; then the producers write the weights of chunk n + 1
// while the matrix wavefronts consume chunk n (two LDS buffers, one barrier
// per chunk, each role in a loop of its own with the same barriers).
template <int OT, int CT, bool FAST>
__global__ __launch_bounds__(kMatrixThreads) void matrix_kernel(MatArgs m) {
  extern __shared__ double smem[];
  constexpr int VS = vb_stride<CT>();
  constexpr int MAXNI = 8 * CT;
  static_assert(MAXNI <= 16 && kMxChunk / 4 * 16 <= kMxProdWaves * 64,
                "one producer lane per (item, 4 steps)");
  __shared__ int item_s[MAXNI];
  __shared__ int tlo_s[MAXNI];
  __shared__ double rf_s[MAXNI];
  __shared__ double tau0_s[MAXNI];
  const SpecArgs& g = m.s;
  const int ni = m.ni;
  const int nsets = g.n_sets;
  const int nterms = g.n_sets * g.n_terms;
  const int n_phi = g.n_phi;
  const int n_t = g.n_t;
  const MxLayout L = mx_layout(n_phi, g.n_sets, g.n_terms);
  const int W = L.W;
  double* tab = smem;                                                    // [ni][TS]
  double* vbuf = tab + static_cast<size_t>(ni) * L.TS;                  // [2][kMxChunk][VS]
  double* tt = vbuf + 2 * kMxChunk * VS;                                // [n_t][1 + nterms]
  uint16_t* kap = reinterpret_cast<uint16_t*>(tt + static_cast<size_t>(1 + nterms) * n_t);  // [ni][KS]

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const bool producer = wave >= kMxMfmaWaves;
  if (tid < MAXNI) item_s[tid] = tid < ni ? m.tiles[static_cast<int64_t>(blockIdx.x) * ni + tid] : -1;
  for (int e = tid; e < 2 * kMxChunk * VS; e += blockDim.x) vbuf[e] = 0.0;  // padding stays 0
  __syncthreads();

  const Item it0 = g.items[item_s[0]];
  const RealInfo ri = g.real[it0.r];
  const double t_a = it0.t_a;
  const int k_a = first_live_step(ri, n_t, t_a);
  // the tile's time tables (every item of the tile has this t_a)
  for (int k = tid; k < n_t; k += blockDim.x) {
    const double t = ri.t_begin + ri.t_step * static_cast<double>(k);
    const double dta = t - t_a;
    double base = 0.0;
    if (dta > 0.0) {
      const double wt = (k == 0 || k == n_t - 1) ? 0.5 : 1.0;
      base = g.kfac * ri.dt * wt * cutoff_weight(g, t) * ((dta * dta) * dta);
    }
    double* tk = tt + static_cast<size_t>(k) * (1 + nterms);
    tk[0] = base;
    for (int q = 0; q < nterms; ++q)
      tk[1 + q] = dta > 0.0 ? g.coef[q] * pow_param(dta, -g.pows[q]) : 0.0;
  }
  const int wbase = static_cast<int>(blockIdx.y) * kMxMfmaWaves * 16 * OT + (producer ? 0 : wave) * 16 * OT;
  double om[OT];
#pragma unroll
  for (int o = 0; o < OT; ++o) {
    const int wi = wbase + o * 16 + (lane & 15);
    om[o] = (!producer && wi < g.n_w) ? g.omega[ri.w_off + wi] : 0.0;
  }
  v4d acc_re[OT][CT], acc_im[OT][CT];
#pragma unroll
  for (int o = 0; o < OT; ++o)
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      acc_re[o][c] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_im[o][c] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  for (int i = 0; i < g.n_c; ++i) {
    __syncthreads();  // the previous ring's tables and buffers are no longer read
    // (1) per item: sorted delays of ring i -> first collided step, weights
    for (int e = tid; e < ni * n_phi; e += blockDim.x) {
      const int j = e / n_phi, r = e - j * n_phi;
      const int item = item_s[j];
      double tau = kInf, gp = 0.0, gx = 0.0;
      if (item >= 0) {
        const int64_t row = (static_cast<int64_t>(item - g.item0) * g.n_c + i) * n_phi;
        tau = g.tau[row + r];
        const int pidx = g.perm[row + r];
        const int surf = surface_of(g, g.items[item]);
        gp = g.gcos[static_cast<int64_t>(surf) * n_phi + pidx];
        gx = g.gsin[static_cast<int64_t>(surf) * n_phi + pidx];
      }
      if (r == 0) tau0_s[j] = tau;
      kap[j * L.KS + r] = static_cast<uint16_t>(first_step_reaching(ri, n_t, t_a, tau));
      double* T = tab + static_cast<size_t>(j) * L.TS;
      T[r * W] = gp;
      T[r * W + 1] = gx;
      const bool fin = tau < kInf;
      for (int q = 0; q < nterms; ++q) {
        const double pw = fin ? pow_param(tau, g.pows[q]) : 0.0;
        T[(r + 1) * W + 2 + 2 * q] = gp * pw;
        T[(r + 1) * W + 3 + 2 * q] = gx * pw;
      }
    }
    for (int j = tid; j < ni; j += blockDim.x) {
      double* T = tab + static_cast<size_t>(j) * L.TS;
      T[n_phi * W] = 0.0;
      T[n_phi * W + 1] = 0.0;
      for (int q = 0; q < nterms; ++q) {
        T[2 + 2 * q] = 0.0;
        T[3 + 2 * q] = 0.0;
      }
      const int item = item_s[j];
      rf_s[j] = item >= 0 ? g.ring_fac[static_cast<int64_t>(surface_of(g, g.items[item])) * g.n_c + i]
                          : 0.0;
    }
    __syncthreads();
    // (2) suffix sums of the uncollided weights, prefix sums of the collided ones
    for (int arr = wave; arr < ni * W; arr += (blockDim.x >> 6)) {
      const int j = arr / W, a = arr - j * W;
      double* T = tab + static_cast<size_t>(j) * L.TS + a;
      if (a < 2) wave_scan_strided(T, n_phi, W, true);
      else wave_scan_strided(T + W, n_phi, W, false);
    }
    __syncthreads();
    // (3) first step with a weight per item; the same rules as the other forms:
    // B = 0 exactly when the ring factor vanishes; a ring whose uncollided sum is
    // rounding noise contributes from its first collision on, and nothing if it
    // never collides
    if (tid < ni) {
      const double* T = tab + static_cast<size_t>(tid) * L.TS;
      const bool noise = fabs(T[0]) <= 1e-12 * n_phi && fabs(T[1]) <= 1e-12 * n_phi;
      const bool drop_early = g.skip_uncollided && noise && !g.has_damping;
      int lo = k_a;
      if (item_s[tid] < 0 || rf_s[tid] == 0.0) lo = n_t;
      else if (drop_early) lo = tau0_s[tid] < kInf ? max(k_a, static_cast<int>(kap[tid * L.KS])) : n_t;
      tlo_s[tid] = lo;
    }
    __syncthreads();
    int kstart = n_t;
    for (int j = 0; j < ni; ++j) kstart = min(kstart, tlo_s[j]);
    kstart = __builtin_amdgcn_readfirstlane(kstart);
    if (kstart >= n_t) continue;
    const int k0 = kstart & ~3;
    const int n_chunks = (n_t - k0 + kMxChunk - 1) / kMxChunk;

    if (producer) {
      const int pl = tid - kMxMfmaWaves * 64;
      produce_chunk<CT, FAST>(g, L, pl, k0, ni, tab, kap, tt, tlo_s, rf_s, vbuf);
      __syncthreads();
      for (int n = 0; n < n_chunks; ++n) {
        if (n + 1 < n_chunks)
          produce_chunk<CT, FAST>(g, L, pl, k0 + (n + 1) * kMxChunk, ni, tab, kap, tt, tlo_s, rf_s,
                                  vbuf + ((n + 1) & 1) * kMxChunk * VS);
        __syncthreads();
      }
    } else {
      int ct_lo[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        int lo = n_t;
        for (int j = 0; j < ni; ++j)
          if ((j * 2 * nsets) / 16 <= c && ((j + 1) * 2 * nsets - 1) / 16 >= c) lo = min(lo, tlo_s[j]);
        ct_lo[c] = __builtin_amdgcn_readfirstlane(lo);
      }
      // A operand: exp(i w 4 dt (1 - c_i)) per lane and omega tile
      const double ci = g.ring_c[i];
      const double ai = 1.0 - ci;
      double sr[OT], si[OT];
#pragma unroll
      for (int o = 0; o < OT; ++o) sincos_f64(om[o] * (4.0 * ri.t_step * ai), &si[o], &sr[o]);
      __syncthreads();
      for (int n = 0; n < n_chunks; ++n) {
        const int kc = k0 + n * kMxChunk;
        const double tk = ri.t_begin + ri.t_step * static_cast<double>(kc + (lane >> 4));
        consume_chunk<OT, CT>(vbuf + (n & 1) * kMxChunk * VS, kc, n_t, lane, om,
                              tk * ai + t_a * ci, sr, si, ct_lo, acc_re, acc_im);
        __syncthreads();
      }
    }
  }
  if (producer) return;

  // C_item(w) = exp(-i w z_a) Y[w, col]: element r of lane L holds row
  // (L / 16) + 4 r, column L % 16 of each 16 x 16 tile
  const int kk = lane >> 4;
  const int64_t nq = static_cast<int64_t>(g.n_sets_out) * g.n_w;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int col = c * 16 + (lane & 15);
    const int j = col / (2 * nsets);
    if (j >= ni) continue;
    const int item = item_s[j];
    if (item < 0) continue;
    const int s = (col >> 1) % nsets, pm = col & 1;
    const double z = g.items[item].z_a;
    double* out = g.partial + static_cast<int64_t>(item - g.item0) * 2 * nq * 2;
#pragma unroll
    for (int o = 0; o < OT; ++o) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int wi = wbase + o * 16 + kk + 4 * r;
        if (wi >= g.n_w) continue;
        const double w = g.omega[ri.w_off + wi];
        double sz, cz;
        sincos_f64(w * z, &sz, &cz);
        const double yr = acc_re[o][c][r], yi = acc_im[o][c][r];
        const int64_t e = ((static_cast<int64_t>(pm) * g.n_sets_out + g.set0 + s) * g.n_w + wi) * 2;
        out[e] = __builtin_fma(yr, cz, yi * sz);
        out[e + 1] = __builtin_fma(yi, cz, -(yr * sz));
      }
    }
  }
}

template <int OT, int CT>
hipError_t launch_matrix_t(const MatArgs& a, int n_tiles, int w_blocks, size_t lds, hipStream_t s) {
  const bool fast = a.s.n_sets == 1 && a.s.n_terms == 1;
  auto k = fast ? matrix_kernel<OT, CT, true> : matrix_kernel<OT, CT, false>;
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_tiles, w_blocks), dim3(kMatrixThreads), lds, s, a);
  return hipGetLastError();
}

}  // namespace

int matrix_items_per_tile(int n_sets, int col_tiles) { return (16 * col_tiles) / (2 * n_sets); }

size_t matrix_lds_bytes(const SpecArgs& a, int ni, int col_tiles) {
  const MxLayout L = mx_layout(a.n_phi, a.n_sets, a.n_terms);
  const size_t vs = col_tiles == 1 ? vb_stride<1>() : vb_stride<2>();
  const size_t nterms = static_cast<size_t>(a.n_sets) * a.n_terms;
  return (static_cast<size_t>(ni) * L.TS + 2 * kMxChunk * vs + (1 + nterms) * a.n_t) *
             sizeof(double) +
         static_cast<size_t>(ni) * L.KS * sizeof(uint16_t);
}

hipError_t launch_matrix(const MatArgs& a, int n_tiles, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  if (a.omega_tiles != 1 && a.omega_tiles != 2) return hipErrorInvalidValue;
  if (a.col_tiles != 1 && a.col_tiles != 2) return hipErrorInvalidValue;
  if (a.ni < 1 || a.ni > matrix_items_per_tile(a.s.n_sets, a.col_tiles)) return hipErrorInvalidValue;
  if (a.s.n_t > 65535 || a.s.has_damping) return hipErrorInvalidValue;
  const size_t lds = matrix_lds_bytes(a.s, a.ni, a.col_tiles);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  const int w_per_block = kMxMfmaWaves * 16 * a.omega_tiles;
  const int w_blocks = (a.s.n_w + w_per_block - 1) / w_per_block;
  if (a.omega_tiles == 1)
    return a.col_tiles == 1 ? launch_matrix_t<1, 1>(a, n_tiles, w_blocks, lds, s)
                            : launch_matrix_t<1, 2>(a, n_tiles, w_blocks, lds, s);
  return a.col_tiles == 1 ? launch_matrix_t<2, 1>(a, n_tiles, w_blocks, lds, s)
                          : launch_matrix_t<2, 2>(a, n_tiles, w_blocks, lds, s);
}

}  // namespace bgw
