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
// Three kernels per batch of tiles (and one per call):
//  * mx_time_kernel: the tile's time tables (K dt w_t cutoff(t) (t - t_a)^3 and
//    coef_q (t - t_a)^-p_q per step), shared by its rings.
//  * mx_weights_kernel, one workgroup per (tile, ring, column tile; both
//    column tiles for rings of at most 64 directions and one set): the
//    items' per-ring tables in LDS -- for every direction in ring order the
//    first time index at which it has collided, the suffix sums of the
//    uncollided weights and the prefix sums of the collided ones -- then the
//    column tile's weights V_i[k, col] of the ring into HBM, row k of 16 CT
//    doubles (rows from the column tile's first weighted step), with the
//    arithmetic of the other forms' weights
//    (compute_b_integral, :481-586, folded with the quadrature weights of
//    :588-664, 775-873).  Latency-bound work, run at full occupancy.
//  * mx_phase_kernel, once per call: per realisation, ring and omega the step
//    and anchor factors below (the same values for every tile).
//  * mx_gemm_kernel, one workgroup per (tile, block of omegas): each wavefront
//    streams the weight rows of every ring as the B operand (loaded four
//    matrix steps ahead, from a pointer with immediate offsets) and generates
//    the A operand, one (omega, step) element per lane advanced four steps at
//    a time by an FP64 complex product with exp(i w 4 dt (1 - c_i)), restarted
//    every 64 steps from an anchor that itself advances by exp(i w 64 dt (1 -
//    c_i)) from the ring's exact FP64 phase; no LDS and no barriers.  FP64
//    matrix instructions share the vector unit's FP64 datapath (they never
//    co-execute), so every vector instruction here costs matrix time: the
//    loop keeps them to the phase products.
// Operand layouts of v_mfma_f64_16x16x4_f64 (measured on the device): A[m][k]
// in lane m + 16 k, B[k][n] in lane n + 16 k, D[m][n] in element m / 4 of lane
// n + 16 (m % 4).  The accumulators stay in VGPRs (this file is compiled with
// -mllvm -amdgpu-mfma-vgpr-form=1; in AGPRs every accumulator would be copied
// out and back around each matrix step).
#include <climits>

#include "bgw_stages.h"

namespace bgw {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// matrix steps (of 4 time steps) from one anchor of the phase factors to the
// next: at most 15 step products in between; the anchors themselves advance
// by one complex product each from the ring's exact FP64 phase at its first
// step (a ring of n_t steps carries at most n_t / 64 of them)
constexpr int kMxAnchorSteps = 16;

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
// summed over the ordered directions >= c) and, per term q, P+_q, Px_q (the
// collided weights times tau^p_q summed over the directions < c).  Records of
// one item are consecutive; items are TS doubles apart (odd, to spread the
// items over the LDS banks).
// The tables are built by kMxSegLanes lanes per item, lane g taking the 2^SLS
// records [g 2^SLS, (g + 1) 2^SLS); record c sits at off(c) = c W + (c >> SLS),
// one double of padding per lane's segment, so the lanes of an item touch
// different bank pairs when they walk their segments together.
constexpr int kMxSegLanes = 32;
constexpr int kMxWeightItems = 256 / kMxSegLanes;  // items of one weights workgroup

// Lanes per item of the weights workgroups: 32, or 16 for one set of one term
// in two column tiles of rings of at most 64 directions -- then a workgroup
// holds 16 items, both column tiles of the tile, in the LDS that 8 items of
// 128 directions take (fewer, fuller workgroups for small grids).
__host__ __device__ inline int mx_seg_lanes(int n_phi, int n_sets, int n_terms, int col_tiles) {
  return (n_sets == 1 && n_terms == 1 && col_tiles == 2 && n_phi <= 64) ? 16 : kMxSegLanes;
}

struct MxLayout {
  int W, TS, KS, SLS;  // record width, item stride (doubles), collided-step row stride
                       // (uint16), log2 of the records per lane segment
  __host__ __device__ int off(int c) const { return c * W + (c >> SLS); }
};

__host__ __device__ inline MxLayout mx_layout(int n_phi, int n_sets, int n_terms, int seg) {
  MxLayout L;
  L.W = 2 + 2 * n_sets * n_terms;
  const int sl = (n_phi + seg - 1) / seg;
  L.SLS = 0;
  while ((1 << L.SLS) < sl) ++L.SLS;
  L.TS = L.off(n_phi) + L.W;
  L.TS += (L.TS & 1) ? 0 : 1;
  const int dw = (n_phi + 1) / 2;  // 32-bit words of one row
  L.KS = 2 * (dw | 1);             // an odd number of words per row
  return L;
}

// The tile's time tables: tt[tile][k][0] = K dt w_t cutoff(t) (t - t_a)^3,
// tt[tile][k][1 + q] = coef_q (t - t_a)^-p_q, zero where t <= t_a (the other
// forms' prepare_item_tables).
__global__ __launch_bounds__(256) void mx_time_kernel(MatArgs m) {
  const SpecArgs& g = m.s;
  const int tile = m.tile0 + static_cast<int>(blockIdx.x);
  const Item it0 = g.items[m.tiles[static_cast<int64_t>(tile) * m.ni]];
  const RealInfo ri = g.real[it0.r];
  const int nterms = g.n_sets * g.n_terms;
  double* tt = m.tt + static_cast<int64_t>(tile) * g.n_t * (1 + nterms);
  for (int k = threadIdx.x; k < g.n_t; k += blockDim.x) {
    const double t = ri.t_begin + ri.t_step * static_cast<double>(k);
    const double dta = t - it0.t_a;
    double base = 0.0;
    if (dta > 0.0) {
      const double wt = (k == 0 || k == g.n_t - 1) ? 0.5 : 1.0;
      base = g.kfac * ri.dt * wt * cutoff_weight(g, t) * ((dta * dta) * dta);
    }
    double* tk = tt + static_cast<int64_t>(k) * (1 + nterms);
    tk[0] = base;
    for (int q = 0; q < nterms; ++q)
      tk[1 + q] = dta > 0.0 ? g.coef[q] * pow_param(dta, -g.pows[q]) : 0.0;
  }
}

// One (tile, ring, CTW column tiles from ct0): the tables of the items whose
// columns meet columns [16 ct0, 16 (ct0 + CTW)) (at most 256 / SEG), then
// those columns of the ring's weight rows k = kbase .. kbase + len - 1 into
// HBM, and each column tile's first weighted step (ctlo).  With 32 lanes per
// item a workgroup takes one column tile (CTW = 1): splitting a tile's columns
// over workgroups halves the LDS of each, so twice as many wavefronts keep
// this latency-bound work in flight; with 16 lanes (small rings) it takes both.
//  * tables: the SEG lanes t = SEG j + g of item j take the ordered directions
//    [g 2^SLS, (g + 1) 2^SLS); each sums its segment, a SEG-lane scan of the
//    segment sums gives every segment's start, and a second walk writes the
//    records.
//  * rows: lane t writes the two columns (+, x) 16 ct0 + 2 (t % CP) and the
//    next (CP = 8 CTW column pairs) of rows from the workgroup's first
//    weighted step (rounded down to 4; the products read nothing of a column
//    tile before its own), 256 / CP rows apart; the collided count is found
//    once by a binary search and carried along the lane's rows.
// FAST: one set of one term.
template <bool FAST, int SEG>
__global__ __launch_bounds__(256) void mx_weights_kernel(MatArgs m) {
  extern __shared__ double smem[];
  constexpr int MAXW = FAST ? 4 : 2 + 2 * kMaxMatrixTerms;
  constexpr int NJ = 256 / SEG;
  constexpr int CTW = SEG == 16 ? 2 : 1;  // column tiles of the workgroup
  __shared__ int item_s[NJ];
  __shared__ int tlo_s[NJ];
  __shared__ double rf_s[NJ];
  __shared__ double tau0_s[NJ];
  __shared__ int lo_s;
  const SpecArgs& g = m.s;
  const int tile = m.tile0 + static_cast<int>(blockIdx.x);
  const int i = blockIdx.y;
  const int ct0 = static_cast<int>(blockIdx.z) * CTW;
  const int ni = m.ni;
  const int ncol = 16 * m.col_tiles;  // doubles per weight row
  const int nsets = FAST ? 1 : g.n_sets;
  const int n_terms = FAST ? 1 : g.n_terms;
  const int nterms = nsets * n_terms;
  const int n_phi = g.n_phi;
  const int n_t = g.n_t;
  const MxLayout L = mx_layout(n_phi, nsets, n_terms, SEG);
  const int W = FAST ? 4 : L.W;
  const int TW = 1 + nterms;
  const MxTileInfo info = m.tinfo[tile];
  // items j0 .. j1 - 1 of the tile have columns in this workgroup's column tiles
  const int cpi = 2 * nsets;
  const int j0 = (16 * ct0) / cpi;
  const int j1 = min(ni, (16 * (ct0 + CTW) - 1) / cpi + 1);
  const int nj = j1 - j0;
  const int nj_cap = min(ni, NJ);  // the tables' item slots (matrix_lds_bytes)
  double* tab = smem;                                                                  // [nj_cap][TS]
  uint16_t* kap = reinterpret_cast<uint16_t*>(tab + static_cast<size_t>(nj_cap) * L.TS);  // [nj_cap][KS]
  const int tid = threadIdx.x;
  if (tid < NJ) item_s[tid] = tid < nj ? m.tiles[static_cast<int64_t>(tile) * ni + j0 + tid] : -1;
  __syncthreads();
  const Item it0 = g.items[m.tiles[static_cast<int64_t>(tile) * ni]];
  const RealInfo ri = g.real[it0.r];
  const double t_a = it0.t_a;
  const double inv_step = ri.t_step > 0.0 ? 1.0 / ri.t_step : 0.0;
  const double* ttile = m.tt + static_cast<int64_t>(tile) * n_t * TW;  // the tile's time tables

  // (1) tables: lane (j, seg) takes directions [seg SL, seg SL + SL) of item j
  const int j = tid / SEG, seg = tid % SEG;
  const int SL = 1 << L.SLS;
  const int r_lo = min(n_phi, seg * SL), r_hi = min(n_phi, r_lo + SL);
  const bool jlive = j < nj && item_s[j] >= 0;
  double* T = tab + static_cast<size_t>(j < nj ? j : 0) * L.TS;
  double sum[MAXW];
#pragma unroll
  for (int a = 0; a < MAXW; ++a) sum[a] = 0.0;
  if (j < nj) {
    const int item = jlive ? item_s[j] : 0;
    const int64_t row = jlive ? (static_cast<int64_t>(item - g.item0) * g.n_c + i) * n_phi : 0;
    const int surf = jlive ? surface_of(g, g.items[item]) : 0;
    // batches of four directions: their loads are issued together
    for (int rb = r_lo; rb < r_hi; rb += 4) {
      double tb4[4], gp4[4], gx4[4];
      int pb4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool ok = jlive && rb + u < r_hi;
        tb4[u] = ok ? g.tau[row + rb + u] : kInf;
        pb4[u] = ok ? g.perm[row + rb + u] : 0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool ok = jlive && rb + u < r_hi;
        gp4[u] = ok ? g.gcos[static_cast<int64_t>(surf) * n_phi + pb4[u]] : 0.0;
        gx4[u] = ok ? g.gsin[static_cast<int64_t>(surf) * n_phi + pb4[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = rb + u;
        if (r >= r_hi) break;
        const double tau = tb4[u], gp = gp4[u], gx = gx4[u];
        if (r == 0) tau0_s[j] = tau;
        kap[j * L.KS + r] = static_cast<uint16_t>(first_step_reaching(ri, n_t, t_a, tau, inv_step));
        T[L.off(r)] = gp;
        T[L.off(r) + 1] = gx;
        sum[0] += gp;
        sum[1] += gx;
        const bool fin = tau < kInf;
        for (int q = 0; q < nterms; ++q) {
          const double pw = fin ? pow_param(tau, g.pows[q]) : 0.0;
          const double vp = gp * pw, vx = gx * pw;
          T[L.off(r + 1) + 2 + 2 * q] = vp;
          T[L.off(r + 1) + 3 + 2 * q] = vx;
          sum[2 + 2 * q] += vp;
          sum[3 + 2 * q] += vx;
        }
      }
    }
  }
  // exclusive prefix of the segment sums over the item's lanes, and their total
  double excl[MAXW], total[MAXW];
#pragma unroll
  for (int a = 0; a < MAXW; ++a) {
    if (a >= W) break;
    double v = sum[a];
#pragma unroll
    for (int off = 1; off < SEG; off <<= 1) {
      const double u = __shfl_up(v, off, SEG);
      if (seg >= off) v += u;
    }
    total[a] = __shfl(v, SEG - 1, SEG);
    excl[a] = v - sum[a];
  }
  if (j < nj) {
    // U: suffix sums (directions >= c), P: prefix sums (directions < c)
    double up = total[0] - excl[0], ux = total[1] - excl[1];
    double pre[MAXW];
#pragma unroll
    for (int a = 2; a < MAXW; ++a) pre[a] = a < W ? excl[a] : 0.0;
    for (int r = r_lo; r < r_hi; ++r) {
      const double gp = T[L.off(r)], gx = T[L.off(r) + 1];
      T[L.off(r)] = up;
      T[L.off(r) + 1] = ux;
      up -= gp;
      ux -= gx;
      for (int a = 2; a < W; ++a) {
        pre[a] += T[L.off(r + 1) + a];
        T[L.off(r + 1) + a] = pre[a];
      }
    }
    if (seg == SEG - 1) {
      T[L.off(n_phi)] = 0.0;
      T[L.off(n_phi) + 1] = 0.0;
      for (int a = 2; a < W; ++a) T[a] = 0.0;
    }
  }
  if (tid < nj) {
    const int item = item_s[tid];
    rf_s[tid] = item >= 0 ? g.ring_fac[static_cast<int64_t>(surface_of(g, g.items[item])) * g.n_c + i]
                          : 0.0;
  }
  __syncthreads();
  // (3) first weighted step per item, the same rules as the other forms: B = 0
  // exactly when the ring factor vanishes; a ring whose uncollided sum is
  // rounding noise contributes from its first collision on, and nothing if it
  // never collides
  // (the tile's first live step k_a comes with its tile info), by the first
  // wavefront, which also takes the minimum over each column tile's items
  if (tid < 64) {
    int lo = n_t, col0 = -1 << 20;
    if (tid < nj) {
      const double* Tj = tab + static_cast<size_t>(tid) * L.TS;
      const bool noise = fabs(Tj[0]) <= 1e-12 * n_phi && fabs(Tj[1]) <= 1e-12 * n_phi;
      const bool drop_early = g.skip_uncollided && noise && !g.has_damping;
      const int k_a = info.k_a;
      lo = k_a;
      if (item_s[tid] < 0 || rf_s[tid] == 0.0) lo = n_t;
      else if (drop_early) lo = tau0_s[tid] < kInf ? max(k_a, static_cast<int>(kap[tid * L.KS])) : n_t;
      tlo_s[tid] = lo;
      col0 = cpi * (j0 + tid);  // the item's first column
    }
    int lo_min = n_t;
#pragma unroll
    for (int c = 0; c < CTW; ++c) {
      // items whose columns [col0, col0 + cpi) meet column tile ct0 + c
      const int c16 = 16 * (ct0 + c);
      int v = col0 < c16 + 16 && col0 + cpi > c16 ? lo : n_t;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
      if (tid == 0) m.ctlo[(static_cast<int64_t>(tile) * g.n_c + i) * 2 + ct0 + c] = v;
      lo_min = min(lo_min, v);
    }
    if (tid == 0) lo_s = lo_min;
  }
  __syncthreads();
  // (4) the column tiles' weight rows from the first one the products read:
  // lane t writes the two columns (+, x) of one (item, set), 16 ct0 + 2 (t % CP)
  // and the next, of rows 256 / CP apart, sharing the collided count, the
  // record and the time-table loads of a row between them
  constexpr int CP = 8 * CTW;    // column pairs of the workgroup
  constexpr int RG = 256 / CP;   // row groups
  const int col = 16 * ct0 + 2 * (tid % CP);
  const int jg = col / cpi;           // the columns' item in the tile
  const int jc = jg - j0;             // and in this workgroup
  const int live0 = lo_s < n_t ? (lo_s & ~3) - info.kbase : info.len;
  const int RP = (info.len - live0 + RG - 1) / RG;
  const int r0 = live0 + (tid / CP) * RP;
  const int r1 = min(info.len, r0 + RP);
  double* vout = m.vbuf + (info.voff - m.voff0) + static_cast<int64_t>(i) * info.len * ncol;
  const bool live = jg < ni && jc < nj && item_s[jc] >= 0;
  const int tlo = live ? tlo_s[jc] : n_t;
  const double rf = live ? rf_s[jc] : 0.0;
  const double* Tc = tab + static_cast<size_t>(live ? jc : 0) * L.TS;
  const uint16_t* kr = kap + (live ? jc : 0) * L.KS;
  const int sc = (col >> 1) % nsets;
  int cnt = -1;
  // groups of four rows: their time-table loads are issued together
  for (int rg = r0; rg < r1; rg += 4) {
    double tb4[4], tf4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = min(info.kbase + rg + u, n_t - 1);
      const double* tk = ttile + static_cast<int64_t>(k) * TW;
      tb4[u] = tk[0];
      tf4[u] = FAST ? tk[1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = rg + u;
      if (r >= r1) break;
      const int k = info.kbase + r;
      double2 v = make_double2(0.0, 0.0);
      if (live && k < n_t && k >= tlo) {
        if (cnt < 0) {
          cnt = count_steps_le(kr, n_phi, k);
        } else {
          while (cnt < n_phi && static_cast<int>(kr[cnt]) <= k) ++cnt;
        }
        const double* rec = Tc + L.off(cnt);
        const double base = tb4[u] * rf;
        double bp = rec[0], bx = rec[1];
        if (FAST) {
          bp = __builtin_fma(tf4[u], rec[2], bp);
          bx = __builtin_fma(tf4[u], rec[3], bx);
        } else {
          const double* tk = ttile + static_cast<int64_t>(k) * TW;
          for (int kt = 0; kt < n_terms; ++kt) {
            const int q = sc * n_terms + kt;
            const double f = tk[1 + q];
            bp = __builtin_fma(f, rec[2 + 2 * q], bp);
            bx = __builtin_fma(f, rec[3 + 2 * q], bx);
          }
        }
        v = make_double2(base * bp, base * bx);
      }
      *reinterpret_cast<double2*>(vout + static_cast<int64_t>(r) * ncol + col) = v;
    }
  }
}

// Per (realisation, ring, omega): the step factor e^{i w 4 dt (1 - c_i)} and
// the anchor factor e^{i w 4 kMxAnchorSteps dt (1 - c_i)} of the products
// kernel, once per call instead of once per tile and ring.
__global__ __launch_bounds__(256) void mx_phase_kernel(MatArgs m, int n_real) {
  const SpecArgs& g = m.s;
  const int64_t n = static_cast<int64_t>(n_real) * g.n_c * g.n_w;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int wi = static_cast<int>(e % g.n_w);
  const int i = static_cast<int>((e / g.n_w) % g.n_c);
  const int r = static_cast<int>(e / (static_cast<int64_t>(g.n_w) * g.n_c));
  const RealInfo ri = g.real[r];
  const double om = g.omega[ri.w_off + wi];
  const double ai = 1.0 - g.ring_c[i];
  double* p = m.phase + e * 4;
  double sr, si, fr, fi;
  sincos_f64(om * (4.0 * ri.t_step * ai), &si, &sr);
  sincos_f64(om * ((4.0 * kMxAnchorSteps) * ri.t_step * ai), &fi, &fr);
  p[0] = sr;
  p[1] = si;
  p[2] = fr;
  p[3] = fi;
}

// One (tile, block of omegas): wavefront w takes omega tiles w OT + [0, OT) of
// the block and every column tile; the weight rows stream from HBM four
// matrix steps ahead.  ct_lo: first weighted step of each column tile.
template <int OT, int CT>
__global__ __launch_bounds__(256) void mx_gemm_kernel(MatArgs m) {
  constexpr int NCOL = 16 * CT;
  const SpecArgs& g = m.s;
  // Workgroups go to the 8 XCDs in turn (dispatch index b to XCD b % 8): the
  // omega blocks of one tile take dispatch indices 8 apart, so they run on the
  // same XCD at about the same time and the second reads the tile's weight
  // rows from that XCD's L2.  Tiles keep their cost order; the last
  // (n_tiles % 8) tiles take the remaining indices in order.
  const int ny = static_cast<int>(gridDim.y);
  const int ntl = static_cast<int>(gridDim.x);
  const int b = static_cast<int>(blockIdx.x) + static_cast<int>(blockIdx.y) * ntl;
  const int full = (ntl / 8) * 8 * ny;
  int tl, by;
  if (b < full) {
    const int slot = b / 8;
    tl = (slot / ny) * 8 + b % 8;
    by = slot % ny;
  } else {
    const int r = b - full;
    tl = (ntl / 8) * 8 + r / ny;
    by = r % ny;
  }
  const int tile = m.tile0 + tl;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kk = lane >> 4;
  const int n_t = g.n_t;
  const int ni = m.ni;
  const int nsets = g.n_sets;
  const Item it0 = g.items[m.tiles[static_cast<int64_t>(tile) * ni]];
  const RealInfo ri = g.real[it0.r];
  const double t_a = it0.t_a;
  const MxTileInfo info = m.tinfo[tile];
  const int wbase = by * (blockDim.x >> 6) * 16 * OT + wave * 16 * OT;
  double om[OT];
#pragma unroll
  for (int o = 0; o < OT; ++o) {
    const int wi = wbase + o * 16 + (lane & 15);
    om[o] = wi < g.n_w ? g.omega[ri.w_off + wi] : 0.0;
  }
  v4d acc_re[OT][CT], acc_im[OT][CT];
#pragma unroll
  for (int o = 0; o < OT; ++o)
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      acc_re[o][c] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_im[o][c] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  const double* vt = m.vbuf + (info.voff - m.voff0) + (lane & 15);
  for (int i = 0; i < g.n_c; ++i) {
    const int64_t ti = static_cast<int64_t>(tile) * g.n_c + i;
    int ct_lo[CT];
    int ks = n_t;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      ct_lo[c] = __builtin_amdgcn_readfirstlane(m.ctlo[ti * 2 + c]);
      ks = min(ks, ct_lo[c]);
    }
    if (ks >= n_t) continue;
    const int k0 = ks & ~3;
    const double ci = g.ring_c[i];
    const double ai = 1.0 - ci;
    // step factor e^{i w 4 dt (1 - c)}, anchor factor e^{i w 4 kMxAnchorSteps dt (1 - c)},
    // and the exact phase of matrix step 0
    double sr[OT], si[OT], fr[OT], fi[OT], er[OT], ei[OT];
    const double tk0 = ri.t_begin + ri.t_step * static_cast<double>(k0 + kk);
    const double* ph = m.phase + (static_cast<int64_t>(it0.r) * g.n_c + i) * g.n_w * 4;
#pragma unroll
    for (int o = 0; o < OT; ++o) {
      const int wi = wbase + o * 16 + (lane & 15);
      sr[o] = 1.0;
      si[o] = 0.0;
      fr[o] = 1.0;
      fi[o] = 0.0;
      if (wi < g.n_w) {
        const double* p = ph + static_cast<int64_t>(wi) * 4;
        sr[o] = p[0];
        si[o] = p[1];
        fr[o] = p[2];
        fi[o] = p[3];
      }
      sincos_f64(om[o] * (tk0 * ai + t_a * ci), &ei[o], &er[o]);
    }
    double xr[OT], xi[OT];  // the anchor of the current group of kMxAnchorSteps steps
#pragma unroll
    for (int o = 0; o < OT; ++o) {
      xr[o] = er[o];
      xi[o] = ei[o];
    }
    const int n_ms = (n_t - k0 + 3) >> 2;
    // this lane's weight row of matrix step 0 (row r <-> step kbase + r); the
    // rows of a block of four matrix steps sit at immediate offsets from it
    const double* q =
        vt + (static_cast<int64_t>(i) * info.len + (k0 - info.kbase) + kk) * NCOL;
    // B operands of four matrix steps, the next four loaded while these run.
    // Loads run up to 28 rows past a ring's last row: into the next ring's
    // rows or the buffer's padding (kMatrixRowPad), never used.
    auto load4 = [&](const double* qb, double (&b)[4][CT]) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < CT; ++c) b[u][c] = qb[u * 4 * NCOL + c * 16];
    };
    auto run4 = [&](int ms0, const double (&b)[4][CT]) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ms = ms0 + u;
        if (ms >= n_ms) break;
        const int kb = k0 + 4 * ms;
        if (ms % kMxAnchorSteps == 0 && ms > 0) {
          // next anchor = anchor * e^{i w 4 kMxAnchorSteps dt (1 - c)}
#pragma unroll
          for (int o = 0; o < OT; ++o) {
            const double nr = __builtin_fma(xr[o], fr[o], -(xi[o] * fi[o]));
            const double nim = __builtin_fma(xr[o], fi[o], xi[o] * fr[o]);
            xr[o] = nr;
            xi[o] = nim;
            er[o] = nr;
            ei[o] = nim;
          }
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          if (kb + 4 <= ct_lo[c]) continue;  // the tile's columns have no weight yet
#pragma unroll
          for (int o = 0; o < OT; ++o) {
            acc_re[o][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(er[o], b[u][c], acc_re[o][c], 0, 0, 0);
            acc_im[o][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei[o], b[u][c], acc_im[o][c], 0, 0, 0);
          }
        }
#pragma unroll
        for (int o = 0; o < OT; ++o) {
          const double nr = __builtin_fma(er[o], sr[o], -(ei[o] * si[o]));
          const double nim = __builtin_fma(er[o], si[o], ei[o] * sr[o]);
          er[o] = nr;
          ei[o] = nim;
        }
      }
    };
    double ba[4][CT], bb[4][CT];
    load4(q, ba);
    for (int ms0 = 0; ms0 < n_ms; ms0 += 8) {
      load4(q + static_cast<int64_t>(ms0 + 4) * 4 * NCOL, bb);
      run4(ms0, ba);
      if (ms0 + 4 >= n_ms) break;
      load4(q + static_cast<int64_t>(ms0 + 8) * 4 * NCOL, ba);
      run4(ms0 + 4, bb);
    }
  }

  // C_item(w) = exp(-i w z_a) Y[w, col]: element r of lane L holds row
  // (L / 16) + 4 r, column L % 16 of each 16 x 16 tile
  const int64_t nq = static_cast<int64_t>(g.n_sets_out) * g.n_w;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int col = c * 16 + (lane & 15);
    const int j = col / (2 * nsets);
    if (j >= ni) continue;
    const int item = m.tiles[static_cast<int64_t>(tile) * ni + j];
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

hipError_t launch_weights(const MatArgs& a, int n_tiles, size_t lds, hipStream_t s) {
  const bool fast = a.s.n_sets == 1 && a.s.n_terms == 1;
  const int seg = mx_seg_lanes(a.s.n_phi, a.s.n_sets, a.s.n_terms, a.col_tiles);
  auto k = !fast ? mx_weights_kernel<false, kMxSegLanes>
                 : (seg == 16 ? mx_weights_kernel<true, 16> : mx_weights_kernel<true, kMxSegLanes>);
  const int ctw = seg == 16 ? 2 : 1;
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_tiles, a.s.n_c, a.col_tiles / ctw), dim3(256), lds, s, a);
  return hipGetLastError();
}

}  // namespace

int matrix_items_per_tile(int n_sets, int col_tiles) { return (16 * col_tiles) / (2 * n_sets); }

size_t matrix_lds_bytes(const SpecArgs& a, int ni, int col_tiles) {
  const int seg = mx_seg_lanes(a.n_phi, a.n_sets, a.n_terms, col_tiles);
  const MxLayout L = mx_layout(a.n_phi, a.n_sets, a.n_terms, seg);
  const int items = 256 / seg;
  const size_t nj = static_cast<size_t>(ni < items ? ni : items);
  return nj * L.TS * sizeof(double) + nj * L.KS * sizeof(uint16_t);
}

namespace {

bool matrix_args_ok(const MatArgs& a) {
  if (a.omega_tiles != 1 && a.omega_tiles != 2 && a.omega_tiles != 4) return false;
  if (a.col_tiles != 1 && a.col_tiles != 2) return false;
  if (a.ni < 1 || a.ni > matrix_items_per_tile(a.s.n_sets, a.col_tiles)) return false;
  if (a.s.n_t > 65535 || a.s.has_damping) return false;
  if (a.s.n_sets * a.s.n_terms > kMaxMatrixTerms) return false;
  return matrix_lds_bytes(a.s, a.ni, a.col_tiles) <= 160 * 1024;
}

}  // namespace

hipError_t launch_matrix_weights(const MatArgs& a, int n_tiles, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  if (!matrix_args_ok(a)) return hipErrorInvalidValue;
  const size_t lds = matrix_lds_bytes(a.s, a.ni, a.col_tiles);
  hipLaunchKernelGGL(mx_time_kernel, dim3(n_tiles), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_weights(a, n_tiles, lds, s);
}

hipError_t launch_matrix_phases(const MatArgs& a, int n_real, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(n_real) * a.s.n_c * a.s.n_w;
  if (n == 0) return hipSuccess;
  if ((n + 255) / 256 > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mx_phase_kernel, dim3(static_cast<unsigned int>((n + 255) / 256)), dim3(256),
                     0, s, a, n_real);
  return hipGetLastError();
}

hipError_t launch_matrix_products(const MatArgs& a, int n_tiles, hipStream_t s) {
  if (n_tiles == 0) return hipSuccess;
  if (!matrix_args_ok(a)) return hipErrorInvalidValue;
  const int w_per_block = 4 * 16 * a.omega_tiles;
  const dim3 grid(n_tiles, (a.s.n_w + w_per_block - 1) / w_per_block);
  if (a.col_tiles == 1) {
    if (a.omega_tiles == 1) hipLaunchKernelGGL((mx_gemm_kernel<1, 1>), grid, dim3(256), 0, s, a);
    else if (a.omega_tiles == 2) hipLaunchKernelGGL((mx_gemm_kernel<2, 1>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((mx_gemm_kernel<4, 1>), grid, dim3(256), 0, s, a);
  } else {
    if (a.omega_tiles == 1) hipLaunchKernelGGL((mx_gemm_kernel<1, 2>), grid, dim3(256), 0, s, a);
    else if (a.omega_tiles == 2) hipLaunchKernelGGL((mx_gemm_kernel<2, 2>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((mx_gemm_kernel<4, 2>), grid, dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace bgw
