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
//  * mx_phase_kernel, once per call: per realisation, ring and omega the
//    phase factors below (the same values for every tile).
//  * mx_gemm_kernel, one workgroup per (tile, block of omegas): the steps of
//    a ring go in blocks of 16 on the tile's grid; about a block's centre its
//    steps pair up as 7 - j and 8 + j, j < 8, with conjugate phases
//    F_j = exp(i w (j + 1/2) dt (1 - c_i)), so the block sum is
//      B(w, col) = sum_j (V_a + V_b) Re F_j + i (V_a - V_b) Im F_j:
//    two K = 4 slices per part, the A operands Re F_j, Im F_j fixed per ring
//    and omega, the B operands the sums and differences of the two weight
//    rows each lane loads (one block ahead, from a pointer with immediate
//    offsets) -- half the matrix instructions of the direct phase sum.  The
//    blocks accumulate by Horner's rule, H <- H conj(E_16) + B_b (four FP64
//    products per accumulator element and block), and the rings through one
//    complex factor per omega: the running sum is A_last H, A_last the exact
//    FP64 phase of the last block centre; no LDS and no barriers.  FP64
//    matrix instructions share the vector unit's FP64 datapath (they never
//    co-execute), so every vector instruction here costs matrix time.
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

// time steps per block of the products: pairs (B/2 - 1 - j, B/2 + j), j < B/2,
// about the block's centre; the tiles' first rows (kbase) and row counts are
// multiples of it
constexpr int kMxBlock = kMatrixBlockSteps;
static_assert(kMxBlock % 8 == 0 && kMxBlock <= kMatrixRowPad, "blocks of whole K slices, prefetched into the padding");

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
//    weighted step (rounded down to its block of 16 steps; the products read
//    nothing of a column tile before its own block), 256 / CP rows apart; the
//    collided count is found once by a binary search and carried along the
//    lane's rows.
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
    int lo = n_t, col0 = -(1 << 20);
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
  const int live0 = lo_s < n_t ? (lo_s & ~(kMxBlock - 1)) - info.kbase : info.len;
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

// Per (realisation, ring, omega), with d = dt (1 - c_i): E_B = e^{i w B d}
// (B = kMxBlock),
// E_{1/2} = e^{i w d / 2}, E_1 = e^{i w d} and E_4 = e^{i w 4 d}, the phase
// factors of the products kernel, once per call instead of once per tile and
// ring.
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
  const double d = ri.t_step * ai;
  double* p = m.phase + e * 8;
  double sn, cs;
  sincos_f64(om * (static_cast<double>(kMxBlock) * d), &sn, &cs);
  p[0] = cs;
  p[1] = sn;
  sincos_f64(om * (0.5 * d), &sn, &cs);
  p[2] = cs;
  p[3] = sn;
  sincos_f64(om * d, &sn, &cs);
  p[4] = cs;
  p[5] = sn;
  sincos_f64(om * (4.0 * d), &sn, &cs);
  p[6] = cs;
  p[7] = sn;
}

// One (tile, block of omegas): wavefront w takes omega tiles w OT + [0, OT) of
// the block and every column tile; the weight rows stream from HBM one block
// of steps ahead.  ct_lo: first weighted step of each column tile.
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
  // the accumulators hold, per (omega, column), H with the running sum of the
  // processed rings equal to A_last H, A_last = exp(i w psi) at the centre of
  // the last processed block; lr/li: A_last for this lane's omega
  double lr[OT], li[OT];
#pragma unroll
  for (int o = 0; o < OT; ++o) {
    lr[o] = 0.0;
    li[o] = 0.0;
  }
  bool any_ring = false;
  // this lane's column of every weight row: lane % 16 of each column tile
  const double* vt = m.vbuf + (info.voff - m.voff0) + (lane & 15);
  // this lane's rows of a block (kMxBlock steps from its start), for the
  // pairs j = 4 sl + kk about the block's centre: a = H + j, b = H - 1 - j,
  // H = kMxBlock / 2
  constexpr int NSL = kMxBlock / 8;  // K slices of four pairs
  int off_a[NSL], off_b[NSL];
#pragma unroll
  for (int sl = 0; sl < NSL; ++sl) {
    off_a[sl] = (kMxBlock / 2 + 4 * sl + kk) * NCOL;
    off_b[sl] = (kMxBlock / 2 - 1 - 4 * sl - kk) * NCOL;
  }
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
    // blocks of kMxBlock steps on the tile's grid (kbase is a multiple of it)
    const int k0 = ks & ~(kMxBlock - 1);
    const int nblk = (n_t - k0 + kMxBlock - 1) / kMxBlock;
    const double ci = g.ring_c[i];
    const double ai = 1.0 - ci;
    const double* ph = m.phase + (static_cast<int64_t>(it0.r) * g.n_c + i) * g.n_w * 8;
    // A operands: F_j = exp(i w (j + 1/2) dt (1 - c)) for j = 4 sl + kk at
    // this lane's omega; E_B = exp(i w kMxBlock dt (1 - c)) (the block's
    // factor); the exact phases of the first and the last block centres
    double far[OT][NSL], fai[OT][NSL], ebr[OT], ebi[OT], a0r[OT], a0i[OT], alr[OT], ali[OT];
    constexpr double kCentre = 0.5 * (kMxBlock - 1);
    const double tc0 = ri.t_begin + ri.t_step * (static_cast<double>(k0) + kCentre);
    const double tcl =
        ri.t_begin + ri.t_step * (static_cast<double>(k0 + kMxBlock * (nblk - 1)) + kCentre);
#pragma unroll
    for (int o = 0; o < OT; ++o) {
      const int wi = wbase + o * 16 + (lane & 15);
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl) far[o][sl] = fai[o][sl] = 0.0;
      ebr[o] = 1.0;
      ebi[o] = 0.0;
      if (wi < g.n_w) {
        const double* p = ph + static_cast<int64_t>(wi) * 8;
        ebr[o] = p[0];
        ebi[o] = p[1];
        double fr = p[2], fi = p[3];  // E_{1/2}
        const double e1r = p[4], e1i = p[5];
        for (int u = 0; u < kk; ++u) {
          const double nr = __builtin_fma(fr, e1r, -(fi * e1i));
          const double nim = __builtin_fma(fr, e1i, fi * e1r);
          fr = nr;
          fi = nim;
        }
#pragma unroll
        for (int sl = 0; sl < NSL; ++sl) {
          far[o][sl] = fr;
          fai[o][sl] = fi;
          const double nr = __builtin_fma(fr, p[6], -(fi * p[7]));  // times E_4
          const double nim = __builtin_fma(fr, p[7], fi * p[6]);
          fr = nr;
          fi = nim;
        }
      }
      sincos_f64(om[o] * (tc0 * ai + t_a * ci), &a0i[o], &a0r[o]);
      sincos_f64(om[o] * (tcl * ai + t_a * ci), &ali[o], &alr[o]);
    }
    // the Horner factor conj(E_B) at the element rows (element r of this lane's
    // accumulators is omega row kk + 4 r of its omega tile, whose factors lane
    // kk + 4 r holds for the A operand), and the carry of the processed rings
    // into this ring's H: H <- H A_last,prev conj(A_0) E_B (the ring's blocks
    // then multiply it by conj(E_B)^nblk)
    double hr[OT][4], hi[OT][4];
#pragma unroll
    for (int o = 0; o < OT; ++o) {
      const double tr = __builtin_fma(lr[o], a0r[o], li[o] * a0i[o]);  // A_last,prev conj(A_0)
      const double tim = __builtin_fma(li[o], a0r[o], -(lr[o] * a0i[o]));
      const double rr = __builtin_fma(tr, ebr[o], -(tim * ebi[o]));
      const double rim = __builtin_fma(tr, ebi[o], tim * ebr[o]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int src = kk + 4 * r;
        hr[o][r] = __shfl(ebr[o], src);
        hi[o][r] = __shfl(ebi[o], src);
        if (any_ring) {
          const double xr = __shfl(rr, src), xi = __shfl(rim, src);
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const double yr = acc_re[o][c][r], yi = acc_im[o][c][r];
            acc_re[o][c][r] = __builtin_fma(yr, xr, -(yi * xi));
            acc_im[o][c][r] = __builtin_fma(yr, xi, yi * xr);
          }
        }
      }
      lr[o] = alr[o];
      li[o] = ali[o];
    }
    any_ring = true;
    // this lane's column of block 0 (row r <-> step kbase + r)
    const double* q = vt + (static_cast<int64_t>(i) * info.len + (k0 - info.kbase)) * NCOL;
    // B operands of one block (the pair rows a and b of every slice), the next
    // block's loaded while this one runs.  The loads of the block after a
    // ring's last run up to kMxBlock rows past it: into the next ring's rows or
    // the buffer's padding (kMatrixRowPad), never used.
    auto loadb = [&](const double* qb, double (&va)[NSL][CT], double (&vb)[NSL][CT]) {
#pragma unroll
      for (int sl = 0; sl < NSL; ++sl)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          va[sl][c] = qb[off_a[sl] + c * 16];
          vb[sl][c] = qb[off_b[sl] + c * 16];
        }
    };
    auto runb = [&](int blk, const double (&va)[NSL][CT], const double (&vb)[NSL][CT]) {
      const int kb = k0 + kMxBlock * blk;
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        // H <- H conj(E_B), every block of the ring
#pragma unroll
        for (int o = 0; o < OT; ++o)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double yr = acc_re[o][c][r], yi = acc_im[o][c][r];
            acc_re[o][c][r] = __builtin_fma(yr, hr[o][r], yi * hi[o][r]);
            acc_im[o][c][r] = __builtin_fma(yi, hr[o][r], -(yr * hi[o][r]));
          }
        // then the block sum about its centre, unless the column tile has no
        // weight in the block (its rows there are not written; adding selected
        // zeros instead measured slower: 11.6 against 11.4 ms at plain_256)
        if (kb + kMxBlock <= ct_lo[c] || ct_lo[c] >= n_t) continue;
        double sv[NSL], dv[NSL];
#pragma unroll
        for (int sl = 0; sl < NSL; ++sl) {
          sv[sl] = va[sl][c] + vb[sl][c];
          dv[sl] = va[sl][c] - vb[sl][c];
        }
#pragma unroll
        for (int sl = 0; sl < NSL; ++sl) {
#pragma unroll
          for (int o = 0; o < OT; ++o) {
            acc_re[o][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(far[o][sl], sv[sl], acc_re[o][c], 0, 0, 0);
            acc_im[o][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(fai[o][sl], dv[sl], acc_im[o][c], 0, 0, 0);
          }
        }
      }
    };
    double xa[NSL][CT], xb[NSL][CT], ya[NSL][CT], yb[NSL][CT];
    loadb(q, xa, xb);
    for (int blk = 0; blk < nblk; blk += 2) {
      loadb(q + static_cast<int64_t>(blk + 1) * kMxBlock * NCOL, ya, yb);
      runb(blk, xa, xb);
      if (blk + 1 >= nblk) break;
      loadb(q + static_cast<int64_t>(blk + 2) * kMxBlock * NCOL, xa, xb);
      runb(blk + 1, ya, yb);
    }
  }
  // the sum of the rings: A_last H at the element rows
  if (any_ring) {
#pragma unroll
    for (int o = 0; o < OT; ++o)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int src = kk + 4 * r;
        const double xr = __shfl(lr[o], src), xi = __shfl(li[o], src);
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          const double yr = acc_re[o][c][r], yi = acc_im[o][c][r];
          acc_re[o][c][r] = __builtin_fma(yr, xr, -(yi * xi));
          acc_im[o][c][r] = __builtin_fma(yr, xi, yi * xr);
        }
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
