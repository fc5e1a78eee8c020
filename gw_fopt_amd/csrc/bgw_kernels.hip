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
//   ring's directions are then ordered for stage 2 by the time step at which
//   they have collided (a bitonic network on integer codes in registers, one
//   ring per wavefront; in LDS for rings of more than 128 slots).  A
//   workgroup takes a range of an item's rings (launch_geometry splits items
//   so that uneven items balance).
//
// Stage 2 (spectrum_kernel): per item, ring by ring, the phi integral
//   B_pm(t) of compute_b_integral (:481-586) is read off prefix sums over the
//   ordered directions -- the directions already collided at time step t are
//   exactly a prefix of the ring -- and the quadrature weights of the A and C
//   integrals (:588-664, 775-873) are folded in.  Lanes own omegas and
//   accumulate C_pm(omega) = sum_t V_pm(t) exp(i omega psi(t)), with
//   psi = (t - z_a) - (t - t_a) cos th, in registers:
//     FP64 form:  one double sincos per (t, omega) term;
//     blocked FP64 form: the phase of a block of R steps is anchor * E_r in
//                 double, about the block's centre (conjugate step pairs
//                 share a phase), blocks combined by Horner's rule, one exact
//                 anchor per chunk of steps;
//     mixed form: psi arguments reduced in double; the phase of a block of
//                 R consecutive steps factorised as anchor * E_r with
//                 E_r = exp(i omega r dt (1 - cos th)), about the block's
//                 centre; one exact double anchor per chunk advanced by E_R in
//                 double, float phase factors, weights and block sums, block
//                 results rotated and accumulated in double.
//
// Stage 3 (reduce_kernel): sum the per-bubble partial C in bubble order
//   (deterministic) and form S = 8 pi w^3 (|C+|^2 + |Cx|^2).
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "bgw_stages.h"

namespace bgw {

// ---------------------------------------------------------------------------
// Stage 1: first collisions
// ---------------------------------------------------------------------------
// SP: ring slots (n_phi rounded up to a power of two) for a register sort,
// -1 for the LDS sort of wider rings.  The kernel is held to 72 VGPRs for 7
// waves per SIMD (7 workgroups of 20 KB of LDS per CU): it is VALU-issue bound
// once the walks are short, and at 64 VGPRs (8 waves) the register spills'
// extra instructions cost more than the eighth wave gains (2.51 against
// 2.49 ms at plain_256; 6 waves 2.53 ms).
template <int SP>
__global__ __launch_bounds__(kGeoThreads) __attribute__((amdgpu_waves_per_eu(7)))
void geometry_kernel(GeoArgs g) {
  __shared__ RotCand cand_lds[kGeoLdsCand];
  __shared__ double sort_key[kSortCap];
  __shared__ int32_t sort_val[kSortCap];
  // workgroup (b * n_split + part): rings [part * rings_per_block, ...) of item b
  const int b = static_cast<int>(blockIdx.x) / g.n_split;
  const int part = static_cast<int>(blockIdx.x) - b * g.n_split;
  const int r0 = part * g.rings_per_block;
  const int r1 = min(g.n_c, r0 + g.rings_per_block);
  const int item_id = g.item0 + b;
  const int64_t row = static_cast<int64_t>(b) * g.n_c * g.n_phi;
  geometry_item<SP>(g, item_id, g.tau + row, g.perm + row, cand_lds, sort_key, sort_val, r0, r1);
}

// ---------------------------------------------------------------------------
// Stage 2: phase sums, one workgroup per item
// ---------------------------------------------------------------------------
template <int NS, int OPL, int MODE, int RB>
__global__ __launch_bounds__(256) void spectrum_kernel(SpecArgs g) {
  extern __shared__ double smem[];
  const int item_id = g.order ? g.order[blockIdx.x] : g.item0 + static_cast<int>(blockIdx.x);
  const int64_t row = item_id - g.item0;
  double res[OPL][NS][4];
  spectrum_item<NS, OPL, MODE, RB>(g, item_id, row, smem, res);
  // partial[item][pm][s][w] complex
  const int64_t nq = static_cast<int64_t>(g.n_sets_out) * g.n_w;
  double* out = g.partial + row * 2 * nq * 2;
#pragma unroll
  for (int q = 0; q < OPL; ++q) {
    const int wi = threadIdx.x + q * blockDim.x;
    if (wi >= g.n_w) continue;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int64_t ip = (static_cast<int64_t>(g.set0 + s) * g.n_w + wi) * 2;
      const int64_t ix = ((nq + static_cast<int64_t>(g.set0 + s) * g.n_w) + wi) * 2;
      out[ip] = res[q][s][0];
      out[ip + 1] = res[q][s][1];
      out[ix] = res[q][s][2];
      out[ix + 1] = res[q][s][3];
    }
  }
}

// ---------------------------------------------------------------------------
// Stage 3: deterministic sum over reference bubbles, spectrum
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_kernel(RedArgs g) {
  int grp, i0, i1;
  if (g.unit_group) {
    const int u = g.u0 + static_cast<int>(blockIdx.x);
    grp = g.unit_group[u];
    i0 = g.unit_off[u];
    i1 = g.unit_off[u + 1];
  } else {
    grp = blockIdx.x;
    i0 = g.group_off[grp];
    i1 = g.group_off[grp + 1];
  }
  const int64_t nq = static_cast<int64_t>(g.n_sets) * g.n_w;
  const RealInfo ri = g.real[g.group_r[grp]];
  const double eight_pi = 8.0 * 3.14159265358979323846;
  for (int64_t q = threadIdx.x; q < nq; q += blockDim.x) {
    double pr = 0.0, pi = 0.0, xr = 0.0, xi = 0.0;
    for (int item = i0; item < i1; ++item) {
      const double* p = g.partial + static_cast<int64_t>(item - g.item0) * 2 * nq * 2;
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
  const int chunk = time_chunk_of(g);
  const RingLds L = carve(smem, g.n_phi, g.n_sets * g.n_terms, g.n_t, false, chunk);
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
    for (int t0 = t_lo; t0 < g.n_t; t0 += chunk) {
      const int len = min(chunk, g.n_t - t0);
      prepare_times<false, NS>(g, L, i, t0, len, it, ri);
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
// step t0 + r is exp(i w psi(t0)) E^r with E = exp(i w step (1 - cos th)), and
// Re(conj(G) exp(i w psi)) follows it by a three-term (Chebyshev) recurrence in
// FP64.  The M partial sums of a block are combined by
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
  static_assert(RB >= 2, "the phase recurrence starts from two steps");
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
            // x_r = Re(conj(G) z E^r) with z = exp(i w psi(t0)), E = exp(i theta):
            // x_0 and x_1 from z and z E, then x_{r+1} = 2 cos(theta) x_r - x_{r-1}
            // (one FMA per step and component instead of a complex product and
            // two FMAs).  The recurrence's rounding grows at most like r^2 eps
            // for theta -> 0, i.e. below 1e-13 over a block.
            double zi, zr;
            sincos_f64(A.om[j] * psi0, &zi, &zr);
            const double er = A.E[2 * j], ei = A.E[2 * j + 1];
            const double z1r = __builtin_fma(zr, er, -(zi * ei));
            const double z1i = __builtin_fma(zr, ei, zi * er);
            const double two_c = 2.0 * er;
#pragma unroll
            for (int c = 0; c < 2 * NS; ++c) {
              const double gr = A.G[(j * 2 * NS + c) * 2];
              const double gi = A.G[(j * 2 * NS + c) * 2 + 1];
              double x0 = __builtin_fma(gr, zr, gi * zi);
              double x1 = __builtin_fma(gr, z1r, gi * z1i);
              w[0][c] += x0;
              w[1][c] += x1;
#pragma unroll
              for (int r = 2; r < RB; ++r) {
                const double x2 = __builtin_fma(two_c, x1, -x0);
                w[r][c] += x2;
                x0 = x1;
                x1 = x2;
              }
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
              const double dmp = pow_param(dta, -g.pows[q]);
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
                const double rp = pow_param(ratio, g.pows[q]) * d;
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
// Workgroups per CU the geometry kernel keeps resident (64 VGPRs, 20 KB of LDS).
constexpr int kGeoResident = 8;

std::mutex& lds_limit_mutex() {
  static std::mutex m;
  return m;
}

int device_cus() {
  // per-device cache; handles on several host threads may fill it at once
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
      n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

hipError_t launch_geometry(const GeoArgs& a_in, int n_items, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  // An item's cost varies by an order of magnitude between bubbles (an early
  // bubble walks long candidate prefixes), so with few workgroups per resident
  // slot the slowest items set the kernel time.  Each item's rings are split
  // over several workgroups (whole patch rows of 64 / min(P, 8) rings) until
  // there are about 64 workgroups per resident slot or a workgroup holds one
  // patch row (measured at plain_256: 4.9 ms unsplit, 4.4 / 4.0 / 3.98 ms with
  // 2 / 4 / 8 workgroups per item; grad_128: 1.35 -> 0.74 ms).
  GeoArgs a = a_in;
  const int ph = 64 / (a.n_phi_pow2 < 8 ? a.n_phi_pow2 : 8);
  const int64_t target = static_cast<int64_t>(device_cus()) * kGeoResident * 64;
  int n_split = 1;
  while (static_cast<int64_t>(n_items) * n_split * 2 <= target && a.n_c / (n_split * 2) >= ph)
    n_split *= 2;
  a.rings_per_block = (a.n_c + n_split - 1) / n_split;
  a.rings_per_block = (a.rings_per_block + ph - 1) / ph * ph;
  a.n_split = (a.n_c + a.rings_per_block - 1) / a.rings_per_block;
  const int64_t blocks = static_cast<int64_t>(n_items) * a.n_split;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned int>(blocks)), block(kGeoThreads);
  switch (a.n_phi_pow2) {
    case 2: hipLaunchKernelGGL(geometry_kernel<2>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(geometry_kernel<4>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(geometry_kernel<8>, grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL(geometry_kernel<16>, grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL(geometry_kernel<32>, grid, block, 0, s, a); break;
    case 64: hipLaunchKernelGGL(geometry_kernel<64>, grid, block, 0, s, a); break;
    case 128: hipLaunchKernelGGL(geometry_kernel<128>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(geometry_kernel<-1>, grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

bool spectrum_time_tables(const SpecArgs& a) {
  return time_table_doubles(a.n_t, a.n_sets * a.n_terms) * sizeof(double) <= 48 * 1024;
}

size_t spectrum_lds_bytes(const SpecArgs& a, int /*mode*/) {
  const int nterms = a.n_sets * a.n_terms;
  const int chunk = time_chunk_of(a);
  size_t doubles = ring_lds_doubles(a.n_phi, nterms) + chunk +
                   static_cast<size_t>(chunk) * 2 * a.n_sets;
  if (a.time_tables) doubles += time_table_doubles(a.n_t, nterms);
  return doubles * sizeof(double);
}

namespace {
int resident_blocks(const void* k, int threads, size_t lds) {
  int nb = 0;
  if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, threads, lds) != hipSuccess)
    return 0;
  return nb;
}

// The chunk length: 512 steps (half the anchors and barriers of 256) when the
// run is longer than 256 steps and the larger chunk costs the phase-sum kernel
// of this shape no resident workgroups; 256 otherwise.  Measured: plain_256
// blocked 27.5 -> 26.8 ms with 512; ensemble_512 (64-lane workgroups,
// LDS-bound occupancy) 6-17 % slower with 512, so it keeps 256.  The caller
// holds the LDS-limit lock.  The fused ensemble kernel takes the same choice
// (spectrum_time_chunk), so every form adds the same terms in the same order.
void choose_time_chunk(const void* kp, int threads, SpecArgs& a) {
  a.time_chunk = kTimeChunkMin;
  if (a.n_t <= kTimeChunkMin) return;
  SpecArgs b = a;
  b.time_chunk = kTimeChunkMax;
  const size_t lds = spectrum_lds_bytes(a, 0), lds_b = spectrum_lds_bytes(b, 0);
  if (lds_b > 160 * 1024) return;
  const int nb = resident_blocks(kp, threads, lds_b);
  if (nb > 0 && nb >= resident_blocks(kp, threads, lds)) a.time_chunk = kTimeChunkMax;
}

template <int NS, int OPL, int MODE, int RB>
hipError_t launch_spec_t(const SpecArgs& a_in, int n_items, int threads, hipStream_t s) {
  auto k = spectrum_kernel<NS, OPL, MODE, RB>;
  const void* kp = reinterpret_cast<const void*>(k);
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
  SpecArgs a = a_in;
  choose_time_chunk(kp, threads, a);
  const size_t lds = spectrum_lds_bytes(a, MODE);
  hipError_t e = hipFuncSetAttribute(kp, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_items), dim3(threads), lds, s, a);
  return hipGetLastError();
}

// Launches the phase-sum kernel of the shape (launch != 0), or only returns
// its chunk length in *chunk (launch == 0, nothing is queued).
template <int NS, int OPL, int MODE, int RB>
hipError_t spec_t(const SpecArgs& a, int n_items, int threads, hipStream_t s, bool launch,
                  int* chunk) {
  if (launch) return launch_spec_t<NS, OPL, MODE, RB>(a, n_items, threads, s);
  const void* kp = reinterpret_cast<const void*>(spectrum_kernel<NS, OPL, MODE, RB>);
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
  SpecArgs b = a;
  choose_time_chunk(kp, threads, b);
  *chunk = b.time_chunk;
  return hipSuccess;
}

template <int NS, int MODE, int RB>
hipError_t spec_opl(const SpecArgs& a, int n_items, int opl, int threads, hipStream_t s,
                    bool launch, int* chunk) {
  switch (opl) {
    case 1: return spec_t<NS, 1, MODE, RB>(a, n_items, threads, s, launch, chunk);
    case 2: return spec_t<NS, 2, MODE, RB>(a, n_items, threads, s, launch, chunk);
    default: return spec_t<NS, 4, MODE, RB>(a, n_items, threads, s, launch, chunk);
  }
}

template <int NS>
hipError_t spec_ns(const SpecArgs& a, int n_items, int mode, int opl, int threads, hipStream_t s,
                   bool launch, int* chunk) {
  if (mode == 0) return spec_opl<NS, 0, 16>(a, n_items, opl, threads, s, launch, chunk);
  const bool rb8 = resolve_block_steps(a.mixed_block, mode) == 8;
  if (mode == 1)
    return rb8 ? spec_opl<NS, 1, 8>(a, n_items, opl, threads, s, launch, chunk)
               : spec_opl<NS, 1, 16>(a, n_items, opl, threads, s, launch, chunk);
  return rb8 ? spec_opl<NS, 2, 8>(a, n_items, opl, threads, s, launch, chunk)
             : spec_opl<NS, 2, 16>(a, n_items, opl, threads, s, launch, chunk);
}

hipError_t spec_dispatch(const SpecArgs& a, int n_items, int mode, hipStream_t s, bool launch,
                         int* chunk) {
  if (mode < 0 || mode > 2) return hipErrorInvalidValue;
  // Lanes own omegas: threads = 64..256 (a multiple of 64), up to 4 omegas
  // per lane.  a.threads / a.mixed_block (0 = default) are the tunable choices.
  const int threads = resolve_threads(a.threads, a.n_w, mode);
  if (threads < 64 || threads > 256 || threads % 64) return hipErrorInvalidValue;
  const int opl_needed = (a.n_w + threads - 1) / threads;
  if (opl_needed > 4) return hipErrorInvalidValue;
  const int opl = opl_needed <= 1 ? 1 : (opl_needed <= 2 ? 2 : 4);
  switch (a.n_sets) {
    case 1: return spec_ns<1>(a, n_items, mode, opl, threads, s, launch, chunk);
    case 2: return spec_ns<2>(a, n_items, mode, opl, threads, s, launch, chunk);
    case 3: return spec_ns<3>(a, n_items, mode, opl, threads, s, launch, chunk);
    case 4: return spec_ns<4>(a, n_items, mode, opl, threads, s, launch, chunk);
    case 5: return spec_ns<5>(a, n_items, mode, opl, threads, s, launch, chunk);
    case 6: return spec_ns<6>(a, n_items, mode, opl, threads, s, launch, chunk);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

int spectrum_time_chunk(const SpecArgs& a, int mode) {
  int chunk = kTimeChunkMin;
  if (spec_dispatch(a, 0, mode, nullptr, false, &chunk) != hipSuccess) return kTimeChunkMin;
  return chunk;
}

hipError_t launch_spectrum(const SpecArgs& a, int n_items, int mode, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  return spec_dispatch(a, n_items, mode, s, true, nullptr);
}

namespace {
template <int NS, int RB>
hipError_t launch_adj_t(const SpecArgs& a, int n_items, int threads, size_t lds, hipStream_t s) {
  auto k = adjoint_kernel<NS, RB>;
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
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
    // blocks of 8 (4) steps: the kernel stays at 3 waves per SIMD
    case 1: return launch_adj_t<1, 8>(a, n_items, threads, lds, s);
    case 2: return launch_adj_t<2, 4>(a, n_items, threads, lds, s);
    default: return hipErrorInvalidValue;
  }
}

namespace {
template <int NS, int OPL>
hipError_t launch_int_t(const SpecArgs& a, int n_items, int threads, size_t lds, hipStream_t s) {
  auto k = integrand_kernel<NS, OPL>;
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
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
