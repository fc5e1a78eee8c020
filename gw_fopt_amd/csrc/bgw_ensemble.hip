// SPDX-License-Identifier: MIT
// Batched ensembles: one persistent launch over every (realisation, k-hat)
// group of a batch of nucleation realisations with uneven bubble counts.
//
// A group is one work unit.  Workgroups take units from a global counter in
// launch order (largest first, so the uneven units balance); for its unit a
// workgroup runs, bubble by bubble, the first-collision geometry (delays into
// its own scratch rows of tau / perm) and the phase sums, adds the bubble's C
// to the unit's running sum in bubble order -- the same additions, in the
// same order, as the per-item kernels followed by reduce_kernel -- and writes
// C and S = 8 pi w^3 (|C+|^2 + |Cx|^2) of the group.  Nothing per item is
// kept in HBM.  The counter is cleared by a kernel (queue_reset_kernel), so
// the whole sequence can be captured in a hipGraph and replayed.
#include "bgw_stages.h"

namespace bgw {

namespace {

template <int NS, int OPL, int MODE, int RB>
__global__ __launch_bounds__(256) void ensemble_kernel(EnsArgs a) {
  __shared__ RotCand cand_lds[kGeoLdsCand];
  __shared__ double sort_key[kSortCap];
  __shared__ int32_t sort_val[kSortCap];
  __shared__ int unit_s;
  extern __shared__ double smem[];
  const SpecArgs& g = a.spec;
  const int64_t n_dir = static_cast<int64_t>(g.n_c) * g.n_phi;
  double* tau_s = a.geo.tau + static_cast<int64_t>(blockIdx.x) * n_dir;
  uint16_t* perm_s = a.geo.perm + static_cast<int64_t>(blockIdx.x) * n_dir;
  const int64_t nq = static_cast<int64_t>(g.n_sets_out) * g.n_w;
  const double eight_pi = 8.0 * 3.14159265358979323846;
  for (;;) {
    if (threadIdx.x == 0) unit_s = atomicAdd(a.counter, 1);
    __syncthreads();
    const int u = unit_s;
    __syncthreads();  // every lane has read unit_s before it is overwritten
    if (u >= a.n_units) break;
    const int grp = a.unit_group[u];
    double gs[OPL][NS][4];
#pragma unroll
    for (int q = 0; q < OPL; ++q)
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) gs[q][s][j] = 0.0;
    const int i1 = a.group_item_off[grp + 1];
    for (int item = a.group_item_off[grp]; item < i1; ++item) {
      geometry_item(a.geo, item, tau_s, perm_s, cand_lds, sort_key, sort_val);
      double res[OPL][NS][4];
      spectrum_item<NS, OPL, MODE, RB>(g, item, blockIdx.x, smem, res);
#pragma unroll
      for (int q = 0; q < OPL; ++q)
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
          for (int j = 0; j < 4; ++j) gs[q][s][j] += res[q][s][j];
      __syncthreads();  // the next item's geometry reuses tau_s and the LDS
    }
    const RealInfo ri = g.real[a.group_r[grp]];
#pragma unroll
    for (int q = 0; q < OPL; ++q) {
      const int wi = threadIdx.x + q * blockDim.x;
      if (wi >= g.n_w) continue;
      const double w = g.omega[ri.w_off + wi];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int64_t e = static_cast<int64_t>(g.set0 + s) * g.n_w + wi;
        double* c = a.c_out + static_cast<int64_t>(grp) * 2 * nq * 2;
        c[e * 2] = gs[q][s][0];
        c[e * 2 + 1] = gs[q][s][1];
        c[(nq + e) * 2] = gs[q][s][2];
        c[(nq + e) * 2 + 1] = gs[q][s][3];
        const double ap = hypot(gs[q][s][0], gs[q][s][1]);
        const double ax = hypot(gs[q][s][2], gs[q][s][3]);
        a.s_out[static_cast<int64_t>(grp) * nq + e] = eight_pi * pow(w, 3.0) * (ap * ap + ax * ax);
      }
    }
  }
}

__global__ void queue_reset_kernel(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// kmean[r][e] = (sum_k S[r][k][e]) / n_k, k ascending.
__global__ __launch_bounds__(256) void kmean_kernel(const double* __restrict__ s, int n_real,
                                                    int n_k, int nq, double* __restrict__ kmean) {
  const int64_t n = static_cast<int64_t>(n_real) * nq;
  for (int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; x < n;
       x += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t r = x / nq, e = x % nq;
    double acc = 0.0;
    for (int k = 0; k < n_k; ++k) acc += s[(r * n_k + k) * nq + e];
    kmean[x] = n_k ? acc / static_cast<double>(n_k) : 0.0;
  }
}

// Mean and population standard deviation over realisations, r ascending.
__global__ __launch_bounds__(256) void ensemble_stats_kernel(const double* __restrict__ kmean,
                                                             int n_real, int nq,
                                                             double* __restrict__ mean,
                                                             double* __restrict__ stdev) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nq; e += gridDim.x * blockDim.x) {
    double m = 0.0;
    for (int r = 0; r < n_real; ++r) m += kmean[static_cast<int64_t>(r) * nq + e];
    m = n_real ? m / static_cast<double>(n_real) : 0.0;
    double v = 0.0;
    for (int r = 0; r < n_real; ++r) {
      const double d = kmean[static_cast<int64_t>(r) * nq + e] - m;
      v += d * d;
    }
    v = n_real ? v / static_cast<double>(n_real) : 0.0;
    mean[e] = m;
    stdev[e] = sqrt(v);
  }
}

using EnsKernel = void (*)(EnsArgs);

template <int NS, int OPL>
EnsKernel pick_mode(int mode, int rb) {
  if (mode == 0) return ensemble_kernel<NS, OPL, 0, 16>;
  if (mode == 1) return rb == 8 ? ensemble_kernel<NS, OPL, 1, 8> : ensemble_kernel<NS, OPL, 1, 16>;
  return rb == 8 ? ensemble_kernel<NS, OPL, 2, 8> : ensemble_kernel<NS, OPL, 2, 16>;
}

template <int NS>
EnsKernel pick_opl(int opl, int mode, int rb) {
  switch (opl) {
    case 1: return pick_mode<NS, 1>(mode, rb);
    case 2: return pick_mode<NS, 2>(mode, rb);
    default: return pick_mode<NS, 4>(mode, rb);
  }
}

// The kernel instance and workgroup size for a launch shape; null if invalid.
EnsKernel ensemble_instance(const EnsArgs& a, int mode, int* threads_out) {
  if (mode < 0 || mode > 2) return nullptr;
  const int threads = resolve_threads(a.spec.threads, a.spec.n_w, mode);
  if (threads < 64 || threads > 256 || threads % 64) return nullptr;
  const int opl_needed = (a.spec.n_w + threads - 1) / threads;
  if (opl_needed > 4) return nullptr;
  const int opl = opl_needed <= 1 ? 1 : (opl_needed <= 2 ? 2 : 4);
  *threads_out = threads;
  switch (a.spec.n_sets) {
    case 1: return pick_opl<1>(opl, mode, resolve_block_steps(a.spec.mixed_block, mode));
    case 2: return pick_opl<2>(opl, mode, resolve_block_steps(a.spec.mixed_block, mode));
    default: return nullptr;
  }
}

}  // namespace

size_t ensemble_lds_bytes(const EnsArgs& a, int mode) { return spectrum_lds_bytes(a.spec, mode); }

int ensemble_blocks_per_cu(const EnsArgs& a, int mode) {
  int threads = 0;
  EnsKernel k = ensemble_instance(a, mode, &threads);
  if (!k) return 0;
  int nb = 0;
  const size_t lds = ensemble_lds_bytes(a, mode);
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          static_cast<int>(lds)) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k), threads,
                                                   lds) != hipSuccess)
    return 0;
  return nb;
}

// Sets the kernel's dynamic-LDS limit again, under the lock, right before the
// launch (it is not a stream operation and stays out of a captured graph).
hipError_t launch_ensemble(const EnsArgs& a, int mode, int n_blocks, hipStream_t s) {
  int threads = 0;
  EnsKernel k = ensemble_instance(a, mode, &threads);
  if (!k || n_blocks <= 0) return hipErrorInvalidValue;
  const size_t lds = ensemble_lds_bytes(a, mode);
  hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, a.counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(lds_limit_mutex());
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(n_blocks), dim3(threads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_ensemble_stats(const double* s_spec, int n_real, int n_k, int nq, double* kmean,
                                 double* mean, double* stdev, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(n_real) * nq;
  if (n == 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(kmean_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, s_spec,
                     n_real, n_k, nq, kmean);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !mean) return e;
  hipLaunchKernelGGL(ensemble_stats_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, kmean, n_real,
                     nq, mean, stdev);
  return hipGetLastError();
}

}  // namespace bgw
