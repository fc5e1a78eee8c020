// SPDX-License-Identifier: MIT
// Bubble nucleation on the device: many independent realisations of the
// reference's Monte Carlo nucleation in one launch, one workgroup each.
//
//   fixed rate   FixedRateNucleation (rust/bubble-gw/src/many_bubbles/
//                bubbles_nucleation/fixed_rate_nucleation.rs): sample points
//                in the lattice, per point the earliest time a bubble (or an
//                exterior image of one) reaches it, the false-vacuum volume
//                as the fraction of points not yet reached; new bubbles at
//                the first sample point still in the false vacuum, either
//                when a uniform draw falls below d_p0 in a time-stepping loop
//                (:188-275) or at the times the expected-count ODE gives
//                (:277-356);
//   spontaneous  SpontaneousNucleation (spontaneous_nucleation.rs:38-104):
//                n bubbles at t0, uniform in the lattice.
//
// Design.  A workgroup of 256 lanes owns a realisation.  Each lane keeps the
// first-reach times of K consecutive sample points in registers (points
// lane * K ... lane * K + K - 1), so the false-vacuum count is a register scan plus one workgroup
// reduction.  The reduction also returns the smallest first-reach time above
// t and the first point still in the false vacuum: while t stays below that
// time and no bubble is added the count cannot change, so the sequential
// time loop runs on uniform scalars and rescans only when it must.  Beyond
// 256 * 40 points per realisation the times live in a global workspace.  The
// kernels are held to 256 VGPRs (two workgroups per CU, 512 realisations at
// once on the device).
//
// Random numbers: Philox4x64-10 (Salmon et al., SC'11), key (seed, stream),
// counter (draw group, purpose, 0, 0); a uniform double is (x >> 11) * 2^-53.
// The reference's own generator (rand's StdRng) is not reproduced.
//
// Arithmetic follows the reference's operation order; this file is compiled
// with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "bgw_ffi.h"
#include "bubble_gw.h"

namespace {

using namespace bgw_ffi;

constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI
constexpr double kEps = 2.220446049250313e-16;  // f64::EPSILON
constexpr int kNucThreads = 256;
constexpr int kNucWaves = kNucThreads / 64;
constexpr int kMaxImages = 6;

// Purposes (second Philox counter word).
constexpr uint64_t kStreamPoints = 0;  // sample points: attempt i -> 3 draws
constexpr uint64_t kStreamSteps = 1;   // time-stepping acceptance draws, 4 per group
constexpr uint64_t kStreamSpont = 2;   // spontaneous candidates: attempt i -> 3 draws

// ---------------------------------------------------------------------------
// Lattice with the derived quantities the kernels use (computed on the host)
// ---------------------------------------------------------------------------
struct DevLattice {
  int kind;     // BgwLatticeKind
  int bc;       // BgwBoundaryCondition
  int n_img;    // exterior images per bubble
  int inv_ok;   // parallelepiped basis invertible
  double o[3];
  double e[3][3];
  double r;
  double volume;
  double inv[3][3];
  double n[3][3];  // unit face normals (parallelepiped reflection)
  double q[3][3];  // far-face points o + e_j
};

struct U4 {
  uint64_t v[4];
};

__device__ inline U4 philox(U4 x, uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t hi0 = __umul64hi(M0, x.v[0]), lo0 = M0 * x.v[0];
    const uint64_t hi1 = __umul64hi(M1, x.v[2]), lo1 = M1 * x.v[2];
    U4 y;
    y.v[0] = hi1 ^ x.v[1] ^ k0;
    y.v[1] = lo1;
    y.v[2] = hi0 ^ x.v[3] ^ k1;
    y.v[3] = lo0;
    x = y;
  }
  return x;
}

__device__ inline double uniform(uint64_t x) { return static_cast<double>(x >> 11) * 0x1.0p-53; }

__host__ __device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// One candidate of sample_points from three uniforms; false if the sphere
// rejects it (parallelepiped_lattice.rs:302-330, spherical_lattice.rs:139-176).
__device__ inline bool sample_point(const DevLattice& L, const U4& x, double* p) {
  const double u = uniform(x.v[0]), v = uniform(x.v[1]), w = uniform(x.v[2]);
  if (L.kind == BgwLatticeKind_Parallelepiped) {
#pragma unroll
    for (int i = 0; i < 3; ++i) p[i] = L.o[i] + u * L.e[0][i] + v * L.e[1][i] + w * L.e[2][i];
    return true;
  }
  const double s[3] = {u * 2.0 - 1.0, v * 2.0 - 1.0, w * 2.0 - 1.0};
  if (dot3(s, s) > 1.0) return false;
#pragma unroll
  for (int i = 0; i < 3; ++i) p[i] = L.o[i] + L.r * s[i];
  return true;
}

// contains (parallelepiped_lattice.rs:161-186, spherical_lattice.rs:54-60)
__device__ inline bool contains(const DevLattice& L, const double* p) {
  if (L.kind == BgwLatticeKind_Parallelepiped) {
    if (!L.inv_ok) return false;
    const double d[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
    bool in = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double u = L.inv[i][0] * d[0] + L.inv[i][1] * d[1] + L.inv[i][2] * d[2];
      in = in && u >= 0.0 && u <= 1.0;
    }
    return in;
  }
  const double v[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
  return dot3(v, v) <= L.r * L.r;
}

// Image k (0 <= k < L.n_img) of one event, as generate_bubbles_exterior
// (parallelepiped_lattice.rs:190-298, spherical_lattice.rs:64-134) lists
// them; one image at a time keeps the register footprint small.
__device__ __attribute__((always_inline)) inline void image(const DevLattice& L, const double* ev, int k,
                                                            double* out) {
  const double p[3] = {ev[1], ev[2], ev[3]};
  out[0] = ev[0];
  const int j = k >> 1;
  if (L.kind == BgwLatticeKind_Parallelepiped && L.bc == BgwBoundaryCondition_Periodic) {
    // +e1, -e1, +e2, -e2, +e3, -e3
#pragma unroll
    for (int i = 0; i < 3; ++i) out[1 + i] = p[i] + ((k & 1) ? -L.e[j][i] : L.e[j][i]);
    return;
  }
  if (L.kind == BgwLatticeKind_Parallelepiped) {
    // faces -e1, +e1, -e2, +e2, -e3, +e3: p - 2 ((p - q) . n) n with the
    // outward normal n and a point q of the face
    double n[3], dv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      n[i] = (k & 1) ? L.n[j][i] : -L.n[j][i];
      dv[i] = p[i] - ((k & 1) ? L.q[j][i] : L.o[i]);
    }
    const double d = dot3(dv, n);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[1 + i] = p[i] - 2.0 * d * n[i];
    return;
  }
  // sphere: reflection across the tangent planes at the two points where the
  // line through the center and p meets the sphere
  const double v[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
  const double d = sqrt(dot3(v, v));
  if (d < kEps) {
    out[1] = k == 0 ? L.o[0] + L.r : L.o[0] - L.r;
    out[2] = L.o[1];
    out[3] = L.o[2];
    return;
  }
  const double u[3] = {v[0] / d, v[1] / d, v[2] / d};
  double q[3], n[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    q[i] = k == 0 ? L.o[i] + L.r * u[i] : L.o[i] - L.r * u[i];
    n[i] = k == 0 ? u[i] : -u[i];
  }
  const double pq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  const double dd = pq[0] * n[0] + pq[1] * n[1] + pq[2] * n[2];
#pragma unroll
  for (int i = 0; i < 3; ++i) out[1 + i] = p[i] - 2.0 * dd * n[i];
}

// Earliest time a bubble at src reaches the point (x, y, z)
// (update_first_collision_times, fixed_rate_nucleation.rs:113-129).
__device__ inline double reach_time(const double* src, double x, double y, double z) {
  const double dx = x - src[1];
  const double dy = y - src[2];
  const double dz = z - src[3];
  return src[0] + sqrt(dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// Workgroup primitives
// ---------------------------------------------------------------------------
struct Scan {
  uint32_t before;  // accepted lanes of lower index in the workgroup
  uint32_t total;   // accepted lanes in the workgroup
};

// Exclusive prefix count of a predicate over the workgroup (one barrier on
// entry to publish, one on exit so the slots can be reused).
__device__ inline Scan block_scan(bool pred, uint32_t* slots) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t mask = __ballot(pred);
  const uint32_t in_wave = __popcll(mask & ((uint64_t(1) << lane) - 1));
  if (lane == 0) slots[wave] = __popcll(mask);
  __syncthreads();
  uint32_t off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kNucWaves; ++w) {
    const uint32_t c = slots[w];
    off += w < wave ? c : 0;
    total += c;
  }
  __syncthreads();
  return Scan{off + in_wave, total};
}

struct FalseVacuum {
  uint32_t count;  // sample points not yet reached at t
  uint32_t first;  // first such point (index), or UINT32_MAX
  double next;     // smallest reach time above t (+inf if none)
};

struct ReduceSlots {
  uint32_t count[2][kNucWaves];
  uint32_t first[2][kNucWaves];
  double next[2][kNucWaves];
};

// Workgroup reduction of the per-lane partials; slots alternate between
// calls, so one barrier suffices (a lane that reaches the next call's write
// has passed this call's barrier, after which nobody reads the other slot).
__device__ inline FalseVacuum block_reduce(FalseVacuum v, ReduceSlots& s, int parity) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    v.count += __shfl_xor(v.count, off, 64);
    v.first = min(v.first, static_cast<uint32_t>(__shfl_xor(v.first, off, 64)));
    const double o = __shfl_xor(v.next, off, 64);
    v.next = o < v.next ? o : v.next;
  }
  if (lane == 0) {
    s.count[parity][wave] = v.count;
    s.first[parity][wave] = v.first;
    s.next[parity][wave] = v.next;
  }
  __syncthreads();
  FalseVacuum r{0u, 0xFFFFFFFFu, INFINITY};
#pragma unroll
  for (int w = 0; w < kNucWaves; ++w) {
    r.count += s.count[parity][w];
    r.first = min(r.first, s.first[parity][w]);
    r.next = s.next[parity][w] < r.next ? s.next[parity][w] : r.next;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Fixed-rate nucleation
// ---------------------------------------------------------------------------
struct NucReal {
  DevLattice lat;
  double beta, gamma0, t0;
  uint64_t seed, stream;
  uint64_t times_off, n_times;  // distribution method: this realisation's times
};

struct FixedRateArgs {
  const NucReal* reals;
  int method;  // BgwFixedRateMethod
  double d_p0;
  uint32_t n_points;
  uint64_t max_iter;
  double cutoff_fraction;
  const double* times;
  uint64_t cap;
  double* px;   // (n_real, n_points) sample points, structure of arrays
  double* py;
  double* pz;
  double* fct;  // (n_real, n_points) reach times when they do not fit in registers
  double* interior;
  double* exterior;
  double* t_hist;
  double* v_hist;
  uint64_t* counts;
  int32_t* status;  // 0 ok, 1 sampling failed, 2 capacity exceeded
};

// Per-lane storage of the reach times: K registers per lane (point lane*K+k),
// or the global workspace when K == 0.
template <int K>
struct ReachTimes {
  double f[K > 0 ? K : 1];
  double* g;
  uint32_t n;
  __device__ void init(double* ws, uint32_t n_points) {
    g = ws;
    n = n_points;
    if constexpr (K > 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        f[k] = uint32_t(threadIdx.x * K + k) < n ? INFINITY : -INFINITY;
    } else {
      for (uint32_t i = threadIdx.x; i < n; i += kNucThreads) g[i] = INFINITY;
    }
  }
  __device__ FalseVacuum scan(double t) const {
    FalseVacuum v{0u, 0xFFFFFFFFu, INFINITY};
    if constexpr (K > 0) {
#pragma unroll
      for (int k = K - 1; k >= 0; --k) {
        const double x = f[k];
        const bool live = x > t;
        v.count += live ? 1u : 0u;
        v.next = (live && x < v.next) ? x : v.next;
        v.first = live ? uint32_t(threadIdx.x * K + k) : v.first;
      }
    } else {
      for (uint32_t i = threadIdx.x; i < n; i += kNucThreads) {
        const double x = g[i];
        if (x > t) {
          v.count += 1u;
          v.next = x < v.next ? x : v.next;
          v.first = min(v.first, i);
        }
      }
    }
    return v;
  }
  // min-update with the new bubble and its images (n_src rows of src, in LDS;
  // the source loop stays a loop so only one source is live at a time)
  __device__ void update(const double (*src)[4], int n_src, const double* px, const double* py,
                         const double* pz) {
    if constexpr (K > 0) {
      // one base address per array, the K points at immediate offsets
      const uint32_t i0 = threadIdx.x * K;
      const double* bx = px + i0;
      const double* by = py + i0;
      const double* bz = pz + i0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (i0 + k < n) {
          const double x = bx[k], y = by[k], z = bz[k];
          double m = f[k];
#pragma unroll 1
          for (int j = 0; j < n_src; ++j) {
            const double tc = reach_time(src[j], x, y, z);
            m = tc < m ? tc : m;
          }
          f[k] = m;
        }
      }
    } else {
      for (uint32_t i = threadIdx.x; i < n; i += kNucThreads) {
        const double x = px[i], y = py[i], z = pz[i];
        double m = g[i];
#pragma unroll 1
        for (int j = 0; j < n_src; ++j) {
          const double tc = reach_time(src[j], x, y, z);
          m = tc < m ? tc : m;
        }
        g[i] = m;
      }
    }
  }
};

template <int K>
__global__ __launch_bounds__(kNucThreads) __attribute__((amdgpu_waves_per_eu(2)))
void fixed_rate_kernel(FixedRateArgs a) {
  __shared__ ReduceSlots slots;
  __shared__ uint32_t scan_slots[kNucWaves];
  __shared__ double s_src[1 + kMaxImages][4];  // the new bubble and its images
  __shared__ uint8_t s_accept[4 * kNucThreads];  // time stepping: draw <= d_p0 per iteration
  const int r = blockIdx.x;
  const NucReal& R = a.reals[r];
  const DevLattice& L = R.lat;
  // scalars of the sequential loop in registers (not re-read from memory every step)
  const double beta = R.beta, gamma0 = R.gamma0, t_start = R.t0;
  const uint64_t seed = R.seed, stream = R.stream;
  const uint32_t np = a.n_points;
  double* px = a.px + size_t(r) * np;
  double* py = a.py + size_t(r) * np;
  double* pz = a.pz + size_t(r) * np;

  // sample_points: the first np accepted candidates in attempt order
  const uint64_t max_att = uint64_t(np) * (L.kind == BgwLatticeKind_Parallelepiped ? 1000 : 10000);
  uint64_t got = 0;
  for (uint64_t base = 0; got < np && base < max_att; base += kNucThreads) {
    const uint64_t att = base + threadIdx.x;
    double p[3] = {0, 0, 0};
    bool ok = false;
    if (att < max_att) ok = sample_point(L, philox(U4{{att, kStreamPoints, 0, 0}}, seed, stream), p);
    const Scan s = block_scan(ok, scan_slots);
    const uint64_t idx = got + s.before;
    if (ok && idx < np) {
      px[idx] = p[0];
      py[idx] = p[1];
      pz[idx] = p[2];
    }
    got += s.total;
  }
  if (got < np) {
    if (threadIdx.x == 0) {
      a.counts[r] = 0;
      a.status[r] = 1;
    }
    return;
  }
  __syncthreads();  // the points are read by other lanes below

  ReachTimes<K> fct;
  fct.init(a.fct + size_t(r) * np, np);
  if constexpr (K == 0) __syncthreads();

  const double volume_lattice = L.volume;
  const double volume_cutoff = a.cutoff_fraction * volume_lattice;
  const size_t cap = a.cap;
  double* out_int = a.interior + size_t(r) * cap * 4;
  double* out_ext = a.exterior ? a.exterior + size_t(r) * cap * L.n_img * 4 : nullptr;
  double* out_t = a.t_hist + size_t(r) * cap;
  double* out_v = a.v_hist + size_t(r) * cap;

  uint64_t nb = 0;
  int32_t status = 0;
  int parity = 0;
  FalseVacuum fv{0u, 0xFFFFFFFFu, INFINITY};
  double t_scan = INFINITY;  // time of the last scan; fv is valid on [t_scan, fv.next)
  bool stale = true;

  // a bubble at the first false-vacuum point: lane 0 records it and its
  // images (also into LDS), then every lane min-updates its points.  s_src is
  // rewritten only after the next scan's barrier, when every lane is done
  // reading it.
  auto nucleate = [&](double t, double vrem, double t_rec) {
    if (threadIdx.x == 0) {
      const uint32_t site = fv.first;
      const double ev[4] = {t, px[site], py[site], pz[site]};
      for (int c = 0; c < 4; ++c) {
        out_int[nb * 4 + c] = ev[c];
        s_src[0][c] = ev[c];
      }
      out_t[nb] = t_rec;
      out_v[nb] = vrem;
#pragma unroll 1
      for (int j = 0; j < L.n_img; ++j) {
        double img[4];
        image(L, ev, j, img);
        for (int c = 0; c < 4; ++c) {
          s_src[1 + j][c] = img[c];
          out_ext[(nb * L.n_img + j) * 4 + c] = img[c];
        }
      }
    }
    __syncthreads();
    fct.update(s_src, 1 + L.n_img, px, py, pz);
    if constexpr (K == 0) __syncthreads();
    ++nb;
    stale = true;
  };
  auto refresh = [&](double t) {
    if (stale || !(t >= t_scan && t < fv.next)) {
      fv = block_reduce(fct.scan(t), slots, parity);
      parity ^= 1;
      t_scan = t;
      stale = false;
    }
  };

  if (a.method == BgwFixedRateMethod_TimeStepping) {
    double t = t_start;
    // The acceptance draws do not depend on the state: the workgroup evaluates
    // them 4 * kNucThreads iterations at a time (one Philox block per lane)
    // into LDS, off the sequential chain.
    uint64_t draw_base = 0;
    // the false-vacuum volume and V / V_fv change only with the count
    uint32_t count_seen = 0xFFFFFFFFu;
    double volume_remaining = 0.0, volume_ratio = 0.0;
    for (uint64_t it = 0; it < a.max_iter; ++it) {
      refresh(t);
      if (fv.count != count_seen) {
        count_seen = fv.count;
        const double fraction = static_cast<double>(fv.count) / static_cast<double>(np);
        volume_remaining = fmax(fraction * volume_lattice, 0.0);
        volume_ratio = volume_lattice / volume_remaining;
      }
      if (volume_remaining < volume_cutoff) break;
      const double exponent = beta * (t - t_start);
      const double gamma_t = gamma0 * exp(exponent);
      if (it == 0 || it - draw_base >= 4 * kNucThreads) {
        draw_base = it;  // a multiple of 4 * kNucThreads
        __syncthreads();  // every lane is done with the previous span
        const U4 x = philox(U4{{(it >> 2) + threadIdx.x, kStreamSteps, 0, 0}}, seed, stream);
#pragma unroll
        for (int q = 0; q < 4; ++q) s_accept[4 * threadIdx.x + q] = uniform(x.v[q]) <= a.d_p0;
        __syncthreads();
      }
      if (s_accept[it - draw_base] && fv.count > 0) {
        if (nb >= cap) {
          status = 2;
          break;
        }
        nucleate(t, volume_remaining, t);
      }
      const double dt = (a.d_p0 / gamma_t) * volume_ratio;
      t += dt;
    }
  } else {
    const double* times = a.times + R.times_off;
    for (uint64_t k = 0; k < R.n_times; ++k) {
      const double t = times[k];
      refresh(t);
      const double fraction = static_cast<double>(fv.count) / static_cast<double>(np);
      const double volume_remaining = fmax(fraction * volume_lattice, 0.0);
      if (volume_remaining < volume_cutoff) break;
      if (fv.count > 0) {
        if (nb >= cap) {
          status = 2;
          break;
        }
        // the reference returns the first n times of the list (:347-348)
        nucleate(t, volume_remaining, times[nb]);
      }
    }
  }
  if (threadIdx.x == 0) {
    a.counts[r] = nb;
    a.status[r] = status;
  }
}

// ---------------------------------------------------------------------------
// Spontaneous nucleation
// ---------------------------------------------------------------------------
struct SpontArgs {
  DevLattice lat;
  uint32_t n_bubbles;
  double t0;
  uint64_t seed, stream0;
  double* interior;  // (n_real, n_bubbles, 4)
  double* exterior;  // (n_real, n_bubbles * n_img, 4)
  int32_t* status;   // 0 ok, 1 insufficient bubbles
  uint32_t* counts;
};

// Candidates come from sample_points(1) (for the sphere: the next in-ball
// cube sample), are kept if inside the lattice; the reference's light-cone
// test against accepted bubbles (spontaneous_nucleation.rs:62-67) never
// rejects here because every candidate has the same time t0, so
// s(d, d) = |d_x|^2 >= 0.  The outer attempt limit n_bubbles * 10000 counts
// sample_points calls.
__global__ __launch_bounds__(kNucThreads) void spontaneous_kernel(SpontArgs a) {
  __shared__ uint32_t scan_slots[kNucWaves];
  const int r = blockIdx.x;
  const DevLattice& L = a.lat;
  const uint64_t stream = a.stream0 + r;
  const uint32_t nb = a.n_bubbles;
  const uint64_t max_outer = uint64_t(nb) * 10000;
  double* out_int = a.interior + size_t(r) * nb * 4;
  double* out_ext = a.exterior ? a.exterior + size_t(r) * nb * L.n_img * 4 : nullptr;
  uint64_t got = 0, outer = 0;
  const uint64_t max_draw = max_outer * (L.kind == BgwLatticeKind_Parallelepiped ? 1 : 10000);
  for (uint64_t base = 0; got < nb && outer < max_outer && base < max_draw; base += kNucThreads) {
    const uint64_t draw = base + threadIdx.x;
    double p[3] = {0, 0, 0};
    bool sampled = false;
    if (draw < max_draw) sampled = sample_point(L, philox(U4{{draw, kStreamSpont, 0, 0}}, a.seed, stream), p);
    const Scan so = block_scan(sampled, scan_slots);
    const bool within = sampled && outer + so.before < max_outer;
    const bool ok = within && contains(L, p);
    const Scan sa = block_scan(ok, scan_slots);
    const uint64_t idx = got + sa.before;
    if (ok && idx < nb) {
      const double ev[4] = {a.t0, p[0], p[1], p[2]};
      for (int c = 0; c < 4; ++c) out_int[idx * 4 + c] = ev[c];
#pragma unroll 1
      for (int j = 0; j < L.n_img; ++j) {
        double img[4];
        image(L, ev, j, img);
        for (int c = 0; c < 4; ++c) out_ext[(idx * L.n_img + j) * 4 + c] = img[c];
      }
    }
    got += sa.total;
    outer += so.total;
  }
  if (threadIdx.x == 0) {
    a.counts[r] = static_cast<uint32_t>(got < nb ? got : nb);
    a.status[r] = got < nb ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int images_per_bubble(int kind, int bc) {
  if (bc == BgwBoundaryCondition_None) return 0;
  if (kind == BgwLatticeKind_Parallelepiped) return 6;
  return bc == BgwBoundaryCondition_Reflection ? 2 : 0;
}

// Matrix3::from_columns(basis): m(row, col) = e_col[row]; nalgebra's 3x3
// determinant and inverse in cofactor form.
DevLattice make_lattice(const BgwNucleationLattice& in, int bc) {
  DevLattice L;
  std::memset(&L, 0, sizeof L);
  if (in.kind != BgwLatticeKind_Parallelepiped && in.kind != BgwLatticeKind_Spherical)
    throw CalcError("lattice kind must be Parallelepiped (0) or Spherical (1)");
  if (bc != BgwBoundaryCondition_Periodic && bc != BgwBoundaryCondition_Reflection &&
      bc != BgwBoundaryCondition_None)
    throw CalcError("Invalid boundary condition. Expected 'periodic', 'reflection', or 'none'.");
  L.kind = in.kind;
  L.bc = bc;
  L.n_img = images_per_bubble(in.kind, bc);
  for (int i = 0; i < 3; ++i) L.o[i] = in.origin[i];
  for (double v : in.origin)
    if (!std::isfinite(v)) throw CalcError("lattice origin must be finite");
  if (in.kind == BgwLatticeKind_Spherical) {
    if (!(std::isfinite(in.radius) && in.radius > 0.0))
      throw CalcError("SphericalLattice radius must be finite and positive");
    L.r = in.radius;
    L.volume = (4.0 / 3.0) * kPi * (L.r * L.r * L.r);
    return L;
  }
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      L.e[j][i] = in.basis[3 * j + i];
      if (!std::isfinite(L.e[j][i])) throw CalcError("lattice basis must be finite");
    }
  auto m = [&](int row, int col) { return L.e[col][row]; };
  const double minor_m12_m23 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
  const double minor_m11_m23 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
  const double minor_m11_m22 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  const double det = m(0, 0) * minor_m12_m23 - m(0, 1) * minor_m11_m23 + m(0, 2) * minor_m11_m22;
  if (std::fabs(det) <= kEps) throw CalcError("Invalid ParallelepipedLattice: degenerate basis");
  L.volume = std::fabs(det);
  L.inv_ok = 1;
  L.inv[0][0] = minor_m12_m23 / det;
  L.inv[0][1] = (m(0, 2) * m(2, 1) - m(2, 2) * m(0, 1)) / det;
  L.inv[0][2] = (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2)) / det;
  L.inv[1][0] = -minor_m11_m23 / det;
  L.inv[1][1] = (m(0, 0) * m(2, 2) - m(2, 0) * m(0, 2)) / det;
  L.inv[1][2] = (m(0, 2) * m(1, 0) - m(1, 2) * m(0, 0)) / det;
  L.inv[2][0] = minor_m11_m22 / det;
  L.inv[2][1] = (m(0, 1) * m(2, 0) - m(2, 1) * m(0, 0)) / det;
  L.inv[2][2] = (m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) / det;
  const double axis[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int j = 0; j < 3; ++j) {
    const double nrm = std::sqrt(dot3(L.e[j], L.e[j]));
    for (int i = 0; i < 3; ++i) {
      L.n[j][i] = nrm > kEps ? L.e[j][i] / nrm : axis[j][i];
      L.q[j][i] = L.o[i] + L.e[j][i];
    }
  }
  return L;
}

// Expected-count ODE of fixed_rate_nucleation.rs:62-73 in tau = beta t:
//   dm0/dtau = exp(ln(gamma0 beta^-4) + tau - (4 pi / 3) m3),
//   dm1/dtau = m0, dm2/dtau = 2 m1, dm3/dtau = 3 m2,
// integrated with the classical fourth-order Runge-Kutta method on the even
// output grid of step 0.001 (the reference integrates with Radau IIA order 5
// at rtol 1e-9, atol 1e-12 and interpolates onto the same grid).  Returns the
// grid points where floor(m0 V beta^3) increases (:160-185).
void solve_distribution(double beta, double gamma0, double t0, double taumax, double volume,
                        std::vector<double>& t_out, std::vector<double>& n_out) {
  t_out.clear();
  n_out.clear();
  const double beta2 = beta * beta;
  const double gamma0_bar = gamma0 * (1.0 / (beta2 * beta2));
  const double log_g = std::log(gamma0_bar);
  const double c43 = 4.0 * kPi / 3.0;
  const double scale = volume * (beta2 * beta);
  auto f = [&](double tau, const double* m, double* dm) {
    dm[0] = std::exp(log_g + tau - c43 * m[3]);
    dm[1] = m[0];
    dm[2] = 2.0 * m[1];
    dm[3] = 3.0 * m[2];
  };
  const double h = 0.001;
  const double tau0 = t0 * beta;
  if (!(taumax > tau0)) return;
  const size_t n_steps = static_cast<size_t>(std::floor((taumax - tau0) / h + 1e-9));
  double m[4] = {0, 0, 0, 0};
  long long prev = static_cast<long long>(std::floor(m[0] * scale));
  for (size_t i = 1; i <= n_steps; ++i) {
    const double tau = tau0 + static_cast<double>(i - 1) * h;
    double k1[4], k2[4], k3[4], k4[4], y[4];
    f(tau, m, k1);
    for (int c = 0; c < 4; ++c) y[c] = m[c] + 0.5 * h * k1[c];
    f(tau + 0.5 * h, y, k2);
    for (int c = 0; c < 4; ++c) y[c] = m[c] + 0.5 * h * k2[c];
    f(tau + 0.5 * h, y, k3);
    for (int c = 0; c < 4; ++c) y[c] = m[c] + h * k3[c];
    f(tau + h, y, k4);
    for (int c = 0; c < 4; ++c) m[c] += h / 6.0 * (k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + k4[c]);
    const double n = m[0] * scale;
    const long long cur = static_cast<long long>(std::floor(n));
    if (cur > prev) {
      t_out.push_back((tau0 + static_cast<double>(i) * h) / beta);
      n_out.push_back(n);
    }
    prev = cur;
  }
}

struct DevMem {
  void* p = nullptr;
  explicit DevMem(size_t bytes) { HIPCHECK(hipMalloc(&p, bytes > 0 ? bytes : 16)); }
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// A timing event, destroyed on every exit path.
struct TimingEvent {
  hipEvent_t e = nullptr;
  TimingEvent() { HIPCHECK(hipEventCreate(&e)); }
  ~TimingEvent() {
    if (e) (void)hipEventDestroy(e);
  }
  TimingEvent(const TimingEvent&) = delete;
  TimingEvent& operator=(const TimingEvent&) = delete;
};

thread_local double t_last_kernel_ms = 0.0;

void need_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw CalcError("no HIP device is available: the nucleation kernels need a gfx950 GPU");
}

template <int K>
void launch_fixed_rate(int n_real, const FixedRateArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(fixed_rate_kernel<K>, dim3(n_real), dim3(kNucThreads), 0, s, a);
}

int registers_per_lane(uint32_t n_points) {
  for (int k : {2, 4, 8, 16, 24, 32, 40})
    if (uint64_t(k) * kNucThreads >= n_points) return k;
  return 0;
}

}  // namespace

extern "C" {

size_t bgw_nucleation_images_per_bubble(int32_t lattice_kind, BgwBoundaryCondition boundary_condition) {
  return static_cast<size_t>(images_per_bubble(lattice_kind, boundary_condition));
}

BgwStatus bgw_nucleation_solve_distribution(double beta, double gamma0, double t0, double taumax,
                                            double volume_lattice, double* t_out, double* n_out,
                                            size_t cap, size_t* n_len) {
  return run_ffi([&] {
    if (!n_len) throw CalcError(null_error("n_len"));
    if (!(std::isfinite(beta) && beta > 0.0)) throw CalcError("beta must be finite and positive");
    if (!(std::isfinite(gamma0) && gamma0 >= 0.0)) throw CalcError("gamma0 must be non-negative");
    if (!std::isfinite(t0) || !std::isfinite(taumax) || !std::isfinite(volume_lattice))
      throw CalcError("Parameters must be finite");
    std::vector<double> t, n;
    solve_distribution(beta, gamma0, t0, taumax, volume_lattice, t, n);
    *n_len = t.size();
    if (t_out) {
      if (!n_out) throw CalcError(null_error("n_out"));
      if (cap < t.size()) throw CalcError("output capacity is smaller than the number of crossings");
      std::copy(t.begin(), t.end(), t_out);
      std::copy(n.begin(), n.end(), n_out);
    }
  });
}

BgwStatus bgw_nucleation_fixed_rate(const BgwFixedRateRealisation* realisations, size_t n_realisations,
                                    BgwBoundaryCondition boundary_condition, BgwFixedRateMethod method,
                                    double d_p0, size_t n_points, size_t max_time_iterations,
                                    double cutoff_fraction, size_t max_bubbles, double* interior,
                                    double* exterior, double* time_history, double* volume_history,
                                    size_t* counts) {
  return run_ffi([&] {
    if (n_realisations == 0) return;
    slice_from_raw(realisations, n_realisations, "realisations");
    if (method != BgwFixedRateMethod_TimeStepping && method != BgwFixedRateMethod_Distribution)
      throw CalcError("method must be TimeStepping (0) or Distribution (1)");
    if (!std::isfinite(d_p0)) throw CalcError("Parameters must be finite");
    if (d_p0 <= 0.0 || d_p0 > 1.0) throw CalcError("d_p0 must be in (0, 1]");
    if (n_points == 0) throw CalcError("n_points must be > 0");
    if (n_points > 0x7FFFFFFFu) throw CalcError("n_points must be < 2^31");
    if (!std::isfinite(cutoff_fraction)) throw CalcError("cutoff_fraction must be finite");
    if (max_bubbles == 0) throw CalcError("max_bubbles must be > 0");
    if (n_realisations > 0x7FFFFFFFu) throw CalcError("too many realisations");
    std::vector<NucReal> reals(n_realisations);
    std::vector<double> times;
    std::map<std::tuple<double, double, double, double>, std::pair<uint64_t, uint64_t>> solved;
    int n_img = -1;
    for (size_t r = 0; r < n_realisations; ++r) {
      const BgwFixedRateRealisation& in = realisations[r];
      if (!std::isfinite(in.beta) || !std::isfinite(in.gamma0) || !std::isfinite(in.t0))
        throw CalcError("Parameters must be finite");
      if (in.gamma0 < 0.0) throw CalcError("gamma0 must be non-negative");
      NucReal& R = reals[r];
      R.lat = make_lattice(in.lattice, boundary_condition);
      if (n_img < 0) n_img = R.lat.n_img;
      if (n_img != R.lat.n_img) throw CalcError("every realisation must use the same lattice kind");
      R.beta = in.beta;
      R.gamma0 = in.gamma0;
      R.t0 = in.t0;
      R.seed = in.seed;
      R.stream = in.stream;
      R.times_off = 0;
      R.n_times = 0;
      if (method == BgwFixedRateMethod_Distribution) {
        if (!(in.beta > 0.0)) throw CalcError("beta must be positive");
        const auto key = std::make_tuple(in.beta, in.gamma0, in.t0, R.lat.volume);
        auto it = solved.find(key);
        if (it == solved.end()) {
          std::vector<double> t, n;
          solve_distribution(in.beta, in.gamma0, in.t0, 40.0, R.lat.volume, t, n);
          it = solved.emplace(key, std::make_pair(uint64_t(times.size()), uint64_t(t.size()))).first;
          times.insert(times.end(), t.begin(), t.end());
        }
        R.times_off = it->second.first;
        R.n_times = it->second.second;
      }
    }
    const size_t cap = max_bubbles;
    const size_t rows = checked_product(n_realisations, cap, "interior");
    slice_from_raw(interior, rows, "interior");
    slice_from_raw(time_history, rows, "time_history");
    slice_from_raw(volume_history, rows, "volume_history");
    slice_from_raw(counts, n_realisations, "counts");
    if (n_img > 0) slice_from_raw(exterior, rows * n_img, "exterior");
    need_device();

    const size_t np = n_points;
    const int K = registers_per_lane(static_cast<uint32_t>(np));
    const size_t pts = checked_product(n_realisations, np, "sample points");
    DevMem d_reals(sizeof(NucReal) * n_realisations);
    DevMem d_times(sizeof(double) * times.size());
    DevMem d_px(sizeof(double) * pts), d_py(sizeof(double) * pts), d_pz(sizeof(double) * pts);
    DevMem d_fct(K == 0 ? sizeof(double) * pts : 16);
    DevMem d_int(sizeof(double) * rows * 4);
    DevMem d_ext(sizeof(double) * rows * std::max(n_img, 1) * 4);
    DevMem d_t(sizeof(double) * rows), d_v(sizeof(double) * rows);
    DevMem d_counts(sizeof(uint64_t) * n_realisations), d_status(sizeof(int32_t) * n_realisations);
    HIPCHECK(hipMemcpy(d_reals.p, reals.data(), sizeof(NucReal) * n_realisations, hipMemcpyHostToDevice));
    if (!times.empty())
      HIPCHECK(hipMemcpy(d_times.p, times.data(), sizeof(double) * times.size(), hipMemcpyHostToDevice));

    FixedRateArgs a;
    a.reals = d_reals.as<NucReal>();
    a.method = method;
    a.d_p0 = d_p0;
    a.n_points = static_cast<uint32_t>(np);
    a.max_iter = max_time_iterations;
    a.cutoff_fraction = cutoff_fraction;
    a.times = d_times.as<double>();
    a.cap = cap;
    a.px = d_px.as<double>();
    a.py = d_py.as<double>();
    a.pz = d_pz.as<double>();
    a.fct = d_fct.as<double>();
    a.interior = d_int.as<double>();
    a.exterior = n_img > 0 ? d_ext.as<double>() : nullptr;
    a.t_hist = d_t.as<double>();
    a.v_hist = d_v.as<double>();
    a.counts = d_counts.as<uint64_t>();
    a.status = d_status.as<int32_t>();

    TimingEvent ev0, ev1;
    const int nr = static_cast<int>(n_realisations);
    HIPCHECK(hipEventRecord(ev0.e, nullptr));
    switch (K) {
      case 2: launch_fixed_rate<2>(nr, a, nullptr); break;
      case 4: launch_fixed_rate<4>(nr, a, nullptr); break;
      case 8: launch_fixed_rate<8>(nr, a, nullptr); break;
      case 16: launch_fixed_rate<16>(nr, a, nullptr); break;
      case 24: launch_fixed_rate<24>(nr, a, nullptr); break;
      case 32: launch_fixed_rate<32>(nr, a, nullptr); break;
      case 40: launch_fixed_rate<40>(nr, a, nullptr); break;
      default: launch_fixed_rate<0>(nr, a, nullptr); break;
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ev1.e, nullptr));
    HIPCHECK(hipEventSynchronize(ev1.e));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, ev0.e, ev1.e));
    t_last_kernel_ms = ms;

    std::vector<uint64_t> cnt(n_realisations);
    std::vector<int32_t> st(n_realisations);
    HIPCHECK(hipMemcpy(cnt.data(), d_counts.p, sizeof(uint64_t) * n_realisations, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(st.data(), d_status.p, sizeof(int32_t) * n_realisations, hipMemcpyDeviceToHost));
    for (size_t r = 0; r < n_realisations; ++r) {
      if (st[r] == 1)
        throw CalcError("Failed to sample " + std::to_string(np) + " points in the lattice (realisation " +
                        std::to_string(r) + ")");
      if (st[r] == 2)
        throw CalcError("realisation " + std::to_string(r) + " nucleates more than max_bubbles = " +
                        std::to_string(cap) + " bubbles");
      counts[r] = static_cast<size_t>(cnt[r]);
    }
    HIPCHECK(hipMemcpy(interior, d_int.p, sizeof(double) * rows * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(time_history, d_t.p, sizeof(double) * rows, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(volume_history, d_v.p, sizeof(double) * rows, hipMemcpyDeviceToHost));
    if (n_img > 0)
      HIPCHECK(hipMemcpy(exterior, d_ext.p, sizeof(double) * rows * n_img * 4, hipMemcpyDeviceToHost));
  });
}

BgwStatus bgw_nucleation_spontaneous(const BgwNucleationLattice* lattice,
                                     BgwBoundaryCondition boundary_condition, size_t n_bubbles,
                                     double t0, uint64_t seed, uint64_t stream0, size_t n_realisations,
                                     double* interior, double* exterior) {
  return run_ffi([&] {
    if (!lattice) throw CalcError(null_error("lattice"));
    if (!std::isfinite(t0)) throw CalcError("t0 must be finite");
    if (n_bubbles > 0x7FFFFFFFu) throw CalcError("n_bubbles must be < 2^31");
    if (n_realisations > 0x7FFFFFFFu) throw CalcError("too many realisations");
    const DevLattice L = make_lattice(*lattice, boundary_condition);
    if (n_realisations == 0 || n_bubbles == 0) return;
    const size_t rows = checked_product(n_realisations, n_bubbles, "interior");
    slice_from_raw(interior, rows, "interior");
    if (L.n_img > 0) slice_from_raw(exterior, rows * L.n_img, "exterior");
    need_device();
    DevMem d_int(sizeof(double) * rows * 4);
    DevMem d_ext(sizeof(double) * rows * std::max(L.n_img, 1) * 4);
    DevMem d_status(sizeof(int32_t) * n_realisations), d_counts(sizeof(uint32_t) * n_realisations);
    SpontArgs a;
    a.lat = L;
    a.n_bubbles = static_cast<uint32_t>(n_bubbles);
    a.t0 = t0;
    a.seed = seed;
    a.stream0 = stream0;
    a.interior = d_int.as<double>();
    a.exterior = L.n_img > 0 ? d_ext.as<double>() : nullptr;
    a.status = d_status.as<int32_t>();
    a.counts = d_counts.as<uint32_t>();
    hipLaunchKernelGGL(spontaneous_kernel, dim3(static_cast<unsigned>(n_realisations)), dim3(kNucThreads), 0,
                       nullptr, a);
    HIPCHECK(hipGetLastError());
    std::vector<int32_t> st(n_realisations);
    std::vector<uint32_t> cnt(n_realisations);
    HIPCHECK(hipMemcpy(st.data(), d_status.p, sizeof(int32_t) * n_realisations, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(cnt.data(), d_counts.p, sizeof(uint32_t) * n_realisations, hipMemcpyDeviceToHost));
    for (size_t r = 0; r < n_realisations; ++r)
      if (st[r] != 0)
        throw CalcError("Insufficient Bubbles: requested " + std::to_string(n_bubbles) + ", generated " +
                        std::to_string(cnt[r]));
    HIPCHECK(hipMemcpy(interior, d_int.p, sizeof(double) * rows * 4, hipMemcpyDeviceToHost));
    if (L.n_img > 0)
      HIPCHECK(hipMemcpy(exterior, d_ext.p, sizeof(double) * rows * L.n_img * 4, hipMemcpyDeviceToHost));
  });
}

BgwStatus bgw_nucleation_last_kernel_ms(double* out) {
  return run_ffi([&] {
    if (!out) throw CalcError(null_error("out"));
    *out = t_last_kernel_ms;
  });
}

}  // extern "C"
