// SPDX-License-Identifier: MIT
// Host side of libbubble_gw_c.so: the reference's C ABI
// (rust/c-bubble-gw/src/lib.rs) over the gfx950 kernels.
//
// Host work is limited to what the reference also does on the host once per
// configuration: copying and validating the bubbles (LatticeBubbles::new,
// rust/bubble-gw/src/many_bubbles/lattice_bubbles.rs:211-308), building the
// angular grids (generalized_bulk_flow.rs:208-270), and sorting each
// reference bubble's partners by a collision-delay lower bound.  Everything
// per k-hat and per time step runs on the device.  There is no CPU fallback:
// without a HIP device every compute call fails with CalculationError.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "bgw_common.h"
#include "bgw_ffi.h"
#include "bgw_kernels.h"
#include "bubble_gw.h"

using bgw::CandEntry;
using bgw::Item;
using bgw::RealInfo;

namespace {

constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI
constexpr size_t kMaxSetsPerLaunch = 6;

// Errors (lib.rs:14-16, 104-134): shared with the other host units.
using namespace bgw_ffi;

// ---------------------------------------------------------------------------
// Arithmetic shared with the kernels (compiled with -ffp-contract=off)
// ---------------------------------------------------------------------------
inline double lscalar(const double* u, const double* v) {
  double a = (-u[0]) * v[0];
  double b = u[1] * v[1];
  double c = u[2] * v[2];
  double d = u[3] * v[3];
  a = a + c;
  b = b + d;
  return a + b;
}

inline double linspace_at(double a, double b, size_t n, size_t i) {
  double step = n > 1 ? (b - a) / static_cast<double>(n - 1) : 0.0;
  return a + step * static_cast<double>(i);
}

// Rotation taking the unit vector k to z-hat: R = R_y(-theta) R_z(-phi); its
// last row is k itself, so the rotated z of a point is k . x.
std::array<double, 9> rotation_to_z(const double* k) {
  double st = std::sqrt(k[0] * k[0] + k[1] * k[1]);
  double cp = 1.0, sp = 0.0;
  if (st > 0.0) {
    cp = k[0] / st;
    sp = k[1] / st;
  }
  double ct = k[2];
  return {ct * cp, ct * sp, -st, -sp, cp, 0.0, k[0], k[1], k[2]};
}

bool valid_precision(BgwPrecision p) {
  return p == BgwPrecision_Fp64 || p == BgwPrecision_Mixed || p == BgwPrecision_Fp64Blocked ||
         p == BgwPrecision_Fp64Matrix;
}

// Phase-sum kernel variant of a precision.
int kernel_mode(BgwPrecision p) {
  switch (p) {
    case BgwPrecision_Fp64: return 0;
    case BgwPrecision_Fp64Blocked: return 1;
    case BgwPrecision_Mixed: return 2;
    case BgwPrecision_Fp64Matrix: return 3;
  }
  throw CalcError("precision is not a BgwPrecision");
}

std::string bubble_index_name(size_t b, size_t n_int) {
  return b < n_int ? "Interior(" + std::to_string(b) + ")"
                   : "Exterior(" + std::to_string(b - n_int) + ")";
}

// ---------------------------------------------------------------------------
// Bubbles of one realisation and their sorted candidate lists
// ---------------------------------------------------------------------------
struct Bubbles {
  size_t n_int = 0, n_ext = 0;
  std::vector<double> xyz;  // (n_int + n_ext, 4): interior rows then exterior rows
  bool nonfinite = false;
  size_t n_tot() const { return n_int + n_ext; }
  const double* row(size_t b) const { return &xyz[b * 4]; }
};

Bubbles make_bubbles(const double* interior, size_t n_int, const double* exterior, size_t n_ext) {
  Bubbles B;
  B.n_int = n_int;
  B.n_ext = n_ext;
  B.xyz.resize((n_int + n_ext) * 4);
  if (n_int) std::memcpy(B.xyz.data(), interior, n_int * 4 * sizeof(double));
  if (n_ext) std::memcpy(B.xyz.data() + n_int * 4, exterior, n_ext * 4 * sizeof(double));
  for (double v : B.xyz)
    if (!std::isfinite(v)) B.nonfinite = true;
  // check_bubble_formed_inside_bubble (lattice_bubbles.rs:71-119), in the
  // reference's loop order: interior-interior, then interior-exterior.
  for (size_t a = 0; a < n_int; ++a)
    for (size_t b = a + 1; b < n_int; ++b) {
      const double* xa = B.row(a);
      const double* xb = B.row(b);
      double d[4] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2], xb[3] - xa[3]};
      if (lscalar(d, d) < 0.0)
        throw CalcError("Bubble " + bubble_index_name(a, n_int) + " is formed inside bubble " +
                        bubble_index_name(b, n_int) +
                        " at initial time (overlapping light cones)");
    }
  for (size_t a = 0; a < n_int; ++a)
    for (size_t e = 0; e < n_ext; ++e) {
      const double* xa = B.row(a);
      const double* xb = B.row(n_int + e);
      double d[4] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2], xb[3] - xa[3]};
      if (lscalar(d, d) < 0.0)
        throw CalcError("Bubble " + bubble_index_name(a, n_int) + " is formed inside bubble " +
                        bubble_index_name(n_int + e, n_int) +
                        " at initial time (overlapping light cones)");
    }
  return B;
}

// Candidate list of reference bubble a: every other bubble b with
// delta = x_b - x_a (bitwise what LatticeBubbles::new stores, :280-297), its
// squared interval, and the delay lower bound (|delta_s| + delta_t) / 2 that
// holds along every direction for pairs well outside the light cone.
void build_candidates(const Bubbles& B, size_t a, CandEntry* out, int32_t* inv) {
  const size_t n_tot = B.n_tot();
  const double* xa = B.row(a);
  size_t m = 0;
  for (size_t b = 0; b < n_tot; ++b) {
    if (b < B.n_int && b == a) continue;
    const double* xb = B.row(b);
    CandEntry e{};
    for (int k = 0; k < 4; ++k) e.d[k] = xb[k] - xa[k];
    e.dsq = lscalar(e.d, e.d);
    const double ns = std::sqrt((e.d[1] * e.d[1] + e.d[2] * e.d[2]) + e.d[3] * e.d[3]);
    e.len = ns;
    const bool finite = std::isfinite(e.d[0]) && std::isfinite(e.d[1]) &&
                        std::isfinite(e.d[2]) && std::isfinite(e.d[3]) && std::isfinite(e.dsq);
    const bool regular = finite && e.dsq >= 1e-6 * (e.d[0] * e.d[0] + ns * ns) && e.dsq > 0.0;
    e.key = regular ? 0.5 * (ns + e.d[0]) : -std::numeric_limits<double>::infinity();
    e.idx = static_cast<int32_t>(b);
    e.flags = regular ? bgw::kCandRegular : 0;
    out[m++] = e;
  }
  std::sort(out, out + m, [](const CandEntry& x, const CandEntry& y) {
    if (x.key != y.key) return x.key < y.key;
    return x.idx < y.idx;
  });
  for (size_t b = 0; b < n_tot; ++b) inv[b] = -1;
  for (size_t s = 0; s < m; ++s) inv[out[s].idx] = static_cast<int32_t>(s);
}

template <class F>
void parallel_for(size_t n, size_t n_threads, F&& f) {
  n_threads = std::max<size_t>(1, std::min(n_threads, n));
  if (n_threads <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (size_t t = 0; t < n_threads; ++t)
    pool.emplace_back([&] {
      for (size_t i = next++; i < n; i = next++) f(i);
    });
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Angular grids and scaling parameters
// ---------------------------------------------------------------------------
struct Resolution {
  bool set = false;
  size_t n_c = 0, n_phi = 0, n_phi_pow2 = 0;
  std::vector<double> cos_t, sin_t, phi, cos_p, sin_p, gcos, gsin, ring_fac;
};

Resolution make_resolution(size_t n_c, size_t n_phi) {
  if (n_c < 2 || n_phi < 2)
    throw CalcError("Invalid resolution: Angular resolutions must be greater than one");
  if (n_phi > 1024) throw CalcError("Invalid resolution: n_phix > 1024 is not supported");
  if (n_c > 65535) throw CalcError("Invalid resolution: n_cos_thetax > 65535 is not supported");
  Resolution R;
  R.set = true;
  R.n_c = n_c;
  R.n_phi = n_phi;
  R.n_phi_pow2 = 1;
  while (R.n_phi_pow2 < n_phi) R.n_phi_pow2 <<= 1;
  // set_resolution (generalized_bulk_flow.rs:221-238)
  R.cos_t.resize(n_c);
  R.sin_t.resize(n_c);
  for (size_t i = 0; i < n_c; ++i) {
    R.cos_t[i] = linspace_at(-1.0, 1.0, n_c, i);
    R.sin_t[i] = std::fabs(std::sqrt(1.0 - R.cos_t[i] * R.cos_t[i]));
  }
  R.phi.resize(n_phi);
  R.cos_p.resize(n_phi);
  R.sin_p.resize(n_phi);
  R.gcos.resize(n_phi);
  R.gsin.resize(n_phi);
  for (size_t j = 0; j < n_phi; ++j) {
    R.phi[j] = linspace_at(0.0, 2.0 * kPi, n_phi, j);
    // phi.sin() and phi.cos() of one angle in one closure (:228-229), which
    // the Rust compiler lowers to a single libm sincos call on GNU targets;
    // sincos and separate sin / cos calls can differ in the last bit.
    ::sincos(R.phi[j], &R.sin_p[j], &R.cos_p[j]);
    // compute_b_integral (:528-539): trapezoid weights 1/2 at both ends
    const double w = (j == 0 || j == n_phi - 1) ? 0.5 : 1.0;
    // sin(2 phi) and cos(2 phi) come from two separate passes in the
    // reference (:528-531): separate libm calls (this file is built without
    // fast-math, so they are not merged into sincos)
    R.gcos[j] = std::cos(2.0 * R.phi[j]) * w;
    R.gsin[j] = std::sin(2.0 * R.phi[j]) * w;
  }
  // per-ring factor dcos * w_i * 0.5 * sin^2 * dphi (:522-523, 633-651)
  const double dphi = 2.0 * kPi / static_cast<double>(n_phi);
  const double dcos = 2.0 / static_cast<double>(n_c - 1);
  R.ring_fac.resize(n_c);
  for (size_t i = 0; i < n_c; ++i) {
    const double sin_sq = 1.0 - R.cos_t[i] * R.cos_t[i];
    const double wi = (i == 0 || i == n_c - 1) ? 0.5 : 1.0;
    R.ring_fac[i] = dcos * wi * (0.5 * sin_sq) * dphi;
  }
  return R;
}

struct Scaling {
  size_t n_sets = 1, n_terms = 1;
  std::vector<double> coef{1.0}, pows{3.0};  // defaults (generalized_bulk_flow.rs:98-99)
  bool has_damping = false;
  double damping = 0.0;
};

Scaling make_scaling(const double* coefficients, const double* powers, size_t n_sets,
                     size_t n_terms, double damping) {
  const size_t len = checked_product(n_sets, n_terms, "gradient scaling");
  slice_from_raw(coefficients, len, "coefficients");
  slice_from_raw(powers, len, "powers");
  if (n_sets > 0 && n_terms == 0)
    throw CalcError("Array shape mismatch: n_terms must be > 0 when n_sets > 0");
  Scaling S;
  S.n_sets = n_sets;
  S.n_terms = n_terms;
  S.coef.assign(coefficients, coefficients + len);
  S.pows.assign(powers, powers + len);
  S.has_damping = !std::isnan(damping);
  S.damping = damping;
  return S;
}

struct Cutoff {
  int kind = 0;
  double t_cut = 0.0, inv_t0_sq = 0.0;
};

// ExponentialTimeCutoff::new (time_cutoff.rs:22-30).
Cutoff make_cutoff(const BgwTimeCutoffConfig& c) {
  Cutoff o;
  if (c.kind == BgwTimeCutoffKind_Exponential) {
    o.kind = 1;
    const double rc = std::isnan(c.ratio_t_cut) ? 0.999999999 : c.ratio_t_cut;
    const double r0 = std::isnan(c.ratio_t_0) ? 0.25 : c.ratio_t_0;
    o.t_cut = rc * c.smax;
    const double t0 = r0 * (c.smax - o.t_cut);
    o.inv_t0_sq = 1.0 / (t0 * t0);
  } else if (c.kind != BgwTimeCutoffKind_Unit) {
    throw CalcError("time_cutoff.kind is not a BgwTimeCutoffKind");
  }
  return o;
}

// ---------------------------------------------------------------------------
// Device state
// ---------------------------------------------------------------------------
// Bump allocator over caller-provided device memory (the scratch of the
// no-allocation entry point): regions 256-byte aligned, never freed.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~size_t(255);
    if (at > cap || bytes > cap - at)
      throw CalcError("the request needs more device memory than the " +
                      std::to_string(cap >> 20) +
                      " MiB scratch of the no-allocation entry point (" +
                      std::to_string(at + bytes) + " bytes so far)");
    used = at + bytes;
    return base + at;
  }
};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  Arena* arena = nullptr;  // set: regions come from the arena, no hipMalloc / hipFree
  void ensure(size_t bytes) {
    if (bytes <= cap && p) return;
    if (arena) {
      p = arena->take(bytes > 0 ? bytes : 16);
      cap = bytes;
      return;
    }
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    HIPCHECK(hipMalloc(&p, bytes > 0 ? bytes : 16));
    cap = bytes;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p && !arena) (void)hipFree(p);
  }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.ensure(v.size() * sizeof(T));
  if (!v.empty())
    HIPCHECK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

struct Device {
  int id = -1;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // second branch of the pipelined ensemble graph
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t dep[4] = {nullptr, nullptr, nullptr, nullptr};  // graph dependencies (no timing)
  static constexpr int kTimedProducts = 16;  // matrix-core launches timed per call
  hipEvent_t mx_ev[2 * kTimedProducts] = {};
  void init() {
    if (stream) {
      HIPCHECK(hipSetDevice(id));
      return;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      throw CalcError("no HIP device is available: the bulk-flow kernels need a gfx950 GPU");
    HIPCHECK(hipGetDevice(&id));
    // `stream` is created last: it marks the device as initialised, so a
    // failure part-way leaves the rest to be created by the next call.
    if (!stream2) HIPCHECK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    for (auto& e : ev)
      if (!e) HIPCHECK(hipEventCreate(&e));
    for (auto& e : dep)
      if (!e) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : mx_ev)
      if (!e) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  }
  ~Device() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : dep)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : mx_ev)
      if (e) (void)hipEventDestroy(e);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// All device buffers of one batch (one or many realisations).
struct Engine {
  Device dev;
  // static (per bubble configuration)
  std::vector<CandEntry> cand;
  std::vector<int32_t> inv;
  std::vector<RealInfo> real;
  std::vector<std::vector<double>> bubbles;  // per realisation (n_tot, 4)
  std::vector<size_t> n_int;
  bool static_uploaded = false;
  DevBuf d_cand, d_inv, d_real;
  // per call
  DevBuf d_items, d_order, d_rot, d_sin_t, d_cos_t, d_cos_p, d_sin_p, d_gcos, d_gsin, d_ring_c,
      d_ring_fac, d_omega, d_coef, d_pows, d_tau, d_perm, d_partial, d_group_off, d_group_r,
      d_c, d_s, d_first, d_dtab, d_gradc, d_gradout, d_iota, d_acc;
  // FP64 matrix form: tiles, their time tables and weight rows
  DevBuf d_tiles, d_tinfo, d_tt, d_ctlo, d_vbuf, d_mxphase;
  // batched ensemble: work units, queue counter, statistics, captured graph
  DevBuf d_units, d_unit_off, d_counter, d_kmean, d_emean, d_estd;
  hipGraphExec_t graph = nullptr;
  std::string graph_key;
  BgwTimings timings{};
  int mx_timed = 0;         // matrix-core launches of this call with events
  bool mx_untimed = false;  // a launch beyond Device::kTimedProducts
  std::vector<DevBuf*> all_buffers() {
    return {&d_cand, &d_inv, &d_real, &d_items, &d_order, &d_rot, &d_sin_t, &d_cos_t,
            &d_cos_p, &d_sin_p, &d_gcos, &d_gsin, &d_ring_c, &d_ring_fac, &d_omega, &d_coef,
            &d_pows, &d_tau, &d_perm, &d_partial, &d_group_off, &d_group_r, &d_c, &d_s,
            &d_first, &d_dtab, &d_gradc, &d_gradout, &d_iota, &d_acc, &d_tiles, &d_tinfo,
            &d_tt, &d_ctlo, &d_vbuf, &d_mxphase, &d_units, &d_unit_off, &d_counter, &d_kmean,
            &d_emean, &d_estd};
  }
  ~Engine() {
    if (graph) (void)hipGraphExecDestroy(graph);
  }
};

// For the lifetime of the scope every buffer of E is carved from the arena
// (the static candidate lists are uploaded again into it); the engine's own
// allocations are put back afterwards, untouched.  The destructor waits for
// the engine's streams, so nothing in flight still uses the arena when the
// call returns, also on an error.
class ArenaScope {
 public:
  ArenaScope(Engine& E, Arena& A) : E_(E), uploaded_(E.static_uploaded) {
    for (DevBuf* b : E.all_buffers()) {
      saved_.push_back({b->p, b->cap, b->arena});
      b->p = nullptr;
      b->cap = 0;
      b->arena = &A;
    }
    E.static_uploaded = false;
  }
  ~ArenaScope() {
    if (E_.dev.stream) (void)hipStreamSynchronize(E_.dev.stream);
    if (E_.dev.stream2) (void)hipStreamSynchronize(E_.dev.stream2);
    const std::vector<DevBuf*> bufs = E_.all_buffers();
    for (size_t i = 0; i < bufs.size(); ++i) {
      bufs[i]->p = saved_[i].p;
      bufs[i]->cap = saved_[i].cap;
      bufs[i]->arena = saved_[i].arena;
    }
    E_.static_uploaded = uploaded_;
  }
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;

 private:
  struct Saved {
    void* p;
    size_t cap;
    Arena* arena;
  };
  Engine& E_;
  bool uploaded_;
  std::vector<Saved> saved_;
};

void build_static(Engine& E, const std::vector<Bubbles>& reals, size_t n_threads) {
  size_t n_cand = 0, n_inv = 0;
  E.real.assign(reals.size(), RealInfo{});
  for (size_t r = 0; r < reals.size(); ++r) {
    const Bubbles& B = reals[r];
    E.real[r].cand_off = static_cast<int64_t>(n_cand);
    E.real[r].inv_off = static_cast<int64_t>(n_inv);
    E.real[r].n_int = static_cast<int32_t>(B.n_int);
    E.real[r].n_tot = static_cast<int32_t>(B.n_tot());
    E.real[r].nonfinite = B.nonfinite ? 1 : 0;
    if (B.n_tot() > 0) n_cand += B.n_int * (B.n_tot() - 1);
    n_inv += B.n_int * B.n_tot();
  }
  E.cand.assign(n_cand, CandEntry{});
  E.inv.assign(n_inv, -1);
  // (realisation, a) pairs in parallel
  std::vector<std::pair<size_t, size_t>> work;
  for (size_t r = 0; r < reals.size(); ++r)
    for (size_t a = 0; a < reals[r].n_int; ++a) work.emplace_back(r, a);
  parallel_for(work.size(), n_threads, [&](size_t w) {
    const size_t r = work[w].first, a = work[w].second;
    const Bubbles& B = reals[r];
    build_candidates(B, a, &E.cand[E.real[r].cand_off + a * (B.n_tot() - 1)],
                     &E.inv[E.real[r].inv_off + a * B.n_tot()]);
  });
  E.bubbles.clear();
  E.n_int.clear();
  for (const auto& B : reals) {
    E.bubbles.push_back(B.xyz);
    E.n_int.push_back(B.n_int);
  }
  E.static_uploaded = false;
}

// One compute request over a batch.
struct Request {
  // shared
  const Resolution* res = nullptr;
  // several surface-weight sets of one grid (the tensor components): k-hat slot
  // k of the request (identity frame) uses surfaces[k]; empty: res's weights
  std::vector<const Resolution*> surfaces;
  const Scaling* sc = nullptr;
  Cutoff cutoff;
  size_t n_w = 0, n_t = 0;
  bool identity = true;                  // no rotation (reference frame)
  std::vector<std::array<double, 3>> k;  // k-hat list (identity: one dummy entry)
  // per realisation
  std::vector<double> omega;             // (n_real, n_w)
  std::vector<double> t_begin, t_end;    // (n_real)
  // items: per (r, k) group the reference bubbles in summation order
  std::vector<std::vector<int32_t>> bubbles_of_real;  // selected a per realisation
  int mode = 0;  // phase-sum kernel: 0 FP64 direct, 1 FP64 blocked, 2 mixed
  int tile_threads = 0;  // phase-sum workgroup size (0: the form's default)
  int tile_rb = 0;       // time steps per factorised phase block (0: the form's default)
  bool want_c = true, want_s = true;
  bool skip_uncollided = true;
  // per-item rows of tau / perm and per-item partial C (false: the fused
  // ensemble kernel, which keeps neither)
  bool item_buffers = true;
  // run(): also the k-hat mean of S per realisation and the mean and spread of
  // those over the realisations (an ensemble), inside the timed region
  bool ensemble_stats = false;
  // setup(): the per-item phase-sum launch order (false: the caller orders the
  // items itself, as the batched ensembles do)
  bool item_order = true;
};

// The indices idx[0, n) (given in ascending order) reordered by cost
// descending: a counting sort on the cost quantised to kCostLevels levels
// between the range's extremes, index order within a level.  The order only
// balances the launches (every item writes its own partial result, reduced in
// a fixed order), so levels of 1/4096 of the range serve as well as an exact
// sort, in linear time (ensemble_512: 14 ms against 175 ms for sorting its
// 2.09 M items in chunks of 65536).
template <typename CostAt>
void order_by_cost(int32_t* idx, size_t n, CostAt cost_at) {
  if (n < 2) return;
  constexpr int kCostLevels = 4096;
  double lo = cost_at(idx[0]), hi = lo;
  for (size_t i = 1; i < n; ++i) {
    const double c = cost_at(idx[i]);
    lo = std::min(lo, c);
    hi = std::max(hi, c);
  }
  const double scale = hi > lo ? kCostLevels / (hi - lo) : 0.0;
  auto level = [&](double c) {
    const int b = static_cast<int>((hi - c) * scale);  // 0: the most expensive
    return b < 0 ? 0 : (b >= kCostLevels ? kCostLevels - 1 : b);
  };
  std::vector<int32_t> start(kCostLevels + 1, 0), out(n);
  std::vector<int16_t> lev(n);
  for (size_t i = 0; i < n; ++i) {
    lev[i] = static_cast<int16_t>(level(cost_at(idx[i])));
    ++start[lev[i] + 1];
  }
  for (int b = 0; b < kCostLevels; ++b) start[b + 1] += start[b];
  for (size_t i = 0; i < n; ++i) out[start[lev[i]]++] = idx[i];
  std::copy(out.begin(), out.end(), idx);
}

struct Result {
  std::vector<double> c;  // (n_groups, 2, n_sets, n_w, 2)
  std::vector<double> s;  // (n_groups, n_sets, n_w)
  size_t n_groups = 0;
  // Request::ensemble_stats: (n_real, n_sets, n_w) k-hat means, (n_sets, n_w) mean and
  // population standard deviation over the realisations
  std::vector<double> kmean, mean, stdev;
};

void ensure_static(Engine& E) {
  E.dev.init();
  if (E.static_uploaded) return;
  upload(E.d_cand, E.cand, E.dev.stream);
  upload(E.d_inv, E.inv, E.dev.stream);
  E.static_uploaded = true;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, a, b));
  return static_cast<double>(ms);
}

// Kernel arguments of one request, after the per-call uploads.
struct Launch {
  bgw::GeoArgs ga{};
  bgw::SpecArgs sa{};
  bgw::RedArgs ra{};
  size_t n_items = 0, n_groups = 0;
  std::vector<Item> items;               // in (realisation, k-hat, bubble) order
  std::vector<double> item_cost;         // relative phase-sum cost of each item
  std::vector<int32_t> group_off, group_r;
  std::vector<RealInfo> real;            // time grids and omega offsets of this request
};

void setup(Engine& E, const Request& q, Launch& Lc) {
  const Resolution& res = *q.res;
  const Scaling& sc = *q.sc;
  const size_t n_real = E.real.size();
  const size_t n_k = q.k.size();
  const size_t n_sets = sc.n_sets;
  if (q.n_w > 1024) throw CalcError("n_w > 1024 is not supported by the GPU kernels");
  if (q.n_t >= static_cast<size_t>(bgw::kSortStepMax))
    throw CalcError("n_t >= " + std::to_string(bgw::kSortStepMax) +
                    " is not supported by the GPU kernels");
  ensure_static(E);
  hipStream_t st = E.dev.stream;

  // per-realisation time grid and omegas
  std::vector<RealInfo> real = E.real;
  for (size_t r = 0; r < n_real; ++r) {
    real[r].t_begin = q.t_begin[r];
    real[r].t_step = q.n_t > 1 ? (q.t_end[r] - q.t_begin[r]) / static_cast<double>(q.n_t - 1) : 0.0;
    const double t1 = real[r].t_begin + real[r].t_step * 1.0;
    real[r].dt = q.n_t > 1 ? t1 - real[r].t_begin : 0.0;
    real[r].w_off = static_cast<int64_t>(r * q.n_w);
  }
  // rotations
  std::vector<double> rot(9 * n_k, 0.0);
  std::vector<std::array<double, 9>> R(n_k);
  for (size_t k = 0; k < n_k; ++k) {
    if (!q.identity) R[k] = rotation_to_z(q.k[k].data());
    for (int j = 0; j < 9; ++j) rot[9 * k + j] = q.identity ? 0.0 : R[k][j];
  }
  // items sorted by (r, k, position in the selection)
  std::vector<Item> items;
  std::vector<int32_t> group_off(1, 0), group_r;
  std::vector<double> cost;
  for (size_t r = 0; r < n_real; ++r) {
    const double* bub = E.bubbles[r].data();
    const double span = q.t_end[r] - q.t_begin[r];
    for (size_t k = 0; k < n_k; ++k) {
      for (int32_t a : q.bubbles_of_real[r]) {
        const double* xa = bub + 4 * static_cast<size_t>(a);
        Item it;
        it.r = static_cast<int32_t>(r);
        it.k = static_cast<int32_t>(k);
        it.a = a;
        it.group = static_cast<int32_t>(r * n_k + k);
        it.t_a = xa[0];
        it.z_a = q.identity ? xa[3] : (R[k][6] * xa[1] + R[k][7] * xa[2]) + R[k][8] * xa[3];
        items.push_back(it);
        cost.push_back(span > 0 ? std::min(1.0, std::max(0.0, (q.t_end[r] - xa[0]) / span)) : 0.0);
      }
      group_off.push_back(static_cast<int32_t>(items.size()));
      group_r.push_back(static_cast<int32_t>(r));
    }
  }
  const size_t n_items = items.size();
  // launch order of the per-item phase-sum kernels: most expensive items first
  // (the matrix form orders its tiles itself and reads no item order)
  const bool item_order = q.item_order && q.mode != 3;
  std::vector<int32_t> order;
  if (item_order) {
    order.resize(n_items);
    std::iota(order.begin(), order.end(), 0);
    order_by_cost(order.data(), n_items, [&](int32_t i) { return cost[i]; });
  }

  const size_t n_dir = res.n_c * res.n_phi;
  const size_t nq = n_sets * q.n_w;
  Lc.n_items = n_items;
  Lc.n_groups = n_real * n_k;
  Lc.items = items;
  Lc.item_cost = cost;
  Lc.group_off = group_off;
  Lc.group_r = group_r;
  Lc.real = real;
  upload(E.d_real, real, st);
  upload(E.d_items, items, st);
  if (item_order) upload(E.d_order, order, st);
  upload(E.d_rot, rot, st);
  upload(E.d_sin_t, res.sin_t, st);
  upload(E.d_cos_t, res.cos_t, st);
  upload(E.d_cos_p, res.cos_p, st);
  upload(E.d_sin_p, res.sin_p, st);
  if (q.surfaces.empty()) {
    upload(E.d_gcos, res.gcos, st);
    upload(E.d_gsin, res.gsin, st);
    upload(E.d_ring_fac, res.ring_fac, st);
  } else {
    if (q.surfaces.size() != n_k) throw CalcError("one surface-weight set per k-hat slot");
    std::vector<double> gc, gs, rf;
    for (const Resolution* S : q.surfaces) {
      if (S->n_c != res.n_c || S->n_phi != res.n_phi)
        throw CalcError("surface-weight sets of different grids");
      gc.insert(gc.end(), S->gcos.begin(), S->gcos.end());
      gs.insert(gs.end(), S->gsin.begin(), S->gsin.end());
      rf.insert(rf.end(), S->ring_fac.begin(), S->ring_fac.end());
    }
    upload(E.d_gcos, gc, st);
    upload(E.d_gsin, gs, st);
    upload(E.d_ring_fac, rf, st);
  }
  upload(E.d_omega, q.omega, st);
  upload(E.d_coef, sc.coef, st);
  upload(E.d_pows, sc.pows, st);
  upload(E.d_group_off, group_off, st);
  upload(E.d_group_r, group_r, st);
  if (q.item_buffers) {
    E.d_tau.ensure(std::max<size_t>(1, n_items * n_dir) * sizeof(double));
    E.d_perm.ensure(std::max<size_t>(1, n_items * n_dir) * sizeof(uint16_t));
    E.d_partial.ensure(std::max<size_t>(1, n_items * 2 * nq) * 2 * sizeof(double));
  }
  E.d_c.ensure(Lc.n_groups * 2 * nq * 2 * sizeof(double));
  E.d_s.ensure(Lc.n_groups * nq * sizeof(double));

  bgw::GeoArgs& ga = Lc.ga;
  ga.cand = E.d_cand.as<CandEntry>();
  ga.inv = E.d_inv.as<int32_t>();
  ga.real = E.d_real.as<RealInfo>();
  ga.items = E.d_items.as<Item>();
  ga.rot = E.d_rot.as<double>();
  ga.identity = q.identity ? 1 : 0;
  ga.sin_t = E.d_sin_t.as<double>();
  ga.cos_t = E.d_cos_t.as<double>();
  ga.cos_p = E.d_cos_p.as<double>();
  ga.sin_p = E.d_sin_p.as<double>();
  ga.n_c = static_cast<int>(res.n_c);
  ga.n_phi = static_cast<int>(res.n_phi);
  ga.n_phi_pow2 = static_cast<int>(res.n_phi_pow2);
  ga.tau = E.d_tau.as<double>();
  ga.perm = E.d_perm.as<uint16_t>();
  ga.n_t = static_cast<int>(q.n_t);

  bgw::SpecArgs& sa = Lc.sa;
  sa.real = ga.real;
  sa.items = ga.items;
  sa.order = E.d_order.as<int32_t>();
  sa.tau = ga.tau;
  sa.perm = ga.perm;
  sa.gcos = E.d_gcos.as<double>();
  sa.gsin = E.d_gsin.as<double>();
  sa.ring_c = E.d_cos_t.as<double>();
  sa.ring_fac = E.d_ring_fac.as<double>();
  sa.omega = E.d_omega.as<double>();
  sa.n_terms = static_cast<int>(sc.n_terms);
  sa.n_sets_out = static_cast<int>(n_sets);
  sa.n_c = ga.n_c;
  sa.n_phi = ga.n_phi;
  sa.n_t = static_cast<int>(q.n_t);
  sa.n_w = static_cast<int>(q.n_w);
  sa.cutoff_kind = q.cutoff.kind;
  sa.t_cut = q.cutoff.t_cut;
  sa.inv_t0_sq = q.cutoff.inv_t0_sq;
  sa.has_damping = sc.has_damping ? 1 : 0;
  sa.damping = sc.damping;
  sa.kfac = 1.0 / (6.0 * kPi);
  sa.skip_uncollided = q.skip_uncollided ? 1 : 0;
  sa.surfaces = q.surfaces.empty() ? 1 : static_cast<int>(q.surfaces.size());
  sa.mixed_block = q.tile_rb;
  sa.threads = q.tile_threads;
  sa.partial = E.d_partial.as<double>();

  bgw::RedArgs& ra = Lc.ra;
  ra.partial = sa.partial;
  ra.group_off = E.d_group_off.as<int32_t>();
  ra.group_r = E.d_group_r.as<int32_t>();
  ra.real = ga.real;
  ra.omega = sa.omega;
  ra.n_sets = static_cast<int>(n_sets);
  ra.n_w = static_cast<int>(q.n_w);
  ra.c_out = E.d_c.as<double>();
  ra.s_out = E.d_s.as<double>();
}

// The FP64 matrix form for the sets of sb: tiles of the items of one
// (realisation, reference bubble) in consecutive k-hat slots, most expensive
// first (earliest nucleation); their weight rows go through HBM in batches of
// at most kMatrixRowBytes.  Launch shape: two column tiles when a bubble has
// more items than one column tile holds; 16-omega tiles per wavefront of the
// matrix kernel from q.tile_rb (1, 2 or 4; 0: 2).
constexpr size_t kMatrixRowBytes = size_t(16) << 30;

// The launches of the matrix form for one set group: the arguments of every
// batch of tiles (the planned buffers already uploaded).
struct MatrixPlan {
  bgw::MatArgs ma{};
  std::vector<std::pair<size_t, size_t>> batches;  // tile ranges
  std::vector<int64_t> voff0;                      // weight-row offset of each batch
};

MatrixPlan plan_matrix_form(Engine& E, const Request& q, const Launch& Lc,
                            const bgw::SpecArgs& sb, hipStream_t st) {
  if (sb.has_damping)
    throw CalcError("the FP64 matrix form does not take a damping width (use fp64_blocked)");
  if (q.n_t > 65535) throw CalcError("the FP64 matrix form supports n_t <= 65535");
  if (q.tile_rb != 0 && q.tile_rb != 1 && q.tile_rb != 2 && q.tile_rb != 4)
    throw CalcError("tiles.block_steps of the FP64 matrix form: 0, 1, 2 or 4 omega tiles per wavefront");
  const size_t n_k = q.k.size();
  bgw::MatArgs ma{};
  ma.s = sb;
  ma.col_tiles = static_cast<size_t>(bgw::matrix_items_per_tile(sb.n_sets, 1)) < n_k ? 2 : 1;
  ma.omega_tiles = q.tile_rb ? q.tile_rb : 2;
  ma.ni = std::min<int>(bgw::matrix_items_per_tile(sb.n_sets, ma.col_tiles), static_cast<int>(n_k));
  // wide rings: fewer items per tile, so that the tables fit the 160 KiB of LDS
  while (ma.ni > 1 && bgw::matrix_lds_bytes(sb, ma.ni, ma.col_tiles) > 160 * 1024) --ma.ni;
  if (bgw::matrix_lds_bytes(sb, ma.ni, ma.col_tiles) > 160 * 1024)
    throw CalcError("resolution too large: the per-ring tables exceed the 160 KiB of LDS");
  const size_t ncol = 16 * static_cast<size_t>(ma.col_tiles);
  std::vector<int32_t> tiles;
  std::vector<std::pair<double, int32_t>> cost;
  std::vector<int32_t> kbase_of, k_a_of;
  const size_t n_real = E.real.size();
  for (size_t r = 0; r < n_real; ++r) {
    const size_t n_sel = q.bubbles_of_real[r].size();
    const RealInfo& ri = Lc.real[r];
    for (size_t p = 0; p < n_sel; ++p)
      for (size_t k0 = 0; k0 < n_k; k0 += static_cast<size_t>(ma.ni)) {
        const int32_t first = Lc.group_off[r * n_k + k0] + static_cast<int32_t>(p);
        cost.emplace_back(Lc.item_cost[first], static_cast<int32_t>(cost.size()));
        for (int j = 0; j < ma.ni; ++j) {
          const size_t k = k0 + static_cast<size_t>(j);
          tiles.push_back(k < n_k ? Lc.group_off[r * n_k + k] + static_cast<int32_t>(p) : -1);
        }
        // first step with t_k - t_a > 0, as the kernels compute t_k
        const double t_a = Lc.items[first].t_a;
        int lo = 0, hi = static_cast<int>(q.n_t);
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((ri.t_begin + ri.t_step * static_cast<double>(mid)) - t_a > 0.0) hi = mid;
          else lo = mid + 1;
        }
        kbase_of.push_back(lo & ~(bgw::kMatrixBlockSteps - 1));
        k_a_of.push_back(lo);
      }
  }
  std::stable_sort(cost.begin(), cost.end(),
                   [](const std::pair<double, int32_t>& x, const std::pair<double, int32_t>& y) {
                     return x.first > y.first;
                   });
  const size_t n_tiles = cost.size();
  std::vector<int32_t> ordered;
  std::vector<bgw::MxTileInfo> info(n_tiles);
  ordered.reserve(tiles.size());
  int64_t voff = 0;
  for (size_t t = 0; t < n_tiles; ++t) {
    const int32_t c = cost[t].second;
    ordered.insert(ordered.end(), tiles.begin() + static_cast<size_t>(c) * ma.ni,
                   tiles.begin() + static_cast<size_t>(c + 1) * ma.ni);
    info[t].kbase = kbase_of[c];
    info[t].k_a = k_a_of[c];
    info[t].len = (static_cast<int32_t>(q.n_t) - kbase_of[c] + bgw::kMatrixBlockSteps - 1) &
                  ~(bgw::kMatrixBlockSteps - 1);
    info[t].voff = voff;
    voff += static_cast<int64_t>(sb.n_c) * info[t].len * static_cast<int64_t>(ncol);
  }
  upload(E.d_tiles, ordered, st);
  upload(E.d_tinfo, info, st);
  const size_t nterms = static_cast<size_t>(sb.n_sets) * sb.n_terms;
  E.d_tt.ensure(std::max<size_t>(1, n_tiles * q.n_t * (1 + nterms)) * sizeof(double));
  E.d_ctlo.ensure(std::max<size_t>(1, n_tiles * sb.n_c * 2) * sizeof(int32_t));
  ma.tiles = E.d_tiles.as<int32_t>();
  ma.tinfo = E.d_tinfo.as<bgw::MxTileInfo>();
  ma.tt = E.d_tt.as<double>();
  ma.ctlo = E.d_ctlo.as<int32_t>();
  E.d_mxphase.ensure(std::max<size_t>(1, n_real * sb.n_c * q.n_w * 8) * sizeof(double));
  ma.phase = E.d_mxphase.as<double>();
  // batches of whole tiles within the row budget (a tile larger than the budget alone)
  std::vector<std::pair<size_t, size_t>> batches;
  size_t max_batch = 1;
  for (size_t t0 = 0; t0 < n_tiles;) {
    size_t t1 = t0 + 1;
    while (t1 < n_tiles &&
           static_cast<size_t>(info[t1].voff + static_cast<int64_t>(sb.n_c) * info[t1].len *
                               static_cast<int64_t>(ncol) - info[t0].voff) * sizeof(double) <=
               kMatrixRowBytes)
      ++t1;
    const int64_t end = t1 < n_tiles ? info[t1].voff : voff;
    max_batch = std::max<size_t>(max_batch, static_cast<size_t>(end - info[t0].voff));
    batches.emplace_back(t0, t1);
    t0 = t1;
  }
  E.d_vbuf.ensure((max_batch + static_cast<size_t>(bgw::kMatrixRowPad) * 32) * sizeof(double));
  ma.vbuf = E.d_vbuf.as<double>();
  MatrixPlan P;
  P.ma = ma;
  P.batches = batches;
  for (const auto& b : batches) P.voff0.push_back(info[b.first].voff);
  return P;
}

// The plan's launches on st: per batch the time tables and weight rows, then
// the products (timed with events unless `timed` is false, as under capture).
void launch_matrix_plan(Engine& E, const MatrixPlan& P, hipStream_t st, bool timed) {
  bgw::MatArgs ma = P.ma;
  HIPCHECK(bgw::launch_matrix_phases(ma, static_cast<int>(E.real.size()), st));
  for (size_t b = 0; b < P.batches.size(); ++b) {
    ma.tile0 = static_cast<int>(P.batches[b].first);
    ma.voff0 = P.voff0[b];
    const int nt = static_cast<int>(P.batches[b].second - P.batches[b].first);
    HIPCHECK(bgw::launch_matrix_weights(ma, nt, st));
    const int slot = E.mx_timed;
    const bool ev = timed && slot < Device::kTimedProducts;
    if (ev) HIPCHECK(hipEventRecord(E.dev.mx_ev[2 * slot], st));
    HIPCHECK(bgw::launch_matrix_products(ma, nt, st));
    if (ev) {
      HIPCHECK(hipEventRecord(E.dev.mx_ev[2 * slot + 1], st));
      ++E.mx_timed;
    } else {
      E.mx_untimed = true;
    }
  }
}

void launch_matrix_form(Engine& E, const Request& q, const Launch& Lc, const bgw::SpecArgs& sb,
                        hipStream_t st) {
  launch_matrix_plan(E, plan_matrix_form(E, q, Lc, sb, st), st, true);
}

Result run(Engine& E, const Request& q) {
  const Resolution& res = *q.res;
  const Scaling& sc = *q.sc;
  const size_t n_sets = sc.n_sets;
  Result out;
  out.n_groups = E.real.size() * q.k.size();
  if (n_sets == 0 || q.n_w == 0) {
    out.c.assign(out.n_groups * 2 * n_sets * q.n_w * 2, 0.0);
    out.s.assign(out.n_groups * n_sets * q.n_w, 0.0);
    return out;
  }
  Launch Lc;
  setup(E, q, Lc);
  hipStream_t st = E.dev.stream;
  const size_t n_items = Lc.n_items;
  const size_t n_dir = res.n_c * res.n_phi;
  const size_t nq = n_sets * q.n_w;
  const bgw::GeoArgs& ga = Lc.ga;
  const bgw::SpecArgs& sa = Lc.sa;
  const bgw::RedArgs& ra = Lc.ra;

  E.mx_timed = 0;
  E.mx_untimed = false;
  HIPCHECK(hipEventRecord(E.dev.ev[0], st));
  HIPCHECK(bgw::launch_geometry(ga, static_cast<int>(n_items), st));
  HIPCHECK(hipEventRecord(E.dev.ev[1], st));
  for (size_t s0 = 0; s0 < n_sets; s0 += kMaxSetsPerLaunch) {
    bgw::SpecArgs sb = sa;
    sb.set0 = static_cast<int>(s0);
    sb.n_sets = static_cast<int>(std::min(kMaxSetsPerLaunch, n_sets - s0));
    sb.coef = E.d_coef.as<double>() + s0 * sc.n_terms;
    sb.pows = E.d_pows.as<double>() + s0 * sc.n_terms;
    sb.time_tables = 0;
    if (q.mode == 3) {
      // sets of one matrix-form launch: at most kMaxMatrixTerms (set, term) pairs
      if (sc.n_terms > static_cast<size_t>(bgw::kMaxMatrixTerms))
        throw CalcError("the FP64 matrix form takes at most " +
                        std::to_string(bgw::kMaxMatrixTerms) + " terms per set");
      const size_t per = std::max<size_t>(1, bgw::kMaxMatrixTerms / sc.n_terms);
      for (size_t s1 = 0; s1 < static_cast<size_t>(sb.n_sets); s1 += per) {
        bgw::SpecArgs sm = sb;
        sm.set0 = sb.set0 + static_cast<int>(s1);
        sm.n_sets = static_cast<int>(std::min(per, static_cast<size_t>(sb.n_sets) - s1));
        sm.coef = sb.coef + s1 * sc.n_terms;
        sm.pows = sb.pows + s1 * sc.n_terms;
        launch_matrix_form(E, q, Lc, sm, st);
      }
      continue;
    }
    if (q.tile_rb != 0 && q.tile_rb != 8 && q.tile_rb != 16)
      throw CalcError("tiles.block_steps must be 0, 8 or 16 for this form");
    sb.time_tables = bgw::spectrum_time_tables(sb) ? 1 : 0;
    if (bgw::spectrum_lds_bytes(sb, q.mode) > 160 * 1024)
      throw CalcError("resolution too large: the per-ring tables exceed the 160 KiB of LDS");
    HIPCHECK(bgw::launch_spectrum(sb, static_cast<int>(n_items), q.mode, st));
  }
  HIPCHECK(hipEventRecord(E.dev.ev[2], st));
  HIPCHECK(bgw::launch_reduce(ra, static_cast<int>(out.n_groups), st));
  const size_t n_real = E.real.size();
  if (q.ensemble_stats) {
    E.d_kmean.ensure(std::max<size_t>(1, n_real * nq) * sizeof(double));
    E.d_emean.ensure(std::max<size_t>(1, nq) * sizeof(double));
    E.d_estd.ensure(std::max<size_t>(1, nq) * sizeof(double));
    HIPCHECK(bgw::launch_ensemble_stats(E.d_s.as<double>(), static_cast<int>(n_real),
                                        static_cast<int>(q.k.size()), static_cast<int>(nq),
                                        E.d_kmean.as<double>(), E.d_emean.as<double>(),
                                        E.d_estd.as<double>(), st));
  }
  HIPCHECK(hipEventRecord(E.dev.ev[3], st));
  out.s.resize(out.n_groups * nq);
  if (q.want_c) {
    out.c.resize(out.n_groups * 2 * nq * 2);
    HIPCHECK(hipMemcpyAsync(out.c.data(), E.d_c.p, out.c.size() * sizeof(double),
                            hipMemcpyDeviceToHost, st));
  }
  HIPCHECK(hipMemcpyAsync(out.s.data(), E.d_s.p, out.s.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  if (q.ensemble_stats) {
    out.kmean.resize(n_real * nq);
    out.mean.resize(nq);
    out.stdev.resize(nq);
    HIPCHECK(hipMemcpyAsync(out.kmean.data(), E.d_kmean.p, out.kmean.size() * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(out.mean.data(), E.d_emean.p, nq * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(out.stdev.data(), E.d_estd.p, nq * sizeof(double),
                            hipMemcpyDeviceToHost, st));
  }
  HIPCHECK(hipStreamSynchronize(st));
  E.timings.geometry_ms = elapsed_ms(E.dev.ev[0], E.dev.ev[1]);
  E.timings.spectrum_ms = elapsed_ms(E.dev.ev[1], E.dev.ev[2]);
  E.timings.reduce_ms = elapsed_ms(E.dev.ev[2], E.dev.ev[3]);
  E.timings.total_ms = elapsed_ms(E.dev.ev[0], E.dev.ev[3]);
  E.timings.n_items = n_items;
  E.timings.n_directions = n_items * n_dir;
  E.timings.n_terms = n_items * res.n_c * q.n_t * q.n_w;
  E.timings.matrix_ms = 0.0;
  for (int k = 0; k < E.mx_timed; ++k)
    E.timings.matrix_ms += elapsed_ms(E.dev.mx_ev[2 * k], E.dev.mx_ev[2 * k + 1]);
  if (E.mx_untimed) E.timings.matrix_ms = std::nan("");
  return out;
}

// Gradient of a loss with respect to the scaling parameters, given the
// cotangent of C (n_groups, 2, n_sets, n_w) complex.  Recomputes the geometry
// (first colliders and delays) and runs the adjoint kernel; the per-item
// gradients are summed on the host in item order.  Returns
// (2, n_sets, n_terms): d/dcoefficients then d/dpowers.
std::vector<double> run_adjoint(Engine& E, const Request& q, const double* grad_c) {
  const Scaling& sc = *q.sc;
  const size_t n_sets = sc.n_sets, n_terms = sc.n_terms;
  std::vector<double> grad(2 * n_sets * n_terms, 0.0);
  if (n_sets == 0 || n_terms == 0 || q.n_w == 0) return grad;
  Launch Lc;
  setup(E, q, Lc);
  hipStream_t st = E.dev.stream;
  const size_t n_items = Lc.n_items;
  const size_t nq = n_sets * q.n_w;
  upload(E.d_gradc, std::vector<double>(grad_c, grad_c + Lc.n_groups * 2 * nq * 2), st);
  E.d_gradout.ensure(std::max<size_t>(1, n_items * 2 * n_sets * n_terms) * sizeof(double));
  bgw::SpecArgs sa = Lc.sa;
  sa.grad_c = E.d_gradc.as<double>();
  sa.grad_out = E.d_gradout.as<double>();
  HIPCHECK(hipEventRecord(E.dev.ev[0], st));
  HIPCHECK(bgw::launch_geometry(Lc.ga, static_cast<int>(n_items), st));
  HIPCHECK(hipEventRecord(E.dev.ev[1], st));
  for (size_t s0 = 0; s0 < n_sets;) {
    bgw::SpecArgs sb = sa;
    sb.set0 = static_cast<int>(s0);
    sb.n_sets = static_cast<int>(std::min<size_t>(bgw::kAdjointMaxSets, n_sets - s0));
    if (bgw::adjoint_lds_bytes(sb, bgw::kAdjointThreads) > 160 * 1024) sb.n_sets = 1;
    if (bgw::adjoint_lds_bytes(sb, bgw::kAdjointThreads) > 160 * 1024)
      throw CalcError("resolution too large: the adjoint's per-ring tables exceed the 160 KiB of LDS");
    sb.coef = E.d_coef.as<double>() + s0 * n_terms;
    sb.pows = E.d_pows.as<double>() + s0 * n_terms;
    HIPCHECK(bgw::launch_adjoint(sb, static_cast<int>(n_items), st));
    s0 += static_cast<size_t>(sb.n_sets);
  }
  HIPCHECK(hipEventRecord(E.dev.ev[2], st));
  std::vector<double> per_item(n_items * 2 * n_sets * n_terms);
  HIPCHECK(hipMemcpyAsync(per_item.data(), E.d_gradout.p, per_item.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  HIPCHECK(hipEventRecord(E.dev.ev[3], st));
  HIPCHECK(hipStreamSynchronize(st));
  for (size_t it = 0; it < n_items; ++it)
    for (size_t e = 0; e < grad.size(); ++e) grad[e] += per_item[it * grad.size() + e];
  E.timings.geometry_ms = elapsed_ms(E.dev.ev[0], E.dev.ev[1]);
  E.timings.spectrum_ms = elapsed_ms(E.dev.ev[1], E.dev.ev[2]);
  E.timings.reduce_ms = 0.0;
  E.timings.matrix_ms = 0.0;
  E.timings.total_ms = elapsed_ms(E.dev.ev[0], E.dev.ev[2]);
  E.timings.n_items = n_items;
  E.timings.n_directions = n_items * q.res->n_c * q.res->n_phi;
  E.timings.n_terms = n_items * q.res->n_c * q.n_t * q.n_w;
  return grad;
}

// Time-resolved integrand, (2, n_sets, n_t, n_w) complex summed over the
// request's bubbles in order.  Items run in chunks whose per-item slabs fit
// kIntegrandScratch bytes; each chunk is added to the total in item order.
constexpr size_t kIntegrandScratch = size_t(1) << 30;

std::vector<double> run_integrand(Engine& E, const Request& q) {
  const size_t n_sets = q.sc->n_sets;
  const size_t n_elem = 2 * n_sets * q.n_t * q.n_w * 2;  // doubles per item
  std::vector<double> out(n_elem, 0.0);
  if (n_sets == 0 || q.n_w == 0) return out;
  Launch Lc;
  setup(E, q, Lc);
  hipStream_t st = E.dev.stream;
  const size_t n_items = Lc.n_items;
  const size_t chunk =
      std::max<size_t>(1, std::min(n_items, kIntegrandScratch / (n_elem * sizeof(double))));
  std::vector<int32_t> iota(n_items);
  std::iota(iota.begin(), iota.end(), 0);
  upload(E.d_iota, iota, st);
  E.d_partial.ensure(chunk * n_elem * sizeof(double));
  E.d_acc.ensure(n_elem * sizeof(double));
  HIPCHECK(hipEventRecord(E.dev.ev[0], st));
  HIPCHECK(bgw::launch_geometry(Lc.ga, static_cast<int>(n_items), st));
  HIPCHECK(hipEventRecord(E.dev.ev[1], st));
  for (size_t c0 = 0; c0 < n_items; c0 += chunk) {
    const size_t nc = std::min(chunk, n_items - c0);
    for (size_t s0 = 0; s0 < n_sets; s0 += kMaxSetsPerLaunch) {
      bgw::SpecArgs sb = Lc.sa;
      sb.order = E.d_iota.as<int32_t>() + c0;
      sb.partial = E.d_partial.as<double>();
      sb.set0 = static_cast<int>(s0);
      sb.n_sets = static_cast<int>(std::min(kMaxSetsPerLaunch, n_sets - s0));
      sb.coef = E.d_coef.as<double>() + s0 * q.sc->n_terms;
      sb.pows = E.d_pows.as<double>() + s0 * q.sc->n_terms;
      sb.time_tables = 0;
      if (bgw::spectrum_lds_bytes(sb, 0) > 160 * 1024)
        throw CalcError("resolution too large: the per-ring tables exceed the 160 KiB of LDS");
      HIPCHECK(bgw::launch_integrand(sb, static_cast<int>(nc), st));
    }
    HIPCHECK(bgw::launch_accumulate(E.d_partial.as<double>(), static_cast<int>(nc),
                                    static_cast<int64_t>(n_elem), c0 == 0 ? 1 : 0,
                                    E.d_acc.as<double>(), st));
  }
  HIPCHECK(hipEventRecord(E.dev.ev[2], st));
  HIPCHECK(hipMemcpyAsync(out.data(), E.d_acc.p, n_elem * sizeof(double), hipMemcpyDeviceToHost,
                          st));
  HIPCHECK(hipEventRecord(E.dev.ev[3], st));
  HIPCHECK(hipStreamSynchronize(st));
  E.timings.geometry_ms = elapsed_ms(E.dev.ev[0], E.dev.ev[1]);
  E.timings.spectrum_ms = elapsed_ms(E.dev.ev[1], E.dev.ev[2]);
  E.timings.reduce_ms = 0.0;
  E.timings.matrix_ms = 0.0;
  E.timings.total_ms = elapsed_ms(E.dev.ev[0], E.dev.ev[2]);
  E.timings.n_items = n_items;
  E.timings.n_directions = n_items * q.res->n_c * q.res->n_phi;
  E.timings.n_terms = n_items * q.res->n_c * q.n_t * q.n_w;
  return out;
}

// ---------------------------------------------------------------------------
// Batched ensemble: the fused persistent kernel, captured in a hipGraph
// ---------------------------------------------------------------------------
struct EnsResult {
  std::vector<double> s, kmean, mean, stdev;
  double ms = 0.0;  // device time of one graph replay
  BgwTileConfig used{0, 0, 0, 0};
};

int device_cus(Engine& E) {
  int n = 0;
  HIPCHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, E.dev.id));
  return n;
}

// Work units, (realisation, k-hat) groups, most expensive first.
std::vector<int32_t> unit_order(const Engine& E, const Launch& Lc) {
  std::vector<double> gcost(Lc.n_groups, 0.0);
  for (size_t g = 0; g < Lc.n_groups; ++g) {
    const size_t r = static_cast<size_t>(Lc.group_r[g]);
    for (int32_t it = Lc.group_off[g]; it < Lc.group_off[g + 1]; ++it)
      gcost[g] += Lc.item_cost[it] + 1e-4 * static_cast<double>(E.real[r].n_tot);
  }
  std::vector<int32_t> order(Lc.n_groups);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return gcost[x] > gcost[y]; });
  return order;
}

// Cache key of the ensemble graph: every field of every argument struct the
// graph's launches capture (sizes, scalars such as the damping width, and
// device addresses), so that a replay never runs with stale arguments after a
// setter or a new k-hat list changed them.  The size checks catch a field
// added to a struct but not listed here.
uint64_t key_bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, sizeof(u));
  return u;
}

void key_args(std::ostringstream& k, const bgw::GeoArgs& a) {
  static_assert(sizeof(bgw::GeoArgs) == 152, "GeoArgs changed: list the new field in key_args");
  k << "|G" << a.cand << ',' << a.inv << ',' << a.real << ',' << a.items << ',' << a.rot << ','
    << a.identity << ',' << a.sin_t << ',' << a.cos_t << ',' << a.cos_p << ',' << a.sin_p << ','
    << a.n_c << ',' << a.n_phi << ',' << a.n_phi_pow2 << ',' << a.tau << ',' << a.perm << ','
    << a.n_t << ','
    << a.first_out << ',' << a.dtab_out << ',' << a.item0 << ',' << a.rings_per_block << ','
    << a.n_split;
}

void key_args(std::ostringstream& k, const bgw::SpecArgs& a) {
  static_assert(sizeof(bgw::SpecArgs) == 232, "SpecArgs changed: list the new field in key_args");
  k << "|S" << a.real << ',' << a.items << ',' << a.order << ',' << a.tau << ',' << a.perm << ','
    << a.gcos << ',' << a.gsin << ',' << a.ring_c << ',' << a.ring_fac << ',' << a.omega << ','
    << a.coef << ',' << a.pows << ',' << a.n_sets << ',' << a.n_terms << ',' << a.set0 << ','
    << a.n_sets_out << ',' << a.n_c << ',' << a.n_phi << ',' << a.n_t << ',' << a.n_w << ','
    << a.cutoff_kind << ',' << key_bits(a.t_cut) << ',' << key_bits(a.inv_t0_sq) << ','
    << a.has_damping << ',' << key_bits(a.damping) << ',' << key_bits(a.kfac) << ','
    << a.skip_uncollided << ',' << a.mixed_block << ',' << a.threads << ',' << a.time_tables
    << ',' << a.time_chunk << ',' << a.surfaces << ',' << a.partial << ',' << a.item0 << ','
    << a.grad_c << ',' << a.grad_out;
}

void key_args(std::ostringstream& k, const bgw::RedArgs& a) {
  static_assert(sizeof(bgw::RedArgs) == 88, "RedArgs changed: list the new field in key_args");
  k << "|R" << a.partial << ',' << a.group_off << ',' << a.group_r << ',' << a.real << ','
    << a.omega << ',' << a.n_sets << ',' << a.n_w << ',' << a.c_out << ',' << a.s_out << ','
    << a.unit_group << ',' << a.unit_off << ',' << a.u0 << ',' << a.item0;
}

void key_args(std::ostringstream& k, const bgw::MatArgs& a) {
  static_assert(sizeof(bgw::MatArgs) == 304, "MatArgs changed: list the new field in key_args");
  key_args(k, a.s);
  k << "|M" << a.tiles << ',' << a.tinfo << ',' << a.ni << ',' << a.omega_tiles << ','
    << a.col_tiles << ',' << a.tile0 << ',' << a.voff0 << ',' << a.vbuf << ',' << a.tt << ','
    << a.ctlo << ',' << a.phase;
}

void key_args(std::ostringstream& k, const bgw::EnsArgs& a) {
  static_assert(sizeof(bgw::EnsArgs) == 440, "EnsArgs changed: list the new field in key_args");
  key_args(k, a.geo);
  key_args(k, a.spec);
  k << "|E" << a.unit_group << ',' << a.group_item_off << ',' << a.group_r << ',' << a.n_units
    << ',' << a.counter << ',' << a.c_out << ',' << a.s_out;
}

// Launch arguments of one chunk of the pipelined form.
struct EnsChunk {
  int item0 = 0, n_items = 0;  // items [item0, item0 + n_items) of the unit-ordered list
  int u0 = 0, n_units = 0;     // units [u0, u0 + n_units)
  int buf = 0;                 // tau / perm buffer (0 or 1)
};

// units: the units to run, in launch order (null: every unit, largest first).
// tiles.items_per_chunk == 0: one fused persistent launch per set chunk;
// otherwise chunks of units, geometry on one stream and phase sums + reduce on
// the other, the delay rows double-buffered.  Either way the whole sequence
// is one captured hipGraph, re-captured only when an argument of one of its
// launches changes (key_args).  replays >= 1 graph launches; ms is their mean.
EnsResult run_ensemble(Engine& E, const Request& q, const BgwTileConfig& tiles,
                       const std::vector<int32_t>* units, bool want_outputs, int replays) {
  const Resolution& res = *q.res;
  const Scaling& sc = *q.sc;
  const size_t n_sets = sc.n_sets;
  const size_t n_real = E.real.size(), n_k = q.k.size();
  EnsResult out;
  const size_t nq = n_sets * q.n_w;
  const bool pipelined = tiles.items_per_chunk > 0;
  Request qq = q;
  qq.item_buffers = false;
  qq.item_order = false;  // the phase-sum order is built per chunk below
  qq.tile_threads = static_cast<int>(tiles.threads);
  qq.tile_rb = bgw::resolve_block_steps(static_cast<int>(tiles.block_steps), q.mode);
  Launch Lc;
  setup(E, qq, Lc);
  hipStream_t st = E.dev.stream;
  const size_t n_dir = res.n_c * res.n_phi;
  const std::vector<int32_t> order = units ? *units : unit_order(E, Lc);

  E.d_counter.ensure(256);
  E.d_c.ensure(std::max<size_t>(1, Lc.n_groups * 2 * nq) * 2 * sizeof(double));
  E.d_s.ensure(std::max<size_t>(1, Lc.n_groups * nq) * sizeof(double));
  E.d_kmean.ensure(std::max<size_t>(1, n_real * nq) * sizeof(double));
  E.d_emean.ensure(std::max<size_t>(1, nq) * sizeof(double));
  E.d_estd.ensure(std::max<size_t>(1, nq) * sizeof(double));
  upload(E.d_units, order, st);

  std::ostringstream key;
  key << (pipelined ? 'P' : 'F') << ':' << q.mode << ':' << order.size() << ':' << n_real << ':'
      << n_k << ':' << n_sets << ':' << q.n_w << ':' << q.n_t << ':' << qq.tile_threads << ':'
      << qq.tile_rb << ':' << tiles.items_per_chunk << ':' << tiles.persistent_blocks;
  std::vector<std::function<hipError_t()>> body;  // graph body, recorded under capture
  int n_blocks = 0;

  if (!pipelined) {
    bgw::EnsArgs ea{};
    ea.geo = Lc.ga;
    ea.geo.first_out = nullptr;
    ea.geo.dtab_out = nullptr;
    ea.spec = Lc.sa;
    ea.unit_group = E.d_units.as<int32_t>();
    ea.group_item_off = E.d_group_off.as<int32_t>();
    ea.group_r = E.d_group_r.as<int32_t>();
    ea.n_units = static_cast<int>(order.size());
    ea.counter = E.d_counter.as<int>();
    ea.c_out = E.d_c.as<double>();
    ea.s_out = E.d_s.as<double>();
    std::vector<bgw::EnsArgs> launches;
    n_blocks = static_cast<int>(tiles.persistent_blocks);
    for (size_t s0 = 0; s0 < n_sets; s0 += bgw::kEnsembleMaxSets) {
      bgw::EnsArgs eb = ea;
      eb.spec.set0 = static_cast<int>(s0);
      eb.spec.n_sets = static_cast<int>(std::min<size_t>(bgw::kEnsembleMaxSets, n_sets - s0));
      eb.spec.coef = E.d_coef.as<double>() + s0 * sc.n_terms;
      eb.spec.pows = E.d_pows.as<double>() + s0 * sc.n_terms;
      eb.spec.time_tables = 0;
      eb.spec.time_tables = bgw::spectrum_time_tables(eb.spec) ? 1 : 0;
      // the chunk length of the per-item phase-sum kernel, so the fused form adds
      // the same terms in the same order as the single-configuration path
      eb.spec.time_chunk = bgw::spectrum_time_chunk(eb.spec, q.mode);
      if (bgw::ensemble_lds_bytes(eb, q.mode) + 24 * 1024 > 160 * 1024)
        throw CalcError("resolution too large: the per-ring tables exceed the 160 KiB of LDS");
      const int per_cu = bgw::ensemble_blocks_per_cu(eb, q.mode);  // sets the LDS limit too
      if (per_cu <= 0) throw CalcError("the ensemble kernel cannot be launched with these tiles");
      if (tiles.persistent_blocks == 0) n_blocks = std::max(n_blocks, per_cu * device_cus(E));
      launches.push_back(eb);
    }
    n_blocks = std::max(n_blocks, 1);
    E.d_tau.ensure(static_cast<size_t>(n_blocks) * n_dir * sizeof(double));
    E.d_perm.ensure(static_cast<size_t>(n_blocks) * n_dir * sizeof(uint16_t));
    for (auto& eb : launches) {
      eb.geo.tau = E.d_tau.as<double>();
      eb.geo.perm = E.d_perm.as<uint16_t>();
      eb.spec.tau = eb.geo.tau;
      eb.spec.perm = eb.geo.perm;
      key_args(key, eb);
      body.push_back([eb, &q, n_blocks, st]() { return bgw::launch_ensemble(eb, q.mode, n_blocks, st); });
    }
    key << ':' << n_blocks;
  } else {
    // items in unit order; chunks of whole units with about items_per_chunk items
    std::vector<Item> items_u;
    std::vector<int32_t> unit_off(1, 0), spec_order;
    const std::vector<Item>& items_orig = Lc.items;
    items_u.reserve(Lc.n_items);
    std::vector<double> cost_u;
    cost_u.reserve(Lc.n_items);
    for (int32_t g : order) {
      for (int32_t it = Lc.group_off[g]; it < Lc.group_off[g + 1]; ++it) {
        items_u.push_back(items_orig[it]);
        cost_u.push_back(Lc.item_cost[it]);
      }
      unit_off.push_back(static_cast<int32_t>(items_u.size()));
    }
    std::vector<EnsChunk> chunks;
    size_t max_chunk = 0;
    for (size_t u = 0; u < order.size();) {
      EnsChunk c;
      c.u0 = static_cast<int>(u);
      c.item0 = unit_off[u];
      while (u < order.size() &&
             (c.n_units == 0 ||
              static_cast<size_t>(unit_off[u + 1] - c.item0) <= tiles.items_per_chunk)) {
        ++u;
        ++c.n_units;
      }
      c.n_items = unit_off[u] - c.item0;
      c.buf = static_cast<int>(chunks.size() & 1);
      max_chunk = std::max(max_chunk, static_cast<size_t>(c.n_items));
      chunks.push_back(c);
    }
    // phase-sum launch order inside each chunk: most expensive items first
    spec_order.resize(items_u.size());
    for (const auto& c : chunks) {
      std::iota(spec_order.begin() + c.item0, spec_order.begin() + c.item0 + c.n_items, c.item0);
      order_by_cost(spec_order.data() + c.item0, static_cast<size_t>(c.n_items),
                    [&](int32_t i) { return cost_u[i]; });
    }
    upload(E.d_items, items_u, st);
    upload(E.d_unit_off, unit_off, st);
    upload(E.d_order, spec_order, st);
    const size_t rows = std::max<size_t>(1, max_chunk);
    E.d_tau.ensure(2 * rows * n_dir * sizeof(double));
    E.d_perm.ensure(2 * rows * n_dir * sizeof(uint16_t));
    E.d_partial.ensure(rows * 2 * nq * 2 * sizeof(double));
    bgw::GeoArgs ga = Lc.ga;
    ga.items = E.d_items.as<Item>();
    ga.first_out = nullptr;
    ga.dtab_out = nullptr;
    bgw::SpecArgs sa = Lc.sa;
    sa.items = ga.items;
    sa.partial = E.d_partial.as<double>();
    bgw::RedArgs ra = Lc.ra;
    ra.partial = sa.partial;
    ra.unit_group = E.d_units.as<int32_t>();
    ra.unit_off = E.d_unit_off.as<int32_t>();
    ra.c_out = E.d_c.as<double>();
    ra.s_out = E.d_s.as<double>();
    for (size_t s0 = 0; s0 < n_sets; s0 += kMaxSetsPerLaunch) {
      bgw::SpecArgs sb = sa;
      sb.n_sets = static_cast<int>(std::min(kMaxSetsPerLaunch, n_sets - s0));
      sb.time_tables = 0;
      sb.time_tables = bgw::spectrum_time_tables(sb) ? 1 : 0;
      if (bgw::spectrum_lds_bytes(sb, q.mode) > 160 * 1024)
        throw CalcError("resolution too large: the per-ring tables exceed the 160 KiB of LDS");
    }
    hipStream_t s2 = E.dev.stream2;
    hipEvent_t fork = E.dev.dep[0], geo_done = E.dev.dep[1];
    hipEvent_t spec_done[2] = {E.dev.dep[2], E.dev.dep[3]};
    double* tau0 = E.d_tau.as<double>();
    uint16_t* perm0 = E.d_perm.as<uint16_t>();
    // stream st: geometry of chunk c (waits for the phase sums of chunk c - 2,
    // the previous user of its buffer); stream s2: phase sums and reduce.
    body.push_back([=]() {
      hipError_t e = hipEventRecord(fork, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(s2, fork, 0);
      return e;
    });
    for (size_t ci = 0; ci < chunks.size(); ++ci) {
      const EnsChunk c = chunks[ci];
      bgw::GeoArgs gc = ga;
      gc.item0 = c.item0;
      gc.tau = tau0 + static_cast<size_t>(c.buf) * rows * n_dir;
      gc.perm = perm0 + static_cast<size_t>(c.buf) * rows * n_dir;
      key_args(key, gc);
      key << ',' << c.n_items;
      body.push_back([=]() {
        hipError_t e = hipSuccess;
        if (ci >= 2) e = hipStreamWaitEvent(st, spec_done[c.buf], 0);
        if (e == hipSuccess) e = bgw::launch_geometry(gc, c.n_items, st);
        if (e == hipSuccess) e = hipEventRecord(geo_done, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(s2, geo_done, 0);
        return e;
      });
      for (size_t s0 = 0; s0 < n_sets; s0 += kMaxSetsPerLaunch) {
        bgw::SpecArgs sb = sa;
        sb.item0 = c.item0;
        sb.order = E.d_order.as<int32_t>() + c.item0;
        sb.tau = gc.tau;
        sb.perm = gc.perm;
        sb.set0 = static_cast<int>(s0);
        sb.n_sets = static_cast<int>(std::min(kMaxSetsPerLaunch, n_sets - s0));
        sb.coef = E.d_coef.as<double>() + s0 * sc.n_terms;
        sb.pows = E.d_pows.as<double>() + s0 * sc.n_terms;
        sb.time_tables = 0;
        sb.time_tables = bgw::spectrum_time_tables(sb) ? 1 : 0;
        const int mode = q.mode;
        key_args(key, sb);
        body.push_back([=]() { return bgw::launch_spectrum(sb, c.n_items, mode, s2); });
      }
      bgw::RedArgs rc = ra;
      rc.u0 = c.u0;
      rc.item0 = c.item0;
      key_args(key, rc);
      key << ',' << c.n_units;
      body.push_back([=]() {
        hipError_t e = bgw::launch_reduce(rc, c.n_units, s2);
        if (e == hipSuccess) e = hipEventRecord(spec_done[c.buf], s2);
        return e;
      });
    }
    key << ':' << chunks.size() << ':' << rows << ':' << static_cast<const void*>(E.d_unit_off.p)
        << ':' << static_cast<const void*>(E.d_partial.p) << ':'
        << static_cast<const void*>(E.d_order.p);
  }
  // statistics on the stream that ran the phase sums, then join
  hipStream_t s_last = pipelined ? E.dev.stream2 : st;
  body.push_back([&, s_last]() {
    hipError_t e = bgw::launch_ensemble_stats(
        E.d_s.as<double>(), static_cast<int>(n_real), static_cast<int>(n_k), static_cast<int>(nq),
        E.d_kmean.as<double>(), E.d_emean.as<double>(), E.d_estd.as<double>(), s_last);
    if (e == hipSuccess && s_last != st) {
      e = hipEventRecord(E.dev.dep[0], s_last);
      if (e == hipSuccess) e = hipStreamWaitEvent(st, E.dev.dep[0], 0);
    }
    return e;
  });
  key << ':' << static_cast<const void*>(E.d_tau.p) << ':' << static_cast<const void*>(E.d_perm.p)
      << ':' << static_cast<const void*>(E.d_items.p) << ':' << static_cast<const void*>(E.d_real.p)
      << ':' << static_cast<const void*>(E.d_omega.p) << ':' << static_cast<const void*>(E.d_coef.p)
      << ':' << static_cast<const void*>(E.d_units.p) << ':'
      << static_cast<const void*>(E.d_group_off.p) << ':' << static_cast<const void*>(E.d_c.p)
      << ':' << static_cast<const void*>(E.d_s.p) << ':' << static_cast<const void*>(E.d_cand.p)
      << ':' << static_cast<const void*>(E.d_rot.p) << ':' << static_cast<const void*>(E.d_kmean.p)
      << ':' << static_cast<const void*>(E.d_emean.p) << ':' << static_cast<const void*>(E.d_estd.p);
  if (!E.graph || E.graph_key != key.str()) {
    if (E.graph) {
      HIPCHECK(hipGraphExecDestroy(E.graph));
      E.graph = nullptr;
    }
    E.graph_key.clear();
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (auto& f : body)
      if (err == hipSuccess) err = f();
    hipError_t end = hipStreamEndCapture(st, &graph);
    HIPCHECK(err);
    HIPCHECK(end);
    hipError_t inst = hipGraphInstantiate(&E.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(inst);
    E.graph_key = key.str();
  }
  HIPCHECK(hipEventRecord(E.dev.ev[0], st));
  for (int rep = 0; rep < std::max(1, replays); ++rep) HIPCHECK(hipGraphLaunch(E.graph, st));
  HIPCHECK(hipEventRecord(E.dev.ev[1], st));
  if (want_outputs) {
    out.s.resize(Lc.n_groups * nq);
    out.kmean.resize(n_real * nq);
    out.mean.resize(nq);
    out.stdev.resize(nq);
    HIPCHECK(hipMemcpyAsync(out.s.data(), E.d_s.p, out.s.size() * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(out.kmean.data(), E.d_kmean.p, out.kmean.size() * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(out.mean.data(), E.d_emean.p, nq * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(out.stdev.data(), E.d_estd.p, nq * sizeof(double),
                            hipMemcpyDeviceToHost, st));
  }
  HIPCHECK(hipStreamSynchronize(st));
  out.ms = elapsed_ms(E.dev.ev[0], E.dev.ev[1]) / std::max(1, replays);
  out.used = tiles;
  out.used.block_steps = static_cast<size_t>(qq.tile_rb);
  out.used.threads = static_cast<size_t>(
      bgw::resolve_threads(static_cast<int>(tiles.threads), static_cast<int>(q.n_w), q.mode));
  out.used.persistent_blocks = pipelined ? 0 : static_cast<size_t>(n_blocks);
  E.timings.geometry_ms = 0.0;
  E.timings.spectrum_ms = out.ms;
  E.timings.reduce_ms = 0.0;
  E.timings.matrix_ms = 0.0;
  E.timings.total_ms = out.ms;
  E.timings.n_items = Lc.n_items;
  E.timings.n_directions = Lc.n_items * n_dir;
  E.timings.n_terms = Lc.n_items * res.n_c * q.n_t * q.n_w;
  return out;
}

// A batched ensemble in the FP64 matrix form: the single-configuration launches
// over every realisation at once (geometry, per batch of tiles the time tables,
// weight rows and products, the reduction, then the k-hat means and the
// ensemble mean and spread), captured as one hipGraph after the per-call
// uploads and replayed while every argument of its launches is unchanged
// (key_args).  Requests whose sets need several matrix launches run uncaptured.
Result run_matrix_ensemble(Engine& E, const Request& q) {
  const Scaling& sc = *q.sc;
  const size_t n_sets = sc.n_sets;
  if (n_sets == 0 || q.n_w == 0 || n_sets > kMaxSetsPerLaunch ||
      n_sets * sc.n_terms > static_cast<size_t>(bgw::kMaxMatrixTerms))
    return run(E, q);
  Launch Lc;
  setup(E, q, Lc);
  hipStream_t st = E.dev.stream;
  const size_t n_real = E.real.size(), n_k = q.k.size();
  const size_t nq = n_sets * q.n_w;
  bgw::SpecArgs sm = Lc.sa;
  sm.set0 = 0;
  sm.n_sets = static_cast<int>(n_sets);
  sm.coef = E.d_coef.as<double>();
  sm.pows = E.d_pows.as<double>();
  sm.time_tables = 0;
  const MatrixPlan P = plan_matrix_form(E, q, Lc, sm, st);
  E.d_kmean.ensure(std::max<size_t>(1, n_real * nq) * sizeof(double));
  E.d_emean.ensure(std::max<size_t>(1, nq) * sizeof(double));
  E.d_estd.ensure(std::max<size_t>(1, nq) * sizeof(double));
  std::ostringstream key;
  key << "matrix:" << Lc.n_items << ':' << Lc.n_groups << ':' << n_real << ':' << n_k << ':' << nq;
  key_args(key, Lc.ga);
  for (size_t b = 0; b < P.batches.size(); ++b) {
    bgw::MatArgs ma = P.ma;
    ma.tile0 = static_cast<int>(P.batches[b].first);
    ma.voff0 = P.voff0[b];
    key << "|b" << (P.batches[b].second - P.batches[b].first);
    key_args(key, ma);
  }
  key_args(key, Lc.ra);
  key << ':' << static_cast<const void*>(E.d_kmean.p) << ':' << static_cast<const void*>(E.d_emean.p)
      << ':' << static_cast<const void*>(E.d_estd.p);
  if (!E.graph || E.graph_key != key.str()) {
    if (E.graph) {
      HIPCHECK(hipGraphExecDestroy(E.graph));
      E.graph = nullptr;
    }
    E.graph_key.clear();
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t err = bgw::launch_geometry(Lc.ga, static_cast<int>(Lc.n_items), st);
    if (err == hipSuccess) {
      try {
        launch_matrix_plan(E, P, st, false);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
    }
    if (err == hipSuccess) err = bgw::launch_reduce(Lc.ra, static_cast<int>(Lc.n_groups), st);
    if (err == hipSuccess)
      err = bgw::launch_ensemble_stats(E.d_s.as<double>(), static_cast<int>(n_real),
                                       static_cast<int>(n_k), static_cast<int>(nq),
                                       E.d_kmean.as<double>(), E.d_emean.as<double>(),
                                       E.d_estd.as<double>(), st);
    hipError_t end = hipStreamEndCapture(st, &graph);
    HIPCHECK(err);
    HIPCHECK(end);
    hipError_t inst = hipGraphInstantiate(&E.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(inst);
    E.graph_key = key.str();
  }
  HIPCHECK(hipEventRecord(E.dev.ev[0], st));
  HIPCHECK(hipGraphLaunch(E.graph, st));
  HIPCHECK(hipEventRecord(E.dev.ev[1], st));
  Result out;
  out.n_groups = Lc.n_groups;
  out.s.resize(Lc.n_groups * nq);
  out.kmean.resize(n_real * nq);
  out.mean.resize(nq);
  out.stdev.resize(nq);
  HIPCHECK(hipMemcpyAsync(out.s.data(), E.d_s.p, out.s.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(out.kmean.data(), E.d_kmean.p, out.kmean.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(out.mean.data(), E.d_emean.p, nq * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(out.stdev.data(), E.d_estd.p, nq * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  const double ms = elapsed_ms(E.dev.ev[0], E.dev.ev[1]);
  const size_t n_dir = q.res->n_c * q.res->n_phi;
  E.timings.geometry_ms = 0.0;
  E.timings.spectrum_ms = ms;
  E.timings.reduce_ms = 0.0;
  E.timings.matrix_ms = std::nan("");
  E.timings.total_ms = ms;
  E.timings.n_items = Lc.n_items;
  E.timings.n_directions = Lc.n_items * n_dir;
  E.timings.n_terms = Lc.n_items * q.res->n_c * q.n_t * q.n_w;
  return out;
}

// Geometry only, for one item: first colliding bubble and delta_tab.
void run_masks(Engine& E, const Resolution& res, int32_t a, const double* k_hat,
               std::vector<int32_t>& first, std::vector<double>& dtab) {
  ensure_static(E);
  hipStream_t st = E.dev.stream;
  std::vector<RealInfo> real = E.real;
  std::vector<Item> items(1);
  items[0].r = 0;
  items[0].k = 0;
  items[0].a = a;
  items[0].group = 0;
  const size_t n_dir = res.n_c * res.n_phi;
  upload(E.d_real, real, st);
  upload(E.d_items, items, st);
  std::vector<double> rot(9, 0.0);
  if (k_hat) {
    auto R = rotation_to_z(k_hat);
    for (int j = 0; j < 9; ++j) rot[j] = R[j];
  }
  upload(E.d_rot, rot, st);
  upload(E.d_sin_t, res.sin_t, st);
  upload(E.d_cos_t, res.cos_t, st);
  upload(E.d_cos_p, res.cos_p, st);
  upload(E.d_sin_p, res.sin_p, st);
  E.d_tau.ensure(n_dir * sizeof(double));
  E.d_perm.ensure(n_dir * sizeof(uint16_t));
  E.d_first.ensure(n_dir * sizeof(int32_t));
  E.d_dtab.ensure(n_dir * sizeof(double));
  bgw::GeoArgs ga{};
  ga.cand = E.d_cand.as<CandEntry>();
  ga.inv = E.d_inv.as<int32_t>();
  ga.real = E.d_real.as<RealInfo>();
  ga.items = E.d_items.as<Item>();
  ga.rot = E.d_rot.as<double>();
  ga.identity = k_hat ? 0 : 1;
  ga.sin_t = E.d_sin_t.as<double>();
  ga.cos_t = E.d_cos_t.as<double>();
  ga.cos_p = E.d_cos_p.as<double>();
  ga.sin_p = E.d_sin_p.as<double>();
  ga.n_c = static_cast<int>(res.n_c);
  ga.n_phi = static_cast<int>(res.n_phi);
  ga.n_phi_pow2 = static_cast<int>(res.n_phi_pow2);
  ga.tau = E.d_tau.as<double>();
  ga.perm = E.d_perm.as<uint16_t>();
  ga.first_out = E.d_first.as<int32_t>();
  ga.dtab_out = E.d_dtab.as<double>();
  HIPCHECK(bgw::launch_geometry(ga, 1, st));
  first.resize(n_dir);
  dtab.resize(n_dir);
  HIPCHECK(hipMemcpyAsync(first.data(), E.d_first.p, n_dir * sizeof(int32_t),
                          hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(dtab.data(), E.d_dtab.p, n_dir * sizeof(double), hipMemcpyDeviceToHost,
                          st));
  HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------
struct bgw_generalized_bulk_flow {
  Bubbles bubbles;
  Resolution res;
  Scaling sc;
  size_t num_threads = 0;  // 0: the hardware's thread count
  BgwTileConfig tiles{0, 0, 0, 0};
  Engine engine;
  bool built = false;      // candidate lists of `bubbles` built into `engine`
};

struct bgw_ensemble {
  std::vector<Bubbles> reals;
  Resolution res;
  Scaling sc;
  Engine engine;
  BgwTileConfig tiles{0, 0, 0, 0};
  BgwTileConfig matrix_tiles{0, 0, 0, 0};  // the FP64 matrix form's (its own autotune)
  bool tuned = false;
};

namespace {

void check_tiles(const BgwTileConfig& t) {
  if (t.block_steps != 0 && t.block_steps != 1 && t.block_steps != 2 && t.block_steps != 4 &&
      t.block_steps != 8 && t.block_steps != 16)
    throw CalcError("tiles.block_steps must be 0, 8 or 16 (1, 2 or 4 for the FP64 matrix form)");
  if (t.threads != 0 && t.threads != 64 && t.threads != 128 && t.threads != 256)
    throw CalcError("tiles.threads must be 0, 64, 128 or 256");
  if (t.persistent_blocks > 65536) throw CalcError("tiles.persistent_blocks must be <= 65536");
  if (t.items_per_chunk > (size_t(1) << 24))
    throw CalcError("tiles.items_per_chunk must be <= 16777216");
}

void apply_tiles(const BgwTileConfig& t, Request& q) {
  q.tile_rb = static_cast<int>(t.block_steps);  // 0: the form's default
  q.tile_threads = static_cast<int>(t.threads);
}

bgw_generalized_bulk_flow* bulk_flow_handle(bgw_generalized_bulk_flow* h) {
  if (!h) throw CalcError(null_error("bulk_flow"));
  return h;
}

size_t host_threads(size_t requested);

// The handle's engine, its per-bubble candidate lists built on first use on
// host_threads(num_threads) threads (the reference's thread pool of
// set_num_threads, lib.rs:432-446).
Engine& built_engine(bgw_generalized_bulk_flow* h) {
  if (!h->built) {
    build_static(h->engine, {h->bubbles}, host_threads(h->num_threads));
    h->built = true;
  }
  return h->engine;
}

const Resolution& need_resolution(const Resolution& r) {
  if (!r.set) throw CalcError("Field 'n_cos_thetax' is not initialized");
  return r;
}

size_t host_threads(size_t requested) {
  if (requested) return requested;
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? hc : 1;
}

std::vector<int32_t> selected_ids(const bgw_generalized_bulk_flow* h, const size_t* selected,
                                  size_t n_selected, bool* empty_selection) {
  std::vector<int32_t> ids;
  *empty_selection = false;
  if (selected) {
    slice_from_raw(selected, n_selected, "selected_bubbles");
    if (n_selected == 0) *empty_selection = true;
    for (size_t q = 0; q < n_selected; ++q) {
      if (selected[q] >= h->bubbles.n_int)
        throw CalcError("Index " + std::to_string(selected[q]) + " out of bounds for max " +
                        std::to_string(h->bubbles.n_int));
      ids.push_back(static_cast<int32_t>(selected[q]));
    }
  } else {
    for (size_t a = 0; a < h->bubbles.n_int; ++a) ids.push_back(static_cast<int32_t>(a));
  }
  return ids;
}

void check_time_range(double t_begin, double t_end) {
  if (t_begin > t_end)
    throw CalcError("Invalid time range: t_begin=" + std::to_string(t_begin) +
                    " > t_end=" + std::to_string(t_end));
}

// compute_c_integral (generalized_bulk_flow.rs:928-984) on the GPU.
void check_unit_vectors(const double* k_hat, size_t n_k) {
  for (size_t k = 0; k < n_k; ++k) {
    const double* kk = k_hat + 3 * k;
    const double nrm = std::sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]);
    if (!(std::fabs(nrm - 1.0) <= 1e-12))
      throw CalcError("k_hat[" + std::to_string(k) + "] is not a unit vector");
  }
}

std::vector<double> compute_c(bgw_generalized_bulk_flow* h, const double* w_arr, size_t n_w,
                              bool has_t_begin, double t_begin, double t_end, size_t n_t,
                              const size_t* selected, size_t n_selected,
                              BgwTimeCutoffConfig cutoff, int mode,
                              const Resolution* surface = nullptr) {
  slice_from_raw(w_arr, n_w, "w_arr");
  const size_t n_sets = h->sc.n_sets;
  std::vector<double> zeros(2 * n_sets * n_w * 2, 0.0);
  if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2 for integration");
  bool empty = false;
  std::vector<int32_t> ids = selected_ids(h, selected, n_selected, &empty);
  if (ids.empty()) return zeros;
  const double tb = has_t_begin ? t_begin : 0.0;
  check_time_range(tb, t_end);
  const Resolution& res = surface ? *surface : need_resolution(h->res);
  Request q;
  q.res = &res;
  q.sc = &h->sc;
  q.cutoff = make_cutoff(cutoff);
  q.n_w = n_w;
  q.n_t = n_t;
  q.identity = true;
  q.k.push_back({0.0, 0.0, 1.0});
  q.omega.assign(w_arr, w_arr + n_w);
  q.t_begin = {tb};
  q.t_end = {t_end};
  q.bubbles_of_real = {ids};
  q.mode = mode;
  q.want_s = false;
  if (surface) q.skip_uncollided = false;  // components such as xz may be small
  apply_tiles(h->tiles, q);
  Result r = run(built_engine(h), q);
  return r.c;
}

// compute_c on several surface-weight sets of the handle's grid in one request:
// the sets take the k-hat slots of an identity-frame request, so their items run
// in the same launches.  Returns each set's C, laid out as compute_c's.
std::vector<std::vector<double>> compute_c_surfaces(bgw_generalized_bulk_flow* h,
                                                    const double* w_arr, size_t n_w,
                                                    bool has_t_begin, double t_begin,
                                                    double t_end, size_t n_t,
                                                    const std::vector<int32_t>& ids, int mode,
                                                    const std::vector<Resolution>& surfaces) {
  const size_t n_sets = h->sc.n_sets;
  Request q;
  q.res = &surfaces.front();
  for (const Resolution& S : surfaces) q.surfaces.push_back(&S);
  q.sc = &h->sc;
  q.cutoff = make_cutoff(BgwTimeCutoffConfig{BgwTimeCutoffKind_Unit, 0.0, NAN, NAN});
  q.n_w = n_w;
  q.n_t = n_t;
  q.identity = true;
  q.k.assign(surfaces.size(), {0.0, 0.0, 1.0});
  q.omega.assign(w_arr, w_arr + n_w);
  q.t_begin = {has_t_begin ? t_begin : 0.0};
  q.t_end = {t_end};
  q.bubbles_of_real = {ids};
  q.mode = mode;
  q.want_s = false;
  q.skip_uncollided = false;  // components such as xz may be small
  apply_tiles(h->tiles, q);
  Result r = run(built_engine(h), q);
  const size_t blk = 2 * n_sets * n_w * 2;  // one group's C
  std::vector<std::vector<double>> out;
  for (size_t k = 0; k < surfaces.size(); ++k)
    out.emplace_back(r.c.begin() + k * blk, r.c.begin() + (k + 1) * blk);
  return out;
}

// Surface weights of the tensor components (new_generalized_bulk_flow.rs:592-727):
// the phi-integral of component c on ring i is
//   sum_j w_j f_j phi_c(phi_j) * theta_c(theta_i) * dphi * dta^3,
// phi_c in {cos^2, cos sin, cos, sin^2, sin, 1}, theta_c in {sin^2, sin^2,
// sin cos, sin^2, sin cos, cos^2}, with sin = sqrt(max(1 - cos^2, 0)).  The
// phase sums take two phi-functions with one theta-factor per weight set, so the
// six components are four weight sets: (xx, xy), (xz, yz), (yy, -), (zz, -).
// The integrand path (compute_c_integrand, generalized_bulk_flow.rs:875-920) on the
// surface weights of `surface` (the handle's own when null).  The two components
// of the run are written to components dst[0], dst[1] (skipped when < 0) of out,
// laid out (n_comp, n_sets, n_w, n_t) complex.  A bubble nucleated at or after
// t_end contributes zeros before the time range is checked (:681-692).
void integrand_into(bgw_generalized_bulk_flow* h, const double* w_arr, size_t n_w,
                    bool has_t_begin, double t_begin, double t_end, size_t n_t,
                    const std::vector<int32_t>& ids, const Resolution* surface, double* out,
                    std::array<int, 2> dst) {
  const size_t n_sets = h->sc.n_sets;
  const double tb = has_t_begin ? t_begin : 0.0;
  bool live = false;
  for (int32_t a : ids) live = live || h->bubbles.row(static_cast<size_t>(a))[0] < t_end;
  std::vector<double> res_t(2 * 2 * n_sets * n_w * n_t, 0.0);  // (2, n_sets, n_t, n_w)
  if (live) {
    check_time_range(tb, t_end);
    Request q;
    q.res = surface ? surface : &need_resolution(h->res);
    q.sc = &h->sc;
    q.cutoff = Cutoff{};
    q.n_w = n_w;
    q.n_t = n_t;
    q.identity = true;
    q.k.push_back({0.0, 0.0, 1.0});
    q.omega.assign(w_arr, w_arr + n_w);
    q.t_begin = {tb};
    q.t_end = {t_end};
    q.bubbles_of_real = {ids};
    q.skip_uncollided = false;
    res_t = run_integrand(built_engine(h), q);
  }
  for (int pm = 0; pm < 2; ++pm) {
    if (dst[pm] < 0) continue;
    for (size_t s = 0; s < n_sets; ++s)
      for (size_t w = 0; w < n_w; ++w)
        for (size_t k = 0; k < n_t; ++k) {
          const size_t src = (((pm * n_sets + s) * n_t + k) * n_w + w) * 2;
          const size_t d = (((dst[pm] * n_sets + s) * n_w + w) * n_t + k) * 2;
          out[d] = res_t[src];
          out[d + 1] = res_t[src + 1];
        }
  }
}

Resolution tensor_surface(const Resolution& R, int pair) {
  Resolution S = R;
  const size_t n_c = R.n_c, n_phi = R.n_phi;
  const double dphi = 2.0 * kPi / static_cast<double>(n_phi);
  const double dcos = 2.0 / static_cast<double>(n_c - 1);
  for (size_t j = 0; j < n_phi; ++j) {
    const double w = (j == 0 || j == n_phi - 1) ? 0.5 : 1.0;
    const double cp = R.cos_p[j], sp = R.sin_p[j];
    double a = 0.0, b = 0.0;
    switch (pair) {
      case 0: a = cp * cp; b = cp * sp; break;
      case 1: a = cp; b = sp; break;
      case 2: a = sp * sp; break;
      default: a = 1.0; break;
    }
    S.gcos[j] = a * w;
    S.gsin[j] = b * w;
  }
  for (size_t i = 0; i < n_c; ++i) {
    const double c = R.cos_t[i];
    const double st = std::sqrt(std::fmax(1.0 - c * c, 0.0));
    const double th = pair == 1 ? st * c : (pair == 3 ? c * c : st * st);
    const double wi = (i == 0 || i == n_c - 1) ? 0.5 : 1.0;
    S.ring_fac[i] = dcos * wi * th * dphi;
  }
  return S;
}

}  // namespace

namespace {

// Request over every realisation of an ensemble (all bubbles, shared k-hat list).
Request ensemble_request(bgw_ensemble* ensemble, const double* w_arr, size_t n_w,
                         const double* k_hat, size_t n_k, const double* t_begin,
                         const double* t_end, size_t n_t, BgwPrecision precision) {
  const size_t n_real = ensemble->reals.size();
  slice_from_raw(w_arr, checked_product(n_real, n_w, "w_arr"), "w_arr");
  slice_from_raw(k_hat, checked_product(n_k, 3, "k_hat"), "k_hat");
  slice_from_raw(t_begin, n_real, "t_begin");
  slice_from_raw(t_end, n_real, "t_end");
  if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2 for integration");
  check_unit_vectors(k_hat, n_k);
  Request q;
  q.res = &need_resolution(ensemble->res);
  q.sc = &ensemble->sc;
  q.n_w = n_w;
  q.n_t = n_t;
  q.identity = false;
  for (size_t k = 0; k < n_k; ++k) q.k.push_back({k_hat[3 * k], k_hat[3 * k + 1], k_hat[3 * k + 2]});
  q.omega.assign(w_arr, w_arr + n_real * n_w);
  q.t_begin.assign(t_begin, t_begin + n_real);
  q.t_end.assign(t_end, t_end + n_real);
  for (size_t r = 0; r < n_real; ++r) {
    check_time_range(q.t_begin[r], q.t_end[r]);
    std::vector<int32_t> ids(ensemble->reals[r].n_int);
    std::iota(ids.begin(), ids.end(), 0);
    q.bubbles_of_real.push_back(ids);
  }
  q.mode = kernel_mode(precision);
  return q;
}

}  // namespace

extern "C" {

const char* bgw_version(void) { return "0.1.0"; }

const char* bgw_last_error_message(void) { return t_last_error.c_str(); }

BgwStatus bgw_generalized_bulk_flow_empty_lattice_new(bgw_generalized_bulk_flow** out,
                                                      const double* bubbles_interior,
                                                      size_t n_interior,
                                                      const double* bubbles_exterior,
                                                      size_t n_exterior) {
  if (!out) {
    set_last_error(null_error("out"));
    return BgwStatus_NullPointer;
  }
  *out = nullptr;
  return run_ffi([&] {
    const size_t ni = checked_product(n_interior, 4, "interior bubbles");
    const size_t ne = checked_product(n_exterior, 4, "exterior bubbles");
    slice_from_raw(bubbles_interior, ni, "bubbles_interior");
    slice_from_raw(bubbles_exterior, ne, "bubbles_exterior");
    if (n_interior + n_exterior > static_cast<size_t>(std::numeric_limits<int32_t>::max() / 4))
      throw CalcError("too many bubbles");
    auto h = std::make_unique<bgw_generalized_bulk_flow>();
    h->bubbles = make_bubbles(bubbles_interior, n_interior, bubbles_exterior, n_exterior);
    *out = h.release();
  });
}

void bgw_generalized_bulk_flow_free(bgw_generalized_bulk_flow* bulk_flow) { delete bulk_flow; }

BgwStatus bgw_generalized_bulk_flow_set_num_threads(bgw_generalized_bulk_flow* bulk_flow,
                                                    size_t num_threads) {
  return run_ffi([&] {
    if (num_threads == 0) throw CalcError("num_threads must be > 0");
    bulk_flow_handle(bulk_flow)->num_threads = num_threads;
  });
}

BgwStatus bgw_generalized_bulk_flow_set_resolution(bgw_generalized_bulk_flow* bulk_flow,
                                                   size_t n_cos_thetax, size_t n_phix,
                                                   bool /*precompute_first_bubbles*/) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    h->res = make_resolution(n_cos_thetax, n_phix);
  });
}

BgwStatus bgw_generalized_bulk_flow_set_gradient_scaling_params(
    bgw_generalized_bulk_flow* bulk_flow, const double* coefficients, const double* powers,
    size_t n_sets, size_t n_terms, double damping_width) {
  return run_ffi([&] {
    Scaling s = make_scaling(coefficients, powers, n_sets, n_terms, damping_width);
    bulk_flow_handle(bulk_flow)->sc = std::move(s);
  });
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point: every buffer of its calls is carved
// from this array, one copy per device, held until the process ends.  Host code
// uses only the address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<char*> g_scratch_ptr;  // by device id

// The device address of `scratch` on device `dev`, which must be current.
static char* no_alloc_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0) throw CalcError("no current HIP device");
  if (g_scratch_ptr.size() <= static_cast<size_t>(dev))
    g_scratch_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw CalcError(std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") +
                      hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

static BgwStatus compute_c_integral_impl(bgw_generalized_bulk_flow* bulk_flow,
                                         const double* w_arr, size_t n_w, bool has_t_begin,
                                         double t_begin, double t_end, size_t n_t,
                                         const size_t* selected, size_t n_selected,
                                         BgwTimeCutoffConfig time_cutoff, int mode,
                                         BgwComplex64* out, size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    std::vector<double> c = compute_c(h, w_arr, n_w, has_t_begin, t_begin, t_end, n_t, selected,
                                      n_selected, time_cutoff, mode);
    const size_t expected = c.size() / 2;
    if (out_len != expected)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(expected));
    slice_from_raw(out, out_len, "out");
    if (out_len) std::memcpy(out, c.data(), c.size() * sizeof(double));
  });
}

BgwStatus bgw_generalized_bulk_flow_compute_c_integral(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t* selected_bubbles,
    size_t n_selected_bubbles, BgwTimeCutoffConfig time_cutoff, BgwComplex64* out,
    size_t out_len) {
  return compute_c_integral_impl(bulk_flow, w_arr, n_w, has_t_begin, t_begin, t_end, n_t,
                                 selected_bubbles, n_selected_bubbles, time_cutoff, 0, out,
                                 out_len);
}

BgwStatus bgw_generalized_bulk_flow_compute_c_integral_ex(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t* selected_bubbles,
    size_t n_selected_bubbles, BgwTimeCutoffConfig time_cutoff, BgwPrecision precision,
    BgwComplex64* out, size_t out_len) {
  if (!valid_precision(precision)) {
    set_last_error("precision is not a BgwPrecision");
    return BgwStatus_InvalidArgument;
  }
  return compute_c_integral_impl(bulk_flow, w_arr, n_w, has_t_begin, t_begin, t_end, n_t,
                                 selected_bubbles, n_selected_bubbles, time_cutoff,
                                 kernel_mode(precision), out, out_len);
}

BgwStatus bgw_generalized_bulk_flow_compute_c_integral_no_alloc(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t* selected_bubbles,
    size_t n_selected_bubbles, BgwTimeCutoffConfig time_cutoff, BgwPrecision precision,
    BgwComplex64* out, size_t out_len) {
  if (!valid_precision(precision)) {
    set_last_error("precision is not a BgwPrecision");
    return BgwStatus_InvalidArgument;
  }
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    const size_t expected = checked_product(2 * h->sc.n_sets, n_w, "out");
    if (out_len != expected)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(expected));
    slice_from_raw(out, out_len, "out");
    Engine& E = built_engine(h);
    E.dev.init();  // makes the handle's device current
    Arena A;
    A.base = no_alloc_scratch(E.dev.id);
    A.cap = kNoAllocScratchBytes;
    std::vector<double> c;
    {
      ArenaScope scope(E, A);
      c = compute_c(h, w_arr, n_w, has_t_begin, t_begin, t_end, n_t, selected_bubbles,
                    n_selected_bubbles, time_cutoff, kernel_mode(precision));
    }
    if (out_len) std::memcpy(out, c.data(), c.size() * sizeof(double));
  });
}

BgwStatus bgw_generalized_bulk_flow_compute_c_tensor(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t* selected_bubbles,
    size_t n_selected_bubbles, BgwPrecision precision, BgwComplex64* out, size_t out_len) {
  if (!valid_precision(precision)) {
    set_last_error("precision is not a BgwPrecision");
    return BgwStatus_InvalidArgument;
  }
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    slice_from_raw(w_arr, n_w, "w_arr");
    const size_t n_sets = h->sc.n_sets;
    const size_t expected = 6 * n_sets * n_w;
    if (out_len != expected)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(expected));
    slice_from_raw(out, out_len, "out");
    if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2 for integration");
    bool empty = false;
    const std::vector<int32_t> ids = selected_ids(h, selected_bubbles, n_selected_bubbles, &empty);
    std::vector<double> res(2 * expected, 0.0);
    if (!ids.empty()) {
      check_time_range(has_t_begin ? t_begin : 0.0, t_end);
      const Resolution& R = need_resolution(h->res);
      // weight set -> (component of C+, component of Cx), -1: unused
      const int comp[4][2] = {{0, 1}, {2, 4}, {3, -1}, {5, -1}};
      const size_t blk = 2 * n_sets * n_w;  // doubles of one component
      std::vector<Resolution> surfaces;
      for (int pair = 0; pair < 4; ++pair) surfaces.push_back(tensor_surface(R, pair));
      // the four weight sets in one request: their items share the launches
      const std::vector<std::vector<double>> cs =
          compute_c_surfaces(h, w_arr, n_w, has_t_begin, t_begin, t_end, n_t, ids,
                             kernel_mode(precision), surfaces);
      for (int pair = 0; pair < 4; ++pair)
        for (int pm = 0; pm < 2; ++pm)
          if (comp[pair][pm] >= 0)
            std::memcpy(&res[comp[pair][pm] * blk], &cs[pair][pm * blk], blk * sizeof(double));
    }
    std::memcpy(out, res.data(), res.size() * sizeof(double));
  });
}

BgwStatus bgw_generalized_bulk_flow_first_colliding_bubble(bgw_generalized_bulk_flow* bulk_flow,
                                                           size_t a_idx, int32_t* out,
                                                           size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    const Resolution& res = need_resolution(h->res);
    if (a_idx >= h->bubbles.n_int)
      throw CalcError("Index " + std::to_string(a_idx) + " out of bounds (max: " +
                      std::to_string(h->bubbles.n_int) + ")");
    const size_t n = res.n_c * res.n_phi;
    if (out_len != n)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(n));
    slice_from_raw(out, out_len, "out");
    std::vector<int32_t> first;
    std::vector<double> dtab;
    run_masks(built_engine(h), res, static_cast<int32_t>(a_idx), nullptr, first, dtab);
    std::memcpy(out, first.data(), n * sizeof(int32_t));
  });
}

BgwStatus bgw_generalized_bulk_flow_first_colliding_bubble_ex(bgw_generalized_bulk_flow* bulk_flow,
                                                              size_t a_idx, const double* k_hat,
                                                              int32_t* first_out,
                                                              double* delta_tab_out,
                                                              size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    const Resolution& res = need_resolution(h->res);
    if (a_idx >= h->bubbles.n_int)
      throw CalcError("Index " + std::to_string(a_idx) + " out of bounds (max: " +
                      std::to_string(h->bubbles.n_int) + ")");
    const size_t n = res.n_c * res.n_phi;
    if (out_len != n)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(n));
    if (k_hat) {
      const double nrm = std::sqrt(k_hat[0] * k_hat[0] + k_hat[1] * k_hat[1] + k_hat[2] * k_hat[2]);
      if (!(std::fabs(nrm - 1.0) <= 1e-12)) throw CalcError("k_hat is not a unit vector");
    }
    std::vector<int32_t> first;
    std::vector<double> dtab;
    run_masks(built_engine(h), res, static_cast<int32_t>(a_idx), k_hat, first, dtab);
    if (first_out) std::memcpy(first_out, first.data(), n * sizeof(int32_t));
    if (delta_tab_out) std::memcpy(delta_tab_out, dtab.data(), n * sizeof(double));
  });
}

BgwStatus bgw_generalized_bulk_flow_delta_tab(bgw_generalized_bulk_flow* bulk_flow, size_t a_idx,
                                              const int32_t* first_bubble, size_t grid_len,
                                              double* out, size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    const Resolution& res = need_resolution(h->res);
    const size_t n = res.n_c * res.n_phi;
    if (grid_len != n || out_len != n)
      throw CalcError("Array shape mismatch: First bubble array must match resolution: expected [" +
                      std::to_string(res.n_c) + ", " + std::to_string(res.n_phi) + "]");
    if (a_idx >= h->bubbles.n_int)
      throw CalcError("Index " + std::to_string(a_idx) + " out of bounds (max: " +
                      std::to_string(h->bubbles.n_int) + ")");
    slice_from_raw(first_bubble, n, "first_bubble");
    slice_from_raw(out, n, "out");
    // compute_delta_tab (generalized_bulk_flow.rs:446-468): one division per
    // direction, no search.
    const Bubbles& B = h->bubbles;
    const double* xa = B.row(a_idx);
    for (size_t i = 0; i < res.n_c; ++i)
      for (size_t j = 0; j < res.n_phi; ++j) {
        const size_t ij = i * res.n_phi + j;
        const int32_t code = first_bubble[ij];
        out[ij] = 0.0;
        if (code == -1) continue;
        size_t b;
        if (code >= 0) b = static_cast<size_t>(code);
        else b = B.n_int + static_cast<size_t>(-code - 2);
        if (b >= B.n_tot())
          throw CalcError("Index " + std::to_string(b) + " out of bounds (max: " +
                          std::to_string(B.n_tot()) + ")");
        const double* xb = B.row(b);
        double d[4] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2], xb[3] - xa[3]};
        const double x[4] = {1.0, res.sin_t[i] * res.cos_p[j], res.sin_t[i] * res.sin_p[j],
                             res.cos_t[i]};
        const double dot = lscalar(d, x);
        if (std::fabs(dot) < 1e-10) continue;
        out[ij] = lscalar(d, d) / (2.0 * dot);
      }
  });
}

BgwStatus bgw_generalized_bulk_flow_collision_status(bgw_generalized_bulk_flow* bulk_flow,
                                                     size_t a_idx, double t,
                                                     const int32_t* first_bubble,
                                                     const double* delta_tab, size_t grid_len,
                                                     int32_t* out, size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    const Resolution& res = need_resolution(h->res);
    const size_t n = res.n_c * res.n_phi;
    if (grid_len != n || out_len != n)
      throw CalcError("Array shape mismatch: Input arrays must match resolution: expected [" +
                      std::to_string(res.n_c) + ", " + std::to_string(res.n_phi) + "]");
    if (a_idx >= h->bubbles.n_int)
      throw CalcError("Index " + std::to_string(a_idx) + " out of bounds (max: " +
                      std::to_string(h->bubbles.n_int) + ")");
    slice_from_raw(first_bubble, n, "first_bubble");
    slice_from_raw(delta_tab, n, "delta_tab");
    slice_from_raw(out, n, "out");
    // compute_collision_status (generalized_bulk_flow.rs:395-411)
    const double delta_ta = t - h->bubbles.row(a_idx)[0];
    for (size_t ij = 0; ij < n; ++ij) {
      if (first_bubble[ij] == -1) out[ij] = 0;
      else if (delta_tab[ij] > 0.0 && delta_ta >= delta_tab[ij]) out[ij] = 1;
      else out[ij] = 2;
    }
  });
}

BgwStatus bgw_generalized_bulk_flow_compute_c_integrand(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t* selected_bubbles,
    size_t n_selected_bubbles, BgwComplex64* out, size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    slice_from_raw(w_arr, n_w, "w_arr");
    const size_t n_sets = h->sc.n_sets;
    // compute_c_integrand (generalized_bulk_flow.rs:875-920)
    if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2");
    bool empty = false;
    std::vector<int32_t> ids = selected_ids(h, selected_bubbles, n_selected_bubbles, &empty);
    const size_t expect = 2 * n_sets * n_w * n_t;
    if (out_len != expect)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(expect));
    slice_from_raw(out, out_len, "out");
    integrand_into(h, w_arr, n_w, has_t_begin, t_begin, t_end, n_t, ids, nullptr,
                   reinterpret_cast<double*>(out), {0, 1});
  });
}

BgwStatus bgw_generalized_bulk_flow_compute_c_tensor_integrand(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t* selected_bubbles,
    size_t n_selected_bubbles, BgwComplex64* out, size_t out_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    slice_from_raw(w_arr, n_w, "w_arr");
    const size_t n_sets = h->sc.n_sets;
    // compute_c_tensor_integrand (new_generalized_bulk_flow.rs:993-1065)
    if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2 for integration");
    bool empty = false;
    std::vector<int32_t> ids = selected_ids(h, selected_bubbles, n_selected_bubbles, &empty);
    const size_t expect = 6 * n_sets * n_w * n_t;
    if (out_len != expect)
      throw CalcError("out_len=" + std::to_string(out_len) + " does not match expected length " +
                      std::to_string(expect));
    slice_from_raw(out, out_len, "out");
    double* o = reinterpret_cast<double*>(out);
    std::fill(o, o + 2 * expect, 0.0);
    const Resolution& R = need_resolution(h->res);
    const int comp[4][2] = {{0, 1}, {2, 4}, {3, -1}, {5, -1}};
    for (int pair = 0; pair < 4; ++pair) {
      const Resolution S = tensor_surface(R, pair);
      integrand_into(h, w_arr, n_w, has_t_begin, t_begin, t_end, n_t, ids, &S, o,
                     {comp[pair][0], comp[pair][1]});
    }
  });
}

BgwStatus bgw_generalized_bulk_flow_compute_spectrum(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, const double* k_hat,
    size_t n_k, bool has_t_begin, double t_begin, double t_end, size_t n_t,
    BgwTimeCutoffConfig time_cutoff, BgwPrecision precision, double* spectrum_out,
    size_t spectrum_len, double* mean_out, size_t mean_len, BgwComplex64* c_out, size_t c_len) {
  if (!valid_precision(precision)) {
    set_last_error("precision is not a BgwPrecision");
    return BgwStatus_InvalidArgument;
  }
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    slice_from_raw(w_arr, n_w, "w_arr");
    slice_from_raw(k_hat, checked_product(n_k, 3, "k_hat"), "k_hat");
    const size_t n_sets = h->sc.n_sets;
    const size_t nq = n_sets * n_w;
    if (spectrum_len != n_k * nq)
      throw CalcError("spectrum_len=" + std::to_string(spectrum_len) +
                      " does not match expected length " + std::to_string(n_k * nq));
    slice_from_raw(spectrum_out, spectrum_len, "spectrum_out");
    if (mean_out && mean_len != nq)
      throw CalcError("mean_len=" + std::to_string(mean_len) + " does not match expected length " +
                      std::to_string(nq));
    if (c_out && c_len != n_k * 2 * nq)
      throw CalcError("c_len=" + std::to_string(c_len) + " does not match expected length " +
                      std::to_string(n_k * 2 * nq));
    if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2 for integration");
    const double tb = has_t_begin ? t_begin : 0.0;
    const Resolution& res = need_resolution(h->res);
    check_unit_vectors(k_hat, n_k);
    std::vector<int32_t> ids;
    for (size_t a = 0; a < h->bubbles.n_int; ++a) ids.push_back(static_cast<int32_t>(a));
    std::vector<double> S(n_k * nq, 0.0), C(n_k * 2 * nq * 2, 0.0);
    if (n_k > 0 && !ids.empty()) {
      check_time_range(tb, t_end);
      Request q;
      q.res = &res;
      q.sc = &h->sc;
      q.cutoff = make_cutoff(time_cutoff);
      q.n_w = n_w;
      q.n_t = n_t;
      q.identity = false;
      for (size_t k = 0; k < n_k; ++k) q.k.push_back({k_hat[3 * k], k_hat[3 * k + 1], k_hat[3 * k + 2]});
      q.omega.assign(w_arr, w_arr + n_w);
      q.t_begin = {tb};
      q.t_end = {t_end};
      q.bubbles_of_real = {ids};
      q.mode = kernel_mode(precision);
      apply_tiles(h->tiles, q);
      Result r = run(built_engine(h), q);
      S = std::move(r.s);
      C = std::move(r.c);
    }
    std::memcpy(spectrum_out, S.data(), S.size() * sizeof(double));
    if (mean_out) {
      for (size_t j = 0; j < nq; ++j) {
        double acc = 0.0;
        for (size_t k = 0; k < n_k; ++k) acc += S[k * nq + j];
        mean_out[j] = n_k ? acc / static_cast<double>(n_k) : 0.0;
      }
    }
    if (c_out) std::memcpy(c_out, C.data(), C.size() * sizeof(double));
  });
}

BgwStatus bgw_generalized_bulk_flow_spectrum_vjp(
    bgw_generalized_bulk_flow* bulk_flow, const double* w_arr, size_t n_w, const double* k_hat,
    size_t n_k, bool has_t_begin, double t_begin, double t_end, size_t n_t,
    BgwTimeCutoffConfig time_cutoff, const BgwComplex64* grad_c, size_t grad_c_len,
    double* grad_coefficients, double* grad_powers, size_t grad_len) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    slice_from_raw(w_arr, n_w, "w_arr");
    slice_from_raw(k_hat, checked_product(n_k, 3, "k_hat"), "k_hat");
    const size_t n_sets = h->sc.n_sets, n_terms = h->sc.n_terms;
    const size_t nq = n_sets * n_w;
    if (grad_c_len != n_k * 2 * nq)
      throw CalcError("grad_c_len=" + std::to_string(grad_c_len) +
                      " does not match expected length " + std::to_string(n_k * 2 * nq));
    slice_from_raw(grad_c, grad_c_len, "grad_c");
    if (grad_len != n_sets * n_terms)
      throw CalcError("grad_len=" + std::to_string(grad_len) + " does not match expected length " +
                      std::to_string(n_sets * n_terms));
    slice_from_raw(grad_coefficients, grad_len, "grad_coefficients");
    slice_from_raw(grad_powers, grad_len, "grad_powers");
    if (n_t < 2) throw CalcError("Invalid resolution: n_t must be >= 2 for integration");
    const double tb = has_t_begin ? t_begin : 0.0;
    const Resolution& res = need_resolution(h->res);
    check_unit_vectors(k_hat, n_k);
    std::vector<double> g(2 * n_sets * n_terms, 0.0);
    if (n_k > 0 && h->bubbles.n_int > 0) {
      check_time_range(tb, t_end);
      Request q;
      q.res = &res;
      q.sc = &h->sc;
      q.cutoff = make_cutoff(time_cutoff);
      q.n_w = n_w;
      q.n_t = n_t;
      q.identity = false;
      for (size_t k = 0; k < n_k; ++k) q.k.push_back({k_hat[3 * k], k_hat[3 * k + 1], k_hat[3 * k + 2]});
      q.omega.assign(w_arr, w_arr + n_w);
      q.t_begin = {tb};
      q.t_end = {t_end};
      std::vector<int32_t> ids(h->bubbles.n_int);
      std::iota(ids.begin(), ids.end(), 0);
      q.bubbles_of_real = {ids};
      g = run_adjoint(built_engine(h), q, reinterpret_cast<const double*>(grad_c));
    }
    std::memcpy(grad_coefficients, g.data(), grad_len * sizeof(double));
    std::memcpy(grad_powers, g.data() + grad_len, grad_len * sizeof(double));
  });
}

BgwStatus bgw_generalized_bulk_flow_set_tiles(bgw_generalized_bulk_flow* bulk_flow,
                                              const BgwTileConfig* tiles) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    BgwTileConfig t{0, 0, 0, 0};
    if (tiles) t = *tiles;
    check_tiles(t);
    h->tiles = t;
  });
}

BgwStatus bgw_generalized_bulk_flow_last_timings(bgw_generalized_bulk_flow* bulk_flow,
                                                 BgwTimings* out) {
  return run_ffi([&] {
    auto* h = bulk_flow_handle(bulk_flow);
    if (!out) throw CalcError(null_error("out"));
    *out = h->engine.timings;
  });
}

// ---- ensemble (filled in below) ----
BgwStatus bgw_ensemble_new(bgw_ensemble** out, const double* bubbles_interior,
                           const size_t* interior_offsets, const double* bubbles_exterior,
                           const size_t* exterior_offsets, size_t n_realisations) {
  if (!out) {
    set_last_error(null_error("out"));
    return BgwStatus_NullPointer;
  }
  *out = nullptr;
  return run_ffi([&] {
    slice_from_raw(interior_offsets, n_realisations + 1, "interior_offsets");
    slice_from_raw(exterior_offsets, n_realisations + 1, "exterior_offsets");
    auto e = std::make_unique<bgw_ensemble>();
    for (size_t r = 0; r < n_realisations; ++r) {
      const size_t i0 = interior_offsets[r], i1 = interior_offsets[r + 1];
      const size_t e0 = exterior_offsets[r], e1 = exterior_offsets[r + 1];
      if (i1 < i0 || e1 < e0) throw CalcError("offsets must be non-decreasing");
      if (i1 > i0) slice_from_raw(bubbles_interior, i1 * 4, "bubbles_interior");
      if (e1 > e0) slice_from_raw(bubbles_exterior, e1 * 4, "bubbles_exterior");
      e->reals.push_back(make_bubbles(bubbles_interior + 4 * i0, i1 - i0,
                                      e1 > e0 ? bubbles_exterior + 4 * e0 : nullptr, e1 - e0));
    }
    build_static(e->engine, e->reals, host_threads(0));
    *out = e.release();
  });
}

void bgw_ensemble_free(bgw_ensemble* ensemble) { delete ensemble; }

BgwStatus bgw_ensemble_set_resolution(bgw_ensemble* ensemble, size_t n_cos_thetax,
                                      size_t n_phix) {
  return run_ffi([&] {
    if (!ensemble) throw CalcError(null_error("ensemble"));
    ensemble->res = make_resolution(n_cos_thetax, n_phix);
  });
}

BgwStatus bgw_ensemble_set_gradient_scaling_params(bgw_ensemble* ensemble,
                                                   const double* coefficients,
                                                   const double* powers, size_t n_sets,
                                                   size_t n_terms, double damping_width) {
  return run_ffi([&] {
    if (!ensemble) throw CalcError(null_error("ensemble"));
    ensemble->sc = make_scaling(coefficients, powers, n_sets, n_terms, damping_width);
  });
}

BgwStatus bgw_ensemble_compute_spectrum(bgw_ensemble* ensemble, const double* w_arr, size_t n_w,
                                        const double* k_hat, size_t n_k, const double* t_begin,
                                        const double* t_end, size_t n_t, BgwPrecision precision,
                                        const BgwTileConfig* tiles, double* spectrum_out,
                                        size_t spectrum_len, double* kmean_out,
                                        size_t kmean_len, double* ens_mean_out,
                                        double* ens_std_out, size_t ens_len) {
  if (!valid_precision(precision)) {
    set_last_error("precision is not a BgwPrecision");
    return BgwStatus_InvalidArgument;
  }
  return run_ffi([&] {
    if (!ensemble) throw CalcError(null_error("ensemble"));
    const size_t n_real = ensemble->reals.size();
    const size_t nq = ensemble->sc.n_sets * n_w;
    if (spectrum_len != n_real * n_k * nq)
      throw CalcError("spectrum_len=" + std::to_string(spectrum_len) +
                      " does not match expected length " + std::to_string(n_real * n_k * nq));
    slice_from_raw(spectrum_out, spectrum_len, "spectrum_out");
    if (kmean_out && kmean_len != n_real * nq) throw CalcError("kmean_len does not match");
    if ((ens_mean_out || ens_std_out) && ens_len != nq) throw CalcError("ens_len does not match");
    BgwTileConfig t = tiles ? *tiles : ensemble->tiles;
    check_tiles(t);
    Request q = ensemble_request(ensemble, w_arr, n_w, k_hat, n_k, t_begin, t_end, n_t, precision);
    EnsResult er;
    if (n_k > 0 && n_real > 0 && nq > 0 && q.mode == 3) {
      // the FP64 matrix form: the single-configuration launches over every
      // realisation at once (tiles per realisation and bubble, so each
      // realisation's spectrum is the one of its own configuration)
      apply_tiles(tiles ? *tiles : ensemble->matrix_tiles, q);
      q.want_c = false;
      q.ensemble_stats = true;
      Result r = run_matrix_ensemble(ensemble->engine, q);
      er.s = std::move(r.s);
      er.kmean = std::move(r.kmean);
      er.mean = std::move(r.mean);
      er.stdev = std::move(r.stdev);
    } else if (n_k > 0 && n_real > 0 && nq > 0) {
      er = run_ensemble(ensemble->engine, q, t, nullptr, true, 1);
    } else {
      er.s.assign(n_real * n_k * nq, 0.0);
      er.kmean.assign(n_real * nq, 0.0);
      er.mean.assign(nq, 0.0);
      er.stdev.assign(nq, 0.0);
    }
    std::memcpy(spectrum_out, er.s.data(), er.s.size() * sizeof(double));
    if (kmean_out) std::memcpy(kmean_out, er.kmean.data(), er.kmean.size() * sizeof(double));
    if (ens_mean_out) std::memcpy(ens_mean_out, er.mean.data(), nq * sizeof(double));
    if (ens_std_out) std::memcpy(ens_std_out, er.stdev.data(), nq * sizeof(double));
  });
}

BgwStatus bgw_ensemble_autotune(bgw_ensemble* ensemble, const double* w_arr, size_t n_w,
                                const double* k_hat, size_t n_k, const double* t_begin,
                                const double* t_end, size_t n_t, BgwPrecision precision,
                                BgwTileConfig* best) {
  if (!valid_precision(precision)) {
    set_last_error("precision is not a BgwPrecision");
    return BgwStatus_InvalidArgument;
  }
  return run_ffi([&] {
    if (!ensemble) throw CalcError(null_error("ensemble"));
    Request q = ensemble_request(ensemble, w_arr, n_w, k_hat, n_k, t_begin, t_end, n_t, precision);
    const size_t n_groups = ensemble->reals.size() * n_k;
    if (n_groups == 0 || n_w == 0 || ensemble->sc.n_sets == 0) {
      if (best) *best = ensemble->tiles;
      return;
    }
    if (q.mode == 3) {
      // The FP64 matrix form: 16-omega tiles per wavefront of the matrix kernel
      // (1, 2 or 4), each timed on every m-th realisation (at most 32).
      const size_t n_real = ensemble->reals.size();
      const size_t m = std::max<size_t>(1, (n_real + 31) / 32);
      Request qs = q;
      for (size_t r = 0; r < n_real; ++r)
        if (r % m != 0) qs.bubbles_of_real[r].clear();
      qs.want_c = false;
      BgwTileConfig bt{0, 0, 0, 0};
      double best_ms = std::numeric_limits<double>::infinity();
      for (size_t ot : {1, 2, 4}) {
        qs.tile_rb = static_cast<int>(ot);
        (void)run(ensemble->engine, qs);
        (void)run(ensemble->engine, qs);
        const double ms = ensemble->engine.timings.total_ms;
        if (ms < best_ms) {
          best_ms = ms;
          bt.block_steps = ot;
        }
      }
      ensemble->matrix_tiles = bt;
      if (best) *best = bt;
      return;
    }
    // Candidates: workgroup size (omegas per lane 1, 2 or 4) x block length;
    // each is timed on an evenly spaced sample of the largest-first units.
    std::vector<size_t> threads;
    for (size_t opl : {1, 2, 4}) {
      const size_t t = ((n_w + opl - 1) / opl + 63) / 64 * 64;
      if (t >= 64 && t <= 256 && (n_w + t - 1) / t <= 4 &&
          std::find(threads.begin(), threads.end(), t) == threads.end())
        threads.push_back(t);
    }
    const std::vector<size_t> blocks = q.mode == 0 ? std::vector<size_t>{16}
                                                   : std::vector<size_t>{8, 16};
    // sample: every m-th unit of the cost order, at most ~4096 units
    std::vector<int32_t> sample;
    {
      Launch Lc;
      Request qq = q;
      qq.item_buffers = false;
      setup(ensemble->engine, qq, Lc);
      std::vector<double> gcost(Lc.n_groups, 0.0);
      for (size_t g = 0; g < Lc.n_groups; ++g)
        for (int32_t it = Lc.group_off[g]; it < Lc.group_off[g + 1]; ++it)
          gcost[g] += Lc.item_cost[it] +
                      1e-4 * static_cast<double>(ensemble->engine.real[Lc.group_r[g]].n_tot);
      std::vector<int32_t> order(Lc.n_groups);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(),
                       [&](int32_t x, int32_t y) { return gcost[x] > gcost[y]; });
      const size_t m = std::max<size_t>(1, order.size() / 4096);
      for (size_t i = 0; i < order.size(); i += m) sample.push_back(order[i]);
    }
    // chunking: 0 = fused persistent kernel; else pipelined chunks of ~N items
    const std::vector<size_t> chunking = {0, 16384, 65536};
    double best_ms = std::numeric_limits<double>::infinity();
    BgwTileConfig win = ensemble->tiles;
    for (size_t ch : chunking)
      for (size_t t : threads)
        for (size_t rb : blocks) {
          BgwTileConfig cand{rb, t, ch, 0};
          (void)run_ensemble(ensemble->engine, q, cand, &sample, false, 1);  // capture + warm-up
          EnsResult r = run_ensemble(ensemble->engine, q, cand, &sample, false, 2);
          if (r.ms < best_ms) {
            best_ms = r.ms;
            win = r.used;
            win.persistent_blocks = 0;
          }
        }
    ensemble->tiles = win;
    ensemble->tuned = true;
    if (best) *best = win;
  });
}

BgwStatus bgw_ensemble_last_timings(bgw_ensemble* ensemble, BgwTimings* out) {
  return run_ffi([&] {
    if (!ensemble) throw CalcError(null_error("ensemble"));
    if (!out) throw CalcError(null_error("out"));
    *out = ensemble->engine.timings;
  });
}

}  // extern "C"
