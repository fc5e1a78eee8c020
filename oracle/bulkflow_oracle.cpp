// SPDX-License-Identifier: MIT
//
// TEST INFRASTRUCTURE ONLY — the CPU oracle for the generalized bulk-flow
// spectrum.  Nothing in the product (gw_fopt_amd/, include/) links, loads or
// calls this file; it is used by tests/, by the smoke check and by the CPU
// baseline leg of bench.py, always as the checker and never as the thing that
// is measured or shipped.
//
// It restates, operation for operation, the reference's Rust implementation
// in rust/bubble-gw/src/many_bubbles/generalized_bulk_flow.rs (all line
// numbers below refer to that file unless another file is named).  Compile
// with -ffp-contract=off: Rust never contracts a*b+c into an FMA, and the
// masks (first colliding bubble, collision status) must come out of the same
// rounded comparisons.  Transcendentals (sin, cos, exp, pow, sqrt) come from
// glibc, which is what Rust's f64::{sin,cos,exp,powf,sqrt} lower to on Linux.
//
// Third-party arithmetic the reference inherits (not under the reference
// tree, so not readable here):
//   * nalgebra 0.34.2 Vector4::dot — restated as the pairwise form
//     (a0*b0 + a2*b2) + (a1*b1 + a3*b3) that nalgebra's small-vector dot uses
//     for dimension 4.  Only mask decisions at exact rounding ties depend on
//     this association.  Parity unpinned (no reference test fixes it).
//   * ndarray 0.17.2 ArrayBase::sum over a contiguous slice — restated as the
//     eight-accumulator unrolled fold (ndarray numeric_util::unrolled_fold).
//     Parity unpinned; effect on C is at the 1e-16 relative level.
//   * ndarray linspace: start + step * i with step = (b - a) / (n - 1).
//   * num-complex 0.4.6: exp(i*y) = (cos y, sin y); Complex*Complex and
//     Complex*f64 in their textbook operation order.
//   * rayon: the time-step fold/reduce order (lines 811-866) is not
//     deterministic in the reference; the oracle sums time steps sequentially
//     in index order, which is what the reference computes with one thread.
//
// Additions beyond the reference (the drop-in's k-hat extension):
//   * rotation of the lattice so that a chosen unit vector k maps to z-hat,
//     applied exactly like LatticeBubbles::transform (lattice_bubbles.rs:
//     144-197): positions and the spatial part of every delta are rotated,
//     delta_squared is kept from the unrotated bubbles;
//   * the spectrum S = 8 pi w^3 (|C+|^2 + |Cx|^2)
//     (src/gw_fopt/bubble_dynamics/utils/data_gw_bulkflow.py:106-111) and its
//     mean over the k-hat list;
//   * forward-mode derivatives of C with respect to the coefficients and
//     powers of the post-collision scaling function (the FP64 reference
//     gradient for the adjoint kernel).

#include <omp.h>

// Test-only switch: a nonzero ORACLE_PLANT compiles a deliberately wrong
// variant (1: sign of the time phase, 2: cos-theta step 2/n instead of
// 2/(n-1), 3: exact ties accepted as first collisions, 4: the k-hat frame
// rotated by the transpose of R, 5: the deltas left unrotated in the k-hat
// frame),
// used by the tests to show that the oracle's checks catch such mistakes.
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI

thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// Spacetime helpers
// ---------------------------------------------------------------------------

// Lorentzian::scalar (spacetime.rs:593-601): u.dual().dot(v), where dual()
// negates the time component of a vector (spacetime.rs:536-548).  The dot
// product is nalgebra's dimension-4 pairwise form (see header).
inline double lscalar(const double* u, const double* v) {
  double a = (-u[0]) * v[0];
  double b = u[1] * v[1];
  double c = u[2] * v[2];
  double d = u[3] * v[3];
  a += c;
  b += d;
  return a + b;
}

// ndarray numeric_util::unrolled_fold for `sum()` of a contiguous f64 slice.
double ndarray_sum(const double* xs, size_t n) {
  double acc = 0.0;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, p4 = 0.0, p5 = 0.0, p6 = 0.0, p7 = 0.0;
  while (n >= 8) {
    p0 = p0 + xs[0];
    p1 = p1 + xs[1];
    p2 = p2 + xs[2];
    p3 = p3 + xs[3];
    p4 = p4 + xs[4];
    p5 = p5 + xs[5];
    p6 = p6 + xs[6];
    p7 = p7 + xs[7];
    xs += 8;
    n -= 8;
  }
  acc = acc + (p0 + p4);
  acc = acc + (p1 + p5);
  acc = acc + (p2 + p6);
  acc = acc + (p3 + p7);
  for (size_t i = 0; i < n && i < 7; ++i) acc = acc + xs[i];
  return acc;
}

// ndarray Array1::linspace(a, b, n)[i].
inline double linspace_at(double a, double b, size_t n, size_t i) {
  double step = n > 1 ? (b - a) / static_cast<double>(n - 1) : 0.0;
  return a + step * static_cast<double>(i);
}

// libm sin / cos as separate calls: the optimiser may fuse std::sin and
// std::cos of one argument into sincos; these out-of-line wrappers keep them
// apart where the reference computes them apart.
__attribute__((noinline)) double sin_separate(double x) { return std::sin(x); }
__attribute__((noinline)) double cos_separate(double x) { return std::cos(x); }

struct Cx {
  double re, im;
};
inline Cx cmul(Cx a, Cx b) {  // num-complex Mul for Complex
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline Cx cscale(Cx a, double s) { return {a.re * s, a.im * s}; }  // Complex * f64
// exp(0 + i y) = from_polar(1, y): cos and sin of one angle together, one
// libm sincos call as the Rust compiler lowers them.
inline Cx cexp_i(double y) {
  double s, c;
  ::sincos(y, &s, &c);
  return {c, s};
}

// ---------------------------------------------------------------------------
// Lattice bubbles (lattice_bubbles.rs:211-308) and their rotation
// ---------------------------------------------------------------------------

struct Lattice {
  size_t n_int = 0, n_ext = 0, n_tot = 0;
  std::vector<double> interior;  // (n_int, 4) rows (t, x, y, z)
  std::vector<double> exterior;  // (n_ext, 4)
  std::vector<double> delta;     // (n_int, n_tot, 4): delta[a][b] = x_b - x_a
  std::vector<double> delta_sq;  // (n_int, n_tot): scalar(delta, delta)
  const double* d(size_t a, size_t b) const { return &delta[(a * n_tot + b) * 4]; }
  double dsq(size_t a, size_t b) const { return delta_sq[a * n_tot + b]; }
  const double* bubble(size_t b) const {
    return b < n_int ? &interior[b * 4] : &exterior[(b - n_int) * 4];
  }
};

std::string index_name(size_t b, size_t n_int) {
  return b < n_int ? "Interior(" + std::to_string(b) + ")"
                   : "Exterior(" + std::to_string(b - n_int) + ")";
}

// LatticeBubbles::new (lattice_bubbles.rs:211-308) with EmptyLattice (no
// containment test), followed by check_bubble_formed_inside_bubble (:71-119).
bool build_lattice(const double* interior, size_t n_int, const double* exterior, size_t n_ext,
                   Lattice& L) {
  L.n_int = n_int;
  L.n_ext = n_ext;
  L.n_tot = n_int + n_ext;
  L.interior.assign(interior, interior + n_int * 4);
  L.exterior.assign(exterior ? exterior : interior, exterior ? exterior + n_ext * 4 : interior);
  L.delta.assign(n_int * L.n_tot * 4, 0.0);
  L.delta_sq.assign(n_int * L.n_tot, 0.0);
  for (size_t a = 0; a < n_int; ++a) {
    const double* xa = &L.interior[a * 4];
    for (size_t b = a; b < n_int; ++b) {  // :282-289, symmetric fill
      const double* xb = &L.interior[b * 4];
      double dab[4] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2], xb[3] - xa[3]};
      double s = lscalar(dab, dab);
      for (int k = 0; k < 4; ++k) {
        L.delta[(a * L.n_tot + b) * 4 + k] = dab[k];
        L.delta[(b * L.n_tot + a) * 4 + k] = -dab[k];
      }
      L.delta_sq[a * L.n_tot + b] = s;
      L.delta_sq[b * L.n_tot + a] = s;
    }
    for (size_t e = 0; e < n_ext; ++e) {  // :291-296
      size_t b = n_int + e;
      const double* xb = &L.exterior[e * 4];
      double dab[4] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2], xb[3] - xa[3]};
      for (int k = 0; k < 4; ++k) L.delta[(a * L.n_tot + b) * 4 + k] = dab[k];
      L.delta_sq[a * L.n_tot + b] = lscalar(dab, dab);
    }
  }
  // check_bubble_formed_inside_bubble (:71-119)
  for (size_t a = 0; a < n_int; ++a)
    for (size_t b = a + 1; b < n_int; ++b)
      if (L.dsq(a, b) < 0.0) {
        g_last_error = "Bubble " + index_name(a, n_int) +
                       " is formed inside bubble " + index_name(b, n_int) +
                       " at initial time (overlapping light cones)";
        return false;
      }
  for (size_t a = 0; a < n_int; ++a)
    for (size_t e = 0; e < n_ext; ++e)
      if (L.dsq(a, n_int + e) < 0.0) {
        g_last_error = "Bubble " + index_name(a, n_int) +
                       " is formed inside bubble " + index_name(n_int + e, n_int) +
                       " at initial time (overlapping light cones)";
        return false;
      }
  return true;
}

// Rotation taking the unit vector k to z-hat: R = R_y(-theta) R_z(-phi) for
// k = (sin theta cos phi, sin theta sin phi, cos theta); its last row is k.
void rotation_to_z(const double* k, double* R) {
  double st = std::sqrt(k[0] * k[0] + k[1] * k[1]);
  double cp = 1.0, sp = 0.0;
  if (st > 0.0) {
    cp = k[0] / st;
    sp = k[1] / st;
  }
  double ct = k[2];
  R[0] = ct * cp; R[1] = ct * sp; R[2] = -st;
  R[3] = -sp;     R[4] = cp;      R[5] = 0.0;
  R[6] = k[0];    R[7] = k[1];    R[8] = k[2];
}

inline void rot3(const double* R, const double* v, double* out) {
  out[0] = (R[0] * v[0] + R[1] * v[1]) + R[2] * v[2];
  out[1] = (R[3] * v[0] + R[4] * v[1]) + R[5] * v[2];
  out[2] = (R[6] * v[0] + R[7] * v[1]) + R[8] * v[2];
}

// LatticeBubbles::transform with a pure rotation (lattice_bubbles.rs:144-197,
// bubbles.rs:47-88): rotate positions and delta spatial parts; delta_squared
// stays as computed from the unrotated bubbles.
void rotate_lattice(Lattice& L, const double* Rk) {
#if ORACLE_PLANT == 4
  // planted mistake: the inverse rotation
  const double R[9] = {Rk[0], Rk[3], Rk[6], Rk[1], Rk[4], Rk[7], Rk[2], Rk[5], Rk[8]};
#else
  const double* R = Rk;
#endif
  double tmp[3];
  for (size_t b = 0; b < L.n_int; ++b) {
    rot3(R, &L.interior[b * 4 + 1], tmp);
    for (int k = 0; k < 3; ++k) L.interior[b * 4 + 1 + k] = tmp[k];
  }
  for (size_t b = 0; b < L.n_ext; ++b) {
    rot3(R, &L.exterior[b * 4 + 1], tmp);
    for (int k = 0; k < 3; ++k) L.exterior[b * 4 + 1 + k] = tmp[k];
  }
#if ORACLE_PLANT == 5
  return;  // planted mistake: the deltas keep the unrotated frame
#endif
  for (size_t i = 0; i < L.n_int * L.n_tot; ++i) {
    rot3(R, &L.delta[i * 4 + 1], tmp);
    for (int k = 0; k < 3; ++k) L.delta[i * 4 + 1 + k] = tmp[k];
  }
}

// ---------------------------------------------------------------------------
// GeneralizedBulkFlow state (generalized_bulk_flow.rs:64-116, 208-270, 315-363)
// ---------------------------------------------------------------------------

struct Flow {
  size_t n_c = 0, n_phi = 0;
  std::vector<double> cos_t, phi;  // grids
  std::vector<double> dirs;        // (n_c, n_phi, 4)
  size_t n_sets = 1, n_terms = 1;
  std::vector<double> coef{1.0}, pows{3.0};  // defaults (:98-99)
  bool has_damping = false;
  double damping = 0.0;
  int n_threads = 1;
};

// set_resolution grid part (:214-240).
bool set_resolution(Flow& F, size_t n_c, size_t n_phi) {
  if (n_c < 2 || n_phi < 2) {
    g_last_error = "Invalid resolution: Angular resolutions must be greater than one";
    return false;
  }
  F.n_c = n_c;
  F.n_phi = n_phi;
  F.cos_t.resize(n_c);
  F.phi.resize(n_phi);
  for (size_t i = 0; i < n_c; ++i) F.cos_t[i] = linspace_at(-1.0, 1.0, n_c, i);
  for (size_t j = 0; j < n_phi; ++j) F.phi[j] = linspace_at(0.0, 2.0 * kPi, n_phi, j);
  F.dirs.resize(n_c * n_phi * 4);
  for (size_t i = 0; i < n_c; ++i) {
    double c = F.cos_t[i];
    double s = std::fabs(std::sqrt(1.0 - c * c));
    for (size_t j = 0; j < n_phi; ++j) {
      // phi.sin() and phi.cos() of one angle in one closure (:228-229): the
      // Rust compiler (LLVM) lowers the pair to a single libm sincos call on
      // GNU targets, whose last bit can differ from separate sin / cos calls.
      double sp, cp;
      ::sincos(F.phi[j], &sp, &cp);
      double* x = &F.dirs[(i * n_phi + j) * 4];
      x[0] = 1.0;
      x[1] = s * cp;
      x[2] = s * sp;
      x[3] = c;
    }
  }
  return true;
}

// check_collision_point (:1003-1018).
inline bool check_collision_point(const double* dba, const double* dca, const double* x) {
  double dba_dot_x = lscalar(dba, x);
  if (dba_dot_x <= 0.0) return false;
  double nba = lscalar(dba, dba);
  double nca = lscalar(dca, dca);
  double cv[4];
  for (int k = 0; k < 4; ++k) cv[k] = dba[k] * nca - dca[k] * nba;
#if ORACLE_PLANT == 3
  return lscalar(cv, x) >= 0.0;  // planted mistake: ties accepted
#else
  return lscalar(cv, x) > 0.0;
#endif
}

// compute_first_colliding_bubble (:127-206).  Returns -1 for None, otherwise
// the total index b (interior b < n_int, exterior b - n_int).
int64_t first_colliding_dir(const Lattice& L, size_t a, const double* x) {
  const double tol = 1e-10;
  int64_t best_idx = -1;
  double best = std::numeric_limits<double>::infinity();
  for (size_t b = 0; b < L.n_tot; ++b) {
    if (b < L.n_int && b == a) continue;
    const double* dba = L.d(a, b);
    double dba_sq = L.dsq(a, b);
    double dot = lscalar(dba, x);
    if (dot <= tol) continue;
    double dtab = dba_sq / (2.0 * dot);
    if (dtab <= 0.0 || dtab >= best) continue;
    bool is_first = true;
    for (size_t c = 0; c < L.n_tot; ++c) {
      if ((c < L.n_int && c == a) || c == b) continue;
      if (!check_collision_point(dba, L.d(a, c), x)) {
        is_first = false;
        break;
      }
    }
    if (is_first) {
      best_idx = static_cast<int64_t>(b);
      best = dtab;
    }
  }
  return best_idx;
}

void first_colliding(const Lattice& L, const Flow& F, size_t a, std::vector<int64_t>& out) {
  out.assign(F.n_c * F.n_phi, -1);
  const long n = static_cast<long>(F.n_c * F.n_phi);
#pragma omp parallel for schedule(dynamic, 16) num_threads(F.n_threads)
  for (long ij = 0; ij < n; ++ij) out[ij] = first_colliding_dir(L, a, &F.dirs[ij * 4]);
}

// compute_delta_tab (:421-471).
void delta_tab(const Lattice& L, const Flow& F, size_t a, const std::vector<int64_t>& first,
               std::vector<double>& out) {
  out.assign(F.n_c * F.n_phi, 0.0);
  for (size_t ij = 0; ij < F.n_c * F.n_phi; ++ij) {
    if (first[ij] < 0) continue;
    size_t b = static_cast<size_t>(first[ij]);
    const double* dba = L.d(a, b);
    double dot = lscalar(dba, &F.dirs[ij * 4]);
    if (std::fabs(dot) < 1e-10) {
      out[ij] = 0.0;
      continue;
    }
    out[ij] = L.dsq(a, b) / (2.0 * dot);
  }
}

enum Status : int32_t { kNever = 0, kAlready = 1, kNotYet = 2 };

// compute_collision_status (:369-414).
inline int32_t collision_status(int64_t first, double dtab, double delta_ta) {
  if (first < 0) return kNever;
  if (dtab > 0.0 && delta_ta >= dtab) return kAlready;
  return kNotYet;
}

// What the scaling function f is replaced by (forward-mode derivatives).
struct FactorMode {
  int kind = 0;     // 0: f itself, 1: d f / d coef[s][k], 2: d f / d pow[s][k]
  size_t term = 0;  // k
};

// compute_b_integral (:481-586) for one cos-theta row; b_plus/b_minus have
// n_sets entries.
void b_integral(const Flow& F, size_t i, const int32_t* status_row, const double* dtab_row,
                double delta_ta, const FactorMode& mode, double* b_plus, double* b_minus) {
  const size_t n_phi = F.n_phi;
  const double cos_t = F.cos_t[i];
  const double sin_sq = 1.0 - cos_t * cos_t;  // powi(2)
  const double dphi = 2.0 * kPi / static_cast<double>(n_phi);
  for (size_t s = 0; s < F.n_sets; ++s) b_plus[s] = b_minus[s] = 0.0;
  if (delta_ta <= 0.0) return;
  std::vector<double> sin2(n_phi), cos2(n_phi), w(n_phi, 1.0), fac(n_phi), cs(n_phi), cc(n_phi);
  // Two separate element-wise passes (azip!, :528-531), i.e. separate sin and
  // cos calls; kept apart here so the compiler cannot merge them into sincos.
  for (size_t j = 0; j < n_phi; ++j) sin2[j] = sin_separate(2.0 * F.phi[j]);
  for (size_t j = 0; j < n_phi; ++j) cos2[j] = cos_separate(2.0 * F.phi[j]);
  w[0] = 0.5;
  w[n_phi - 1] = 0.5;
  const double dta3 = (delta_ta * delta_ta) * delta_ta;  // powi(3)
  for (size_t s = 0; s < F.n_sets; ++s) {
    for (size_t j = 0; j < n_phi; ++j) {
      if (status_row[j] != kAlready) {
        fac[j] = mode.kind == 0 ? 1.0 : 0.0;
        continue;
      }
      double ratio = dtab_row[j] / delta_ta;
      double f = 0.0;
      if (mode.kind == 0) {
        for (size_t k = 0; k < F.n_terms; ++k)
          f += F.coef[s * F.n_terms + k] * std::pow(ratio, F.pows[s * F.n_terms + k]);
      } else if (mode.kind == 1) {
        f = std::pow(ratio, F.pows[s * F.n_terms + mode.term]);
      } else {
        size_t k = mode.term;
        f = F.coef[s * F.n_terms + k] * std::pow(ratio, F.pows[s * F.n_terms + k]) *
            std::log(ratio);
      }
      if (F.has_damping) {
        double damp = std::exp((-delta_ta) * (1.0 - ratio) / F.damping);
        f *= damp;
      }
      fac[j] = f;
    }
    for (size_t j = 0; j < n_phi; ++j) {
      cs[j] = sin2[j] * fac[j] * w[j];
      cc[j] = cos2[j] * fac[j] * w[j];
    }
    double integral_minus = ndarray_sum(cs.data(), n_phi) * dphi * dta3;
    double integral_plus = ndarray_sum(cc.data(), n_phi) * dphi * dta3;
    b_plus[s] += 0.5 * sin_sq * integral_plus;
    b_minus[s] += 0.5 * sin_sq * integral_minus;
  }
}

// compute_a_integral (:588-664).  a_plus/a_minus are (n_sets, n_w).
void a_integral(const Lattice& L, const Flow& F, size_t a, const double* w_arr, size_t n_w,
                double t, const std::vector<int64_t>& first, const std::vector<double>& dtab,
                const FactorMode& mode, Cx* a_plus, Cx* a_minus) {
  const size_t n_c = F.n_c, n_phi = F.n_phi, ns = F.n_sets;
  const double ta = L.interior[a * 4 + 0];
  const double delta_ta = t - ta;
  std::vector<int32_t> status(n_c * n_phi);
  for (size_t ij = 0; ij < n_c * n_phi; ++ij)
    status[ij] = collision_status(first[ij], dtab[ij], delta_ta);
  for (size_t q = 0; q < ns * n_w; ++q) a_plus[q] = a_minus[q] = {0.0, 0.0};
#if ORACLE_PLANT == 2
  const double dcos = 2.0 / static_cast<double>(n_c);  // planted mistake: off-by-one step
#else
  const double dcos = 2.0 / static_cast<double>(n_c - 1);
#endif
  std::vector<double> bp(ns), bm(ns);
  std::vector<Cx> pf(n_w);
  for (size_t i = 0; i < n_c; ++i) {
    b_integral(F, i, &status[i * n_phi], &dtab[i * n_phi], delta_ta, mode, bp.data(), bm.data());
    const double base_im = (-delta_ta) * F.cos_t[i];
    const double weight = (i == 0 || i == n_c - 1) ? 0.5 : 1.0;
    for (size_t q = 0; q < n_w; ++q) {
      Cx ph = cexp_i(w_arr[q] * base_im);  // (w * phase_base).exp(), re part w*0 = 0
      pf[q] = cscale(cscale(ph, dcos), weight);
    }
    for (size_t s = 0; s < ns; ++s) {
      Cx bps = {bp[s], 0.0}, bms = {bm[s], 0.0};
      for (size_t q = 0; q < n_w; ++q) {
        Cx up = cmul(bps, pf[q]);
        Cx um = cmul(bms, pf[q]);
        a_plus[s * n_w + q].re += up.re;
        a_plus[s * n_w + q].im += up.im;
        a_minus[s * n_w + q].re += um.re;
        a_minus[s * n_w + q].im += um.im;
      }
    }
  }
}

// TimeCutoff (time_cutoff.rs:1-43).
struct Cutoff {
  int kind = 0;  // 0 unit, 1 exponential
  double t_cut = 0.0, t_0 = 0.0, inv_t0_sq = 0.0;
  double eval(double t) const {
    if (kind == 0) return 1.0;
    if (t < t_cut) return 1.0;
    double d = t - t_cut;
    double exponent = -(d * d) * inv_t0_sq;
    return std::exp(exponent);
  }
};

Cutoff make_cutoff(int kind, double smax, double ratio_t_cut, double ratio_t_0) {
  Cutoff c;
  c.kind = kind;
  if (kind == 1) {
    double rc = std::isnan(ratio_t_cut) ? 0.999999999 : ratio_t_cut;
    double r0 = std::isnan(ratio_t_0) ? 0.25 : ratio_t_0;
    c.t_cut = rc * smax;
    c.t_0 = r0 * (smax - c.t_cut);
    c.inv_t0_sq = 1.0 / (c.t_0 * c.t_0);
  }
  return c;
}

// compute_c_integral_fixed_bubble (:775-873); out is (2, n_sets, n_w).
// The rayon fold over time steps is summed sequentially in index order.
void c_integral_fixed(const Lattice& L, const Flow& F, size_t a, const double* w_arr, size_t n_w,
                      double t_begin, double t_end, size_t n_t, const Cutoff& cut,
                      const FactorMode& mode, Cx* out) {
  const size_t ns = F.n_sets, nq = ns * n_w;
  std::vector<double> t_arr(n_t);
  for (size_t k = 0; k < n_t; ++k) t_arr[k] = linspace_at(t_begin, t_end, n_t, k);
  const double dt = n_t > 1 ? t_arr[1] - t_arr[0] : 0.0;
  const double z_a = L.interior[a * 4 + 3];
  const double t_nuc = L.interior[a * 4 + 0];
  std::vector<int64_t> first;
  first_colliding(L, F, a, first);
  std::vector<double> dtab;
  delta_tab(L, F, a, first, dtab);
  // per-time-step contributions, then a sequential sum over t
  std::vector<Cx> contrib(n_t * 2 * nq, Cx{0.0, 0.0});
  std::vector<char> active(n_t, 0);
  const long nt = static_cast<long>(n_t);
#pragma omp parallel num_threads(F.n_threads)
  {
    std::vector<Cx> ap(nq), am(nq);
#pragma omp for schedule(dynamic, 1)
    for (long k = 0; k < nt; ++k) {
      const double t = t_arr[k];
      if (!(t_nuc <= t)) continue;
      active[k] = 1;
      const double cutoff_weight = cut.eval(t);
      a_integral(L, F, a, w_arr, n_w, t, first, dtab, mode, ap.data(), am.data());
      Cx* cp_out = &contrib[(k * 2 + 0) * nq];
      Cx* cm_out = &contrib[(k * 2 + 1) * nq];
      const double weight = (k == 0 || k == nt - 1) ? 0.5 : 1.0;
      for (size_t s = 0; s < ns; ++s) {
        for (size_t q = 0; q < n_w; ++q) {
#if ORACLE_PLANT == 1
          Cx phase = cexp_i(-w_arr[q] * (t - z_a));  // planted mistake: phase sign
#else
          Cx phase = cexp_i(w_arr[q] * (t - z_a));
#endif
          Cx vp = cscale(cscale(cmul(ap[s * n_w + q], phase), weight), cutoff_weight);
          Cx vm = cscale(cscale(cmul(am[s * n_w + q], phase), weight), cutoff_weight);
          cp_out[s * n_w + q] = {0.0 + vp.re, 0.0 + vp.im};
          cm_out[s * n_w + q] = {0.0 + vm.re, 0.0 + vm.im};
        }
      }
      const Cx dtc = {dt, 0.0};
      for (size_t q = 0; q < nq; ++q) {
        cp_out[q] = cmul(cp_out[q], dtc);
        cm_out[q] = cmul(cm_out[q], dtc);
      }
    }
  }
  std::vector<Cx> ip(nq, Cx{0.0, 0.0}), im(nq, Cx{0.0, 0.0});
  for (size_t k = 0; k < n_t; ++k) {
    if (!active[k]) continue;
    const Cx* cp_in = &contrib[(k * 2 + 0) * nq];
    const Cx* cm_in = &contrib[(k * 2 + 1) * nq];
    for (size_t q = 0; q < nq; ++q) {
      ip[q].re += cp_in[q].re;
      ip[q].im += cp_in[q].im;
      im[q].re += cm_in[q].re;
      im[q].im += cm_in[q].im;
    }
  }
  const Cx factor = {1.0 / (6.0 * kPi), 0.0};
  for (size_t q = 0; q < nq; ++q) {
    out[q] = cmul(ip[q], factor);
    out[nq + q] = cmul(im[q], factor);
  }
}

// compute_c_integrand_fixed_bubble (:666-767); out is (2, n_sets, n_w, n_t)
// and is added into (the caller zero-fills it).
bool c_integrand_fixed(const Lattice& L, const Flow& F, size_t a, const double* w_arr, size_t n_w,
                       double t_begin, double t_end, size_t n_t, Cx* out) {
  const size_t ns = F.n_sets;
  const double t_nuc = L.interior[a * 4 + 0];
  if (t_nuc >= t_end) return true;  // zeros (:681-684)
  if (t_begin > t_end) {            // (:687-692)
    g_last_error = "Invalid time range: t_begin=" + std::to_string(t_begin) +
                   " > t_end=" + std::to_string(t_end);
    return false;
  }
  std::vector<double> t_arr(n_t);
  for (size_t k = 0; k < n_t; ++k) t_arr[k] = linspace_at(t_begin, t_end, n_t, k);
  const double dt = n_t > 1 ? t_arr[1] - t_arr[0] : 0.0;
  const double z_a = L.interior[a * 4 + 3];
  std::vector<int64_t> first;
  first_colliding(L, F, a, first);
  std::vector<double> dtab;
  delta_tab(L, F, a, first, dtab);
  const Cx factor = {1.0 / (6.0 * kPi), 0.0};
  const long nt = static_cast<long>(n_t);
  FactorMode mode;
#pragma omp parallel num_threads(F.n_threads)
  {
    std::vector<Cx> ap(ns * n_w), am(ns * n_w);
#pragma omp for schedule(dynamic, 1)
    for (long k = 0; k < nt; ++k) {
      const double t = t_arr[k];
      if (t <= t_nuc) continue;  // zeros (:715-719)
      a_integral(L, F, a, w_arr, n_w, t, first, dtab, mode, ap.data(), am.data());
      const double weight = (k == 0 || k == nt - 1) ? 0.5 : 1.0;
      const Cx dtw = {dt * weight, 0.0};
      for (size_t s = 0; s < ns; ++s)
        for (size_t q = 0; q < n_w; ++q) {
          Cx phase = cexp_i(w_arr[q] * (t - z_a));
          Cx vp = cmul(cmul(ap[s * n_w + q], phase), dtw);
          Cx vm = cmul(cmul(am[s * n_w + q], phase), dtw);
          // result[...] = 0 + contribution, then *= factor (:753-764)
          vp = cmul(Cx{0.0 + vp.re, 0.0 + vp.im}, factor);
          vm = cmul(Cx{0.0 + vm.re, 0.0 + vm.im}, factor);
          Cx& op = out[((0 * ns + s) * n_w + q) * n_t + k];
          Cx& om = out[((1 * ns + s) * n_w + q) * n_t + k];
          op.re += vp.re;
          op.im += vp.im;
          om.re += vm.re;
          om.im += vm.im;
        }
    }
  }
  return true;
}

bool validate_selected(const size_t* selected, size_t n_selected, bool has_selected, size_t n_int,
                       std::vector<size_t>& ids) {
  ids.clear();
  if (has_selected) {
    for (size_t q = 0; q < n_selected; ++q) {
      if (selected[q] >= n_int) {
        g_last_error = "Index " + std::to_string(selected[q]) + " out of bounds for max " +
                       std::to_string(n_int);
        return false;
      }
      ids.push_back(selected[q]);
    }
  } else {
    for (size_t a = 0; a < n_int; ++a) ids.push_back(a);
  }
  return true;
}

struct Common {
  Lattice L;
  Flow F;
};

bool setup(Common& c, const double* interior, size_t n_int, const double* exterior, size_t n_ext,
           size_t n_c, size_t n_phi, const double* coef, const double* pows, size_t n_sets,
           size_t n_terms, double damping, int n_threads) {
  if (!build_lattice(interior, n_int, exterior, n_ext, c.L)) return false;
  if (!set_resolution(c.F, n_c, n_phi)) return false;
  c.F.n_sets = n_sets;
  c.F.n_terms = n_terms;
  c.F.coef.assign(coef, coef + n_sets * n_terms);
  c.F.pows.assign(pows, pows + n_sets * n_terms);
  c.F.has_damping = !std::isnan(damping);
  c.F.damping = damping;
  c.F.n_threads = n_threads > 0 ? n_threads : omp_get_max_threads();
  return true;
}

// compute_c_integral (:928-984) with optional k-hat rotation and factor mode.
bool c_integral(Common& c, const double* w_arr, size_t n_w, double t_begin, double t_end,
                size_t n_t, const std::vector<size_t>& ids, const Cutoff& cut,
                const FactorMode& mode, Cx* out) {
  const size_t nq = c.F.n_sets * n_w;
  for (size_t q = 0; q < 2 * nq; ++q) out[q] = {0.0, 0.0};
  if (ids.empty()) return true;  // (:955-957, and no bubbles at all)
  if (t_begin > t_end) {  // raised by the first compute_c_integral_fixed_bubble (:792-797)
    g_last_error = "Invalid time range: t_begin=" + std::to_string(t_begin) +
                   " > t_end=" + std::to_string(t_end);
    return false;
  }
  std::vector<Cx> one(2 * nq);
  for (size_t a : ids) {
    c_integral_fixed(c.L, c.F, a, w_arr, n_w, t_begin, t_end, n_t, cut, mode, one.data());
    for (size_t q = 0; q < 2 * nq; ++q) {
      out[q].re += one[q].re;
      out[q].im += one[q].im;
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Tensor form of the experimental NewGeneralizedBulkFlow
// (new_generalized_bulk_flow.rs); components in the order xx, xy, xz, yy, yz, zz.
// ---------------------------------------------------------------------------

// compute_b_tensor (:592-727) for one cos-theta row; b is (6, n_sets).
void b_tensor(const Flow& F, size_t i, const int32_t* status_row, const double* dtab_row,
              double delta_ta, double* b) {
  const size_t n_phi = F.n_phi, ns = F.n_sets;
  const double cos_t = F.cos_t[i];
  const double sin_t = std::sqrt(std::fmax(1.0 - cos_t * cos_t, 0.0));
  const double sin_sq = sin_t * sin_t;
  const double cos_sq = cos_t * cos_t;
  const double sin_cos = sin_t * cos_t;
  const double dphi = 2.0 * kPi / static_cast<double>(n_phi);
  std::vector<double> basis[6];
  for (auto& v : basis) v.resize(n_phi);
  for (size_t j = 0; j < n_phi; ++j) {
    double sp, cp;
    ::sincos(F.phi[j], &sp, &cp);  // p.cos(), p.sin() in one closure
    basis[0][j] = (cp * cp) * sin_sq;
    basis[1][j] = (cp * sp) * sin_sq;
    basis[2][j] = cp * sin_cos;
    basis[3][j] = (sp * sp) * sin_sq;
    basis[4][j] = sp * sin_cos;
    basis[5][j] = cos_sq;
  }
  std::vector<double> w(n_phi, 1.0), fac(n_phi), tmp(n_phi);
  w[0] = 0.5;
  w[n_phi - 1] = 0.5;
  const double dta3 = delta_ta > 0.0 ? (delta_ta * delta_ta) * delta_ta : 0.0;
  for (size_t s = 0; s < ns; ++s) {
    for (size_t j = 0; j < n_phi; ++j) {
      if (status_row[j] != kAlready) {
        fac[j] = 1.0;
        continue;
      }
      const double ratio = dtab_row[j] / delta_ta;
      double f = 0.0;
      for (size_t k = 0; k < F.n_terms; ++k)
        f += F.coef[s * F.n_terms + k] * std::pow(ratio, F.pows[s * F.n_terms + k]);
      if (F.has_damping) f *= std::exp((-delta_ta) * (1.0 - ratio) / F.damping);
      fac[j] = f;
    }
    for (int c = 0; c < 6; ++c) {
      for (size_t j = 0; j < n_phi; ++j) tmp[j] = basis[c][j] * fac[j] * w[j];
      b[c * ns + s] = ndarray_sum(tmp.data(), n_phi) * dphi * dta3;
    }
  }
}

// compute_a_tensor (:729-830); a is (6, n_sets, n_w).
void a_tensor(const Lattice& L, const Flow& F, size_t a_idx, const double* w_arr, size_t n_w,
              double t, const std::vector<int64_t>& first, const std::vector<double>& dtab, Cx* a) {
  const size_t n_c = F.n_c, n_phi = F.n_phi, ns = F.n_sets;
  const double delta_ta = t - L.interior[a_idx * 4 + 0];
  std::vector<int32_t> status(n_c * n_phi);
  for (size_t ij = 0; ij < n_c * n_phi; ++ij)
    status[ij] = collision_status(first[ij], dtab[ij], delta_ta);
  for (size_t q = 0; q < 6 * ns * n_w; ++q) a[q] = {0.0, 0.0};
  const double dcos = 2.0 / static_cast<double>(n_c - 1);
  std::vector<double> b(6 * ns);
  std::vector<Cx> ph(n_w);
  for (size_t i = 0; i < n_c; ++i) {
    b_tensor(F, i, &status[i * n_phi], &dtab[i * n_phi], delta_ta, b.data());
    const double base_im = (-delta_ta) * F.cos_t[i];
    const double weight = (i == 0 || i == n_c - 1) ? 0.5 : 1.0;
    const double scalar = dcos * weight;
    for (size_t q = 0; q < n_w; ++q) ph[q] = cscale(cexp_i(w_arr[q] * base_im), scalar);
    for (int c = 0; c < 6; ++c)
      for (size_t s = 0; s < ns; ++s) {
        const Cx bv = {b[c * ns + s], 0.0};
        for (size_t q = 0; q < n_w; ++q) {
          const Cx u = cmul(bv, ph[q]);
          Cx& dst = a[(c * ns + s) * n_w + q];
          dst.re += u.re;
          dst.im += u.im;
        }
      }
  }
}

// compute_c_tensor_fixed_bubble (:1067-1203); out is (6, n_sets, n_w).  No
// time cutoff; steps with t <= t_a are skipped; the rayon fold over t is
// summed sequentially in index order.
void c_tensor_fixed(const Lattice& L, const Flow& F, size_t a_idx, const double* w_arr, size_t n_w,
                    double t_begin, double t_end, size_t n_t, Cx* out) {
  const size_t nq = 6 * F.n_sets * n_w;
  std::vector<double> t_arr(n_t);
  for (size_t k = 0; k < n_t; ++k) t_arr[k] = linspace_at(t_begin, t_end, n_t, k);
  const double dt = n_t > 1 ? t_arr[1] - t_arr[0] : 0.0;
  const double z_a = L.interior[a_idx * 4 + 3];
  const double t_nuc = L.interior[a_idx * 4 + 0];
  std::vector<int64_t> first;
  first_colliding(L, F, a_idx, first);
  std::vector<double> dtab;
  delta_tab(L, F, a_idx, first, dtab);
  std::vector<Cx> contrib(n_t * nq, Cx{0.0, 0.0});
  std::vector<char> active(n_t, 0);
  const long nt = static_cast<long>(n_t);
#pragma omp parallel num_threads(F.n_threads)
  {
    std::vector<Cx> at(nq);
#pragma omp for schedule(dynamic, 1)
    for (long k = 0; k < nt; ++k) {
      const double t = t_arr[k];
      if (t <= t_nuc) continue;
      active[k] = 1;
      a_tensor(L, F, a_idx, w_arr, n_w, t, first, dtab, at.data());
      const double weight = (k == 0 || k == nt - 1) ? 0.5 : 1.0;
      const Cx dtc = {dt * weight, 0.0};
      Cx* o = &contrib[k * nq];
      for (size_t cs = 0; cs < 6 * F.n_sets; ++cs)
        for (size_t q = 0; q < n_w; ++q) {
          const Cx factor = cmul(cexp_i(w_arr[q] * (t - z_a)), dtc);
          o[cs * n_w + q] = cmul(at[cs * n_w + q], factor);
        }
    }
  }
  std::vector<Cx> acc(nq, Cx{0.0, 0.0});
  for (size_t k = 0; k < n_t; ++k) {
    if (!active[k]) continue;
    for (size_t q = 0; q < nq; ++q) {
      acc[q].re += contrib[k * nq + q].re;
      acc[q].im += contrib[k * nq + q].im;
    }
  }
  const Cx factor = {1.0 / (6.0 * kPi), 0.0};
  for (size_t q = 0; q < nq; ++q) out[q] = cmul(acc[q], factor);
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface used by the tests through ctypes.  Arrays are row-major;
// complex outputs are interleaved (re, im) doubles.  Every function returns 0
// on success and a nonzero code on error (message: oracle_last_error()).
// ---------------------------------------------------------------------------
extern "C" {

const char* oracle_last_error() { return g_last_error.c_str(); }

// Grids and direction vectors exactly as set_resolution builds them.
int oracle_grids(size_t n_c, size_t n_phi, double* cos_t, double* phi, double* dirs) {
  Flow F;
  if (!set_resolution(F, n_c, n_phi)) return 3;
  std::memcpy(cos_t, F.cos_t.data(), n_c * sizeof(double));
  std::memcpy(phi, F.phi.data(), n_phi * sizeof(double));
  std::memcpy(dirs, F.dirs.data(), n_c * n_phi * 4 * sizeof(double));
  return 0;
}

int oracle_rotation_to_z(const double* k, double* R) {
  rotation_to_z(k, R);
  return 0;
}

// Deltas of LatticeBubbles::new (validation included).
int oracle_lattice(const double* interior, size_t n_int, const double* exterior, size_t n_ext,
                   double* delta, double* delta_sq) {
  Lattice L;
  if (!build_lattice(interior, n_int, exterior, n_ext, L)) return 3;
  std::memcpy(delta, L.delta.data(), L.delta.size() * sizeof(double));
  std::memcpy(delta_sq, L.delta_sq.data(), L.delta_sq.size() * sizeof(double));
  return 0;
}

// First colliding bubble (Python int32 encoding: -1 None, i interior,
// -(j+2) exterior j), delta_tab, for bubble a, in the frame where k_hat -> z
// (k_hat == NULL: no rotation).
int oracle_first_colliding(const double* interior, size_t n_int, const double* exterior,
                           size_t n_ext, const double* k_hat, size_t n_c, size_t n_phi,
                           size_t a, int n_threads, int32_t* first_out, double* dtab_out) {
  Common c;
  double one = 1.0, three = 3.0;
  if (!setup(c, interior, n_int, exterior, n_ext, n_c, n_phi, &one, &three, 1, 1, NAN, n_threads))
    return 3;
  if (a >= n_int) {
    g_last_error = "Index " + std::to_string(a) + " out of bounds for max " + std::to_string(n_int);
    return 3;
  }
  if (k_hat) {
    double R[9];
    rotation_to_z(k_hat, R);
    rotate_lattice(c.L, R);
  }
  std::vector<int64_t> first;
  first_colliding(c.L, c.F, a, first);
  std::vector<double> dtab;
  delta_tab(c.L, c.F, a, first, dtab);
  for (size_t ij = 0; ij < first.size(); ++ij) {
    int64_t b = first[ij];
    first_out[ij] = b < 0 ? -1
                          : (static_cast<size_t>(b) < n_int ? static_cast<int32_t>(b)
                                                            : -static_cast<int32_t>(b - n_int) - 2);
    if (dtab_out) dtab_out[ij] = dtab[ij];
  }
  return 0;
}

// compute_collision_status for a given (first, delta_tab) grid.
int oracle_collision_status(double t_a, double t, size_t n, const int32_t* first,
                            const double* dtab, int32_t* out) {
  for (size_t q = 0; q < n; ++q)
    out[q] = collision_status(first[q] == -1 ? -1 : 0, dtab[q], t - t_a);
  return 0;
}

// compute_c_integral (optionally in the k_hat frame).  out: (2, n_sets, n_w)
// complex.  mode_kind/mode_term select forward-mode derivatives (0: value).
int oracle_c_integral(const double* interior, size_t n_int, const double* exterior, size_t n_ext,
                      const double* k_hat, size_t n_c, size_t n_phi, const double* coef,
                      const double* pows, size_t n_sets, size_t n_terms, double damping,
                      const double* w_arr, size_t n_w, int has_t_begin, double t_begin,
                      double t_end, size_t n_t, const size_t* selected, size_t n_selected,
                      int has_selected, int cutoff_kind, double smax, double ratio_t_cut,
                      double ratio_t_0, int mode_kind, size_t mode_term, int n_threads,
                      double* out) {
  Common c;
  if (!setup(c, interior, n_int, exterior, n_ext, n_c, n_phi, coef, pows, n_sets, n_terms, damping,
             n_threads))
    return 3;
  if (n_t < 2) {
    g_last_error = "Invalid resolution: n_t must be >= 2 for integration";
    return 3;
  }
  std::vector<size_t> ids;
  if (!validate_selected(selected, n_selected, has_selected != 0, n_int, ids)) return 3;
  if (k_hat) {
    double R[9];
    rotation_to_z(k_hat, R);
    rotate_lattice(c.L, R);
  }
  FactorMode mode;
  mode.kind = mode_kind;
  mode.term = mode_term;
  Cutoff cut = make_cutoff(cutoff_kind, smax, ratio_t_cut, ratio_t_0);
  return c_integral(c, w_arr, n_w, has_t_begin ? t_begin : 0.0, t_end, n_t, ids, cut, mode,
                    reinterpret_cast<Cx*>(out))
             ? 0
             : 3;
}

// NewGeneralizedBulkFlow::compute_c_tensor (new_generalized_bulk_flow.rs:1205-1275).
// out: (6, n_sets, n_w) complex, components xx, xy, xz, yy, yz, zz.
int oracle_c_tensor(const double* interior, size_t n_int, const double* exterior, size_t n_ext,
                    size_t n_c, size_t n_phi, const double* coef, const double* pows,
                    size_t n_sets, size_t n_terms, double damping, const double* w_arr,
                    size_t n_w, int has_t_begin, double t_begin, double t_end, size_t n_t,
                    const size_t* selected, size_t n_selected, int has_selected, int n_threads,
                    double* out) {
  Common c;
  if (!setup(c, interior, n_int, exterior, n_ext, n_c, n_phi, coef, pows, n_sets, n_terms, damping,
             n_threads))
    return 3;
  if (n_t < 2) {
    g_last_error = "Invalid resolution: n_t must be >= 2 for integration";
    return 3;
  }
  std::vector<size_t> ids;
  if (!validate_selected(selected, n_selected, has_selected != 0, n_int, ids)) return 3;
  const size_t nq = 6 * n_sets * n_w;
  Cx* res = reinterpret_cast<Cx*>(out);
  for (size_t q = 0; q < nq; ++q) res[q] = {0.0, 0.0};
  const double tb = has_t_begin ? t_begin : 0.0;
  if (!ids.empty() && tb > t_end) {
    g_last_error = "Invalid time range: t_begin=" + std::to_string(tb) +
                   " > t_end=" + std::to_string(t_end);
    return 3;
  }
  std::vector<Cx> one(nq);
  for (size_t a : ids) {
    c_tensor_fixed(c.L, c.F, a, w_arr, n_w, tb, t_end, n_t, one.data());
    for (size_t q = 0; q < nq; ++q) {
      res[q].re += one[q].re;
      res[q].im += one[q].im;
    }
  }
  return 0;
}

// compute_c_integrand (:875-920).  out: (2, n_sets, n_w, n_t) complex.
int oracle_c_integrand(const double* interior, size_t n_int, const double* exterior, size_t n_ext,
                       size_t n_c, size_t n_phi, const double* coef, const double* pows,
                       size_t n_sets, size_t n_terms, double damping, const double* w_arr,
                       size_t n_w, int has_t_begin, double t_begin, double t_end, size_t n_t,
                       const size_t* selected, size_t n_selected, int has_selected, int n_threads,
                       double* out) {
  Common c;
  if (!setup(c, interior, n_int, exterior, n_ext, n_c, n_phi, coef, pows, n_sets, n_terms, damping,
             n_threads))
    return 3;
  if (n_t < 2) {
    g_last_error = "Invalid resolution: n_t must be >= 2";
    return 3;
  }
  std::vector<size_t> ids;
  if (!validate_selected(selected, n_selected, has_selected != 0, n_int, ids)) return 3;
  double tb = has_t_begin ? t_begin : 0.0;
  Cx* o = reinterpret_cast<Cx*>(out);
  std::fill(o, o + 2 * n_sets * n_w * n_t, Cx{0.0, 0.0});
  for (size_t a : ids)
    if (!c_integrand_fixed(c.L, c.F, a, w_arr, n_w, tb, t_end, n_t, o)) return 3;
  return 0;
}

// Spectrum over a list of k-hat directions: S[k][s][w] = 8 pi w^3 (|C+|^2 +
// |Cx|^2) with C computed in the frame where k-hat -> z.  c_out (optional):
// (n_k, 2, n_sets, n_w) complex.
int oracle_spectrum(const double* interior, size_t n_int, const double* exterior, size_t n_ext,
                    const double* k_hat, size_t n_k, size_t n_c, size_t n_phi,
                    const double* coef, const double* pows, size_t n_sets, size_t n_terms,
                    double damping, const double* w_arr, size_t n_w, int has_t_begin,
                    double t_begin, double t_end, size_t n_t, int cutoff_kind, double smax,
                    double ratio_t_cut, double ratio_t_0, int n_threads, double* s_out,
                    double* c_out) {
  const size_t nq = n_sets * n_w;
  std::vector<double> cbuf(2 * nq * 2);
  for (size_t kk = 0; kk < n_k; ++kk) {
    int rc = oracle_c_integral(interior, n_int, exterior, n_ext, &k_hat[kk * 3], n_c, n_phi, coef,
                               pows, n_sets, n_terms, damping, w_arr, n_w, has_t_begin, t_begin,
                               t_end, n_t, nullptr, 0, 0, cutoff_kind, smax, ratio_t_cut,
                               ratio_t_0, 0, 0, n_threads, cbuf.data());
    if (rc) return rc;
    const Cx* C = reinterpret_cast<const Cx*>(cbuf.data());
    for (size_t s = 0; s < n_sets; ++s)
      for (size_t q = 0; q < n_w; ++q) {
        Cx cp = C[s * n_w + q], cm = C[nq + s * n_w + q];
        double ap = std::hypot(cp.re, cp.im), am = std::hypot(cm.re, cm.im);
        double w = w_arr[q];
        s_out[(kk * n_sets + s) * n_w + q] = 8.0 * kPi * std::pow(w, 3.0) * (ap * ap + am * am);
      }
    if (c_out) std::memcpy(&c_out[kk * 2 * nq * 2], cbuf.data(), cbuf.size() * sizeof(double));
  }
  return 0;
}

}  // extern "C"
