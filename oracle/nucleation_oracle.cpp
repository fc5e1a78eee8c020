// SPDX-License-Identifier: MIT
//
// TEST INFRASTRUCTURE ONLY — the CPU oracle for bubble nucleation.  Nothing in
// the product (gw_fopt_amd/, include/) links, loads or calls this file; tests/
// and the CPU baseline leg of bench.py use it as the checker.
//
// It restates, operation for operation, the reference's nucleation:
//   rust/bubble-gw/src/many_bubbles/bubbles_nucleation/fixed_rate_nucleation.rs
//     (update_first_collision_times :113-129, volume_remaining :131-142,
//      nucleate_fixed_probability_time_stepping :188-275,
//      the site loop of nucleate_fixed_probability_distribution :277-356),
//   rust/bubble-gw/src/many_bubbles/bubbles_nucleation/spontaneous_nucleation.rs
//     (sample_interior :38-87),
//   rust/bubble-gw/src/many_bubbles/lattice/parallelepiped_lattice.rs
//     (volume :144-147, contains :161-186, generate_bubbles_exterior :190-298,
//      sample_points :302-330),
//   rust/bubble-gw/src/many_bubbles/lattice/spherical_lattice.rs
//     (volume :41-43, contains :54-60, generate_bubbles_exterior :64-134,
//      sample_points :139-176),
// with one deliberate difference: the random stream.  The reference draws
// from rand's StdRng (ChaCha12 seeded through seed_from_u64), which cannot be
// run or pinned here (no Rust toolchain, no reference vectors).  The oracle and
// the GPU kernels both draw from Philox4x64-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), counter-based so that
// every draw is addressable: key = (seed, stream id), counter =
// (draw group, purpose, 0, 0).  A uniform double is (x >> 11) * 2^-53, the
// conversion rand's StandardUniform uses for f64.  Given the same draws the
// oracle makes every decision the reference makes, in the same order.
//
// Compiled with -ffp-contract=off (Rust never fuses a*b+c).  Transcendentals
// (exp, sqrt) come from glibc, as Rust's f64::{exp,sqrt} do on Linux.
//
// Third-party arithmetic the reference inherits (not readable here; parity
// unpinned, effect limited to rounding at region boundaries): nalgebra's 3x3
// determinant and inverse (cofactor forms, restated below), Vector3 dot and
// norm_squared (left-to-right sums of products).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI
constexpr double kEps = 2.220446049250313e-16;  // f64::EPSILON
thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// Philox4x64-10
// ---------------------------------------------------------------------------
struct U4 {
  uint64_t v[4];
};

inline void mulhilo(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  const __uint128_t p = static_cast<__uint128_t>(a) * b;
  *hi = static_cast<uint64_t>(p >> 64);
  *lo = static_cast<uint64_t>(p);
}

U4 philox(U4 x, uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo(M0, x.v[0], &hi0, &lo0);
    mulhilo(M1, x.v[2], &hi1, &lo1);
    U4 y;
    y.v[0] = hi1 ^ x.v[1] ^ k0;
    y.v[1] = lo1;
    y.v[2] = hi0 ^ x.v[3] ^ k1;
    y.v[3] = lo0;
    x = y;
  }
  return x;
}

inline double uniform(uint64_t x) { return static_cast<double>(x >> 11) * 0x1.0p-53; }

// Purposes (second counter word).
constexpr uint64_t kStreamPoints = 0;  // sample points: attempt i -> 3 draws
constexpr uint64_t kStreamSteps = 1;   // time-stepping acceptance draws, 4 per group
constexpr uint64_t kStreamSpont = 2;   // spontaneous candidates: attempt i -> 3 draws

// ---------------------------------------------------------------------------
// Lattices
// ---------------------------------------------------------------------------
struct Lattice {
  int kind = 0;  // 0 parallelepiped (also the Cartesian box), 1 sphere
  double o[3] = {0, 0, 0};
  double e[3][3] = {{0}};  // basis vectors e1, e2, e3
  double r = 0;
  double volume = 0;
  bool inv_ok = false;
  double inv[3][3] = {{0}};
};

// Matrix3::from_columns(basis): m(row, col) = e[col][row].
inline double mat(const Lattice& L, int row, int col) { return L.e[col][row]; }

// nalgebra's 3x3 determinant (cofactor expansion along the first row).
double det3(const Lattice& L) {
  const double m11 = mat(L, 0, 0), m12 = mat(L, 0, 1), m13 = mat(L, 0, 2);
  const double m21 = mat(L, 1, 0), m22 = mat(L, 1, 1), m23 = mat(L, 1, 2);
  const double m31 = mat(L, 2, 0), m32 = mat(L, 2, 1), m33 = mat(L, 2, 2);
  const double minor_m12_m23 = m22 * m33 - m32 * m23;
  const double minor_m11_m23 = m21 * m33 - m31 * m23;
  const double minor_m11_m22 = m21 * m32 - m31 * m22;
  return m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
}

// nalgebra's 3x3 try_inverse (adjugate over the determinant).
bool inverse3(const Lattice& L, double out[3][3]) {
  auto m = [&](int i, int j) { return mat(L, i, j); };
  const double minor_m12_m23 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
  const double minor_m11_m23 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
  const double minor_m11_m22 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  const double det =
      m(0, 0) * minor_m12_m23 - m(0, 1) * minor_m11_m23 + m(0, 2) * minor_m11_m22;
  if (det == 0.0) return false;
  out[0][0] = minor_m12_m23 / det;
  out[0][1] = (m(0, 2) * m(2, 1) - m(2, 2) * m(0, 1)) / det;
  out[0][2] = (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2)) / det;
  out[1][0] = -minor_m11_m23 / det;
  out[1][1] = (m(0, 0) * m(2, 2) - m(2, 0) * m(0, 2)) / det;
  out[1][2] = (m(0, 2) * m(1, 0) - m(1, 2) * m(0, 0)) / det;
  out[2][0] = minor_m11_m22 / det;
  out[2][1] = (m(0, 1) * m(2, 0) - m(2, 1) * m(0, 0)) / det;
  out[2][2] = (m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) / det;
  return true;
}

Lattice make_lattice(const double* params, int kind) {
  Lattice L;
  L.kind = kind;
  if (kind == 0) {
    for (int i = 0; i < 3; ++i) L.o[i] = params[i];
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) L.e[j][i] = params[3 + 3 * j + i];
    L.volume = std::fabs(det3(L));  // parallelepiped_lattice.rs:144-147
    L.inv_ok = inverse3(L, L.inv);
  } else {
    for (int i = 0; i < 3; ++i) L.o[i] = params[i];
    L.r = params[3];
    L.volume = (4.0 / 3.0) * kPi * (L.r * L.r * L.r);  // powi(3)
  }
  return L;
}

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// contains (parallelepiped_lattice.rs:161-186, spherical_lattice.rs:54-60)
bool contains(const Lattice& L, const double* p) {
  if (L.kind == 0) {
    if (!L.inv_ok) return false;
    const double d[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
    for (int i = 0; i < 3; ++i) {
      const double u = L.inv[i][0] * d[0] + L.inv[i][1] * d[1] + L.inv[i][2] * d[2];
      if (!(u >= 0.0 && u <= 1.0)) return false;
    }
    return true;
  }
  const double v[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
  return dot3(v, v) <= L.r * L.r;
}

// One candidate of the lattice's sample_points from three uniforms; false if
// the sphere rejects it (parallelepiped_lattice.rs:302-330,
// spherical_lattice.rs:139-176).
bool sample_point(const Lattice& L, const U4& x, double* p) {
  const double u = uniform(x.v[0]), v = uniform(x.v[1]), w = uniform(x.v[2]);
  if (L.kind == 0) {
    for (int i = 0; i < 3; ++i) p[i] = L.o[i] + u * L.e[0][i] + v * L.e[1][i] + w * L.e[2][i];
    return true;
  }
  const double s[3] = {u * 2.0 - 1.0, v * 2.0 - 1.0, w * 2.0 - 1.0};
  if (dot3(s, s) > 1.0) return false;
  for (int i = 0; i < 3; ++i) p[i] = L.o[i] + L.r * s[i];
  return true;
}

// Number of exterior images per interior bubble.
int images_per_bubble(const Lattice& L, int bc) {
  if (bc == 2) return 0;                      // None
  if (L.kind == 0) return 6;                  // periodic or reflection: six faces
  return bc == 1 ? 2 : 0;                     // sphere: reflection only
}

// generate_bubbles_exterior for one interior event (t, x, y, z).
void images(const Lattice& L, int bc, const double* ev, double* out) {
  const double t = ev[0];
  const double p[3] = {ev[1], ev[2], ev[3]};
  if (bc == 2) return;
  if (L.kind == 0 && bc == 0) {  // parallelepiped_lattice.rs:197-216
    int k = 0;
    for (int j = 0; j < 3; ++j)
      for (int sgn = 0; sgn < 2; ++sgn) {
        double* o = out + 4 * k++;
        o[0] = t;
        for (int i = 0; i < 3; ++i) o[1 + i] = p[i] + (sgn ? -L.e[j][i] : L.e[j][i]);
      }
    return;
  }
  if (L.kind == 0) {  // reflection across the six faces (parallelepiped_lattice.rs:217-295)
    const double axis[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    int k = 0;
    for (int j = 0; j < 3; ++j) {
      const double nrm = std::sqrt(dot3(L.e[j], L.e[j]));
      double n[3];
      for (int i = 0; i < 3; ++i) n[i] = nrm > kEps ? L.e[j][i] / nrm : axis[j][i];
      double q[3];  // the far face passes through origin + e_j
      for (int i = 0; i < 3; ++i) q[i] = L.o[i] + L.e[j][i];
      {  // -e_j face through the origin, outward normal -n
        const double nn[3] = {-n[0], -n[1], -n[2]};
        const double dv[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
        const double d = dot3(dv, nn);
        double* o = out + 4 * k++;
        o[0] = t;
        for (int i = 0; i < 3; ++i) o[1 + i] = p[i] - 2.0 * d * nn[i];
      }
      {  // +e_j face through origin + e_j, outward normal +n
        const double dv[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
        const double d = dot3(dv, n);
        double* o = out + 4 * k++;
        o[0] = t;
        for (int i = 0; i < 3; ++i) o[1 + i] = p[i] - 2.0 * d * n[i];
      }
    }
    return;
  }
  if (bc == 1) {  // sphere, reflection across the two tangent planes (spherical_lattice.rs:72-131)
    const double v[3] = {p[0] - L.o[0], p[1] - L.o[1], p[2] - L.o[2]};
    const double d = std::sqrt(dot3(v, v));
    if (d < kEps) {
      const double a[4] = {t, L.o[0] + L.r, L.o[1], L.o[2]};
      const double b[4] = {t, L.o[0] - L.r, L.o[1], L.o[2]};
      std::memcpy(out, a, sizeof a);
      std::memcpy(out + 4, b, sizeof b);
      return;
    }
    const double u[3] = {v[0] / d, v[1] / d, v[2] / d};
    for (int side = 0; side < 2; ++side) {
      double q[3], n[3];
      for (int i = 0; i < 3; ++i) {
        q[i] = side == 0 ? L.o[i] + L.r * u[i] : L.o[i] - L.r * u[i];
        n[i] = side == 0 ? u[i] : -u[i];
      }
      const double pq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
      const double dd = pq[0] * n[0] + pq[1] * n[1] + pq[2] * n[2];
      double* o = out + 4 * side;
      o[0] = t;
      for (int i = 0; i < 3; ++i) o[1 + i] = p[i] - 2.0 * dd * n[i];
    }
  }
}

// update_first_collision_times (fixed_rate_nucleation.rs:113-129)
void update_fct(const std::vector<double>& pts, std::vector<double>& fct, const double* b) {
  const size_t n = fct.size();
  for (size_t i = 0; i < n; ++i) {
    const double dx = pts[3 * i] - b[1];
    const double dy = pts[3 * i + 1] - b[2];
    const double dz = pts[3 * i + 2] - b[3];
    const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double tc = b[0] + dist;
    if (tc < fct[i]) fct[i] = tc;
  }
}

// sample_points(n_points) of the realisation with key (seed, stream).
bool sample_points(const Lattice& L, uint64_t seed, uint64_t stream, size_t n_points,
                   std::vector<double>& pts) {
  pts.assign(3 * n_points, 0.0);
  const size_t max_attempts = n_points * (L.kind == 0 ? 1000 : 10000);
  size_t got = 0;
  for (size_t a = 0; a < max_attempts && got < n_points; ++a) {
    const U4 x = philox(U4{{a, kStreamPoints, 0, 0}}, seed, stream);
    double p[3];
    if (sample_point(L, x, p)) {
      std::memcpy(&pts[3 * got], p, sizeof p);
      ++got;
    }
  }
  return got == n_points;
}

}  // namespace

extern "C" {

const char* oracle_nuc_last_error(void) { return g_last_error.c_str(); }

// Philox4x64-10 block at counter ctr with key (k0, k1) -> out[4].
void oracle_philox4x64(const uint64_t* ctr, uint64_t k0, uint64_t k1, uint64_t* out) {
  const U4 y = philox(U4{{ctr[0], ctr[1], ctr[2], ctr[3]}}, k0, k1);
  std::memcpy(out, y.v, sizeof y.v);
}

// Lattice volume (params: parallelepiped origin[3] + basis rows[9]; sphere center[3] + radius).
double oracle_lattice_volume(const double* params, int kind) { return make_lattice(params, kind).volume; }

int oracle_images_per_bubble(int kind, int bc) {
  Lattice L;
  L.kind = kind;
  return images_per_bubble(L, bc);
}

// Exterior images of n interior events, in generate_bubbles_exterior order.
void oracle_lattice_images(const double* params, int kind, int bc, const double* interior, size_t n,
                           double* out) {
  const Lattice L = make_lattice(params, kind);
  const int ni = images_per_bubble(L, bc);
  for (size_t b = 0; b < n; ++b) images(L, bc, interior + 4 * b, out + 4 * ni * b);
}

void oracle_lattice_contains(const double* params, int kind, const double* pts, size_t n, int32_t* out) {
  const Lattice L = make_lattice(params, kind);
  for (size_t i = 0; i < n; ++i) out[i] = contains(L, pts + 3 * i) ? 1 : 0;
}

// The sample points of one realisation (n_points x 3); returns 0 on success.
int oracle_sample_points(const double* params, int kind, uint64_t seed, uint64_t stream, size_t n_points,
                         double* out) {
  const Lattice L = make_lattice(params, kind);
  std::vector<double> pts;
  if (!sample_points(L, seed, stream, n_points, pts)) {
    g_last_error = "Failed to sample points";
    return 1;
  }
  std::memcpy(out, pts.data(), pts.size() * sizeof(double));
  return 0;
}

// Fixed-rate nucleation of one realisation.
//   method 0: fixed-probability time stepping (fixed_rate_nucleation.rs:188-275)
//   method 1: fixed-probability distribution: the site loop over the given
//             times (fixed_rate_nucleation.rs:277-356)
// Writes up to cap bubbles: interior (cap, 4), exterior (cap * n_img, 4),
// the nucleation times and false-vacuum volumes at each acceptance; *n_out =
// bubbles accepted.  Returns 0, or 1 with a message (sampling failure,
// capacity exceeded).
int oracle_fixed_rate(const double* params, int kind, int bc, int method, double beta, double gamma0,
                      double t0, double d_p0, uint64_t seed, uint64_t stream, size_t n_points,
                      size_t max_time_iterations, double cutoff_fraction, const double* times,
                      size_t n_times, size_t cap, double* interior, double* exterior, double* t_hist,
                      double* v_hist, size_t* n_out) {
  const Lattice L = make_lattice(params, kind);
  const int n_img = images_per_bubble(L, bc);
  const double volume_lattice = L.volume;
  const double volume_cutoff = cutoff_fraction * volume_lattice;
  std::vector<double> pts;
  std::vector<double> fct;
  *n_out = 0;
  if (n_points > 0) {
    if (!sample_points(L, seed, stream, n_points, pts)) {
      g_last_error = "Failed to sample points";
      return 1;
    }
    fct.assign(n_points, std::numeric_limits<double>::infinity());
  }
  auto volume_remaining = [&](double t) {
    if (n_points == 0) return volume_lattice;
    size_t count_valid = 0;
    for (double tc : fct)
      if (tc > t) ++count_valid;
    const double fraction = static_cast<double>(count_valid) / static_cast<double>(n_points);
    return std::fmax(fraction * volume_lattice, 0.0);  // f64::max: NaN-ignoring
  };
  size_t nb = 0;
  std::vector<double> img(4 * 6);
  auto accept = [&](double t, double vrem) -> bool {
    // candidate: the first sample point still in the false vacuum at t
    size_t site = n_points;
    for (size_t i = 0; i < n_points; ++i)
      if (fct[i] > t) {
        site = i;
        break;
      }
    if (site == n_points) return true;
    if (nb >= cap) {
      g_last_error = "bubble capacity exceeded";
      return false;
    }
    const double b[4] = {t, pts[3 * site], pts[3 * site + 1], pts[3 * site + 2]};
    std::memcpy(interior + 4 * nb, b, sizeof b);
    t_hist[nb] = t;
    v_hist[nb] = vrem;
    images(L, bc, b, img.data());
    update_fct(pts, fct, b);
    for (int k = 0; k < n_img; ++k) update_fct(pts, fct, &img[4 * k]);
    std::memcpy(exterior + 4 * n_img * nb, img.data(), sizeof(double) * 4 * n_img);
    ++nb;
    return true;
  };
  if (method == 0) {
    double t = t0;
    for (size_t it = 0; it < max_time_iterations; ++it) {
      const double volume_remaining_t = volume_remaining(t);
      if (volume_remaining_t < volume_cutoff) break;
      const double exponent = beta * (t - t0);
      const double gamma_t = gamma0 * std::exp(exponent);
      const U4 xr = philox(U4{{it >> 2, kStreamSteps, 0, 0}}, seed, stream);
      const double x = uniform(xr.v[it & 3]);
      if (x <= d_p0) {
        if (!accept(t, volume_remaining_t)) {
          *n_out = nb;
          return 1;
        }
      }
      const double dt = (d_p0 / gamma_t) * (volume_lattice / volume_remaining_t);
      t += dt;
    }
  } else {
    for (size_t k = 0; k < n_times; ++k) {
      const double t = times[k];
      const double volume_remaining_t = volume_remaining(t);
      if (volume_remaining_t < volume_cutoff) break;
      if (!accept(t, volume_remaining_t)) {
        *n_out = nb;
        return 1;
      }
    }
    // the reference returns the first n times of the list (:347-348)
    for (size_t i = 0; i < nb; ++i) t_hist[i] = times[i];
  }
  *n_out = nb;
  return 0;
}

// Spontaneous nucleation of one realisation (spontaneous_nucleation.rs:38-87):
// candidates (t0, p) from the lattice's sample_points, kept if inside the
// lattice and outside the light cone of every accepted bubble.  Writes
// interior (n_bubbles, 4) and exterior (n_bubbles * n_img, 4).
int oracle_spontaneous(const double* params, int kind, int bc, size_t n_bubbles, double t0, uint64_t seed,
                       uint64_t stream, double* interior, double* exterior) {
  const Lattice L = make_lattice(params, kind);
  const int n_img = images_per_bubble(L, bc);
  std::vector<double> acc;
  const size_t max_attempts = n_bubbles * 10000;
  uint64_t draw = 0;  // candidate counter of the sample_points stream
  for (size_t attempt = 0; attempt < max_attempts; ++attempt) {
    if (acc.size() / 4 >= n_bubbles) break;
    double p[3];
    // sample_points(1): rejection inside the sampler for the sphere
    const size_t inner_max = L.kind == 0 ? 1000 : 10000;
    bool ok = false;
    for (size_t a = 0; a < inner_max; ++a) {
      const U4 x = philox(U4{{draw++, kStreamSpont, 0, 0}}, seed, stream);
      if (sample_point(L, x, p)) {
        ok = true;
        break;
      }
    }
    if (!ok) {
      g_last_error = "Failed to sample points";
      return 1;
    }
    const double cand[4] = {t0, p[0], p[1], p[2]};
    if (!contains(L, p)) continue;
    bool conflict = false;
    for (size_t j = 0; j < acc.size() / 4 && !conflict; ++j) {
      const double d[4] = {cand[0] - acc[4 * j], cand[1] - acc[4 * j + 1], cand[2] - acc[4 * j + 2],
                           cand[3] - acc[4 * j + 3]};
      // Lorentzian scalar s(d, d) in nalgebra's dimension-4 pairwise order
      const double s = ((-d[0]) * d[0] + d[2] * d[2]) + (d[1] * d[1] + d[3] * d[3]);
      conflict = s < 0.0;
    }
    if (!conflict) acc.insert(acc.end(), cand, cand + 4);
  }
  if (acc.size() / 4 != n_bubbles) {
    g_last_error = "Insufficient Bubbles: requested " + std::to_string(n_bubbles) + ", generated " +
                   std::to_string(acc.size() / 4);
    return 1;
  }
  std::memcpy(interior, acc.data(), acc.size() * sizeof(double));
  for (size_t b = 0; b < n_bubbles; ++b) images(L, bc, &acc[4 * b], exterior + 4 * n_img * b);
  return 0;
}

}  // extern "C"
