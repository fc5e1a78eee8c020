/*
 * bubble_gw.h — C ABI of the MI355X (gfx950) generalized bulk-flow spectrum.
 *
 * Drop-in for the reference's C ABI (rust/c-bubble-gw/src/lib.rs, header
 * conventions of rust/c-bubble-gw/cbindgen.toml: include guard BUBBLE_GW_H,
 * C with C++ guards, usize -> size_t, enum variants prefixed with the type
 * name).  The first block repeats the reference's bulk-flow symbols with the
 * same names, argument meaning, layouts, ownership and error behaviour; the
 * second block adds entry points with the same conventions.  The library is
 * libbubble_gw_c.so (the reference's cdylib name, rust/c-bubble-gw/Cargo.toml:10-12).
 *
 * Conventions (rust/c-bubble-gw/README.md:39-47):
 *  - all arrays are caller-owned, C-contiguous, row-major; inputs are copied;
 *  - handles are opaque and released with their *_free function;
 *  - functions return BgwStatus, details via bgw_last_error_message()
 *    (thread-local, cleared by every successful call);
 *  - NAN is the "None" sentinel for optional floating-point arguments.
 */
#ifndef BUBBLE_GW_H
#define BUBBLE_GW_H

#pragma once

#include <stdarg.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

/* Reference: BgwStatus (lib.rs:18-26). */
typedef enum BgwStatus {
    BgwStatus_Ok = 0,
    BgwStatus_NullPointer = 1,
    BgwStatus_InvalidArgument = 2,
    BgwStatus_CalculationError = 3,
    BgwStatus_Panic = 4,
} BgwStatus;

/* Reference: BgwTimeCutoffKind (lib.rs:63-68). */
typedef enum BgwTimeCutoffKind {
    BgwTimeCutoffKind_Unit = 0,
    BgwTimeCutoffKind_Exponential = 1,
} BgwTimeCutoffKind;

/* Reference: BgwTimeCutoffConfig (lib.rs:70-77); NAN ratios take the
 * defaults of ExponentialTimeCutoff::new (time_cutoff.rs:22-30). */
typedef struct BgwTimeCutoffConfig {
    BgwTimeCutoffKind kind;
    double smax;
    double ratio_t_cut;
    double ratio_t_0;
} BgwTimeCutoffConfig;

/* Reference: BgwComplex64 (lib.rs:79-84). */
typedef struct BgwComplex64 {
    double re;
    double im;
} BgwComplex64;

/* Reference: opaque bgw_generalized_bulk_flow (lib.rs:91-94). */
typedef struct bgw_generalized_bulk_flow bgw_generalized_bulk_flow;

/* ---- additions: types ---------------------------------------------------- */

/* Arithmetic of the phase sums.
 *  Fp64:        every phase exp(i w psi) evaluated with its own double sincos
 *               (the drop-in's parity form).
 *  Mixed:       collision tests and phase-argument reduction in double; the
 *               phases of R consecutive time steps factorised as anchor x
 *               table, sincos, tables, weights and block sums in float, block
 *               results rotated and accumulated in double.
 *  Fp64Blocked: the same block factorisation entirely in double.
 *  Fp64Matrix:  double throughout, on the matrix cores: the phase sums of the
 *               items of one bubble in several k-hat frames (same t_a) are one
 *               matrix product per ring (half-step phase factors x the sums
 *               and differences of the weights paired about the centres of
 *               16-step blocks), accumulated by FP64 MFMA; fast when many
 *               k-hat share the call (spectra); no damping width;
 *               n_t <= 65535; single configurations and the batched
 *               ensembles. */
typedef enum BgwPrecision {
    BgwPrecision_Fp64 = 0,
    BgwPrecision_Mixed = 1,
    BgwPrecision_Fp64Blocked = 2,
    BgwPrecision_Fp64Matrix = 3,
} BgwPrecision;

/* Device times of the most recent compute call on a handle (HIP events on
 * the handle's stream), in milliseconds. */
typedef struct BgwTimings {
    double geometry_ms;  /* first-collision kernel */
    double spectrum_ms;  /* phase-sum kernel (or adjoint kernel) */
    double reduce_ms;    /* per-(k-hat) reduction and spectrum */
    double total_ms;     /* first kernel start to last kernel end */
    size_t n_items;      /* (realisation, k-hat, bubble) work items */
    size_t n_terms;      /* (item, ring, time step, omega) phase terms evaluated */
    size_t n_directions; /* (item, surface direction) collision searches */
    double matrix_ms;    /* FP64 matrix form: the matrix-core kernel's launches (within
                            spectrum_ms); NAN when not every launch was timed; 0 otherwise */
} BgwTimings;

/* A batch of independent nucleation realisations (uneven bubble counts). */
typedef struct bgw_ensemble bgw_ensemble;

/* Launch shape of the kernels (chosen by bgw_ensemble_autotune; 0 = automatic). */
typedef struct BgwTileConfig {
    size_t block_steps;       /* time steps per factorised phase block: 8 or 16 */
    size_t threads;           /* phase-sum workgroup size: 64, 128 or 256 (lanes own omegas) */
    size_t items_per_chunk;   /* ensembles: 0 = one fused persistent launch (geometry and
                                 phase sums per (realisation, k-hat) unit); N > 0 = chunks of
                                 about N (realisation, k-hat, bubble) items whose geometry and
                                 phase-sum kernels run pipelined on two streams */
    size_t persistent_blocks; /* workgroups of the persistent launch (0: all that fit on the
                                 device at once) */
} BgwTileConfig;

/* ---- additions: nucleation types ------------------------------------------ */

/* Lattice of a nucleation run (rust/bubble-gw/src/many_bubbles/lattice/):
 * Parallelepiped (also the Cartesian box): origin + u e1 + v e2 + w e3,
 * 0 <= u, v, w <= 1, basis rows e1, e2, e3; Spherical: |x - origin| <= radius. */
typedef enum BgwLatticeKind {
    BgwLatticeKind_Parallelepiped = 0,
    BgwLatticeKind_Spherical = 1,
} BgwLatticeKind;

/* BoundaryConditions (lattice.rs): exterior images of every interior bubble.
 * Parallelepiped: Periodic = six face translates (+e1, -e1, +e2, -e2, +e3, -e3),
 * Reflection = six face mirror images (-e1, +e1, -e2, +e2, -e3, +e3 faces);
 * Spherical: Reflection = two tangent-plane mirror images, Periodic = none. */
typedef enum BgwBoundaryCondition {
    BgwBoundaryCondition_Periodic = 0,
    BgwBoundaryCondition_Reflection = 1,
    BgwBoundaryCondition_None = 2,
} BgwBoundaryCondition;

typedef struct BgwNucleationLattice {
    int32_t kind;       /* BgwLatticeKind */
    double origin[3];   /* parallelepiped origin / sphere center */
    double basis[9];    /* parallelepiped basis rows e1, e2, e3 */
    double radius;      /* sphere radius */
} BgwNucleationLattice;

/* FixedRateNucleationMethod (fixed_rate_nucleation.rs:35-40). */
typedef enum BgwFixedRateMethod {
    BgwFixedRateMethod_TimeStepping = 0,  /* fixed-probability time stepping with d_p0 */
    BgwFixedRateMethod_Distribution = 1,  /* nucleation times from the expected-count ODE */
} BgwFixedRateMethod;

/* One fixed-rate realisation: Gamma(t) = gamma0 exp(beta (t - t0)).  Its random
 * stream is Philox4x64-10 with key (seed, stream). */
typedef struct BgwFixedRateRealisation {
    BgwNucleationLattice lattice;
    double beta;
    double gamma0;
    double t0;
    uint64_t seed;
    uint64_t stream;
} BgwFixedRateRealisation;

#ifdef __cplusplus
extern "C" {
#endif  // __cplusplus

/* ---- reference bulk-flow symbols (rust/c-bubble-gw/src/lib.rs) ----------- */

/* Reference: bgw_version (lib.rs:219-223). */
const char *bgw_version(void);

/* Reference: bgw_last_error_message (lib.rs:225-229). */
const char *bgw_last_error_message(void);

/* Reference: bgw_generalized_bulk_flow_empty_lattice_new (lib.rs:386-422).
 * Bubble arrays are (n, 4) rows (t_c, x_c, y_c, z_c). */
BgwStatus bgw_generalized_bulk_flow_empty_lattice_new(bgw_generalized_bulk_flow **out,
                                                      const double *bubbles_interior,
                                                      size_t n_interior,
                                                      const double *bubbles_exterior,
                                                      size_t n_exterior);

/* Reference: bgw_generalized_bulk_flow_free (lib.rs:424-430); NULL is allowed. */
void bgw_generalized_bulk_flow_free(bgw_generalized_bulk_flow *bulk_flow);

/* Reference: bgw_generalized_bulk_flow_set_num_threads (lib.rs:432-446).
 * num_threads (> 0) sizes the host thread pool of the per-configuration
 * preparation -- each bubble's partners sorted by a collision-delay bound --
 * which runs at the handle's first compute call; the default is the
 * hardware's thread count.  Everything per k-hat runs on the GPU. */
BgwStatus bgw_generalized_bulk_flow_set_num_threads(bgw_generalized_bulk_flow *bulk_flow,
                                                    size_t num_threads);

/* Reference: bgw_generalized_bulk_flow_set_resolution (lib.rs:448-461). */
BgwStatus bgw_generalized_bulk_flow_set_resolution(bgw_generalized_bulk_flow *bulk_flow,
                                                   size_t n_cos_thetax,
                                                   size_t n_phix,
                                                   bool precompute_first_bubbles);

/* Reference: bgw_generalized_bulk_flow_set_gradient_scaling_params
 * (lib.rs:463-494); row-major (n_sets, n_terms); NAN damping = None. */
BgwStatus bgw_generalized_bulk_flow_set_gradient_scaling_params(
    bgw_generalized_bulk_flow *bulk_flow, const double *coefficients, const double *powers,
    size_t n_sets, size_t n_terms, double damping_width);

/* Reference: bgw_generalized_bulk_flow_compute_c_integral (lib.rs:526-574).
 * out: (2, n_sets, n_w) = [C+, Cx], out_len must be 2 * n_sets * n_w. */
BgwStatus bgw_generalized_bulk_flow_compute_c_integral(bgw_generalized_bulk_flow *bulk_flow,
                                                       const double *w_arr,
                                                       size_t n_w,
                                                       bool has_t_begin,
                                                       double t_begin,
                                                       double t_end,
                                                       size_t n_t,
                                                       const size_t *selected_bubbles,
                                                       size_t n_selected_bubbles,
                                                       BgwTimeCutoffConfig time_cutoff,
                                                       BgwComplex64 *out,
                                                       size_t out_len);

/* ---- additions --------------------------------------------------------------- */

/* Same as bgw_generalized_bulk_flow_compute_c_integral with a precision choice.
 * No reference counterpart (the reference computes in double only). */
BgwStatus bgw_generalized_bulk_flow_compute_c_integral_ex(bgw_generalized_bulk_flow *bulk_flow,
                                                          const double *w_arr, size_t n_w,
                                                          bool has_t_begin, double t_begin,
                                                          double t_end, size_t n_t,
                                                          const size_t *selected_bubbles,
                                                          size_t n_selected_bubbles,
                                                          BgwTimeCutoffConfig time_cutoff,
                                                          BgwPrecision precision,
                                                          BgwComplex64 *out, size_t out_len);

/* Same as bgw_generalized_bulk_flow_compute_c_integral_ex, but allocates no
 * device memory: every device buffer of the call (bubbles, candidate lists,
 * collision delays, partial sums, C) is carved from a 64 MiB scratch array
 * that the library holds on each device it runs on until the process ends.
 * A request needing more fails with CalculationError and a message giving the
 * limit.  Calls that overlap on one device share that scratch and race; the
 * call returns after its device work has finished.  No reference counterpart:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
BgwStatus bgw_generalized_bulk_flow_compute_c_integral_no_alloc(
    bgw_generalized_bulk_flow *bulk_flow, const double *w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t *selected_bubbles,
    size_t n_selected_bubbles, BgwTimeCutoffConfig time_cutoff, BgwPrecision precision,
    BgwComplex64 *out, size_t out_len);

/* First colliding bubble of reference bubble a_idx on the (n_cos_thetax,
 * n_phix) grid, int32 encoding -1 None, i Interior(i), -(j+2) Exterior(j).
 * Replaces GeneralizedBulkFlow::compute_first_colliding_bubble
 * (generalized_bulk_flow.rs:127-206) as exposed by the Python binding
 * (rust/py-bubble-gw/src/py_many_bubbles/py_generalized_bulk_flow.rs:235-247). */
BgwStatus bgw_generalized_bulk_flow_first_colliding_bubble(bgw_generalized_bulk_flow *bulk_flow,
                                                           size_t a_idx, int32_t *out,
                                                           size_t out_len);

/* First colliding bubble and collision delay (the GPU's own values) of
 * reference bubble a_idx in the frame where the unit vector k_hat -> z (NULL:
 * the reference frame); either output may be NULL.  No reference counterpart:
 * the reference computes these per rotated LatticeBubbles
 * (lattice_bubbles.rs:144-197) with compute_first_colliding_bubble and
 * compute_delta_tab. */
BgwStatus bgw_generalized_bulk_flow_first_colliding_bubble_ex(bgw_generalized_bulk_flow *bulk_flow,
                                                              size_t a_idx, const double *k_hat,
                                                              int32_t *first_out,
                                                              double *delta_tab_out,
                                                              size_t out_len);

/* Collision delays t_c - t_a for a first-bubble grid.
 * Replaces compute_delta_tab (generalized_bulk_flow.rs:421-471;
 * py_generalized_bulk_flow.rs:249-266). */
BgwStatus bgw_generalized_bulk_flow_delta_tab(bgw_generalized_bulk_flow *bulk_flow, size_t a_idx,
                                              const int32_t *first_bubble, size_t grid_len,
                                              double *out, size_t out_len);

/* Collision status 0 NeverCollided, 1 AlreadyCollided, 2 NotYetCollided.
 * Replaces compute_collision_status (generalized_bulk_flow.rs:369-414;
 * py_generalized_bulk_flow.rs:268-291). */
BgwStatus bgw_generalized_bulk_flow_collision_status(bgw_generalized_bulk_flow *bulk_flow,
                                                     size_t a_idx, double t,
                                                     const int32_t *first_bubble,
                                                     const double *delta_tab, size_t grid_len,
                                                     int32_t *out, size_t out_len);

/* Time-resolved integrand (2, n_sets, n_w, n_t), no time cutoff.
 * Replaces compute_c_integrand (generalized_bulk_flow.rs:875-920;
 * py_generalized_bulk_flow.rs:343-379). */
BgwStatus bgw_generalized_bulk_flow_compute_c_integrand(bgw_generalized_bulk_flow *bulk_flow,
                                                        const double *w_arr, size_t n_w,
                                                        bool has_t_begin, double t_begin,
                                                        double t_end, size_t n_t,
                                                        const size_t *selected_bubbles,
                                                        size_t n_selected_bubbles,
                                                        BgwComplex64 *out, size_t out_len);

/* Spectrum over a list of unit k-hat directions (n_k, 3).  For each k-hat the
 * bubbles are rotated so that k-hat -> z (the reference's route,
 * LatticeBubbles::transform, lattice_bubbles.rs:144-197) and
 *   S[k][s][w] = 8 pi w^3 (|C+|^2 + |Cx|^2)
 * (src/gw_fopt/bubble_dynamics/utils/data_gw_bulkflow.py:106-111).
 * spectrum_out: (n_k, n_sets, n_w); mean_out (optional): (n_sets, n_w) mean
 * over k-hat; c_out (optional): (n_k, 2, n_sets, n_w).  No reference
 * counterpart: the reference rotates and calls compute_c_integral per k-hat. */
BgwStatus bgw_generalized_bulk_flow_compute_spectrum(bgw_generalized_bulk_flow *bulk_flow,
                                                     const double *w_arr, size_t n_w,
                                                     const double *k_hat, size_t n_k,
                                                     bool has_t_begin, double t_begin,
                                                     double t_end, size_t n_t,
                                                     BgwTimeCutoffConfig time_cutoff,
                                                     BgwPrecision precision,
                                                     double *spectrum_out, size_t spectrum_len,
                                                     double *mean_out, size_t mean_len,
                                                     BgwComplex64 *c_out, size_t c_len);

/* Adjoint: given the cotangent dL/dC (n_k, 2, n_sets, n_w) (complex,
 * dL/dRe + i dL/dIm) of the C of bgw_generalized_bulk_flow_compute_spectrum,
 * returns dL/dcoefficients and dL/dpowers, each (n_sets, n_terms).  The
 * collision masks are recomputed, not stored.  FP64.  No reference
 * counterpart (the reference has no derivatives). */
BgwStatus bgw_generalized_bulk_flow_spectrum_vjp(bgw_generalized_bulk_flow *bulk_flow,
                                                 const double *w_arr, size_t n_w,
                                                 const double *k_hat, size_t n_k,
                                                 bool has_t_begin, double t_begin, double t_end,
                                                 size_t n_t, BgwTimeCutoffConfig time_cutoff,
                                                 const BgwComplex64 *grad_c, size_t grad_c_len,
                                                 double *grad_coefficients,
                                                 double *grad_powers, size_t grad_len);

/* Launch shape used by this handle's compute calls (NULL: automatic). */
BgwStatus bgw_generalized_bulk_flow_set_tiles(bgw_generalized_bulk_flow *bulk_flow,
                                              const BgwTileConfig *tiles);

/* Device timings of the last compute call on this handle. */
BgwStatus bgw_generalized_bulk_flow_last_timings(bgw_generalized_bulk_flow *bulk_flow,
                                                 BgwTimings *out);

/* Reference: NewGeneralizedBulkFlow::compute_c_tensor (rust/bubble-gw/src/many_bubbles/
 * new_generalized_bulk_flow.rs:1205-1275, the experimental tensor form): C_ij of the
 * bubbles in the reference frame, out (6, n_sets, n_w) in the order xx, xy, xz, yy, yz,
 * zz (out_len = 6 n_sets n_w); no time cutoff.  C+ = (C_xx - C_yy) / 2 and Cx = C_xy. */
BgwStatus bgw_generalized_bulk_flow_compute_c_tensor(bgw_generalized_bulk_flow *bulk_flow,
                                                     const double *w_arr, size_t n_w,
                                                     bool has_t_begin, double t_begin,
                                                     double t_end, size_t n_t,
                                                     const size_t *selected_bubbles,
                                                     size_t n_selected_bubbles,
                                                     BgwPrecision precision, BgwComplex64 *out,
                                                     size_t out_len);

/* Reference: NewGeneralizedBulkFlow::compute_c_tensor_integrand (new_generalized_bulk_flow.rs:
 * 993-1065): the per-step integrand of compute_c_tensor, out (6, n_sets, n_w, n_t). */
BgwStatus bgw_generalized_bulk_flow_compute_c_tensor_integrand(
    bgw_generalized_bulk_flow *bulk_flow, const double *w_arr, size_t n_w, bool has_t_begin,
    double t_begin, double t_end, size_t n_t, const size_t *selected_bubbles,
    size_t n_selected_bubbles, BgwComplex64 *out, size_t out_len);

/* ---- additions: batched ensemble ------------------------------------------- */

/* Realisation r owns rows [interior_offsets[r], interior_offsets[r+1]) of
 * bubbles_interior and [exterior_offsets[r], exterior_offsets[r+1]) of
 * bubbles_exterior (both (rows, 4), offsets have n_realisations + 1 entries). */
BgwStatus bgw_ensemble_new(bgw_ensemble **out, const double *bubbles_interior,
                           const size_t *interior_offsets, const double *bubbles_exterior,
                           const size_t *exterior_offsets, size_t n_realisations);

void bgw_ensemble_free(bgw_ensemble *ensemble);

BgwStatus bgw_ensemble_set_resolution(bgw_ensemble *ensemble, size_t n_cos_thetax, size_t n_phix);

BgwStatus bgw_ensemble_set_gradient_scaling_params(bgw_ensemble *ensemble,
                                                   const double *coefficients,
                                                   const double *powers, size_t n_sets,
                                                   size_t n_terms, double damping_width);

/* Spectra of every realisation over a shared k-hat list.  w_arr: (n_real,
 * n_w), each realisation with its own omegas; t_begin/t_end: (n_real); all
 * realisations share n_t.  spectrum_out: (n_real, n_k, n_sets, n_w);
 * kmean_out (optional): (n_real, n_sets, n_w) mean over k-hat;
 * ens_mean_out / ens_std_out (optional): (n_sets, n_w) mean and population
 * standard deviation over realisations of kmean.  The whole batch runs as
 * one captured hipGraph (persistent kernels over a work queue, or pipelined
 * chunks); BgwPrecision_Fp64Matrix runs the single-configuration matrix-form
 * launches over every realisation at once instead (captured as one hipGraph
 * too, unless its sets need several matrix launches; each realisation's
 * spectrum equals its own configuration's).  tiles NULL uses the
 * autotuned (or default) launch shape of the form. */
BgwStatus bgw_ensemble_compute_spectrum(bgw_ensemble *ensemble, const double *w_arr, size_t n_w,
                                        const double *k_hat, size_t n_k, const double *t_begin,
                                        const double *t_end, size_t n_t,
                                        BgwPrecision precision, const BgwTileConfig *tiles,
                                        double *spectrum_out, size_t spectrum_len,
                                        double *kmean_out, size_t kmean_len,
                                        double *ens_mean_out, double *ens_std_out,
                                        size_t ens_len);

/* Time the candidate launch shapes on this ensemble's workload and keep the
 * fastest (returned in *best). */
BgwStatus bgw_ensemble_autotune(bgw_ensemble *ensemble, const double *w_arr, size_t n_w,
                                const double *k_hat, size_t n_k, const double *t_begin,
                                const double *t_end, size_t n_t, BgwPrecision precision,
                                BgwTileConfig *best);

BgwStatus bgw_ensemble_last_timings(bgw_ensemble *ensemble, BgwTimings *out);

/* ---- additions: nucleation on the device ------------------------------------ */

/* Exterior images per interior bubble for a lattice kind and boundary condition. */
size_t bgw_nucleation_images_per_bubble(int32_t lattice_kind, BgwBoundaryCondition boundary_condition);

/* Reference: FixedRateNucleation::nucleate (fixed_rate_nucleation.rs:358-377) with
 * nucleate_fixed_probability_time_stepping (:188-275) or
 * nucleate_fixed_probability_distribution (:277-356), for n_realisations
 * independent realisations in one launch (one workgroup each).  Shared
 * parameters: d_p0 (time stepping), n_points (Monte Carlo sample points,
 * > 0), max_time_iterations, cutoff_fraction (false-vacuum fraction that ends
 * the run).  Realisation r writes at most max_bubbles bubbles:
 * interior rows [r * max_bubbles, ...) of (n_real * max_bubbles, 4),
 * exterior rows [r * max_bubbles * n_img, ...) of (n_real * max_bubbles * n_img, 4)
 * (images of bubble b at b * n_img, in generate_bubbles_exterior order),
 * time_history and volume_history (n_real * max_bubbles) at each nucleation,
 * counts[r] = bubbles nucleated.  max_bubbles >= n_points can never overflow
 * (every bubble takes a distinct sample point); a realisation that would
 * overflow fails the call.  exterior may be NULL when n_img == 0. */
BgwStatus bgw_nucleation_fixed_rate(const BgwFixedRateRealisation *realisations,
                                    size_t n_realisations,
                                    BgwBoundaryCondition boundary_condition,
                                    BgwFixedRateMethod method, double d_p0, size_t n_points,
                                    size_t max_time_iterations, double cutoff_fraction,
                                    size_t max_bubbles, double *interior, double *exterior,
                                    double *time_history, double *volume_history,
                                    size_t *counts);

/* Reference: FixedRateNucleation::solve_bubbles_distribution
 * (fixed_rate_nucleation.rs:144-186): integrates the expected-count ODE from
 * tau0 = t0 beta to taumax on the even grid of step 0.001 in tau = beta t and
 * returns the grid points where floor(n) increases (t = tau / beta) with n
 * there.  With t_out == NULL only *n_len is set; otherwise up to cap entries
 * are written. */
BgwStatus bgw_nucleation_solve_distribution(double beta, double gamma0, double t0, double taumax,
                                            double volume_lattice, double *t_out, double *n_out,
                                            size_t cap, size_t *n_len);

/* Reference: SpontaneousNucleation::nucleate (spontaneous_nucleation.rs:89-104):
 * n_bubbles bubbles at t0, uniform in the lattice, for n_realisations
 * realisations (realisation r uses key (seed, stream0 + r)).  interior:
 * (n_real * n_bubbles, 4); exterior: (n_real * n_bubbles * n_img, 4). */
BgwStatus bgw_nucleation_spontaneous(const BgwNucleationLattice *lattice,
                                     BgwBoundaryCondition boundary_condition, size_t n_bubbles,
                                     double t0, uint64_t seed, uint64_t stream0,
                                     size_t n_realisations, double *interior, double *exterior);

/* Device time of the most recent bgw_nucleation_fixed_rate kernel on this thread (ms). */
BgwStatus bgw_nucleation_last_kernel_ms(double *out);

#ifdef __cplusplus
}  // extern "C"
#endif  // __cplusplus

#endif  /* BUBBLE_GW_H */
