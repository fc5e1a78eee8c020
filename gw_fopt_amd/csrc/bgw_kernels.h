// SPDX-License-Identifier: MIT
// Launch interfaces of the bulk-flow kernels (host side includes this).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "bgw_common.h"

namespace bgw {

// time steps of a grid at most (exclusive): the geometry orders each ring's
// directions by the step at which they have collided, in 21 bits
constexpr int kSortStepMax = (1 << 21) - 1;

struct GeoArgs {
  const CandEntry* cand;
  const int32_t* inv;
  const RealInfo* real;
  const Item* items;
  const double* rot;      // (n_k, 9) rotation taking k-hat to z-hat
  int identity;           // 1: no rotation (the reference's own frame)
  const double* sin_t;    // (n_c)
  const double* cos_t;    // (n_c)
  const double* cos_p;    // (n_phi)
  const double* sin_p;    // (n_phi)
  int n_c, n_phi, n_phi_pow2;
  double* tau;            // (n_items, n_c * n_phi), each ring's delays ordered by the first
                          //   time step at which the direction has collided (n_t > 0)
  uint16_t* perm;         // (n_items, n_c * n_phi), phi index of each ordered entry
  int n_t;                // time steps of the item's grid (0: rings in phi order)
  int32_t* first_out;     // optional (n_items, n_c * n_phi), unsorted
  double* dtab_out;       // optional (n_items, n_c * n_phi), unsorted
  int item0;              // first item of the launch; tau / perm row b is item item0 + b
  int rings_per_block;    // set by launch_geometry: workgroups split an item's rings
  int n_split;            //   into n_split ranges of rings_per_block rings
};

struct SpecArgs {
  const RealInfo* real;
  const Item* items;
  const int32_t* order;   // launch order (block b works on item order[b])
  const double* tau;
  const uint16_t* perm;
  const double* gcos;     // (n_phi) w_j cos(2 phi_j)
  const double* gsin;     // (n_phi) w_j sin(2 phi_j)
  const double* ring_c;   // (n_c) cos theta_i
  const double* ring_fac; // (n_c) dcos * w_i * 0.5 * sin^2 theta_i * dphi
  const double* omega;    // omegas of every realisation (RealInfo::w_off)
  const double* coef;     // (n_sets, n_terms)
  const double* pows;     // (n_sets, n_terms)
  int n_sets, n_terms;    // sets handled by this launch (<= 6)
  int set0, n_sets_out;   // first set of this launch, total sets in `partial`
  int n_c, n_phi, n_t, n_w;
  int cutoff_kind;        // 0 unit, 1 exponential
  double t_cut, inv_t0_sq;
  int has_damping;
  double damping;
  double kfac;            // 1 / (6 pi)
  int skip_uncollided;    // drop rings whose uncollided part is rounding noise
  int mixed_block;        // time steps per factorised phase block (8 or 16; 0: default)
  int threads;            // workgroup size (0: default); lanes own omegas
  int time_tables;        // per-item time tables in LDS (spectrum_time_tables)
  int time_chunk;         // time steps per chunk of weights in LDS: 512, else 256
                          //   (launch_spectrum picks it)
  int surfaces;           // > 1: that many surface-weight sets stacked in gcos / gsin
                          //   (n_phi each) and ring_fac (n_c each), item k uses set k
  double* partial;        // (n_items, 2, n_sets_out, n_w) complex
  int item0;              // tau / perm / partial rows are item - item0
  // adjoint only
  const double* grad_c;   // (n_groups, 2, n_sets, n_w) complex cotangent of C
  double* grad_out;       // (n_items, 2, n_sets_out, n_terms): d/dcoef, d/dpow
};

struct RedArgs {
  const double* partial;
  const int32_t* group_off;  // (n_groups + 1) item ranges, items sorted by group
  const int32_t* group_r;    // realisation of each group
  const RealInfo* real;
  const double* omega;
  int n_sets, n_w;
  double* c_out;             // (n_groups, 2, n_sets, n_w) complex
  double* s_out;             // (n_groups, n_sets, n_w)
  // unit mode (unit_group != null): workgroup b reduces unit u0 + b, i.e. group
  // unit_group[u0 + b] with items [unit_off[u0 + b], unit_off[u0 + b + 1]);
  // partial rows are item - item0
  const int32_t* unit_group;
  const int32_t* unit_off;
  int u0, item0;
};

// Batched ensemble: one persistent launch over (realisation, k-hat) groups.
struct EnsArgs {
  GeoArgs geo;                    // tau / perm: per-workgroup scratch rows (n_blocks, n_c * n_phi)
  SpecArgs spec;                  // tau / perm as geo; partial unused; n_sets <= 2
  const int32_t* unit_group;      // (n_units) groups in launch order
  const int32_t* group_item_off;  // (n_groups + 1) items of each group, in bubble order
  const int32_t* group_r;         // (n_groups) realisation of each group
  int n_units;
  int* counter;                   // work-queue head, cleared by launch_ensemble
  double* c_out;                  // (n_groups, 2, n_sets_out, n_w) complex
  double* s_out;                  // (n_groups, n_sets_out, n_w)
};
constexpr int kEnsembleMaxSets = 2;
// Resident workgroups per CU for this launch shape (0: cannot launch); also
// sets the kernel's dynamic-LDS limit, so call it before launch_ensemble.
int ensemble_blocks_per_cu(const EnsArgs& a, int mode);
// Clears the queue counter (a kernel, so the pair can be graph-captured) and
// launches n_blocks persistent workgroups.
hipError_t launch_ensemble(const EnsArgs& a, int mode, int n_blocks, hipStream_t s);
size_t ensemble_lds_bytes(const EnsArgs& a, int mode);
// kmean (n_real, nq) = mean over k of s (n_real, n_k, nq); if mean: mean and
// population standard deviation over realisations of kmean.
hipError_t launch_ensemble_stats(const double* s_spec, int n_real, int n_k, int nq, double* kmean,
                                 double* mean, double* stdev, hipStream_t s);

// The FP64 matrix form of the phase sums (bgw_matrix.hip), over tiles of `ni`
// items of one (realisation, reference bubble); every item of a tile has the
// same t_a.  launch_matrix_weights and launch_matrix_products run tiles
// [tile0, tile0 + n_tiles) of the arrays
// below: their time tables, their weight rows (vbuf, rows of 16 col_tiles
// doubles from tinfo[tile].kbase to the end of the time grid, ring after
// ring, at tinfo[tile].voff - voff0), and the matrix products into partial.
struct MxTileInfo {
  int64_t voff;           // first double of the tile's weight rows
  int32_t kbase;          // step of row 0: the first weighted step of the bubble, rounded down to 16
  int32_t len;            // rows per ring: n_t - kbase rounded up to 4
  int32_t k_a;            // the first step with t_k - t_a > 0 (n_t if none)
};
struct MatArgs {
  SpecArgs s;             // the phase-sum arguments (no damping; n_t <= 65535)
  const int32_t* tiles;   // (n_tiles, ni) item indices, -1 padding
  const MxTileInfo* tinfo;
  int ni;                 // items per tile, <= matrix_items_per_tile(s.n_sets, col_tiles)
  int omega_tiles;        // 16-omega tiles per wavefront of the matrix kernel: 1, 2 or 4
  int col_tiles;          // 16-column tiles: 1 or 2
  int tile0;              // first tile of the launch
  int64_t voff0;          // tinfo offset of vbuf[0]
  double* vbuf;           // weight rows of the launch's tiles
  double* tt;             // (n_tiles, n_t, 1 + n_sets n_terms) time tables
  int32_t* ctlo;          // (n_tiles, n_c, 2) first weighted step of each column tile
                          //   (written by the weights kernel, one workgroup per column tile)
  double* phase;          // (n_real, n_c, n_w, 4) per ring and omega the step factor
                          //   e^{i w 4 dt (1 - c_i)} and the anchor factor (cos, sin each)
};
// n_sets * n_terms of one matrix-form launch
constexpr int kMaxMatrixTerms = 12;
// rows of padding (of 32 doubles) after the weight rows of a launch: the
// products kernel prefetches up to one block of rows past a ring's last row
constexpr int kMatrixRowPad = 32;
// time steps per block of the matrix form's products: a tile's first weight
// row (kbase) and its row count per ring are multiples of it
constexpr int kMatrixBlockSteps = 16;
int matrix_items_per_tile(int n_sets, int col_tiles);
size_t matrix_lds_bytes(const SpecArgs& a, int ni, int col_tiles);
// the time tables and weight rows of the launch's tiles, then their matrix products
hipError_t launch_matrix_weights(const MatArgs& a, int n_tiles, hipStream_t s);
hipError_t launch_matrix_products(const MatArgs& a, int n_tiles, hipStream_t s);
// the phase factors of every realisation's rings and omegas (MatArgs::phase)
hipError_t launch_matrix_phases(const MatArgs& a, int n_real, hipStream_t s);

// Default launch shape of the phase sums (0 in SpecArgs::mixed_block /
// threads): time steps per factorised block, 8 for the FP64 blocked form and
// 16 otherwise; workgroups of 64..256 lanes holding one omega each in the
// FP64 form and two in the blocked forms.
inline int resolve_block_steps(int block_steps, int mode) {
  return block_steps ? block_steps : 8;
}
inline int resolve_threads(int threads, int n_w, int mode) {
  if (threads > 0) return threads;
  // omegas per lane by default, from a sweep of (block steps, lanes) at
  // plain_256 (256 omegas, phase-sum kernel per step): 2 for the blocked FP64
  // form (8/128: 22.0 ms; 16/128: 22.1; 8/256: 28.6; 16/256: 33.2), 1
  // otherwise (mixed 8/256: 18.7 ms; 16/256: 19.8)
  const int per_lane = mode == 1 ? 2 : 1;
  int t = ((n_w + per_lane - 1) / per_lane + 63) / 64 * 64;
  return t < 64 ? 64 : (t > 256 ? 256 : t);
}

// The dynamic-LDS limit (hipFuncSetAttribute) is per-function process state:
// each launcher sets it, queries occupancy under it and launches while holding
// this lock, so host threads driving separate handles cannot interleave them.
// Launchers take it themselves; callers must not hold it.
std::mutex& lds_limit_mutex();

hipError_t launch_geometry(const GeoArgs& a, int n_items, hipStream_t s);
hipError_t launch_spectrum(const SpecArgs& a, int n_items, int mixed, hipStream_t s);
hipError_t launch_reduce(const RedArgs& a, int n_groups, hipStream_t s);
// Adjoint: a.n_sets <= kAdjointMaxSets, workgroups of kAdjointThreads lanes.
constexpr int kAdjointMaxSets = 2;
constexpr int kAdjointThreads = 256;
hipError_t launch_adjoint(const SpecArgs& a, int n_items, hipStream_t s);
size_t adjoint_lds_bytes(const SpecArgs& a, int threads);
// Time-resolved integrand: items [0, n_items) of a.order write
// a.partial[block][pm][set][t][w] (complex, a.n_sets_out sets); FP64.
hipError_t launch_integrand(const SpecArgs& a, int n_items, hipStream_t s);
// acc[e] = (first ? 0 : acc[e]) + sum_{it < n_chunk_items} partial[it][e], in order.
hipError_t launch_accumulate(const double* partial, int n_chunk_items, int64_t n_elem, int first,
                             double* acc, hipStream_t s);
size_t spectrum_lds_bytes(const SpecArgs& a, int mode);
bool spectrum_time_tables(const SpecArgs& a);
// Time steps per chunk (SpecArgs::time_chunk) that launch_spectrum takes for
// the phase-sum kernel of this shape.
int spectrum_time_chunk(const SpecArgs& a, int mode);

}  // namespace bgw
