// SPDX-License-Identifier: MIT
// Plain-old-data layouts shared by the host library and the gfx950 kernels.
//
// HBM layout (one "batch" = one or more nucleation realisations r, a list of
// k-hat directions k and, per (r, k), the reference bubbles a that contribute):
//
//   CandEntry   [sum_r n_int(r) * (n_tot(r) - 1)]  per (r, a): the other
//               bubbles b != a sorted by a lower bound of their collision delay
//   inv         [sum_r n_int(r) * n_tot(r)]  position of b in (r, a)'s list
//   Item        [n_items]  one (r, k, a) triple, sorted by (r, k, a)
//   tau, perm   [n_items, n_c * n_phi]  per ring: collision delays sorted
//               ascending (+inf: never collides) and their phi indices
//   partial     [n_items, 2, n_sets, n_w] complex  C contribution of one a
//   C, S        [n_groups, ...]  per (r, k) after the deterministic reduce
#pragma once

#include <stdint.h>

namespace bgw {

// One entry of a reference bubble's candidate list (64 bytes, one scalar
// load).  `key` is a lower bound of the collision delay of b with a along any
// direction, (|d_spatial| + d_t) / 2, for "regular" pairs (spacelike well away
// from the light cone); irregular pairs get key = -inf so every scan that
// relies on the bound looks at them unconditionally.
struct alignas(16) CandEntry {
  double d[4];       // delta = x_b - x_a  (t, x, y, z), unrotated
  double dsq;        // delta_squared = scalar(delta, delta), unrotated
  double key;        // collision-delay lower bound, or -inf
  int32_t idx;       // total bubble index b (interior first, then exterior)
  int32_t flags;     // bit 0: regular pair
  double len;        // |d_spatial|
};
static_assert(sizeof(CandEntry) == 64, "CandEntry must be 64 bytes");

constexpr int32_t kCandRegular = 1;

// Per realisation.
struct RealInfo {
  int64_t cand_off;  // index of the first CandEntry of reference bubble 0
  int64_t inv_off;   // index of the first inverse-permutation entry of bubble 0
  int64_t w_off;     // index of the realisation's first omega
  int32_t n_int, n_tot;
  int32_t nonfinite;  // some coordinate is not finite: use the reference scan
  int32_t pad;
  double t_begin, t_step, dt;  // t_k = t_begin + t_step * k, dt = t_1 - t_0
};

// One (realisation, k-hat, reference bubble) unit of work.
struct Item {
  int32_t r, k, a, group;  // group: index of the (r, k) pair
  double t_a;              // nucleation time of a
  double z_a;              // z of a in the k-hat frame (k-hat . x_a)
};

}  // namespace bgw
