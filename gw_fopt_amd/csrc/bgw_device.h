// SPDX-License-Identifier: MIT
// Device arithmetic for the bulk-flow kernels (gfx950).
//
// Every file that includes this header is compiled with -ffp-contract=off:
// the geometry (collision tests, collision delays, rotations) must round
// exactly like the reference's Rust code, which never fuses a*b+c.  Where a
// fused multiply-add is wanted (polynomials, accumulations) it is written
// explicitly with __builtin_fma / __builtin_fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bgw {

// Lorentzian scalar product s(u, v) = -u0 v0 + u1 v1 + u2 v2 + u3 v3, with the
// pairwise association of a dimension-4 dot product:
// ((-u0 v0) + u2 v2) + (u1 v1 + u3 v3).  The CPU oracle uses the same order,
// so collision decisions agree bit for bit.
__device__ __forceinline__ double lscalar(double u0, double u1, double u2, double u3, double v0,
                                          double v1, double v2, double v3) {
  double a = (-u0) * v0;
  double b = u1 * v1;
  double c = u2 * v2;
  double d = u3 * v3;
  a = a + c;
  b = b + d;
  return a + b;
}

// ---------------------------------------------------------------------------
// sin/cos.  Argument reduction by x - q*pi/2 with pi/2 split into three
// doubles (Cody & Waite), evaluated with fused multiply-adds so that each
// step rounds once; the split constants are pi/2 rounded to double and the
// two successive remainders (computed from a 100-digit value of pi).  The
// reduced argument |r| <= pi/4 (plus rounding) goes into truncated Taylor
// series through r^17 for sin and r^16 for cos (truncation below 0.1 ulp at
// pi/4).  The quadrant q mod 4 picks and negates the two results.
// ---------------------------------------------------------------------------

// a * b + c as one v_fma_f64 with the addend c in an SGPR pair; c must be
// uniform (here: a polynomial coefficient).  In Horner steps
// p = fma(p, r2, coeff) hipcc otherwise selects the two-address v_fmac_f64,
// which overwrites its addend, and copies the coefficient into a VGPR first, a
// v_mov_b64 per step.  The three-address form needs no copy and no VGPR for the
// coefficient: FP64 form's phase sums 177 -> 142 ms at plain_256, adjoint
// 3.26 -> 3.06 ms at grad_128, blocked and mixed forms unchanged.  Same rounding
// as __builtin_fma (one fused operation).
__device__ __forceinline__ double fma_coeff(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
}

__device__ __forceinline__ double reduce_half_pi(double x, int* quadrant) {
  const double two_over_pi = 0x1.45f306dc9c883p-1;
  const double p1 = 0x1.921fb54442d18p+0;
  const double p2 = 0x1.1a62633145c07p-54;
  const double p3 = -0x1.f1976b7ed8fbcp-110;
  double q = __builtin_rint(x * two_over_pi);
  double r = __builtin_fma(-q, p1, x);
  r = __builtin_fma(-q, p2, r);
  r = __builtin_fma(-q, p3, r);
  *quadrant = static_cast<int>(static_cast<long long>(q) & 3);
  return r;
}

template <typename T>
__device__ __forceinline__ void apply_quadrant(int n, T s, T c, T* sin_out, T* cos_out) {
  T ss = (n & 1) ? c : s;
  T cc = (n & 1) ? s : c;
  *sin_out = (n & 2) ? -ss : ss;
  *cos_out = ((n + 1) & 2) ? -cc : cc;
}

__device__ __forceinline__ void sincos_f64(double x, double* sin_out, double* cos_out) {
  int n;
  double r = reduce_half_pi(x, &n);
  double r2 = r * r;
  double ps = 0x1.952c77030ad4ap-49;                    //  1/17!
  ps = fma_coeff(ps, r2, -0x1.ae7f3e733b81fp-41);  // -1/15!
  ps = fma_coeff(ps, r2, 0x1.6124613a86d09p-33);   //  1/13!
  ps = fma_coeff(ps, r2, -0x1.ae64567f544e4p-26);  // -1/11!
  ps = fma_coeff(ps, r2, 0x1.71de3a556c734p-19);   //  1/9!
  ps = fma_coeff(ps, r2, -0x1.a01a01a01a01ap-13);  // -1/7!
  ps = fma_coeff(ps, r2, 0x1.1111111111111p-7);    //  1/5!
  ps = fma_coeff(ps, r2, -0x1.5555555555555p-3);   // -1/3!
  double s = __builtin_fma(r * r2, ps, r);
  double pc = 0x1.ae7f3e733b81fp-45;                    //  1/16!
  pc = fma_coeff(pc, r2, -0x1.93974a8c07c9dp-37);  // -1/14!
  pc = fma_coeff(pc, r2, 0x1.1eed8eff8d898p-29);   //  1/12!
  pc = fma_coeff(pc, r2, -0x1.27e4fb7789f5cp-22);  // -1/10!
  pc = fma_coeff(pc, r2, 0x1.a01a01a01a01ap-16);   //  1/8!
  pc = fma_coeff(pc, r2, -0x1.6c16c16c16c17p-10);  // -1/6!
  pc = fma_coeff(pc, r2, 0x1.5555555555555p-5);    //  1/4!
  pc = fma_coeff(pc, r2, -0.5);                    // -1/2!
  double c = __builtin_fma(r2, pc, 1.0);
  apply_quadrant(n, s, c, sin_out, cos_out);
}

}  // namespace bgw
