// Correctly rounded FP32 reciprocal, quotient and square root for the step
// kernels, built from the gfx950 hardware approximations and FMA residuals
// without the scaling and fix-up steps that only extreme exponents need.  The
// kernels' operands are distances, coordinate differences, masses and forces,
// far from the exponent limits, so these give the results of IEEE division
// and square root there.
//
// tests/recip_check.hip includes this file and checks each function on the
// GPU against the exact value (a result is correctly rounded iff the exact
// value lies strictly between the midpoints to its neighbours, compared
// exactly in double):
//   qrcp   every significand at every exponent in [-100, 100]
//   qdiv   2^34 random signed pairs, exponents in [-60, 60]
//   qsqrt  every x in [2^-100, FLT_MAX] (below about 2^-104 its residual
//          underflows and it can miss by an ulp)
// and tests/markstein_check.c checks qdiv_by exhaustively on the CPU for the
// pairwise pass's fixed divisors.
#pragma once
#include <hip/hip_runtime.h>

namespace cc {

// 1/b correctly rounded for a normal b with a normal reciprocal: the hardware
// reciprocal (within an ulp) refined once, y + y (1 - b y), the residual exact
// by FMA.  On gfx950 the refined value is the correctly rounded reciprocal for
// every significand (the one case a single refinement can miss, a reciprocal
// just above a rounding midpoint reached from the wrong side, does not arise
// from the hardware's first approximation).
__device__ __forceinline__ float qrcp(float b) {
  const float y = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, y, 1.0f);
  return fmaf(e, y, y);
}

// a / b correctly rounded for a divisor whose correctly rounded reciprocal y
// is at hand: one residual correction of a * y (Markstein: with y the
// correctly rounded 1/b, q + (a - b q) y rounds to a / b), valid while the
// residual a - b q does not underflow, i.e. for dividends above about 2^-100
// (or 0) and normal quotients.  A zero dividend of either sign gives +0.
__device__ __forceinline__ float qdiv_by(float a, float b, float y) {
  const float q = a * y;
  const float r = fmaf(-b, q, a);
  return fmaf(r, y, q);
}

// a / b correctly rounded: the reciprocal above, then the one correction.
__device__ __forceinline__ float qdiv(float a, float b) { return qdiv_by(a, b, qrcp(b)); }

// The correctly rounded square root of x >= 2^-100 (or of 0): the hardware
// root, then the neighbour above or below where its residual says the root
// lies.
__device__ __forceinline__ float qsqrt(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u), su = __uint_as_float(__float_as_uint(s) + 1u);
  const float rd = fmaf(-sd, s, x), ru = fmaf(-su, s, x);
  s = rd <= 0.0f ? sd : s;
  return ru > 0.0f ? su : s;
}

}  // namespace cc
