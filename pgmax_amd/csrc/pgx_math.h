// FP32 exp / log / log1p / expm1 used by the BP kernels.
//
// Written from the textbook constructions so that a NumPy float32 restatement
// of the same operation sequence reproduces them bit for bit:
//   exp:   x = k ln2 + r with k = rint(x / ln2), ln2 split into a 9-bit head
//          (k * head is exact) and a tail (Cody-Waite reduction, two fused
//          multiply-adds), then a degree-6 polynomial 1 + r + r^2 q(r) on
//          |r| <= ln2/2 in Horner form of fused multiply-adds (each rounds
//          once), scaled by 2^k. q's five coefficients are the FP32
//          roundings of the fit of least maximum relative error with the
//          first two terms held at 1, so e^0 = 1 exactly (FP32 evaluation
//          within 0.71 ulp of e^r on the reduced range; one multiply-add
//          fewer than the degree-7 Taylor polynomial, which is 0.76).
//   log:   x = m 2^e with m in [sqrt(1/2), sqrt(2)); with f = m - 1 and
//          s = f / (2 + f), log m = 2 atanh(s) = 2s + s R,
//          R = 2s^2/3 + 2s^4/5 + ... + 2s^10/11 (|s| <= 0.172, truncation
//          < 1e-9 relative), evaluated as f - s (f - R); plus e ln2.
//          The quotient s is computed without the division sequence but
//          rounds exactly as the division does (div_2pf).
//   log1p: u = 1 + x; log(u) * x / (u - 1), exact-argument correction
//          (x itself when u rounds to 1).
//   expm1: u = exp(x); (u - 1) * x / log(u), the same correction.
// All four stay within 2.5 ulp of the exact values over the arguments BP
// produces; tests/test_oracle.py measures this against float64.
// The file is compiled with -ffp-contract=off so every operation rounds.
#ifndef PGX_MATH_H_
#define PGX_MATH_H_

#include <hip/hip_runtime.h>

namespace pgx {

constexpr float kLn2Hi = 0.693359375f;                 // 355 / 512
constexpr float kLn2Lo = -2.12194440054690583e-4f;     // ln 2 - 355/512
constexpr float kInvLn2 = 1.44269504088896341f;        // 1 / ln 2
constexpr float kExpHi = 88.7228317f;                  // above: +inf
constexpr float kExpLo = -103.972084f;                 // below: 0
constexpr float kSqrtHalf = 0.707106781186547524f;
// e^r = 1 + r + r^2 (c2 + c3 r + c4 r^2 + c5 r^3 + c6 r^4) on |r| <= ln2/2
constexpr float kExpC2 = 0x1.fffffcp-2f;
constexpr float kExpC3 = 0x1.555452p-3f;
constexpr float kExpC4 = 0x1.555748p-5f;
constexpr float kExpC5 = 0x1.12609cp-7f;
constexpr float kExpC6 = 0x1.6c07b2p-10f;

// f / (2 + f) for f = m - 1, m in [sqrt(1/2), sqrt(2)): a reciprocal with one
// Newton step, the quotient, and one correction from the exact residual
// (fused multiply-adds). For every such f it gives the correctly rounded
// quotient, the bits of the IEEE division; pgx_check_numerics (which = 1)
// compares the two for all 8,388,640 values of m in and around that range.
__device__ __forceinline__ float div_2pf(float f) {
  const float d = 2.f + f;
  float y = __builtin_amdgcn_rcpf(d);
  y = __builtin_fmaf(__builtin_fmaf(-d, y, 1.f), y, y);
  const float q = f * y;
  return __builtin_fmaf(__builtin_fmaf(-q, d, f), y, q);
}

// log of a positive finite x (the main evaluation of pgx_log).
__device__ __forceinline__ float pgx_log_pos(float x) {
  int e;
  float m = frexpf(x, &e);
  if (m < kSqrtHalf) {
    m = m * 2.f;
    e = e - 1;
  }
  const float f = m - 1.f;
  const float s = div_2pf(f);
  const float z = s * s;
  float q = 2.f / 11.f;
  q = q * z + 2.f / 9.f;
  q = q * z + 2.f / 7.f;
  q = q * z + 2.f / 5.f;
  q = q * z + 2.f / 3.f;
  // log(1+f) = 2s + s R with R = 2 z/3 + 2 z^2/5 + ...; since 2s = f - s f,
  // log(1+f) = f - s (f - R), which never rounds 2s.
  const float logm = f - s * (f - q * z);
  const float ef = float(e);
  return (ef * kLn2Lo + logm) + ef * kLn2Hi;
}

// e^xc for an in-range xc (kExpLo <= xc <= kExpHi).
__device__ __forceinline__ float pgx_exp_core(float xc) {
  const float k = rintf(xc * kInvLn2);
  float r = __builtin_fmaf(-k, kLn2Hi, xc);  // exact product: xc - k * head
  r = __builtin_fmaf(-k, kLn2Lo, r);
  float p = kExpC6;
  p = __builtin_fmaf(p, r, kExpC5);
  p = __builtin_fmaf(p, r, kExpC4);
  p = __builtin_fmaf(p, r, kExpC3);
  p = __builtin_fmaf(p, r, kExpC2);
  p = __builtin_fmaf(p, r, 1.f);
  p = __builtin_fmaf(p, r, 1.f);
  return ldexpf(p, int(k));
}

// The special cases (NaN, out of range, non-positive logarithms) are
// selected after the main evaluation, which runs on an in-range stand-in,
// so a wavefront never branches; in-range results are unchanged.
__device__ __forceinline__ float pgx_exp(float x) {
  const float xc = fminf(fmaxf(x, kExpLo), kExpHi);  // == x in range
  float y = pgx_exp_core(xc);
  y = (x > kExpHi) ? __builtin_inff() : y;
  y = (x < kExpLo) ? 0.f : y;
  return (x != x) ? x : y;
}

// pgx_exp for arguments <= 0 or NaN (a value minus the maximum it was taken
// with, as in a log-sum-exp): the same result in fewer instructions, since
// the upper clamp and the overflow case cannot apply.
__device__ __forceinline__ float pgx_exp_nonpos(float x) {
  float y = pgx_exp_core(fmaxf(x, kExpLo));
  y = (x < kExpLo) ? 0.f : y;
  return (x != x) ? x : y;
}

// pgx_exp for arguments <= 0 or -inf, not NaN, without any select: below
// kExpLo the argument is raised to -104 instead of kExpLo, where the core's
// result k = -150, p = e^r <= 1 scales to at most 2^-150, which rounds to 0
// -- the value pgx_exp returns there. (On [-104, kExpLo) the core sees the
// argument itself: k = -150 and r <= 0 again, so the result is 0 as well.)
__device__ __forceinline__ float pgx_exp_shift(float x) {
  return pgx_exp_core(fmaxf(x, -104.f));
}

__device__ __forceinline__ float pgx_log(float x) {
  const bool finite_pos = x > 0.f && x < __builtin_inff();  // false for NaN
  const float y = pgx_log_pos(finite_pos ? x : 1.f);
  float z = (x == __builtin_inff()) ? x : y;
  z = (x == 0.f) ? -__builtin_inff() : z;
  z = (x < 0.f) ? __builtin_nanf("") : z;
  return (x != x) ? x : z;
}

__device__ __forceinline__ float pgx_log1p(float x) {
  if (x != x) return x;
  if (x == __builtin_inff()) return x;
  const float u = 1.f + x;
  if (u == 1.f) return x;
  return pgx_log(u) * (x / (u - 1.f));
}

__device__ __forceinline__ float pgx_expm1(float x) {
  if (x != x) return x;
  const float u = pgx_exp(x);
  if (u == 1.f) return x;
  const float um1 = u - 1.f;
  if (um1 == -1.f) return -1.f;
  if (u == __builtin_inff()) return u;
  return um1 * (x / pgx_log(u));
}

}  // namespace pgx

#endif  // PGX_MATH_H_
