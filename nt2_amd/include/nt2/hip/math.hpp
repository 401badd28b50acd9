// Device math kernels of the HIP architecture site, for gfx950.
//
// Each function restates, for one lane, the Boost.SIMD / NT2 algorithm the
// reference evaluates on CPU packs (file:line cited per function, paths
// relative to the reference tree).  Two evaluation policies share the code:
//
//  * exact_policy: every NT2 fma(a,b,c) is a*b+c with two roundings, as NT2
//    computes it on x86 AVX (boost/simd/arithmetic/functions/generic/fma.hpp:32),
//    and no other operation is contracted.  Results are bit-identical to the
//    CPU evaluation for the same inputs.
//  * fast_policy: NT2's fma/fnms call sites (Horner steps included) use the
//    hardware fused multiply-add; everything else is unchanged.  Results stay
//    within each function's documented ULP bound.
//
// Every function body opens with NT2H_STRICT so that the compiler never fuses
// an operation on its own, whatever -ffp-contract the including file uses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "detail/pio2_tables.hpp"

#define NT2H_FN __device__ __forceinline__
#define NT2H_STRICT _Pragma("clang fp contract(off)")

namespace nt2 { namespace hip {

struct exact_policy {
  static constexpr bool fused = false;
};
struct fast_policy {
  static constexpr bool fused = true;
};

namespace dm {

NT2H_FN uint32_t bits(float f) { return __builtin_bit_cast(uint32_t, f); }
NT2H_FN uint64_t bits(double d) { return __builtin_bit_cast(uint64_t, d); }
NT2H_FN float as_f(uint32_t b) { return __builtin_bit_cast(float, b); }
NT2H_FN double as_d(uint64_t b) { return __builtin_bit_cast(double, b); }

#define NT2H_F(b) __builtin_bit_cast(float, (uint32_t)(b##u))
#define NT2H_D(b) __builtin_bit_cast(double, (uint64_t)(b##ull))

// NT2 fma(a,b,c) and fnms(a,b,c) = c - a*b under the policy.
template <class P> NT2H_FN float fma(float a, float b, float c) {
  NT2H_STRICT
  if constexpr (P::fused) return __builtin_fmaf(a, b, c);
  else return a * b + c;
}
template <class P> NT2H_FN double fma(double a, double b, double c) {
  NT2H_STRICT
  if constexpr (P::fused) return __builtin_fma(a, b, c);
  else return a * b + c;
}
template <class P> NT2H_FN float fnms(float a, float b, float c) {
  NT2H_STRICT
  if constexpr (P::fused) return __builtin_fmaf(-a, b, c);
  else return c - a * b;
}
template <class P> NT2H_FN double fnms(double a, double b, double c) {
  NT2H_STRICT
  if constexpr (P::fused) return __builtin_fma(-a, b, c);
  else return c - a * b;
}

// Static Horner, coefficients in decreasing powers; each step is NT2's
// fma(x, acc, c): modules/core/polynomials/include/nt2/polynomials/functions/
// scalar/impl/horner.hpp:13-45.
template <class P, class T, int K>
NT2H_FN T horner(const T (&c)[K], T x) {
  NT2H_STRICT
  T acc = c[0];
#pragma unroll
  for (int i = 1; i < K; ++i) acc = fma<P>(x, acc, c[i]);
  return acc;
}

// ---------------------------------------------------------------------------
// exp.  f32: modules/core/exponential/include/nt2/exponential/functions/scalar/
// impl/expo/f_expo_reduction.hpp:58-79, expo_scale.hpp:26-35,
// simd/common/impl/expo/expo_base.hpp:47-60.
// ---------------------------------------------------------------------------
template <class P>
NT2H_FN float exp_f(float a0) {
  NT2H_STRICT
  constexpr float cExp[5] = {NT2H_F(0x3ab778cf), NT2H_F(0x3c098d8b), NT2H_F(0x3d2aa957),
                             NT2H_F(0x3e2aa9a5), NT2H_F(0x3f000000)};
  const float k = __builtin_rintf(NT2H_F(0x3fb8aa3b) * a0);
  float x = fnms<P>(k, NT2H_F(0x3f318000), a0);
  x = fnms<P>(k, NT2H_F(0xb95e8083), x);
  const float y = horner<P>(cExp, x);
  float r = fma<P>(y, x * x, x) + 1.0f;
  const float kc = __builtin_fminf(__builtin_fmaxf(k, -200.0f), 200.0f);
  r = r * as_f((uint32_t)((int32_t)kc + 127) << 23);
  r = (a0 <= NT2H_F(0xc2b0c0a5)) ? 0.0f : r;
  r = (a0 >= NT2H_F(0x42b0c0a5)) ? __builtin_inff() : r;
  return (a0 != a0) ? a0 : r;
}

// f64: d_expo_reduction.hpp:57-84, simd/common/impl/expo/expo_base.hpp:31-42.
template <class P>
NT2H_FN double exp_d(double a0) {
  NT2H_STRICT
  constexpr double cExp[5] = {NT2H_D(0x3e66376972bea4d0), NT2H_D(0xbebbbd41c5d26bf1),
                              NT2H_D(0x3f11566aaf25de2c), NT2H_D(0xbf66c16c16bebd93),
                              NT2H_D(0x3fc555555555553e)};
  const double k = __builtin_rint(NT2H_D(0x3ff71547652b82fe) * a0);
  const double hi = fnms<P>(k, NT2H_D(0x3fe62e42fee00000), a0);
  const double lo = k * NT2H_D(0x3dea39ef35793c76);
  const double x = hi - lo;
  const double t = x * x;
  double c = fnms<P>(t, horner<P>(cExp, t), x);
  c = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
  const double kc = __builtin_fmin(__builtin_fmax(k, -2000.0), 2000.0);
  c = c * as_d((uint64_t)((int64_t)(int32_t)kc + 1023) << 52);  // |kc| <= 2000: an int32 conversion
  c = (a0 <= NT2H_D(0xc086232bdd7abcd2)) ? 0.0 : c;
  c = (a0 >= NT2H_D(0x40862e42fefa39ef)) ? __builtin_inf() : c;
  return (a0 != a0) ? a0 : c;
}

// ---------------------------------------------------------------------------
// log.  f32: simd/common/impl/logs/f_log.hpp:101-125, :173-186 (finalize),
// scalar/impl/logs/f_kernel.hpp:47-78; fast_frexp:
// boost/simd/ieee/functions/simd/common/fast_frexp.hpp:33-40.
// ---------------------------------------------------------------------------
template <class P>
NT2H_FN float log_f(float a0) {
  NT2H_STRICT
  constexpr float cLog[8] = {NT2H_F(0xbda5dff0), NT2H_F(0x3e0229f9), NT2H_F(0xbe04d6b7),
                             NT2H_F(0x3e116e80), NT2H_F(0xbe2a6aa0), NT2H_F(0x3e4cd0a3),
                             NT2H_F(0xbe800064), NT2H_F(0x3eaaaaa9)};
  float z = a0, t = 0.0f;
  const bool den = __builtin_fabsf(z) < NT2H_F(0x00800000);
  z = den ? z * NT2H_F(0x4B000000) : z;
  t = den ? NT2H_F(0xc17f1402) : t;
  const uint32_t b = bits(z);
  const int32_t e = (int32_t)((b & 0x7f800000u) >> 23) - 126;
  const float m = as_f((b & ~0x7f800000u) | 0x3f000000u);
  const bool lt = m < NT2H_F(0x3f3504f3);
  const float fe = lt ? (float)e + (-1.0f) : (float)e;
  const float x = (lt ? m + m : m) - 1.0f;
  const float x2 = x * x;
  float y = (horner<P>(cLog, x) * x) * x2;
  y = fma<P>(fe, NT2H_F(0xb95e8083), y);
  y = fma<P>(-0.5f, x2, y);
  float r = fma<P>(NT2H_F(0x3f318000), fe, (x + y) + t);
  r = (a0 < 0.0f || a0 != a0) ? __builtin_nanf("") : r;
  r = (a0 == __builtin_inff()) ? a0 : r;
  return (a0 == 0.0f) ? -__builtin_inff() : r;
}

// f64: simd/common/impl/logs/d_log.hpp:60-97.
template <class P>
NT2H_FN double log_d(double a0) {
  NT2H_STRICT
  constexpr double cT1[3] = {NT2H_D(0x3fc39a09d078c69f), NT2H_D(0x3fcc71c51d8e78af),
                             NT2H_D(0x3fd999999997fa04)};
  constexpr double cT2[4] = {NT2H_D(0x3fc2f112df3e5244), NT2H_D(0x3fc7466496cb03de),
                             NT2H_D(0x3fd2492494229359), NT2H_D(0x3fe5555555555593)};
  const uint64_t b = bits(a0);
  int64_t k = (int64_t)((b & 0x7ff0000000000000ull) >> 52) - 1022;
  const double x = as_d((b & ~0x7ff0000000000000ull) | 0x3fe0000000000000ull);
  const bool lt = NT2H_D(0x3FE6A09E667F3BCD) > x;
  k = k + (lt ? -1 : 0);
  const double f = (x + (lt ? x : 0.0)) - 1.0;
  const double dk = (double)k;
  const double s = f / (2.0 + f);
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * horner<P>(cT1, w);
  const double t2 = z * horner<P>(cT2, w);
  const double R = t2 + t1;
  const double hfsq = 0.5 * (f * f);
  double r = dk * NT2H_D(0x3fe62e42fee00000) -
             ((hfsq - (s * (hfsq + R) + dk * NT2H_D(0x3dea39ef35793c76))) - f);
  r = (a0 < 0.0 || a0 != a0) ? __builtin_nan("") : r;
  r = (a0 == __builtin_inf()) ? a0 : r;
  return (a0 == 0.0) ? -__builtin_inf() : r;
}

// ---------------------------------------------------------------------------
// exp2, exp10 (NT2's expa with the two_tag / ten_tag reductions):
// f32 scalar/impl/expo/f_expo_reduction.hpp:90-167, f64 d_expo_reduction.hpp:87-166,
// both finished as simd/common/impl/expo/expo_base.hpp:31-61 (the scale of
// expo_scale.hpp:26-35, zero at or below Minlog*, Inf at or above Maxlog*).
// ---------------------------------------------------------------------------
NT2H_FN float expa_finish_f(float a0, float r, float k, float minlog, float maxlog) {
  NT2H_STRICT
  const float kc = __builtin_fminf(__builtin_fmaxf(k, -200.0f), 200.0f);
  r = r * as_f((uint32_t)((int32_t)kc + 127) << 23);
  r = (a0 <= minlog) ? 0.0f : r;
  r = (a0 >= maxlog) ? __builtin_inff() : r;
  return (a0 != a0) ? a0 : r;
}
NT2H_FN double expa_finish_d(double a0, double c, double k, double minlog, double maxlog) {
  NT2H_STRICT
  const double kc = __builtin_fmin(__builtin_fmax(k, -2000.0), 2000.0);
  c = c * as_d((uint64_t)((int64_t)(int32_t)kc + 1023) << 52);  // |kc| <= 2000: an int32 conversion
  c = (a0 <= minlog) ? 0.0 : c;
  c = (a0 >= maxlog) ? __builtin_inf() : c;
  return (a0 != a0) ? a0 : c;
}
template <class P>
NT2H_FN float exp2_f(float a0) {
  NT2H_STRICT
  constexpr float c[5] = {NT2H_F(0x3922c8c4), NT2H_F(0x3ab01218), NT2H_F(0x3c1d9422),
                          NT2H_F(0x3d6356eb), NT2H_F(0x3e75fdf1)};
  const float k = __builtin_rintf(a0);
  const float x = a0 - k;
  const float y = horner<P>(c, x);
  const float r = fma<P>(y, x * x, x * NT2H_F(0x3f317218)) + 1.0f;
  return expa_finish_f(a0, r, k, NT2H_F(0xc2fe0000), NT2H_F(0x42fe0000));
}
template <class P>
NT2H_FN float exp10_f(float a0) {
  NT2H_STRICT
  constexpr float c[6] = {NT2H_F(0x3e54dff1), NT2H_F(0x3f0aacef), NT2H_F(0x3f95eb4c),
                          NT2H_F(0x400237da), NT2H_F(0x4029a926), NT2H_F(0x40135d8e)};
  const float k = __builtin_rintf(NT2H_F(0x40549a78) * a0);
  float x = fnms<P>(k, NT2H_F(0x3e9a0000), a0);
  x = x - k * NT2H_F(0x39826a14);
  const float r = horner<P>(c, x) * x + 1.0f;
  return expa_finish_f(a0, r, k, NT2H_F(0xc2179999), NT2H_F(0x4218ec59));
}
template <class P>
NT2H_FN double exp2_d(double a0) {
  NT2H_STRICT
  constexpr double cExp[5] = {NT2H_D(0x3e66376972bea4d0), NT2H_D(0xbebbbd41c5d26bf1),
                              NT2H_D(0x3f11566aaf25de2c), NT2H_D(0xbf66c16c16bebd93),
                              NT2H_D(0x3fc555555555553e)};
  const double k = __builtin_rint(a0);
  const double x = (a0 - k) * NT2H_D(0x3fe62e42fefa39ef);
  const double t = x * x;
  double c = fnms<P>(t, horner<P>(cExp, t), x);
  c = 1.0 - ((-(x * c) / (2.0 - c)) - x);
  return expa_finish_d(a0, c, k, NT2H_D(0xc08ff00000000000), NT2H_D(0x408ff80000000000));
}
template <class P>
NT2H_FN double exp10_d(double a0) {
  NT2H_STRICT
  constexpr double cP[4] = {NT2H_D(0x3fa4fd75f3062dd4), NT2H_D(0x40277d9474c55934),
                            NT2H_D(0x40796b7a050349e4), NT2H_D(0x40a2b4798e134a01)};
  constexpr double cQ[4] = {NT2H_D(0x3ff0000000000000), NT2H_D(0x405545fdce51ca08),
                            NT2H_D(0x4093e05eefd67782), NT2H_D(0x40a03f37650df6e2)};
  const double k = __builtin_rint(NT2H_D(0x400a934f0979a372) * a0);
  double x = fnms<P>(k, NT2H_D(0x3fd3440000000000), a0);
  x = fnms<P>(k, NT2H_D(0x3ed3509f79fef312), x);
  const double xx = x * x;
  const double px = x * horner<P>(cP, xx);
  const double x2 = px / (horner<P>(cQ, xx) - px);
  const double c = 1.0 + (x2 + x2);
  return expa_finish_d(a0, c, k, NT2H_D(0xc0734413509f79fe), NT2H_D(0x40734413509f79fe));
}

// ---------------------------------------------------------------------------
// log2, log10.  f32 simd/common/impl/logs/f_log.hpp:128-174 on the kernel of
// scalar/impl/logs/f_kernel.hpp:47-78 (denormals prescaled by 2^23 with the
// matching constant added); f64 simd/common/impl/logs/d_log.hpp:58-110.
// The results are finished like log (NaN below 0 and for NaN, Inf at Inf,
// -Inf at 0).
// ---------------------------------------------------------------------------
template <class P>
NT2H_FN void log_kernel_f(float z, float& fe, float& x, float& x2, float& y) {
  NT2H_STRICT
  constexpr float cLog[8] = {NT2H_F(0xbda5dff0), NT2H_F(0x3e0229f9), NT2H_F(0xbe04d6b7),
                             NT2H_F(0x3e116e80), NT2H_F(0xbe2a6aa0), NT2H_F(0x3e4cd0a3),
                             NT2H_F(0xbe800064), NT2H_F(0x3eaaaaa9)};
  const uint32_t b = bits(z);
  const int32_t e = (int32_t)((b & 0x7f800000u) >> 23) - 126;
  const float m = as_f((b & ~0x7f800000u) | 0x3f000000u);
  const bool lt = m < NT2H_F(0x3f3504f3);
  fe = lt ? (float)e + (-1.0f) : (float)e;
  x = (lt ? m + m : m) - 1.0f;
  x2 = x * x;
  y = (horner<P>(cLog, x) * x) * x2;
}
NT2H_FN float log_finish_f(float a0, float r) {
  r = (a0 < 0.0f || a0 != a0) ? __builtin_nanf("") : r;
  r = (a0 == __builtin_inff()) ? a0 : r;
  return (a0 == 0.0f) ? -__builtin_inff() : r;
}
template <class P>
NT2H_FN float log2_f(float a0) {
  NT2H_STRICT
  const bool den = __builtin_fabsf(a0) < NT2H_F(0x00800000);
  const float z0 = den ? a0 * NT2H_F(0x4B000000) : a0;
  const float t = den ? NT2H_F(0xc1b80000) : 0.0f;
  float fe, x, x2, y;
  log_kernel_f<P>(z0, fe, x, x2, y);
  y = fma<P>(-0.5f, x2, y);
  const float z = fma<P>(x, NT2H_F(0x3ee2a8ed), y * NT2H_F(0x3ee2a8ed));
  return log_finish_f(a0, (((z + y) + x) + fe) + t);
}
template <class P>
NT2H_FN float log10_f(float a0) {
  NT2H_STRICT
  const bool den = __builtin_fabsf(a0) < NT2H_F(0x00800000);
  const float z0 = den ? a0 * NT2H_F(0x4B000000) : a0;
  const float t = den ? NT2H_F(0xc0dd8ede) : 0.0f;
  float fe, x, x2, y;
  log_kernel_f<P>(z0, fe, x, x2, y);
  y = fma<P>(x2, -0.5f, y);
  float z = (x + y) * NT2H_F(0x3a37b152);
  z = fma<P>(y, NT2H_F(0x3ede0000), z);
  z = fma<P>(x, NT2H_F(0x3ede0000), z);
  z = fma<P>(fe, NT2H_F(0x3e9a0000), z);
  return log_finish_f(a0, fma<P>(fe, NT2H_F(0x39826a14), z + t));
}
// f64: the kernel of log_d, then -(hfsq - s (hfsq + R) - f) scaled
template <class P>
NT2H_FN double log_poly_d(double a0, double& dk) {
  NT2H_STRICT
  constexpr double cT1[3] = {NT2H_D(0x3fc39a09d078c69f), NT2H_D(0x3fcc71c51d8e78af),
                             NT2H_D(0x3fd999999997fa04)};
  constexpr double cT2[4] = {NT2H_D(0x3fc2f112df3e5244), NT2H_D(0x3fc7466496cb03de),
                             NT2H_D(0x3fd2492494229359), NT2H_D(0x3fe5555555555593)};
  const uint64_t b = bits(a0);
  int64_t k = (int64_t)((b & 0x7ff0000000000000ull) >> 52) - 1022;
  const double x = as_d((b & ~0x7ff0000000000000ull) | 0x3fe0000000000000ull);
  const bool lt = NT2H_D(0x3FE6A09E667F3BCD) > x;
  k = k + (lt ? -1 : 0);
  const double f = (x + (lt ? x : 0.0)) - 1.0;
  dk = (double)k;
  const double s = f / (2.0 + f);
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * horner<P>(cT1, w);
  const double t2 = z * horner<P>(cT2, w);
  const double R = t2 + t1;
  const double hfsq = 0.5 * (f * f);
  return -((hfsq - (s * (hfsq + R))) - f);
}
NT2H_FN double log_finish_d(double a0, double r) {
  r = (a0 < 0.0 || a0 != a0) ? __builtin_nan("") : r;
  r = (a0 == __builtin_inf()) ? a0 : r;
  return (a0 == 0.0) ? -__builtin_inf() : r;
}
template <class P>
NT2H_FN double log2_d(double a0) {
  NT2H_STRICT
  double dk;
  const double u = log_poly_d<P>(a0, dk);
  return log_finish_d(a0, u * NT2H_D(0x3ff71547652b82fe) + dk);
}
template <class P>
NT2H_FN double log10_d(double a0) {
  NT2H_STRICT
  double dk;
  const double u = log_poly_d<P>(a0, dk);
  return log_finish_d(a0, u * NT2H_D(0x3fdbcb7b1526e50d) + dk * NT2H_D(0x3fd34413509f79ff));
}

// ---------------------------------------------------------------------------
// rounding and sign.  floor, ceil, trunc, round2even are the IEEE operations;
// round is half away from zero as boost/simd/arithmetic/functions/simd/common/round.hpp:45-52
// computes it (c = ceil(|a|), c - 1 when c - 1/2 > |a|, the sign of a);
// sign is (a > 0) - (a < 0), NaN for NaN
// (boost/simd/ieee/functions/simd/common/sign.hpp:55-66).
// ---------------------------------------------------------------------------
NT2H_FN float round_nt2(float a0) {
  NT2H_STRICT
  const float v = __builtin_fabsf(a0);
  const float c = __builtin_ceilf(v);
  const float r = (c - 0.5f > v) ? c - 1.0f : c;
  return __builtin_copysignf(r, a0);
}
NT2H_FN double round_nt2(double a0) {
  NT2H_STRICT
  const double v = __builtin_fabs(a0);
  const double c = __builtin_ceil(v);
  const double r = (c - 0.5 > v) ? c - 1.0 : c;
  return __builtin_copysign(r, a0);
}
NT2H_FN float sign_nt2(float a0) {
  const float r = (float)(a0 > 0.0f) - (float)(a0 < 0.0f);
  return (a0 != a0) ? as_f(0xffffffffu) : r;
}
NT2H_FN double sign_nt2(double a0) {
  const double r = (double)(a0 > 0.0) - (double)(a0 < 0.0);
  return (a0 != a0) ? as_d(0xffffffffffffffffull) : r;
}

// ---------------------------------------------------------------------------
// expm1: modules/core/exponential/include/nt2/exponential/functions/details/
// expm1_kernel.hpp:26-50 (f32) and :53-85 (f64), with
// simd/common/expm1.hpp:29-37 (-1 below Logeps, Inf above Maxlog).  The scale
// factors 2^-k and 2^k are built in the exponent field as NT2 does
// (shl(Maxexponent - k, Nbmantissabits), fast_ldexp), including its values at
// the top of the range (k = 128 / 1024 makes 2^-k -Inf and the result Inf).
// k is clamped before the integer conversion: lanes where that matters are
// replaced by the -1 / Inf / NaN selects.
// ---------------------------------------------------------------------------
template <class P>
NT2H_FN float expm1_f(float a0) {
  NT2H_STRICT
  constexpr float c[3] = {NT2H_F(0x3acf6db4), NT2H_F(0xbd08887f), NT2H_F(0x3f800000)};
  const float k = __builtin_rintf(NT2H_F(0x3fb8aa3b) * a0);
  float x = fnms<P>(k, NT2H_F(0x3f318000), a0);
  x = fnms<P>(k, NT2H_F(0xb95e8083), x);
  const float hx = x * 0.5f;
  const float hxs = x * hx;
  const float r1 = horner<P>(c, hxs);
  const float t = fnms<P>(r1, hx, 3.0f);
  float e = hxs * ((r1 - t) / (6.0f - x * t));
  e = fma<P>(x, e, -hxs);
  const int32_t ik = (k == k) ? (int32_t)__builtin_fminf(__builtin_fmaxf(k, -200.0f), 200.0f) : 0;
  const float two2mk = as_f((uint32_t)(127 - ik) << 23);
  const float y = (1.0f - two2mk) - (e - x);
  float r = y * as_f((uint32_t)(ik + 127) << 23);
  r = (a0 > NT2H_F(0x42b0c0a5)) ? __builtin_inff() : r;
  return (a0 < NT2H_F(0xc17f1402)) ? -1.0f : r;
}
template <class P>
NT2H_FN double expm1_d(double a0) {
  NT2H_STRICT
  constexpr double c[6] = {NT2H_D(0xBE8AFDB76E09C32D), NT2H_D(0x3ED0CFCA86E65239),
                           NT2H_D(0xBF14CE199EAADBB7), NT2H_D(0x3F5A01A019FE5585),
                           NT2H_D(0xBFA11111111110F4), NT2H_D(0x3ff0000000000000)};
  const double k = __builtin_rint(NT2H_D(0x3ff71547652b82fe) * a0);
  const double hi = fnms<P>(k, NT2H_D(0x3fe62e42fee00000), a0);
  const double lo = k * NT2H_D(0x3dea39ef35793c76);
  const double x = hi - lo;
  const double hxs = (x * x) * 0.5;
  const double r1 = horner<P>(c, hxs);
  const double t = 3.0 - (r1 * 0.5) * x;
  double e = hxs * ((r1 - t) / (6.0 - x * t));
  const double cc = (hi - x) - lo;
  e = (x * (e - cc) - cc) - hxs;
  const int64_t ik = (k == k) ? (int64_t)(int32_t)__builtin_fmin(__builtin_fmax(k, -2000.0), 2000.0) : 0;
  const double two2mk = as_d((uint64_t)(1023 - ik) << 52);
  const double ct1 = (1.0 - two2mk) - (e - x);
  const double ct2 = 1.0 + (x - (e + two2mk));
  const double y = (k < 20.0) ? ct1 : ct2;
  double r = y * as_d((uint64_t)(ik + 1023) << 52);
  r = (a0 > NT2H_D(0x40862e42fefa39ef)) ? __builtin_inf() : r;
  return (a0 < NT2H_D(0xc04205966f2b4f19)) ? -1.0 : r;
}

// log1p: simd/common/log1p.hpp:26-37, u = 1 + a0 and log(u) corrected by
// (a0 - (u - 1))/u where u != 0, Inf at Inf (the scalar form,
// scalar/log1p.hpp:27-37, returns the same values)
template <class P>
NT2H_FN float log1p_f(float a0) {
  NT2H_STRICT
  const float u = 1.0f + a0;
  const float l = log_f<P>(u);
  const float r = (u != 0.0f) ? l + (a0 - (u - 1.0f)) / u : l;
  return (u == __builtin_inff()) ? u : r;
}
template <class P>
NT2H_FN double log1p_d(double a0) {
  NT2H_STRICT
  const double u = 1.0 + a0;
  const double l = log_d<P>(u);
  const double r = (u != 0.0) ? l + (a0 - (u - 1.0)) / u : l;
  return (u == __builtin_inf()) ? u : r;
}

// tanh: modules/core/hyperbolic/include/nt2/hyperbolic/functions/simd/common/tanh.hpp:27-45:
// |a0| < 5/8 by the Cephes kernel (details/tanh_kernel.hpp:24-36 f32
// polynomial, :44-71 f64 rational) on |a0|, else 1 - 2/(1 + exp(2|a0|)); the
// sign of a0 is xored in.  The choice is per lane, so a lane's bits do not
// depend on its pack.
// the kernels on |x| < 5/8 (x2 = x*x): tanh_kernel<float>::tanh, and the f64
// rational's numerator x2 p(x2) and denominator q(x2), shared with coth
template <class P>
NT2H_FN float tanh_small_f(float x, float x2) {
  NT2H_STRICT
  constexpr float c[5] = {NT2H_F(0xbbbaf0ea), NT2H_F(0x3ca9134e), NT2H_F(0xbd5c1e2d),
                          NT2H_F(0x3e088393), NT2H_F(0xbeaaaa99)};
  return fma<P>(horner<P>(c, x2) * x2, x, x);
}
template <class P>
NT2H_FN void tanh_pq_d(double x2, double& p_, double& q_) {
  NT2H_STRICT
  constexpr double p[3] = {NT2H_D(0xbfeedc5baafd6f4b), NT2H_D(0xc058d26a0e26682d),
                           NT2H_D(0xc0993ac030580563)};
  constexpr double q[4] = {NT2H_D(0x3FF0000000000000), NT2H_D(0x405c33f28a581b86),
                           NT2H_D(0x40a176fa0e5535fa), NT2H_D(0x40b2ec102442040c)};
  p_ = horner<P>(p, x2);
  q_ = horner<P>(q, x2);
}
template <class P>
NT2H_FN float tanh_f(float a0) {
  NT2H_STRICT
  const float x = __builtin_fabsf(a0);
  float r;
  if (x < 0.625f) {
    r = tanh_small_f<P>(x, x * x);
  } else {
    r = fma<P>(-2.0f, 1.0f / (1.0f + exp_f<P>(x + x)), 1.0f);
  }
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
template <class P>
NT2H_FN double tanh_d(double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0);
  double r;
  if (x < 0.625) {
    const double x2 = x * x;
    double p, q;
    tanh_pq_d<P>(x2, p, q);
    r = fma<P>((x2 * p) / q, x, x);
  } else {
    r = fma<P>(-2.0, 1.0 / (1.0 + exp_d<P>(x + x)), 1.0);
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}

// sinh: modules/core/hyperbolic/include/nt2/hyperbolic/functions/simd/common/sinh.hpp:27-47
// (|a0| < 1: details/sinh_kernel.hpp:25-37 f32 polynomial, :44-66 f64
// rational; above: (e - 1/e)/2 with e = exp(|a0|), or (e/2) e with
// e = exp(|a0|/2) above Maxlog - Log_2), sign xored in; the scalar form
// (scalar/sinh.hpp:27-56) gives the same values.
// cosh: generic/cosh.hpp:27-45, average(e, 1/e) = e/2 + (1/e)/2, or (e/2) e.
template <class P>
NT2H_FN float sinh_f(float a0) {
  NT2H_STRICT
  constexpr float c[4] = {NT2H_F(0x39559e2f), NT2H_F(0x3c087bbe), NT2H_F(0x3e2aaacc), NT2H_F(0x3f800000)};
  const float x = __builtin_fabsf(a0);
  float r;
  if (x < 1.0f) {
    r = horner<P>(c, x * x) * x;
  } else {
    const bool big = x > NT2H_F(0x42b0c0a5) - NT2H_F(0x3f317218);
    const float e = exp_f<P>(x * (big ? 0.5f : 1.0f));
    const float h = 0.5f * e;
    r = big ? h * e : h - 0.5f * (1.0f / e);
  }
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
template <class P>
NT2H_FN double sinh_d(double a0) {
  NT2H_STRICT
  constexpr double p[4] = {NT2H_D(0xbfe9435fe8bb3cd6), NT2H_D(0xc064773a398ff4fe),
                           NT2H_D(0xc0c694b8c71d6182), NT2H_D(0xc115782bdbf6ab05)};
  constexpr double q[4] = {NT2H_D(0x3ff0000000000000), NT2H_D(0xc0715b6096e96484),
                           NT2H_D(0x40e1a7ba7ed72245), NT2H_D(0xc1401a20e4f90044)};
  const double x = __builtin_fabs(a0);
  double r;
  if (x < 1.0) {
    const double x2 = x * x;
    r = fma<P>(x, (x2 * horner<P>(p, x2)) / horner<P>(q, x2), x);
  } else {
    const bool big = x > NT2H_D(0x40862e42fefa39ef) - NT2H_D(0x3fe62e42fefa39ef);
    const double e = exp_d<P>(x * (big ? 0.5 : 1.0));
    const double h = 0.5 * e;
    r = big ? h * e : h - 0.5 * (1.0 / e);
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
template <class P>
NT2H_FN float cosh_f(float a0) {
  NT2H_STRICT
  const float x = __builtin_fabsf(a0);
  const bool big = x > NT2H_F(0x42b0c0a5) - NT2H_F(0x3f317218);
  const float e = exp_f<P>(x * (big ? 0.5f : 1.0f));
  return big ? (0.5f * e) * e : fma<P>(e, 0.5f, (1.0f / e) * 0.5f);
}
template <class P>
NT2H_FN double cosh_d(double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0);
  const bool big = x > NT2H_D(0x40862e42fefa39ef) - NT2H_D(0x3fe62e42fefa39ef);
  const double e = exp_d<P>(x * (big ? 0.5 : 1.0));
  return big ? (0.5 * e) * e : fma<P>(e, 0.5, (1.0 / e) * 0.5);
}

// atan: modules/core/trigonometric/include/nt2/trigonometric/functions/
// simd/common/impl/invtrig/f_invtrig.hpp:98-133 and d_invtrig.hpp:128-166 (packs):
// |a0| below tan(pi/8) directly, up to tan(3pi/8) by (x-1)/(x+1) + pi/4, above
// by -1/x + pi/2, the pi/4 and pi/2 low parts added to the polynomial's
// result; sign xored in.  The scalar forms that NT2 uses for a pack tail
// (scalar/impl/invtrig/f_invtrig.hpp:92-128, d_invtrig.hpp:112-159) take the
// boundaries as strict and, for f64, switch to (x-1)/(x+1) only above 0.66:
// Scalar selects them.
template <class P, bool Scalar>
NT2H_FN float atan_f(float a0) {
  NT2H_STRICT
  constexpr float c[4] = {NT2H_F(0x3da4f0d1), NT2H_F(0xbe0e1b85), NT2H_F(0x3e4c925f), NT2H_F(0xbeaaaa2a)};
  const float t3 = NT2H_F(0x401a827a), t1 = NT2H_F(0x3ed413cd);
  const float pio2 = NT2H_F(0x3fc90fdb), pio4 = NT2H_F(0x3f490fdb);
  const float pio2lo = NT2H_F(0xb33bbd2e), pio4lo = NT2H_F(0xb2bbbd2e);
  const float x = __builtin_fabsf(a0);
  float r;
  if constexpr (Scalar) {
    if (a0 == 0.0f) {
      r = 0.0f;
    } else if (x == __builtin_inff()) {
      r = pio2;
    } else {
      float y = 0.0f, more = 0.0f, xx = x;
      if (x > t3) {
        y = pio2; more = pio2lo; xx = -(1.0f / x);
      } else if (x > t1) {
        y = pio4; more = pio4lo; xx = (x - 1.0f) / (x + 1.0f);
      }
      const float z = xx * xx;
      const float z1 = fma<P>(xx, horner<P>(c, z) * z, xx);
      r = y + (z1 + more);
    }
  } else {
    const bool f1 = x < t3;
    const bool f2 = (x >= t1) && f1;
    const float yy = f2 ? pio4 : (f1 ? 0.0f : pio2);
    const float xx = f2 ? (x - 1.0f) / (x + 1.0f) : (f1 ? x : -(1.0f / x));
    const float z = xx * xx;
    float z1 = fma<P>(xx, horner<P>(c, z) * z, xx);
    z1 = f2 ? z1 + pio4lo : z1;
    z1 = f1 ? z1 : z1 + pio2lo;
    r = yy + z1;
  }
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
template <class P, bool Scalar>
NT2H_FN double atan_d(double a0) {
  NT2H_STRICT
  constexpr double p[5] = {NT2H_D(0xbfec007fa1f72594), NT2H_D(0xc03028545b6b807a),
                           NT2H_D(0xc052c08c36880273), NT2H_D(0xc05eb8bf2d05ba25),
                           NT2H_D(0xc0503669fd28ec8e)};
  constexpr double q[6] = {NT2H_D(0x3ff0000000000000), NT2H_D(0x4038dbc45b14603c),
                           NT2H_D(0x4064a0dd43b8fa25), NT2H_D(0x407b0e18d2e2be3b),
                           NT2H_D(0x407e563f13b049ea), NT2H_D(0x4068519efbbd62ec)};
  const double t3 = NT2H_D(0x4003504f333f9de6), t1 = NT2H_D(0x3fda827999fcef31);
  const double pio2 = NT2H_D(0x3ff921fb54442d18), pio4 = NT2H_D(0x3fe921fb54442d18);
  const double pio2lo = NT2H_D(0x3c91a62633145c07), pio4lo = NT2H_D(0x3c81a62633145c07);
  const double x = __builtin_fabs(a0);
  double r;
  if constexpr (Scalar) {
    if (a0 == 0.0) {
      r = 0.0;
    } else if (x == __builtin_inf()) {
      r = pio2;
    } else {
      double y, flag = (x > t3) ? 1.0 : 0.0, xx = x;
      if (x > t3) {
        y = pio2; xx = -(1.0 / x);
      } else if (x <= NT2H_D(0x3fe51eb851eb851f)) {
        y = 0.0;
      } else {
        y = pio4; flag = 0.5; xx = (x - 1.0) / (x + 1.0);
      }
      const double z = xx * xx;
      double zz = (z * horner<P>(p, z)) / horner<P>(q, z);
      zz = fma<P>(xx, zz, xx);
      zz = zz + flag * pio2lo;
      r = y + zz;
    }
  } else {
    const bool f1 = x < t3;
    const bool f2 = (x >= t1) && f1;
    const double yy = f2 ? pio4 : (f1 ? 0.0 : pio2);
    const double xx = f2 ? (x - 1.0) / (x + 1.0) : (f1 ? x : -(1.0 / x));
    const double z = xx * xx;
    double zz = (z * horner<P>(p, z)) / horner<P>(q, z);
    zz = fma<P>(xx, zz, xx);
    zz = f2 ? zz + pio4lo : zz;
    zz = zz + (f1 ? 0.0 : pio2lo);
    r = yy + zz;
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}

// atan2(a0, a1): simd/common/atan2.hpp:30-54 (packs) and scalar/atan2.hpp:29-45:
// both infinite -> the unit signs; z = atan(|a0/a1|) by the kernel above,
// pi - z when a1 <= 0, the sign of a0 (signnz: -0 counts as negative);
// a0 = +-0 -> pi when a1 < 0, else +0 (both forms drop the sign of a0
// there); NaN in either -> NaN.
template <class P, bool Scalar, class T>
NT2H_FN T atan2_t(T a0, T a1) {
  NT2H_STRICT
  const T pi = sizeof(T) == 4 ? (T)NT2H_F(0x40490fdb) : (T)NT2H_D(0x400921fb54442d18);
  if (a0 != a0 || a1 != a1) return a0 + a1;
  if (__builtin_isinf(a0) && __builtin_isinf(a1)) {
    a0 = __builtin_copysign((T)1, a0);
    a1 = __builtin_copysign((T)1, a1);
  }
  T z;
  if constexpr (sizeof(T) == 4) z = atan_f<P, Scalar>(__builtin_fabsf((float)(a0 / a1)));
  else z = atan_d<P, Scalar>(__builtin_fabs((double)(a0 / a1)));
  z = (a1 > (T)0) ? z : pi - z;
  z = __builtin_signbit(a0) ? -z : z;
  return (a0 == (T)0) ? ((a1 < (T)0) ? pi : (T)0) : z;
}
template <class P, bool Scalar> NT2H_FN float atan2_f(float a0, float a1) { return atan2_t<P, Scalar, float>(a0, a1); }
template <class P, bool Scalar> NT2H_FN double atan2_d(double a0, double a1) { return atan2_t<P, Scalar, double>(a0, a1); }

// pow(a0, a1): modules/core/exponential/include/nt2/exponential/functions/
// simd/common/pow.hpp:38-46, pow_abs.hpp:65-133 (packs), scalar/pow.hpp:30-36,
// pow_abs.hpp:51-111 (pack tail), details/pow_kernel.hpp:30-228: Cephes' pow.
// log2 of the significand against 2^(-i/16) (a 17-entry table with a 9-entry
// continuation), the product with a1 split into sixteenths, 2^fraction by a
// polynomial, ldexp as boost.simd does it (two products).  Differences of the
// forms: at a1 = +Inf the tail gives Inf for a NaN base (packs 0), at an
// infinite base the tail gives Inf for a NaN exponent (packs 1).
template <class T> struct powk;
template <> struct powk<float> {
  using I = int;
  using U = unsigned;
  static constexpr int nbm = 23, maxe = 127, mine = -126, lime = 128;
  static constexpr unsigned A[17] = {0x3f800000, 0x3f75257d, 0x3f6ac0c7, 0x3f60ccdf, 0x3f5744fd, 0x3f4e248c,
                                     0x3f45672a, 0x3f3d08a4, 0x3f3504f3, 0x3f2d583f, 0x3f25fed7, 0x3f1ef532,
                                     0x3f1837f0, 0x3f11c3d3, 0x3f0b95c2, 0x3f05aac3, 0x3f000000};
  static constexpr unsigned B[9] = {0x00000000, 0xb1c116de, 0xb254a58a, 0x318aa837, 0x324fe77a,
                                    0xb2ac9d5e, 0x32a31b71, 0xb1e0aba1, 0x00000000};
  NT2H_FN static float a(int i) { return as_f(A[i]); }
  NT2H_FN static float b(int i) { return as_f(B[i]); }
  template <class P> NT2H_FN static float pow1(float x, float z) {
    NT2H_STRICT
    constexpr float c[4] = {NT2H_F(0xbe2a61b2), NT2H_F(0x3e4d2fa3), NT2H_F(0xbe800015), NT2H_F(0x3eaaaaa3)};
    return (z * x) * horner<P>(c, x);
  }
  template <class P> NT2H_FN static float pow2(float x) {
    constexpr float c[4] = {NT2H_F(0x3c1a49bc), NT2H_F(0x3d634d38), NT2H_F(0x3e75fde1), NT2H_F(0x3f317218)};
    return horner<P>(c, x);
  }
  NT2H_FN static float log2em1() { return NT2H_F(0x3ee2a8ed); }
  NT2H_FN static float large() { return NT2H_F(0x44ffe000); }  // 2047 sixteenths
  NT2H_FN static float low() { return NT2H_F(0xc5160000); }    // -2400 sixteenths
  NT2H_FN static float tiny() { return NT2H_F(0x00800000); }
  NT2H_FN static float from_bits(U u) { return as_f(u); }
  NT2H_FN static I toint(float x) { return (x > -2147483648.0f && x < 2147483648.0f) ? (I)x : (I)0x80000000u; }
  NT2H_FN static float frexp(float x, I& e) { return __builtin_frexpf(x, &e); }
};
template <> struct powk<double> {
  using I = long long;
  using U = unsigned long long;
  static constexpr int nbm = 52, maxe = 1023, mine = -1022, lime = 1024;
  static constexpr unsigned long long A[17] = {
      0x3ff0000000000000, 0x3feea4afa2a490da, 0x3fed5818dcfba487, 0x3fec199bdd85529c, 0x3feae89f995ad3ad,
      0x3fe9c49182a3f090, 0x3fe8ace5422aa0db, 0x3fe7a11473eb0187, 0x3fe6a09e667f3bcd, 0x3fe5ab07dd485429,
      0x3fe4bfdad5362a27, 0x3fe3dea64c123422, 0x3fe306fe0a31b715, 0x3fe2387a6e756238, 0x3fe172b83c7d517b,
      0x3fe0b5586cf9890f, 0x3fe0000000000000};
  static constexpr unsigned long long B[9] = {0x0000000000000000, 0x3c72ed02d75b3707, 0x3c87a1cd345dcc81,
                                              0x3c86e9f156864b27, 0xbc8bdd3413b26456, 0x3c6d4397afec42e2,
                                              0x3c76f46ad23182e4, 0xbc719041b9d78a76, 0x0000000000000000};
  NT2H_FN static double a(int i) { return as_d(A[i]); }
  NT2H_FN static double b(int i) { return as_d(B[i]); }
  template <class P> NT2H_FN static double pow1(double x, double z) {
    NT2H_STRICT
    constexpr double p[4] = {NT2H_D(0x3fdfdb997f5b5cf0), NT2H_D(0x400dddefea9edf15), NT2H_D(0x401eccbd7f78eb6f),
                             NT2H_D(0x4012aa83b65c9b74)};
    constexpr double q[5] = {NT2H_D(0x3ff0000000000000), NT2H_D(0x4022aab49b20914e), NT2H_D(0x403bffff41c1c9f5),
                             NT2H_D(0x4040ccbc1b176402), NT2H_D(0x402bffc5918ae92e)};
    return x * ((z * horner<P>(p, x)) / horner<P>(q, x));
  }
  template <class P> NT2H_FN static double pow2(double x) {
    constexpr double c[7] = {NT2H_D(0x3eef6307d7f2937f), NT2H_D(0x3f242fbe60fc9259), NT2H_D(0x3f55d87ec84aef1d),
                             NT2H_D(0x3f83b2ab6ef133b7), NT2H_D(0x3fac6b08d7041a92), NT2H_D(0x3fcebfbdff82c56d),
                             NT2H_D(0x3fe62e42fefa39ef)};
    return horner<P>(c, x);
  }
  NT2H_FN static double log2em1() { return NT2H_D(0x3fdc551d94ae0bf8); }
  NT2H_FN static double large() { return NT2H_D(0x40cfff8000000000); }  // 16383 sixteenths
  NT2H_FN static double low() { return NT2H_D(0xc0d0c7c000000000); }    // -17183 sixteenths
  NT2H_FN static double tiny() { return NT2H_D(0x0010000000000000); }
  NT2H_FN static double from_bits(U u) { return as_d(u); }
  NT2H_FN static I toint(double x) {
    return (x > -9223372036854775808.0 && x < 9223372036854775808.0) ? (I)x : (I)0x8000000000000000ull;
  }
  NT2H_FN static double frexp(double x, I& e) {
    int ei;
    const double m = __builtin_frexp(x, &ei);
    e = ei;
    return m;
  }
};
NT2H_FN float pw_floor(float x) { return __builtin_floorf(x); }
NT2H_FN double pw_floor(double x) { return __builtin_floor(x); }
NT2H_FN float pw_trunc(float x) { return __builtin_truncf(x); }
NT2H_FN double pw_trunc(double x) { return __builtin_trunc(x); }
NT2H_FN float pw_abs(float x) { return __builtin_fabsf(x); }
NT2H_FN double pw_abs(double x) { return __builtin_fabs(x); }
template <class T> NT2H_FN T pw_reduc(T x) { return T(0.0625) * pw_floor(T(16) * x); }
template <class T> NT2H_FN bool pw_flint(T x) { return __builtin_isfinite(x) && (x - pw_trunc(x)) == T(0); }
template <class T> NT2H_FN bool pw_odd(T x) {
  NT2H_STRICT
  return pw_flint((x - T(1)) * T(0.5));
}
// boost.simd's ldexp: 2^e built in the exponent field, below Minexponent by
// a second product with 2^Minexponent, at Limitexponent by a second product with 2
template <class T> NT2H_FN T pw_ldexp(T a, typename powk<T>::I e) {
  NT2H_STRICT
  using K = powk<T>;
  using U = typename K::U;
  T f = T(1);
  if (e < K::mine) { e -= K::mine; f = K::tiny(); }
  if (e == K::lime) { f = f + T(1); e -= 1; }
  return (a * K::from_bits((U)(e + K::maxe) << K::nbm)) * f;
}
// the arithmetic shared by both forms; w receives the result's exponent in sixteenths
template <class P, class T> NT2H_FN T pow_core(T ax, T a1, T& wout) {
  NT2H_STRICT
  using K = powk<T>;
  using I = typename K::I;
  I e = 0;
  T x = ax;
  if (ax != T(0) && __builtin_isfinite(ax)) x = K::frexp(ax, e);
  else x = ax + T(0.5);  // boost.simd's frexp for Inf and NaN (0 is replaced below)
  int i = 1;
  if (x <= K::a(9)) i = 9;
  if (x <= K::a(i + 4)) i += 4;
  if (x <= K::a(i + 2)) i += 2;
  if (x >= K::a(1)) i = -1;
  i = i + 1;
  const T tmp = K::a(i);
  x = x - tmp;
  x = x - K::b(i >> 1);
  x = x / tmp;
  T z = x * x;
  T w = K::template pow1<P>(x, z);
  w = fma<P>(T(-0.5), z, w);
  w = fma<P>(K::log2em1(), w, w);
  z = fma<P>(K::log2em1(), x, w);
  z = z + x;
  w = fma<P>((T)(-i), T(0.0625), (T)e);
  const T ya = pw_reduc(a1);
  const T yb = a1 - ya;
  T W = fma<P>(z, a1, w * yb);
  T Wa = pw_reduc(W);
  T Wb = W - Wa;
  W = fma<P>(w, ya, Wa);
  Wa = pw_reduc(W);
  const T u = W - Wa;
  W = Wb + u;
  Wb = pw_reduc(W);
  w = T(16) * (Wa + Wb);
  wout = w;
  e = K::toint(w);
  Wb = W - Wb;
  if (Wb > T(0)) { e = e + 1; Wb = Wb - T(0.0625); }
  z = K::template pow2<P>(Wb) * Wb;
  const I ii = e / 16 + (e > 0 ? 1 : 0);
  e = ii * 16 - e;
  // e lies in [0, 16] wherever the result is kept; elsewhere read entry 0
  const T wt = K::a((e >= 0 && e <= 16) ? (int)e : 0);
  z = fma<P>(wt, z, wt);
  return pw_ldexp(z, ii);
}
template <class P, bool Scalar, class T> NT2H_FN T pow_t(T a0, T a1) {
  NT2H_STRICT
  using K = powk<T>;
  const T inf = (T)__builtin_inf(), nan = (T)__builtin_nan("");
  const T x = pw_abs(a0);
  const bool ltz = a0 < T(0);
  T z;
  if constexpr (Scalar) {
    if (ltz && !pw_flint(a1)) return nan;
    if (x == T(1)) z = x;
    else if (x == T(0)) z = (a1 == T(0)) ? T(1) : T(0);
    else if (x == a1 && a1 == inf) z = inf;
    else if (x == inf && a1 == -inf) z = T(0);
    else if (a1 == inf) z = (x < T(1)) ? T(0) : inf;
    else if (a1 == -inf) z = (x > T(1)) ? T(0) : inf;
    else if (x == inf) z = (a1 < T(0)) ? T(0) : ((a1 == T(0)) ? T(1) : inf);
    else if (a0 != a0 || a1 != a1) z = nan;
    else {
      T w;
      z = pow_core<P>(x, a1, w);
      if (w > K::large()) z = inf;
      else if (w < K::low()) z = T(0);
    }
  } else {
    T w;
    z = pow_core<P>(x, a1, w);
    const bool gtax1 = x > T(1);
    if (a1 == inf) z = gtax1 ? inf : T(0);
    if (a1 == -inf) z = gtax1 ? T(0) : inf;
    if (x == inf) z = (a1 > T(0)) ? inf : ((a1 < T(0)) ? T(0) : T(1));
    if (w < K::low()) z = T(0);
    if (w > K::large()) z = inf;
    if (x == T(1)) z = x;
    if (a0 == T(0)) z = (a1 == T(0)) ? T(1) : T(0);
  }
  if (ltz && pw_odd(a1)) z = -z;
  if (ltz && !pw_flint(a1)) z = nan;
  return z;
}
template <class P, bool Scalar> NT2H_FN float pow_f(float a0, float a1) { return pow_t<P, Scalar, float>(a0, a1); }
template <class P, bool Scalar> NT2H_FN double pow_d(double a0, double a1) { return pow_t<P, Scalar, double>(a0, a1); }

// asin, acos (f32): simd/common/impl/invtrig/f_invtrig.hpp:55-97.  |x| up to
// 1/2 by x + x^3 p(x^2); above by pi/2 - 2 asin(sqrt((1-|x|)/2)); acos from
// asin of sqrt((1-|x|)/2) or of x.  NT2's scalar forms
// (scalar/impl/invtrig/f_invtrig.hpp:49-90) give the same values, so a pack's
// tail needs no separate form.
template <class P>
NT2H_FN float asin_f(float a0) {
  NT2H_STRICT
  constexpr float c[5] = {NT2H_F(0x3d2cb352), NT2H_F(0x3cc617e3), NT2H_F(0x3d3a3ec7),
                          NT2H_F(0x3d9980f6), NT2H_F(0x3e2aaae4)};
  float x = __builtin_fabsf(a0);
  const bool big = x > 0.5f;
  const float z = big ? 0.5f * (1.0f - x) : x * x;
  if (big) x = __builtin_sqrtf(z);
  float z1 = fma<P>(horner<P>(c, z), z * x, x);
  if (big) z1 = NT2H_F(0x3fc90fdb) - (z1 + z1);
  return as_f(bits(z1) ^ (bits(a0) & 0x80000000u));
}
template <class P>
NT2H_FN float acos_f(float a0) {
  NT2H_STRICT
  float x = __builtin_fabsf(a0);
  const bool big = x > 0.5f;
  x = big ? __builtin_sqrtf(fma<P>(-0.5f, x, 0.5f)) : a0;
  x = asin_f<P>(x);
  if (big) x = x + x;
  if (a0 < -0.5f) x = NT2H_F(0x40490fdb) - x;
  return big ? x : NT2H_F(0x3fc90fdb) - x;
}

// asin, acos (f64).  Packs: simd/common/impl/invtrig/d_invtrig.hpp:57-127,
// Cephes' rational forms, |x| above 0.625 through sqrt(2(1-|x|)) with the low
// part of pi/2 subtracted, acos below -1/2 as pi - 2 asin(sqrt((1-|x|)/2)).
// The scalar forms that NT2 uses for a pack's tail
// (scalar/impl/invtrig/d_invtrig.hpp:46-111) add that low part instead and
// take acos for x <= 1/2 as (pi/4 - asin(x) + pi/4lo) + pi/4: Scalar selects
// them.
template <class P, bool Scalar>
NT2H_FN double asin_d(double a0) {
  NT2H_STRICT
  constexpr double rr[5] = {NT2H_D(0x3f684fc3988e9f08), NT2H_D(0xbfe2079259f9290f), NT2H_D(0x401bdff5baf33e6a),
                            NT2H_D(0xc03991aaac01ab68), NT2H_D(0x403c896240f3081d)};
  constexpr double ss[5] = {NT2H_D(0x3ff0000000000000), NT2H_D(0xc035f2a2b6bf5d8c), NT2H_D(0x40626219af6a7f42),
                            NT2H_D(0xc077fe08959063ee), NT2H_D(0x40756709b0b644be)};
  constexpr double pp[6] = {NT2H_D(0x3f716b9b0bd48ad3), NT2H_D(0xbfe34341333e5c16), NT2H_D(0x4015c74b178a2dd9),
                            NT2H_D(0xc0304331de27907b), NT2H_D(0x40339007da779259), NT2H_D(0xc020656c06ceafd5)};
  constexpr double qq[6] = {NT2H_D(0x3ff0000000000000), NT2H_D(0xc02d7b590b5e0eab), NT2H_D(0x40519fc025fe9054),
                            NT2H_D(0xc06265bb6d3576d7), NT2H_D(0x4061705684ffbf9d), NT2H_D(0xc04898220a3607ac)};
  const double pio4 = NT2H_D(0x3fe921fb54442d18), pio2lo = NT2H_D(0x3c91a62633145c07);
  const double sqrteps = NT2H_D(0x3e50000000000000), ct1 = NT2H_D(0x3fe4000000000000);
  const double x = __builtin_fabs(a0);
  double r;
  if constexpr (Scalar) {
    if (x > 1.0) return __builtin_nan("");
    if (x < sqrteps) return a0;
    if (x > ct1) {
      double zz = 1.0 - x;
      const double vp = (zz * horner<P>(rr, zz)) / horner<P>(ss, zz);
      zz = __builtin_sqrt(zz + zz);
      const double z = pio4 - zz;
      r = (z - fma<P>(zz, vp, pio2lo)) + pio4;
    } else {
      const double zz = x * x;
      const double z = (zz * horner<P>(pp, zz)) / horner<P>(qq, zz);
      r = x * z + x;
    }
  } else {
    if (x > 1.0) return __builtin_nan("");
    if (x < sqrteps) {
      r = x;
    } else if (x > ct1) {
      double zz1 = 1.0 - x;
      const double vp = (zz1 * horner<P>(rr, zz1)) / horner<P>(ss, zz1);
      zz1 = __builtin_sqrt(zz1 + zz1);
      const double z = pio4 - zz1;
      r = (z - fma<P>(zz1, vp, -pio2lo)) + pio4;  // NT2's fms(zz1, vp, pio2lo)
    } else {
      const double zz2 = x * x;
      const double z = (zz2 * horner<P>(pp, zz2)) / horner<P>(qq, zz2);
      r = x * z + x;
    }
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
template <class P, bool Scalar>
NT2H_FN double acos_d(double a0) {
  NT2H_STRICT
  const double pi = NT2H_D(0x400921fb54442d18), pio2 = NT2H_D(0x3ff921fb54442d18);
  const double pio4 = NT2H_D(0x3fe921fb54442d18), pio4lo = NT2H_D(0x3c81a62633145c07);
  if constexpr (Scalar) {
    if (__builtin_fabs(a0) > 1.0) return __builtin_nan("");
    if (a0 > 0.5) return 2.0 * asin_d<P, true>(__builtin_sqrt(fma<P>(-0.5, a0, 0.5)));
    return ((pio4 - asin_d<P, true>(a0)) + pio4lo) + pio4;
  } else {
    double x = __builtin_fabs(a0);
    const bool big = x > 0.5;
    x = big ? __builtin_sqrt(fma<P>(-0.5, x, 0.5)) : a0;
    x = asin_d<P, false>(x);
    if (big) x = x + x;
    if (a0 < -0.5) x = pi - x;
    return big ? x : pio2 - x;
  }
}

// ---------------------------------------------------------------------------
// erfc.  f32: modules/core/euler/include/nt2/euler/functions/simd/common/erfc.hpp:92-123
// (the scalar form returns the same values), kernels erf_kernel.hpp:46-85.
// erfc_abs_f returns erfc(|a|) before the sign fold, so erfc(a) and erfc(-a)
// can share one evaluation.
// ---------------------------------------------------------------------------
// erfc(z/(1-z)) exp((z/(1-z))^2) for z = x/(1+x) - 0.4: erf_kernel.hpp:48-70
template <class P>
NT2H_FN float erfc2_kernel_f(float z) {
  NT2H_STRICT
  constexpr float c2[11] = {NT2H_F(0xc0283611), NT2H_F(0x40d67e3b), NT2H_F(0xc0cd1a85),
                            NT2H_F(0x4050b063), NT2H_F(0xbfaea865), NT2H_F(0x3e2006f0),
                            NT2H_F(0x3e1175c7), NT2H_F(0x3ec4ca6e), NT2H_F(0x3e243828),
                            NT2H_F(0xbf918a62), NT2H_F(0x3f0a0e8b)};
  return horner<P>(c2, z);
}
template <class P>
NT2H_FN float erfc_abs_f(float x) {
  NT2H_STRICT
  constexpr float c3[9] = {NT2H_F(0x41aa8e55), NT2H_F(0x401b5680), NT2H_F(0xc010d956),
                           NT2H_F(0x3f2cff3b), NT2H_F(0xc016c985), NT2H_F(0xbffc9284),
                           NT2H_F(0xbfa11698), NT2H_F(0xbe036d7e), NT2H_F(0x3f7ffffe)};
  const float z = x / (1.0f + x);
  float r;
  if (x < NT2H_F(0x3f2aaaab)) {
    r = (1.0f - z) * horner<P>(c3, z);
  } else {
    r = exp_f<P>(-(x * x)) * erfc2_kernel_f<P>(z - 0.4f);
  }
  return (x == __builtin_inff()) ? 0.0f : r;
}
NT2H_FN float erfc_fold_f(float a0, float r) {
  NT2H_STRICT
  return (a0 < 0.0f) ? 2.0f - r : r;
}
template <class P>
NT2H_FN float erfc_f(float a0) {
  return erfc_fold_f(a0, erfc_abs_f<P>(__builtin_fabsf(a0)));
}

// f64 rational kernels, erf_kernel.hpp:101-166: erf(x)/x in x^2 on [0, 0.65),
// erfc(x) exp(x^2) on [0.65, 2.2) and on [2.2, 6)
template <class P>
NT2H_FN double erf1_kernel_d(double xx) {
  NT2H_STRICT
  constexpr double e1p[5] = {NT2H_D(0x3f110512d5b20332), NT2H_D(0x3f53b7664358865a),
                             NT2H_D(0x3fa4a59a4f02579c), NT2H_D(0x3fc16500f106c0a5),
                             NT2H_D(0x3ff20dd750429b61)};
  constexpr double e1q[5] = {NT2H_D(0x3f37ea4332348252), NT2H_D(0x3f8166f75999dbd1),
                             NT2H_D(0x3fb64536ca92ea2f), NT2H_D(0x3fdd0a84eb1ca867),
                             NT2H_D(0x3ff0000000000000)};
  return horner<P>(e1p, xx) / horner<P>(e1q, xx);
}
template <class P>
NT2H_FN double erfc2_kernel_d(double x) {
  NT2H_STRICT
  constexpr double e2p[7] = {0.0, NT2H_D(0x3f7cf4cfe0aacbb4), NT2H_D(0x3fb2488a6b5cb5e5),
                             NT2H_D(0x3fd53dd7a67c7e9f), NT2H_D(0x3fec1986509e687b),
                             NT2H_D(0x3ff54dfe9b258a60), NT2H_D(0x3feffffffbbb552b)};
  constexpr double e2q[7] = {NT2H_D(0x3f89a996639b0d00), NT2H_D(0x3fc033c113a7deee),
                             NT2H_D(0x3fe307622fcff772), NT2H_D(0x3ff9e677c2777c3c),
                             NT2H_D(0x40053b1052dca8bd), NT2H_D(0x4003adeae79b9708),
                             NT2H_D(0x3ff0000000000000)};
  return horner<P>(e2p, x) / horner<P>(e2q, x);
}
template <class P>
NT2H_FN double erfc3_kernel_d(double x) {
  NT2H_STRICT
  constexpr double e3p[7] = {0.0, NT2H_D(0x3f971d0907ea7a92), NT2H_D(0x3fc42210f88b9d43),
                             NT2H_D(0x3fe29be1cff90d94), NT2H_D(0x3ff44744306832ae),
                             NT2H_D(0x3ff9fa202deb88e5), NT2H_D(0x3fefff5a9e697ae2)};
  constexpr double e3q[7] = {NT2H_D(0x3fa47bd61bbb3843), NT2H_D(0x3fd1d7ab774bb837),
                             NT2H_D(0x3ff0cfd4cb6cde9f), NT2H_D(0x400315ffdfd5ce91),
                             NT2H_D(0x400afd487397568f), NT2H_D(0x400602f24bf3fdb6),
                             NT2H_D(0x3ff0000000000000)};
  return horner<P>(e3p, x) / horner<P>(e3q, x);
}

// erfc(x) of the f64 pack form for x = |a| >= 0, before the sign fold
template <class P>
NT2H_FN double erfc_abs_d(double x) {
  NT2H_STRICT
  const double xx = x * x;
  double r;
  if (x < 0.65) r = 1.0 - x * erf1_kernel_d<P>(xx);
  else if (x < 2.2) r = exp_d<P>(-xx) * erfc2_kernel_d<P>(x);
  else r = exp_d<P>(-xx) * erfc3_kernel_d<P>(x);
  return (x == __builtin_inf()) ? 0.0 : r;
}

// f64 SIMD: simd/common/erfc.hpp:46-89; scalar (pack tail): scalar/erfc.hpp:38-70.
template <class P, bool Scalar>
NT2H_FN double erfc_d(double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0), xx = x * x;
  double r;
  if (Scalar) {
    if (a0 != a0) return a0;
    r = 0.0;
    if (x <= 0.0000000001) r = 1.0 - (x * 2.0) / __builtin_sqrt(NT2H_D(0x400921fb54442d18));
    else if (x < 0.65) r = 1.0 - x * erf1_kernel_d<P>(xx);
    else if (x < 2.2) r = exp_d<P>(-xx) * erfc2_kernel_d<P>(x);
    else if (x < 6.0) r = exp_d<P>(-xx) * erfc3_kernel_d<P>(x);
    return (a0 < 0.0) ? 2.0 - r : r;
  }
  r = erfc_abs_d<P>(x);
  return (a0 < 0.0) ? 2.0 - r : r;
}

// ---------------------------------------------------------------------------
// erf.  f32: simd/common/erf.hpp:94-124, a0 erf1(x^2) below 2/3, else
// 1 - exp(-x^2) erfc2(x/(1+x) - 0.4) with the sign of a0, +-1 at +-Inf; the
// scalar form (scalar/erf.hpp:76-104) gives the same values.
// ---------------------------------------------------------------------------
template <class P>
NT2H_FN float erf_f(float a0) {
  NT2H_STRICT
  constexpr float c1[6] = {NT2H_F(0xba1fc83b), NT2H_F(0x3ba4468d), NT2H_F(0xbcdb61f4),
                           NT2H_F(0x3de70f22), NT2H_F(0xbec0937e), NT2H_F(0x3f906eba)};
  const float x = __builtin_fabsf(a0);
  float r;
  if (x < NT2H_F(0x3f2aaaab)) {
    r = a0 * horner<P>(c1, x * x);
  } else {
    const float z = x / (1.0f + x) - 0.4f;
    r = 1.0f - exp_f<P>(-(x * x)) * erfc2_kernel_f<P>(z);
    r = (a0 < 0.0f) ? -r : r;
  }
  return (x == __builtin_inff()) ? (a0 > 0.0f ? 1.0f : -1.0f) : r;
}

// f64 packs: simd/common/erf.hpp:47-92, regions split at 0.65 and 2.2 as in
// erfc, +-1 at +-Inf.  The scalar form NT2 uses for a pack's tail
// (scalar/erf.hpp:42-74) returns a0 for NaN, 2 a0 / sqrt(pi) up to 1e-10,
// closes its regions at 0.65, 2.2 and 6 (<=), and returns +1 beyond 6 for
// either sign: erf(-7) is +1 in a tail.  Scalar selects it.
template <class P, bool Scalar>
NT2H_FN double erf_d(double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0), xx = x * x;
  if constexpr (Scalar) {
    if (a0 != a0) return a0;
    if (x <= 0.0000000001) return (a0 * 2.0) / __builtin_sqrt(NT2H_D(0x400921fb54442d18));
    if (x <= 0.65) return a0 * erf1_kernel_d<P>(xx);
    if (x <= 6.0) {
      const double k = (x <= 2.2) ? erfc2_kernel_d<P>(x) : erfc3_kernel_d<P>(x);
      const double z = 1.0 - exp_d<P>(-xx) * k;
      return (a0 < 0.0) ? -z : z;
    }
    return 1.0;
  } else {
    if (x < 0.65) return a0 * erf1_kernel_d<P>(xx);
    const double k = (x < 2.2) ? erfc2_kernel_d<P>(x) : erfc3_kernel_d<P>(x);
    double z = 1.0 - exp_d<P>(-xx) * k;
    z = (a0 < 0.0) ? -z : z;
    return (x == __builtin_inf()) ? (a0 > 0.0 ? 1.0 : -1.0) : z;
  }
}

// ---------------------------------------------------------------------------
// cbrt: exponential/functions/simd/common/cbrt.hpp:52-166 (the scalar forms,
// scalar/cbrt.hpp:42-156, give the same values).  Denormals scaled by
// 2^nmb, frexp mantissa m in [1/2, 1) through a degree-4 polynomial, times
// 2^(e/3) as cbrt(2)^(e mod 3) and a power of two, then Newton steps
// x -= (x - z/x^2)/3 (one in f32, two in f64); zero and Inf unchanged.
// ---------------------------------------------------------------------------
template <class P>
NT2H_FN float cbrt_f(float a0) {
  NT2H_STRICT
  constexpr float c[5] = {NT2H_F(0xbe09e49a), NT2H_F(0x3f0bf0fe), NT2H_F(0xbf745265),
                          NT2H_F(0x3f91eb77), NT2H_F(0x3ece0609)};
  float z = __builtin_fabsf(a0);
  const bool den = z < NT2H_F(0x00800000);
  z = den ? z * NT2H_F(0x4B000000) : z;
  const float f = den ? NT2H_F(0x3ba14518) : 1.0f;
  const uint32_t b = bits(z);
  const int e = (int)((b & 0x7f800000u) >> 23) - 126;
  float x = horner<P>(c, as_f((b & ~0x7f800000u) | 0x3f000000u));
  const bool pos = e >= 0;
  const int ae = pos ? e : -e, q = ae / 3, rem = ae - 3 * q;
  const float c2 = pos ? NT2H_F(0x3fa14518) : NT2H_F(0x3f4b2ff5);
  const float c4 = pos ? NT2H_F(0x3fcb2ff5) : NT2H_F(0x3f214518);
  const float fact = (rem == 2) ? c4 : ((rem == 1) ? c2 : 1.0f);
  x = (x * fact) * as_f((uint32_t)((pos ? q : -q) + 127) << 23);
  x = x - (x - z / (x * x)) * NT2H_F(0x3eaaaaab);
  x = as_f(bits(x) | (bits(a0) & 0x80000000u)) * f;
  return (a0 == 0.0f || __builtin_isinf(a0)) ? a0 : x;
}
template <class P>
NT2H_FN double cbrt_d(double a0) {
  NT2H_STRICT
  constexpr double c[5] = {NT2H_D(0xbfc13c93386fdff6), NT2H_D(0x3fe17e1fc7e59d58), NT2H_D(0xbfee8a4ca3ba37b8),
                           NT2H_D(0x3ff23d6ee505873a), NT2H_D(0x3fd9c0c12122a4fe)};
  double z = __builtin_fabs(a0);
  const bool den = z < NT2H_D(0x0010000000000000);
  z = den ? z * NT2H_D(0x4330000000000000) : z;
  const double f = den ? NT2H_D(0x3ed428a2f98d7286) : 1.0;
  const uint64_t b = bits(z);
  const int e = (int)((b & 0x7ff0000000000000ull) >> 52) - 1022;
  double x = horner<P>(c, as_d((b & ~0x7ff0000000000000ull) | 0x3fe0000000000000ull));
  const bool pos = e >= 0;
  const int ae = pos ? e : -e, q = ae / 3, rem = ae - 3 * q;
  const double c2 = pos ? NT2H_D(0x3ff428a2f98d728b) : NT2H_D(0x3fe965fea53d6e3d);
  const double c4 = pos ? NT2H_D(0x3ff965fea53d6e3d) : NT2H_D(0x3fe428a2f98d728b);
  const double fact = (rem == 2) ? c4 : ((rem == 1) ? c2 : 1.0);
  x = (x * fact) * as_d((uint64_t)((pos ? q : -q) + 1023) << 52);
  x = x - (x - z / (x * x)) * NT2H_D(0x3fd5555555555555);
  x = x - (x - z / (x * x)) * NT2H_D(0x3fd5555555555555);
  x = as_d(bits(x) | (bits(a0) & 0x8000000000000000ull)) * f;
  return (a0 == 0.0 || __builtin_isinf(a0)) ? a0 : x;
}

// ---------------------------------------------------------------------------
// Inverse hyperbolic functions (hyperbolic/functions/{simd/common,scalar}/).
// hypot(1, x) is evaluated as sqrt(1 + x*x): NT2's hypot
// (boost/simd/arithmetic/functions/simd/common/hypot.hpp:24-45) scales both
// operands by 2^-e, e = exponent(max(1, x)), which is exact and commutes with
// every rounding while no result leaves the normal range, and that holds for
// each x whose hypot is used here (x <= 2^26 in f64, 2^12 in f32).
// ---------------------------------------------------------------------------
// asinh f32 packs: simd/common/asinh.hpp:66-106, x + x^3 p(x^2) below 1/2,
// else log(average(x, hypot(1, x))) + log 2 (log(x) + log 2 above
// 1/sqrt(eps)).  The scalar form (scalar/asinh.hpp:68-111) takes
// log(x + hypot(1, x)) in the middle range: Scalar selects it.
template <class P, bool Scalar>
NT2H_FN float asinh_f(float a0) {
  NT2H_STRICT
  constexpr float c[5] = {NT2H_F(0x3ca4d6e6), NT2H_F(0xbd2ee581), NT2H_F(0x3d9949b1),
                          NT2H_F(0xbe2aa9ad), NT2H_F(0x3f800000)};
  const float x = __builtin_fabsf(a0), ln2 = NT2H_F(0x3f317218);
  float z;
  if (x < 0.5f) {
    z = horner<P>(c, x * x) * x;
  } else if (x > NT2H_F(0x453504f3)) {
    z = log_f<P>(x) + ln2;
  } else {
    const float h = __builtin_sqrtf(1.0f + x * x);
    z = Scalar ? log_f<P>(x + h) : log_f<P>(fma<P>(x, 0.5f, h * 0.5f)) + ln2;
  }
  return as_f(bits(z) ^ (bits(a0) & 0x80000000u));
}
// asinh f64 packs: simd/common/asinh.hpp:46-64, log1p(x + x^2/(1 + hypot(1, x)))
// up to 1/sqrt(eps) = 2^26, log1p(x - 1) + log 2 above.  The scalar form
// (scalar/asinh.hpp:32-66): a0 below sqrt(eps), log1p(x + x/(1/x +
// sqrt(1/x^2 + 1))) below 1/2, log(x) + log 2 above 2^26, else
// log(x + hypot(1, x)).
template <class P, bool Scalar>
NT2H_FN double asinh_d(double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0), ln2 = NT2H_D(0x3fe62e42fefa39ef);
  double z;
  if constexpr (Scalar) {
    if (x < NT2H_D(0x3e50000000000000)) return a0;
    if (x < 0.5) {
      const double ix = 1.0 / x;
      z = log1p_d<P>(x + x / (ix + __builtin_sqrt(fma<P>(ix, ix, 1.0))));
    } else if (x > NT2H_D(0x4190000000000000)) {
      z = log_d<P>(x) + ln2;
    } else {
      z = log_d<P>(x + __builtin_sqrt(1.0 + x * x));
    }
  } else {
    const bool big = x > NT2H_D(0x4190000000000000);
    double u = big ? x - 1.0 : x + (x * x) / (1.0 + __builtin_sqrt(1.0 + x * x));
    u = (x == __builtin_inf()) ? x : u;
    z = log1p_d<P>(u);
    z = big ? z + ln2 : z;
  }
  return as_d(bits(z) ^ (bits(a0) & 0x8000000000000000ull));
}

// acosh: simd/common/acosh.hpp:26-42 for packs, t = a0 - 1,
// log1p(t + sqrt(2t + t^2)) up to 1/(2 eps), log1p(a0) + log 2 above; the
// scalar form (scalar/acosh.hpp:22-38) takes log(t) + log 2 above.
template <class P, bool Scalar>
NT2H_FN float acosh_f(float a0) {
  NT2H_STRICT
  const float t = a0 - 1.0f, ln2 = NT2H_F(0x3f317218);
  if (Scalar && !(t <= NT2H_F(0x4A800000))) return log_f<P>(t) + ln2;
  const bool big = t > NT2H_F(0x4A800000);
  const float z = big ? a0 : t + __builtin_sqrtf((t + t) + t * t);
  const float r = log1p_f<P>(z);
  return big ? r + ln2 : r;
}
template <class P, bool Scalar>
NT2H_FN double acosh_d(double a0) {
  NT2H_STRICT
  const double t = a0 - 1.0, ln2 = NT2H_D(0x3fe62e42fefa39ef);
  if (Scalar && !(t <= NT2H_D(0x4320000000000000))) return log_d<P>(t) + ln2;
  const bool big = t > NT2H_D(0x4320000000000000);
  const double z = big ? a0 : t + __builtin_sqrt((t + t) + t * t);
  const double r = log1p_d<P>(z);
  return big ? r + ln2 : r;
}

// atanh: simd/common/atanh.hpp:28-46 (scalar/atanh.hpp:22-44 gives the same
// values), 1/2 log1p(2x + 2x x/(1-x)) below 1/2, 1/2 log1p(2x/(1-x)) above,
// with the sign of a0.
template <class P>
NT2H_FN float atanh_f(float a0) {
  NT2H_STRICT
  const float x = __builtin_fabsf(a0), t = x + x, z1 = 1.0f - x;
  const bool small = x < 0.5f;
  const float q = (small ? x : t) / z1;
  const float r = 0.5f * log1p_f<P>(small ? fma<P>(t, q, t) : q);
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
template <class P>
NT2H_FN double atanh_d(double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0), t = x + x, z1 = 1.0 - x;
  const bool small = x < 0.5;
  const double q = (small ? x : t) / z1;
  const double r = 0.5 * log1p_d<P>(small ? fma<P>(t, q, t) : q);
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}

// normcdf(x) = Half*erfc(-Sqrt_2o_2*x):
// modules/core/statistics/include/nt2/statistics/functions/generic/normcdf.hpp:33
// erfc's magnitude is taken at c|x|, which is |(-c) x| exactly (rounding is
// symmetric in sign): normcdf(x) and normcdf(-x) then share one evaluation
// wherever both appear with the same x.
// normcdf_mag(|x|) is that shared part, normcdf_join(x, m) the rest.
template <class P>
NT2H_FN float normcdf_mag_f(float ax) {
  NT2H_STRICT
  return erfc_abs_f<P>(NT2H_F(0x3f3504f3) * ax);
}
NT2H_FN float normcdf_join_f(float a0, float m) {
  NT2H_STRICT
  return 0.5f * erfc_fold_f((-NT2H_F(0x3f3504f3)) * a0, m);
}
template <class P>
NT2H_FN float normcdf_f(float a0) {
  return normcdf_join_f(a0, normcdf_mag_f<P>(__builtin_fabsf(a0)));
}
template <class P>
NT2H_FN double normcdf_mag_d(double ax) {
  NT2H_STRICT
  return erfc_abs_d<P>(NT2H_D(0x3FE6A09E667F3BCD) * ax);
}
NT2H_FN double normcdf_join_d(double a0, double m) {
  NT2H_STRICT
  const double t = (-NT2H_D(0x3FE6A09E667F3BCD)) * a0;
  return 0.5 * ((t < 0.0) ? 2.0 - m : m);
}
template <class P, bool Scalar>
NT2H_FN double normcdf_d(double a0) {
  NT2H_STRICT
  if constexpr (Scalar) {
    return 0.5 * erfc_d<P, true>((-NT2H_D(0x3FE6A09E667F3BCD)) * a0);
  } else {
    return normcdf_join_d(a0, normcdf_mag_d<P>(__builtin_fabs(a0)));
  }
}

// fastnormcdf (single only): generic/fastnormcdf.hpp:39-54
template <class P>
NT2H_FN float fastnormcdf_f(float a0) {
  NT2H_STRICT
  constexpr float c[5] = {NT2H_F(0x3faa466f), NT2H_F(0xbfe91eea), NT2H_F(0x3fe40778),
                          NT2H_F(0xbeb68f87), NT2H_F(0x3ea385fa)};
  const float l = __builtin_fabsf(a0);
  const float k = 1.0f / fma<P>(0.2316419f, l, 1.0f);
  float w = horner<P>(c, k);
  w = w * ((k * NT2H_F(0x3ecc422a)) * exp_f<P>((-(l * l)) * 0.5f));
  return (a0 > 0.0f) ? 1.0f - w : w;
}

// ---------------------------------------------------------------------------
// Trigonometric range reduction.
// fdlibm kernel (single: 8-bit chunks, jk = 9; double: 24-bit chunks, jk = 4;
// prec = 2) as NT2 embeds it:
// modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:234-402, :690-905.
// Rarely reached (|x| > 2^18*pi single, |x| > 2^19*pi/2 double); runs per lane.
// ---------------------------------------------------------------------------
template <class T, int B>
__device__ __noinline__ int kernel_rem_pio2(const T* x, T* y, int e0, int nx) {
  NT2H_STRICT
  const int jk = (B == 8) ? 9 : 4;
  const int jp = jk;
  const T two_b = (T)(1 << B);
  const T twon_b = (T)1 / two_b;
  const int32_t full = 1 << B;
  const int32_t maskall = full - 1;
  int32_t jz, jx, jv, carry, n, iq[20], i, j, k, m, q0, ih;
  T z, fw, f[20], fq[20], q[20], PIo2[(B == 8) ? 11 : 8];
  for (i = 0; i < ((B == 8) ? 11 : 8); ++i)
    PIo2[i] = (B == 8) ? (T)as_f(nt2_hip_tables::pio2_8_bits[i])
                       : (T)as_d(nt2_hip_tables::pio2_24_bits[i]);
  jx = nx - 1;
  jv = (e0 - 3) / B;
  if (jv < 0) jv = 0;
  q0 = e0 - B * (jv + 1);
  j = jv - jx;
  m = jx + jk;
  for (i = 0; i <= m; i++, j++) {
    const int32_t d = (j < 0) ? 0 : ((B == 8) ? nt2_hip_tables::two_over_pi_8[j]
                                              : nt2_hip_tables::two_over_pi_24[j]);
    f[i] = (T)d;
  }
  for (i = 0; i <= jk; i++) {
    for (j = 0, fw = 0; j <= jx; j++) fw += x[j] * f[jx + i - j];
    q[i] = fw;
  }
  jz = jk;
  for (;;) {
    for (i = 0, j = jz, z = q[jz]; j > 0; i++, j--) {
      fw = (T)((int32_t)(twon_b * z));
      iq[i] = (int32_t)(z - two_b * fw);
      z = q[j - 1] + fw;
    }
    if constexpr (B == 8) {
      z = __builtin_ldexpf(z, q0);
      z -= 8.0f * __builtin_floorf(z * 0.125f);
    } else {
      z = __builtin_ldexp(z, q0);
      z -= 8.0 * __builtin_floor(z * 0.125);
    }
    n = (int32_t)z;
    z -= (T)n;
    ih = 0;
    if (q0 > 0) {
      i = (iq[jz - 1] >> (B - q0));
      n += i;
      iq[jz - 1] -= i << (B - q0);
      ih = iq[jz - 1] >> (B - 1 - q0);
    } else if (q0 == 0) {
      // rem_pio2.hpp:290 shifts the single-precision chunk by 8, :791 the
      // double one by 23; kept as NT2 computes it.
      ih = iq[jz - 1] >> ((B == 8) ? 8 : 23);
    } else if (z >= (T)0.5) {
      ih = 2;
    }
    if (ih > 0) {
      n += 1;
      carry = 0;
      for (i = 0; i < jz; i++) {
        j = iq[i];
        if (carry == 0) {
          if (j != 0) { carry = 1; iq[i] = full - j; }
        } else {
          iq[i] = maskall - j;
        }
      }
      if (q0 == 1) iq[jz - 1] &= (maskall >> 1);
      else if (q0 == 2) iq[jz - 1] &= (maskall >> 2);
      if (ih == 2) {
        z = (T)1 - z;
        if (carry != 0) z -= (T)__builtin_ldexp(1.0, q0);  // exact power of two
      }
    }
    bool again = false;
    if (z == (T)0) {
      j = 0;
      for (i = jz - 1; i >= jk; i--) j |= iq[i];
      if (j == 0) {
        for (k = 1; iq[jk - k] == 0; k++) {
        }
        for (i = jz + 1; i <= jz + k; i++) {
          f[jx + i] = (T)((B == 8) ? nt2_hip_tables::two_over_pi_8[jv + i]
                                   : nt2_hip_tables::two_over_pi_24[jv + i]);
          for (j = 0, fw = 0; j <= jx; j++) fw += x[j] * f[jx + i - j];
          q[i] = fw;
        }
        jz += k;
        again = true;
      }
    }
    if (!again) break;
  }
  if (z == (T)0) {
    jz -= 1;
    q0 -= B;
    while (iq[jz] == 0) { jz--; q0 -= B; }
  } else {
    if constexpr (B == 8) z = __builtin_ldexpf(z, -q0);
    else z = __builtin_ldexp(z, -q0);
    if (z >= two_b) {
      fw = (T)((int32_t)(twon_b * z));
      iq[jz] = (int32_t)(z - two_b * fw);
      jz += 1;
      q0 += B;
      iq[jz] = (int32_t)fw;
    } else {
      iq[jz] = (int32_t)z;
    }
  }
  fw = (T)__builtin_ldexp(1.0, q0);
  for (i = jz; i >= 0; i--) { q[i] = fw * (T)iq[i]; fw *= twon_b; }
  for (i = jz; i >= 0; i--) {
    for (fw = 0, k = 0; k <= jp && k <= jz - i; k++) fw += PIo2[k] * q[i + k];
    fq[jz - i] = fw;
  }
  fw = 0;
  for (i = jz; i >= 0; i--) fw += fq[i];
  y[0] = (ih == 0) ? fw : -fw;
  fw = fq[0] - fw;
  for (i = 1; i <= jz; i++) fw += fq[i];
  y[1] = (ih == 0) ? fw : -fw;
  return n & 7;
}

// __ieee754_rem_pio2f (rem_pio2.hpp:114-229) with NT2's n&3 wrapper (:37-44).
__device__ __noinline__ int rem_pio2f_fdlibm(float x, float* xr) {
  NT2H_STRICT
  const float two8 = 2.5600000000e+02f, invpio2 = NT2H_F(0x3f22f984),
              pio2_1 = NT2H_F(0x3fc90f80), pio2_1t = NT2H_F(0x37354443),
              pio2_2 = NT2H_F(0x37354400), pio2_2t = NT2H_F(0x2e85a308),
              pio2_3 = NT2H_F(0x2e85a300), pio2_3t = NT2H_F(0x248d3132);
  float y0, z, w, t, r, fn, tx[3];
  int32_t e0, i, j, nx, n, ix, hx;
  hx = (int32_t)bits(x);
  ix = hx & 0x7fffffff;
  if (ix <= 0x3f490fd8) { *xr = x; return 0; }
  if (ix < 0x4016cbe4) {
    if (hx > 0) {
      z = x - pio2_1;
      if ((ix & 0xfffffff0) != 0x3fc90fd0) { y0 = z - pio2_1t; }
      else { z -= pio2_2; y0 = z - pio2_2t; }
      *xr = y0; return 1;
    }
    z = x + pio2_1;
    if ((ix & 0xfffffff0) != 0x3fc90fd0) { y0 = z + pio2_1t; }
    else { z += pio2_2; y0 = z + pio2_2t; }
    *xr = y0; return 3;
  }
  if (ix <= 0x43490f80) {
    t = __builtin_fabsf(x);
    n = (int32_t)(t * invpio2 + 0.5f);
    fn = (float)n;
    r = t - fn * pio2_1;
    w = fn * pio2_1t;
    if (n < 32 && (int32_t)(ix & 0xffffff00) != nt2_hip_tables::npio2_hw_f[n - 1]) {
      y0 = r - w;
    } else {
      j = ix >> 23;
      y0 = r - w;
      uint32_t high = bits(y0);
      i = j - ((high >> 23) & 0xff);
      if (i > 8) {
        t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y0 = r - w;
        high = bits(y0);
        i = j - ((high >> 23) & 0xff);
        if (i > 25) {
          t = r;
          w = fn * pio2_3;
          r = t - w;
          w = fn * pio2_3t - ((t - r) - w);
          y0 = r - w;
        }
      }
    }
    if (hx < 0) { *xr = -y0; return (-n) & 3; }
    *xr = y0; return n & 3;
  }
  if (ix >= 0x7f800000) { *xr = x - x; return 0; }
  e0 = (ix >> 23) - 134;
  z = as_f((uint32_t)(ix - (int32_t)((uint32_t)e0 << 23)));
  for (i = 0; i < 2; i++) { tx[i] = (float)((int32_t)(z)); z = (z - tx[i]) * two8; }
  tx[2] = z;
  nx = 3;
  while (tx[nx - 1] == 0.0f) nx--;
  float y[2];
  n = kernel_rem_pio2<float, 8>(tx, y, e0, nx);
  if (hx < 0) { *xr = -y[0]; return (-n) & 3; }
  *xr = y[0];
  return n & 3;
}

// __ieee754_rem_pio2 (double): rem_pio2.hpp:521-621, NT2 wrapper :414-427.
__device__ __noinline__ int rem_pio2d_fdlibm(double x, double* xr) {
  NT2H_STRICT
  if (x == __builtin_inf()) { *xr = __builtin_nan(""); return 0; }
  const double two24 = 1.67772160000000000000e+07, invpio2 = NT2H_D(0x3FE45F306DC9C883),
               pio2_1 = NT2H_D(0x3FF921FB54400000), pio2_1t = NT2H_D(0x3DD0B4611A626331),
               pio2_2 = NT2H_D(0x3DD0B4611A600000), pio2_2t = NT2H_D(0x3BA3198A2E037073),
               pio2_3 = NT2H_D(0x3BA3198A2E000000), pio2_3t = NT2H_D(0x397B839A252049C1);
  double z, w, t, r, fn, y0, tx[3];
  int32_t e0, i, j, nx, n, ix, hx;
  const uint64_t xb = bits(x);
  hx = (int32_t)(xb >> 32);
  ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) { *xr = x; return 0; }
  if (ix < 0x4002d97c) {
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) { y0 = z - pio2_1t; }
      else { z -= pio2_2; y0 = z - pio2_2t; }
      *xr = y0; return 1;
    }
    z = x + pio2_1;
    if (ix != 0x3ff921fb) { y0 = z + pio2_1t; }
    else { z += pio2_2; y0 = z + pio2_2t; }
    *xr = y0; return 3;
  }
  if (ix <= 0x413921fb) {
    t = __builtin_fabs(x);
    n = (int32_t)(t * invpio2 + 0.5);
    fn = (double)n;
    r = t - fn * pio2_1;
    w = fn * pio2_1t;
    if ((n < 32) && (n > 0) && (ix != nt2_hip_tables::npio2_hw_d[n - 1])) {
      y0 = r - w;
    } else {
      j = ix >> 20;
      y0 = r - w;
      uint32_t high = (uint32_t)(bits(y0) >> 32);
      i = j - ((high >> 20) & 0x7ff);
      if (i > 16) {
        t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y0 = r - w;
        high = (uint32_t)(bits(y0) >> 32);
        i = j - ((high >> 20) & 0x7ff);
        if (i > 49) {
          t = r;
          w = fn * pio2_3;
          r = t - w;
          w = fn * pio2_3t - ((t - r) - w);
          y0 = r - w;
        }
      }
    }
    if (hx < 0) { *xr = -y0; return (-n) & 3; }
    *xr = y0; return n & 3;
  }
  if (ix >= 0x7ff00000) { *xr = x - x; return 0; }
  const uint32_t low = (uint32_t)xb;
  e0 = (ix >> 20) - 1046;
  z = as_d(((uint64_t)(uint32_t)(ix - (int32_t)((uint32_t)e0 << 20)) << 32) | low);
  for (i = 0; i < 2; i++) { tx[i] = (double)((int32_t)(z)); z = (z - tx[i]) * two24; }
  tx[2] = z;
  nx = 3;
  while (tx[nx - 1] == 0.0) nx--;
  double y[2];
  n = kernel_rem_pio2<double, 24>(tx, y, e0, nx);
  if (hx < 0) { *xr = -y[0]; return (-n) & 3; }
  *xr = y[0];
  return n & 3;
}

NT2H_FN bool ngt(float a, float b) { return !(a > b); }
NT2H_FN bool ngt(double a, double b) { return !(a > b); }

// The per-lane fdlibm reductions of a whole group behind ONE out-of-line call:
// the path is rare (arguments beyond 2^18*pi), and a single call site keeps
// the call's register saves and stack traffic out of the hot loops that
// inline the reduction cascade.
__device__ __noinline__ void huge_d(const double* x, double* xr, int* n, int g) {
  for (int i = 0; i < g; ++i) n[i] = rem_pio2d_fdlibm(x[i], &xr[i]);
}
__device__ __noinline__ void huge_f(const float* x, float* xr, int* n, int g) {
  for (int i = 0; i < g; ++i) n[i] = rem_pio2f_fdlibm(x[i], &xr[i]);
}

// f64 reduction cascade over a group of G lanes (G = 4: one NT2 AVX pack;
// G = 1: the scalar tail): scalar/impl/trigo/trig_reduction.hpp:94-215;
// straight rem_pio2_straight.hpp:58-63, cephes rem_pio2_cephes.hpp:65-72,
// medium rem_pio2_medium.hpp:68-83; beyond Medium_pi (2^18*pi) fdlibm per lane.
// BAIL: instead of the out-of-line fdlibm path, set *bail and return zeros
// (the caller redoes its work with BAIL false); keeps every call out of hot loops.
template <int G, bool BAIL = false>
NT2H_FN void reduce_d(const double* x, double* xr, int* n, bool* bail = nullptr) {
  NT2H_STRICT
  bool all4 = true, all2 = true, all20 = true, allm = true;
#pragma unroll
  for (int i = 0; i < G; ++i) {
    all4 = all4 && ngt(x[i], NT2H_D(0x3fe921fb54442d18));
    all2 = all2 && ngt(x[i], NT2H_D(0x3ff921fb54442d18));
    all20 = all20 && ngt(x[i], NT2H_D(0x404f6a7a2955385e));
    allm = allm && ngt(x[i], NT2H_D(0x412921fb54442d18));
  }
  if (all4) {
#pragma unroll
    for (int i = 0; i < G; ++i) { xr[i] = x[i]; n[i] = 0; }
  } else if (all2) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      double r = x[i] - NT2H_D(0x3FF921FB54400000);
      r -= NT2H_D(0x3DD0B4611A600000);
      r -= NT2H_D(0x3BA3198A2E000000);
      xr[i] = r; n[i] = 1;
    }
  } else if (all20) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const double xi = __builtin_rint(x[i] * NT2H_D(0x3fe45f306dc9c883));
      double r = x[i] - xi * NT2H_D(0x3FF921FB54400000);
      r -= xi * NT2H_D(0x3DD0B4611A600000);
      r -= xi * NT2H_D(0x3BA3198A2E000000);
      xr[i] = r; n[i] = ((int)xi) & 3;  // |xi| <= 13: the int32 conversion gives NT2's int64 quadrant bits
    }
  } else if (allm) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const double t = x[i];
      const double fn = __builtin_rint(t * NT2H_D(0x3FE45F306DC9C883));
      double r = t - fn * NT2H_D(0x3FF921FB54400000);
      double w = fn * NT2H_D(0x3DD0B4611A626331);
      double t2 = r;
      w = fn * NT2H_D(0x3DD0B4611A600000);
      r = t2 - w;
      w = fn * NT2H_D(0x3BA3198A2E037073) - ((t2 - r) - w);
      t2 = r;
      w = fn * NT2H_D(0x3BA3198A2E000000);
      r = t2 - w;
      w = fn * NT2H_D(0x397B839A252049C1) - ((t2 - r) - w);
      xr[i] = r - w;
      n[i] = ((int)fn) & 3;  // |fn| <= 2^19: as NT2's int64 conversion, in one instruction
    }
  } else if constexpr (BAIL) {
    *bail = true;
#pragma unroll
    for (int i = 0; i < G; ++i) { xr[i] = 0.0; n[i] = 0; }
  } else {
    huge_d(x, xr, n, G);
  }
}

// f32 reduction cascade over G lanes (G = 8: one NT2 AVX pack; 1: scalar
// tail): trig_reduction.hpp:94-215; when every lane is within 2^18*pi the
// pack is split into halves of 4 doubles that each run the f64 cascade
// (use_conversion, :190-215); beyond that, fdlibm single per lane.
template <int G, bool BAIL = false>
NT2H_FN void reduce_f(const float* x, float* xr, int* n, bool* bail = nullptr) {
  NT2H_STRICT
  bool all4 = true, all2 = true, all20 = true, allm = true, alld = true;
#pragma unroll
  for (int i = 0; i < G; ++i) {
    all4 = all4 && ngt(x[i], NT2H_F(0x3f490fdb));
    all2 = all2 && ngt(x[i], NT2H_F(0x3fc90fdb));
    all20 = all20 && ngt(x[i], NT2H_F(0x427b53d1));
    allm = allm && ngt(x[i], NT2H_F(0x43490fdb));
    alld = alld && ngt(x[i], NT2H_F(0x49490fdb));
  }
  if (all4) {
#pragma unroll
    for (int i = 0; i < G; ++i) { xr[i] = x[i]; n[i] = 0; }
  } else if (all2) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      float r = x[i] - NT2H_F(0x3fc90f80);
      r -= NT2H_F(0x37354400);
      r -= NT2H_F(0x2e85a300);
      xr[i] = r; n[i] = 1;
    }
  } else if (all20) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const float xi = __builtin_rintf(x[i] * NT2H_F(0x3f22f983));
      float r = x[i] - xi * NT2H_F(0x3fc90f80);
      r -= xi * NT2H_F(0x37354400);
      r -= xi * NT2H_F(0x2e85a300);
      xr[i] = r; n[i] = ((int)xi) & 3;
    }
  } else if (allm) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const float t = x[i];
      const float fn = __builtin_rintf(t * NT2H_F(0x3f22f984));
      float r = t - fn * NT2H_F(0x3fc90f80);
      float w = fn * NT2H_F(0x37354443);
      float t2 = r;
      w = fn * NT2H_F(0x37354400);
      r = t2 - w;
      w = fn * NT2H_F(0x2e85a308) - ((t2 - r) - w);
      t2 = r;
      w = fn * NT2H_F(0x2e85a300);
      r = t2 - w;
      w = fn * NT2H_F(0x248d3132) - ((t2 - r) - w);
      xr[i] = r - w;
      n[i] = ((int)fn) & 3;
    }
  } else if (alld) {
    constexpr int H = (G >= 4) ? 4 : G;
#pragma unroll
    for (int h = 0; h < G; h += H) {
      double ux[H], uxr[H];
      int un[H];
#pragma unroll
      for (int i = 0; i < H; ++i) ux[i] = (double)x[h + i];
      reduce_d<H, BAIL>(ux, uxr, un, bail);
#pragma unroll
      for (int i = 0; i < H; ++i) { xr[h + i] = (float)uxr[i]; n[h + i] = un[i]; }
    }
  } else if constexpr (BAIL) {
    *bail = true;
#pragma unroll
    for (int i = 0; i < G; ++i) { xr[i] = 0.0f; n[i] = 0; }
  } else {
    huge_f(x, xr, n, G);
  }
}

// evaluation polynomials: simd/common/impl/trigo/f_trig_evaluation.hpp:35-46,
// d_trig_evaluation.hpp:33-50
template <class P>
NT2H_FN float cos_eval_f(float z) {
  NT2H_STRICT
  constexpr float c[3] = {NT2H_F(0x37ccf5ce), NT2H_F(0xbab60619), NT2H_F(0x3d2aaaa5)};
  const float y = horner<P>(c, z);
  return 1.0f + fma<P>(z, -0.5f, y * (z * z));
}
template <class P>
NT2H_FN float sin_eval_f(float z, float x) {
  NT2H_STRICT
  constexpr float c[3] = {NT2H_F(0xb94ca1f9), NT2H_F(0x3c08839d), NT2H_F(0xbe2aaaa2)};
  const float y1 = horner<P>(c, z);
  return fma<P>(y1 * z, x, x);
}
template <class P>
NT2H_FN double cos_eval_d(double z) {
  NT2H_STRICT
  constexpr double c[7] = {NT2H_D(0x3da8ff831ad9b219), NT2H_D(0xbe21eea7c1e514d4),
                           NT2H_D(0x3e927e4f8e06d9a5), NT2H_D(0xbefa01a019ddbcd9),
                           NT2H_D(0x3f56c16c16c15d47), NT2H_D(0xbfa5555555555551),
                           NT2H_D(0x3fe0000000000000)};
  return 1.0 - horner<P>(c, z) * z;
}
template <class P>
NT2H_FN double sin_eval_d(double z, double x) {
  NT2H_STRICT
  constexpr double c[6] = {NT2H_D(0x3de5d8fd1fcf0ec1), NT2H_D(0xbe5ae5e5a9291691),
                           NT2H_D(0x3ec71de3567d4896), NT2H_D(0xbf2a01a019bfdf03),
                           NT2H_D(0x3f8111111110f7d0), NT2H_D(0xbfc5555555555548)};
  const double y1 = horner<P>(c, z);
  return fma<P>(y1 * z, x, x);
}

enum trig_kind { k_cos = 0, k_sin = 1, k_sincos = 2, k_tan = 3, k_cot = 4, k_sec = 5, k_csc = 6 };

// tan on the reduced argument: f_trig_evaluation.hpp:47-64 (f32 polynomial in
// z^2, z + z^3 P) and d_trig_evaluation.hpp:58-78 (f64 rational); n odd gives
// -1/y (simd/common/impl/trigo/trig_base.hpp:115-124)
template <class P>
NT2H_FN float tan_eval_f(float z, int n) {
  NT2H_STRICT
  constexpr float c[6] = {NT2H_F(0x3c19c53b), NT2H_F(0x3b4c779c), NT2H_F(0x3cc821b5),
                          NT2H_F(0x3d5ac5c9), NT2H_F(0x3e0896dd), NT2H_F(0x3eaaaa6f)};
  const float zz = z * z;
  const float y = ((horner<P>(c, zz) * zz) * z) + z;
  return (n & 1) ? -(1.0f / y) : y;
}
template <class P>
NT2H_FN double tan_eval_d(double z, int n) {
  NT2H_STRICT
  constexpr double num[3] = {NT2H_D(0xc0c992d8d24f3f38), NT2H_D(0x413199eca5fc9ddd),
                             NT2H_D(0xc1711fead3299176)};
  constexpr double den[5] = {NT2H_D(0x3ff0000000000000), NT2H_D(0x40cab8a5eeb36572),
                             NT2H_D(0xc13427bc582abc96), NT2H_D(0x4177d98fc2ead8ef),
                             NT2H_D(0xc189afe03cbe5a31)};
  const double zz = z * z;
  const double y = z + z * (zz * (horner<P>(num, zz) / horner<P>(den, zz)));
  return (n & 1) ? -(1.0 / y) : y;
}
// cot on the reduced argument: the same base y, 1/y for even n, -y for odd
// (f_trig_evaluation.hpp / d_trig_evaluation.hpp cot_eval)
template <class P>
NT2H_FN float cot_eval_f(float z, int n) {
  const float y = tan_eval_f<P>(z, 0);
  return (n & 1) ? -y : 1.0f / y;
}
template <class P>
NT2H_FN double cot_eval_d(double z, int n) {
  const double y = tan_eval_d<P>(z, 0);
  return (n & 1) ? -y : 1.0 / y;
}

// cos / sin / sincos over a group of G lanes: SIMD trig_base
// simd/common/impl/trigo/trig_base.hpp:59-71 (cos), :80-91 (sin), :140-156
// (sincos); scalar trig_base scalar/impl/trigo/trig_base.hpp:126-207 (G == 1:
// NaN for Inf/NaN inputs first).  For sincos, out0 = sin and out1 = cos.
template <class P, int G, bool BAIL = false>
NT2H_FN void trig_group_f(int kind, const float* a0, float* out0, float* out1, bool* bail = nullptr) {
  NT2H_STRICT
  float x[G], xr[G];
  int n[G];
#pragma unroll
  for (int i = 0; i < G; ++i) x[i] = __builtin_fabsf(a0[i]);
  reduce_f<G, BAIL>(x, xr, n, bail);
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const uint32_t swap = (uint32_t)n[i] & 1u;
    const float z = xr[i] * xr[i];
    const float se = sin_eval_f<P>(z, xr[i]);
    const float ce = cos_eval_f<P>(z);
    const uint32_t cos_sign = (swap ^ (((uint32_t)n[i] & 2u) >> 1)) << 31;
    const uint32_t sin_sign = (((uint32_t)n[i] & 2u) << 30) ^ (bits(a0[i]) & 0x80000000u);
    float c = as_f(bits(swap ? se : ce) ^ cos_sign);
    float s = as_f(bits(swap ? ce : se) ^ sin_sign);
    if (kind == k_tan) {
      // scalar form (scalar/impl/trigo/trig_base.hpp:165-176): NaN for
      // Inf/NaN, a0 itself for +-0
      float t = as_f(bits(tan_eval_f<P>(xr[i], n[i])) ^ (bits(a0[i]) & 0x80000000u));
      if (G == 1 && !__builtin_isfinite(a0[i])) t = __builtin_nanf("");
      else if (G == 1 && a0[i] == 0.0f) t = a0[i];
      out0[i] = t;
      continue;
    }
    if (kind == k_cot) {
      // packs: trig_base.hpp:126-138, rec(a0) at +-0; scalar: :178-188
      float t = as_f(bits(cot_eval_f<P>(xr[i], n[i])) ^ (bits(a0[i]) & 0x80000000u));
      if (G == 1 && !__builtin_isfinite(a0[i])) t = __builtin_nanf("");
      else if (a0[i] == 0.0f) t = 1.0f / a0[i];
      out0[i] = t;
      continue;
    }
    if (G == 1 && !__builtin_isfinite(a0[i])) { c = __builtin_nanf(""); s = c; }
    if (kind == k_cos) out0[i] = c;
    else if (kind == k_sin) out0[i] = s;
    else if (kind == k_sec) out0[i] = 1.0f / c;  // generic/sec.hpp: rec(cos(a0))
    else if (kind == k_csc) out0[i] = 1.0f / s;  // generic/csc.hpp: rec(sin(a0))
    else { out0[i] = s; out1[i] = c; }
  }
}

template <class P, int G>
NT2H_FN void trig_group_d(int kind, const double* a0, double* out0, double* out1) {
  NT2H_STRICT
  double x[G], xr[G];
  int n[G];
#pragma unroll
  for (int i = 0; i < G; ++i) x[i] = __builtin_fabs(a0[i]);
  reduce_d<G>(x, xr, n);
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const uint64_t swap = (uint64_t)n[i] & 1u;
    const double z = xr[i] * xr[i];
    const double se = sin_eval_d<P>(z, xr[i]);
    const double ce = cos_eval_d<P>(z);
    const uint64_t cos_sign = (swap ^ (((uint64_t)n[i] & 2u) >> 1)) << 63;
    const uint64_t sin_sign = (((uint64_t)n[i] & 2u) << 62) ^ (bits(a0[i]) & 0x8000000000000000ull);
    double c = as_d(bits(swap ? se : ce) ^ cos_sign);
    double s = as_d(bits(swap ? ce : se) ^ sin_sign);
    if (kind == k_tan) {
      double t = as_d(bits(tan_eval_d<P>(xr[i], n[i])) ^ (bits(a0[i]) & 0x8000000000000000ull));
      if (G == 1 && !__builtin_isfinite(a0[i])) t = __builtin_nan("");
      else if (G == 1 && a0[i] == 0.0) t = a0[i];
      out0[i] = t;
      continue;
    }
    if (kind == k_cot) {
      double t = as_d(bits(cot_eval_d<P>(xr[i], n[i])) ^ (bits(a0[i]) & 0x8000000000000000ull));
      if (G == 1 && !__builtin_isfinite(a0[i])) t = __builtin_nan("");
      else if (a0[i] == 0.0) t = 1.0 / a0[i];
      out0[i] = t;
      continue;
    }
    if (G == 1 && !__builtin_isfinite(a0[i])) { c = __builtin_nan(""); s = c; }
    if (kind == k_cos) out0[i] = c;
    else if (kind == k_sin) out0[i] = s;
    else if (kind == k_sec) out0[i] = 1.0 / c;
    else if (kind == k_csc) out0[i] = 1.0 / s;
    else { out0[i] = s; out1[i] = c; }
  }
}

// ---------------------------------------------------------------------------
// Trigonometric functions of pi x and of x degrees: sinpi, cospi, tanpi,
// cotpi, sind, cosd, tand, cotd (generic/sinpi.hpp ... cotd.hpp: trig_base
// with pi_tag / degree_tag).  Their reductions work lane by lane
// (scalar/impl/trigo/trig_reduction.hpp:262-318), so packs and the scalar
// tail give the same values:
//   pi:     xi = round2even(2|x|), xr = (|x| - xi/2) * Pi
//   degree: xi = round2even(|x| * (1/90)), xr = (|x| - 90 xi) * Deginrad
// n = xi converted as x86 converts (cvtt: beyond the int range, or NaN, the
// most negative integer).  Then sin/cos/tan/cot of xr with the quadrant's
// swap and sign as in the radian functions.  tan is NaN where x - 1/2 (pi) or
// (x - 90)/180 (degree, times 1/180 rounded) is an integer; cot is NaN at
// nonzero integers (multiples of 180) and 1/x at +-0.
// ---------------------------------------------------------------------------
enum trig_unit { u_pi = 0, u_deg = 1 };
NT2H_FN bool is_flint(float v) { return v - __builtin_truncf(v) == 0.0f; }
NT2H_FN bool is_flint(double v) { return v - __builtin_trunc(v) == 0.0; }
template <class P, int U>
NT2H_FN float trigu_f(int kind, float a0) {
  NT2H_STRICT
  const float x = __builtin_fabsf(a0);
  float xi, xr;
  if (U == u_pi) {
    xi = __builtin_rintf(x * 2.0f);
    xr = (x - xi * 0.5f) * NT2H_F(0x40490fdb);
  } else {
    xi = __builtin_rintf(x * NT2H_F(0x3c360b61));
    xr = (x - xi * 90.0f) * NT2H_F(0x3c8efa35);
  }
  const int n = (xi < 2147483648.0f) ? (int)xi : (int)0x80000000u;
  const uint32_t sa = bits(a0) & 0x80000000u;
  if (kind == k_tan || kind == k_cot) {
    float t = as_f(bits(kind == k_tan ? tan_eval_f<P>(xr, n) : cot_eval_f<P>(xr, n)) ^ sa);
    bool bad;
    if (kind == k_tan) bad = is_flint(U == u_pi ? a0 - 0.5f : (a0 - 90.0f) * NT2H_F(0x3bb60b61));
    else bad = a0 != 0.0f && is_flint(U == u_pi ? a0 : a0 * NT2H_F(0x3bb60b61));
    t = bad ? __builtin_nanf("") : t;
    return (kind == k_cot && a0 == 0.0f) ? 1.0f / a0 : t;
  }
  const uint32_t swap = (uint32_t)n & 1u;
  const float z = xr * xr;
  const float se = sin_eval_f<P>(z, xr), ce = cos_eval_f<P>(z);
  if (kind == k_cos) return as_f(bits(swap ? se : ce) ^ ((swap ^ (((uint32_t)n & 2u) >> 1)) << 31));
  return as_f(bits(swap ? ce : se) ^ ((((uint32_t)n & 2u) << 30) ^ sa));
}
template <class P, int U>
NT2H_FN double trigu_d(int kind, double a0) {
  NT2H_STRICT
  const double x = __builtin_fabs(a0);
  double xi, xr;
  if (U == u_pi) {
    xi = __builtin_rint(x * 2.0);
    xr = (x - xi * 0.5) * NT2H_D(0x400921fb54442d18);
  } else {
    xi = __builtin_rint(x * NT2H_D(0x3f86c16c16c16c17));
    xr = (x - xi * 90.0) * NT2H_D(0x3f91df46a2529d39);
  }
  const int64_t n = (xi < 9223372036854775808.0) ? (int64_t)xi : (int64_t)0x8000000000000000ull;
  const uint64_t sa = bits(a0) & 0x8000000000000000ull;
  if (kind == k_tan || kind == k_cot) {
    double t = as_d(bits(kind == k_tan ? tan_eval_d<P>(xr, (int)(n & 1)) : cot_eval_d<P>(xr, (int)(n & 1))) ^ sa);
    bool bad;
    if (kind == k_tan) bad = is_flint(U == u_pi ? a0 - 0.5 : (a0 - 90.0) * NT2H_D(0x3f76c16c16c16c17));
    else bad = a0 != 0.0 && is_flint(U == u_pi ? a0 : a0 * NT2H_D(0x3f76c16c16c16c17));
    t = bad ? __builtin_nan("") : t;
    return (kind == k_cot && a0 == 0.0) ? 1.0 / a0 : t;
  }
  const uint64_t swap = (uint64_t)n & 1u;
  const double z = xr * xr;
  const double se = sin_eval_d<P>(z, xr), ce = cos_eval_d<P>(z);
  if (kind == k_cos) return as_d(bits(swap ? se : ce) ^ ((swap ^ (((uint64_t)n & 2u) >> 1)) << 63));
  return as_d(bits(swap ? ce : se) ^ ((((uint64_t)n & 2u) << 62) ^ sa));
}

// ---------------------------------------------------------------------------
// uint8 saturated / selection ops (bit-exact):
// adds: boost/simd/arithmetic/functions/scalar/adds.hpp:54-67 (min(255, a+b));
// subs: max(0, a-b); selinc/seldec: boolean/functions/scalar/selinc.hpp:18-29,
// seldec.hpp:18-29; if_zero_else_one: boolean/functions/scalar/if_zero_else_one.hpp:21-52.
// ---------------------------------------------------------------------------
NT2H_FN uint32_t adds_u8(uint32_t a, uint32_t b) {
  const uint32_t s = a + b;
  return s > 255u ? 255u : s;
}
NT2H_FN uint32_t subs_u8(uint32_t a, uint32_t b) { return a > b ? a - b : 0u; }

}  // namespace dm
}}  // namespace nt2::hip
