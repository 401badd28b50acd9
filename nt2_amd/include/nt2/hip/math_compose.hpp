// NT2's composed elementwise functions on the device: the reciprocal
// hyperbolic functions and their inverses, the inverse trigonometric
// functions of reciprocals, results in degrees or in units of pi, the
// reciprocal trigonometric functions of degree and pi arguments, and sinc.
// Each follows NT2's own definition (file:line below), whose operation order
// fixes the bits.  P is the precision policy of the inner kernels; the outer
// divisions, products and differences are single IEEE operations under both
// policies.  S selects NT2's scalar form (the tail of a table) where the
// inner kernel or the definition has one.
#pragma once

#include "math.hpp"

namespace nt2 { namespace hip { namespace dm {

// ---- type dispatch onto the float / double kernels ---------------------------
template <class P, bool S> NT2H_FN float atan_t(float x) { return atan_f<P, S>(x); }
template <class P, bool S> NT2H_FN double atan_t(double x) { return atan_d<P, S>(x); }
template <class P, bool S> NT2H_FN float asin_t(float x) { return asin_f<P>(x); }
template <class P, bool S> NT2H_FN double asin_t(double x) { return asin_d<P, S>(x); }
template <class P, bool S> NT2H_FN float acos_t(float x) { return acos_f<P>(x); }
template <class P, bool S> NT2H_FN double acos_t(double x) { return acos_d<P, S>(x); }
template <class P, bool S> NT2H_FN float asinh_t(float x) { return asinh_f<P, S>(x); }
template <class P, bool S> NT2H_FN double asinh_t(double x) { return asinh_d<P, S>(x); }
template <class P, bool S> NT2H_FN float acosh_t(float x) { return acosh_f<P, S>(x); }
template <class P, bool S> NT2H_FN double acosh_t(double x) { return acosh_d<P, S>(x); }
template <class P> NT2H_FN float atanh_t(float x) { return atanh_f<P>(x); }
template <class P> NT2H_FN double atanh_t(double x) { return atanh_d<P>(x); }
template <class P> NT2H_FN float sinh_t(float x) { return sinh_f<P>(x); }
template <class P> NT2H_FN double sinh_t(double x) { return sinh_d<P>(x); }
template <class P> NT2H_FN float cosh_t(float x) { return cosh_f<P>(x); }
template <class P> NT2H_FN double cosh_t(double x) { return cosh_d<P>(x); }
template <class P> NT2H_FN float exp_t(float x) { return exp_f<P>(x); }
template <class P> NT2H_FN float expm1_t(float x) { return expm1_f<P>(x); }
template <class P> NT2H_FN double expm1_t(double x) { return expm1_d<P>(x); }
template <class P> NT2H_FN float log_t(float x) { return log_f<P>(x); }
template <class P> NT2H_FN double log_t(double x) { return log_d<P>(x); }
template <class P> NT2H_FN float log1p_t(float x) { return log1p_f<P>(x); }
template <class P> NT2H_FN double log1p_t(double x) { return log1p_d<P>(x); }
template <class P> NT2H_FN double exp_t(double x) { return exp_d<P>(x); }
template <class P, int U> NT2H_FN float trigu_t(int kind, float x) { return trigu_f<P, U>(kind, x); }
template <class P, int U> NT2H_FN double trigu_t(int kind, double x) { return trigu_d<P, U>(kind, x); }
template <class P, bool S> NT2H_FN float atan2_tt(float a, float b) { return atan2_f<P, S>(a, b); }
template <class P, bool S> NT2H_FN double atan2_tt(double a, double b) { return atan2_d<P, S>(a, b); }

// NT2's constants, by their float and double bit patterns
template <class T> NT2H_FN T cst(uint32_t f, uint64_t d) {
  if constexpr (sizeof(T) == 4) return as_f(f);
  else return as_d(d);
}
template <class T> NT2H_FN T c_pio2() { return cst<T>(0x3fc90fdbu, 0x3ff921fb54442d18ull); }
template <class T> NT2H_FN T c_pi() { return cst<T>(0x40490fdbu, 0x400921fb54442d18ull); }
template <class T> NT2H_FN T c_invpi() { return cst<T>(0x3ea2f983u, 0x3fd45f306dc9c883ull); }
template <class T> NT2H_FN T c_oneo180() { return cst<T>(0x3bb60b61u, 0x3f76c16c16c16c17ull); }
template <class T> NT2H_FN T c_eps() { return cst<T>(0x34000000u, 0x3cb0000000000000ull); }
template <class T> NT2H_FN T c_nan() { return cst<T>(0x7fc00000u, 0x7ff8000000000000ull); }

template <class T> NT2H_FN T absv(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_fabsf(x);
  else return __builtin_fabs(x);
}

template <class T> NT2H_FN T sqrtv(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_sqrtf(x);
  else return __builtin_sqrt(x);
}
template <class T> NT2H_FN T floorv(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_floorf(x);
  else return __builtin_floor(x);
}

// r with a0's sign bit xored in (b_xor(r, bitofsign(a0))) or or-ed in (b_or)
template <class T> NT2H_FN T xor_sign(T r, T a0) {
  if constexpr (sizeof(T) == 4) return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
  else return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
template <class T> NT2H_FN T or_sign(T r, T a0) {
  if constexpr (sizeof(T) == 4) return as_f(bits(r) | (bits(a0) & 0x80000000u));
  else return as_d(bits(r) | (bits(a0) & 0x8000000000000000ull));
}

// indeg: trigonometric/functions/generic/indeg.hpp:20-30, a0*Radindeg -
// a0*Radindegr (constants/radindeg.hpp:29-30, radindegr.hpp:28-29);
// inrad: generic/inrad.hpp:17-29, a0*Deginrad (constants/deginrad.hpp:29-30)
template <class T> NT2H_FN T indeg(T a0) {
  NT2H_STRICT
  return a0 * cst<T>(0x42652ee1u, 0x404ca5dc1a63c1f8ull) - a0 * cst<T>(0xb53387c1u, 0xbce1e7ab456405f8ull);
}
template <class T> NT2H_FN T inrad(T a0) {
  NT2H_STRICT
  return a0 * cst<T>(0x3c8efa35u, 0x3f91df46a2529d39ull);
}

// ---- hyperbolic ----------------------------------------------------------------
// coth: hyperbolic/functions/simd/common/coth.hpp:35-63 (scalar/coth.hpp:29-55
// gives the same value per lane): |a0| < 5/8 by tanh_kernel's coth
// (details/tanh_kernel.hpp:40-43 f32: rec of the tanh polynomial; :55-59 f64:
// q/(x fma(p, x2, q))), else fma(2, rec(exp(2|a0|) - 1), 1); a0's sign xored in.
template <class P, class T> NT2H_FN T coth_t(T a0) {
  NT2H_STRICT
  const T x = absv(a0);
  T r;
  if (x < T(0.625)) {
    const T x2 = x * x;
    if constexpr (sizeof(T) == 4) {
      r = 1.0f / tanh_small_f<P>(x, x2);
    } else {
      double p, q;
      tanh_pq_d<P>(x2, p, q);
      r = q / (x * fma<P>(p, x2, q));
    }
  } else {
    r = fma<P>(T(2), T(1) / (exp_t<P>(x + x) - T(1)), T(1));
  }
  return xor_sign(r, a0);
}
// sech: generic/sech.hpp:19-29, rec(cosh); csch: generic/csch.hpp:18-30, rec(sinh)
template <class P, class T> NT2H_FN T sech_t(T a0) { NT2H_STRICT return T(1) / cosh_t<P>(a0); }
template <class P, class T> NT2H_FN T csch_t(T a0) { NT2H_STRICT return T(1) / sinh_t<P>(a0); }
// acoth: generic/acoth.hpp:18-28, atanh(rec); asech: generic/asech.hpp:18-28,
// acosh(rec); acsch: generic/acsch.hpp:18-28, asinh(rec)
template <class P, class T> NT2H_FN T acoth_t(T a0) { NT2H_STRICT return atanh_t<P>(T(1) / a0); }
template <class P, bool S, class T> NT2H_FN T asech_t(T a0) { NT2H_STRICT return acosh_t<P, S>(T(1) / a0); }
template <class P, bool S, class T> NT2H_FN T acsch_t(T a0) { NT2H_STRICT return asinh_t<P, S>(T(1) / a0); }

// ---- inverse trigonometric -------------------------------------------------------
// acot: trigonometric/functions/generic/acot.hpp:18-28, atan(rec);
// acsc: generic/acsc.hpp:17-27, asin(rec)
template <class P, bool S, class T> NT2H_FN T acot_t(T a0) { NT2H_STRICT return atan_t<P, S>(T(1) / a0); }
template <class P, bool S, class T> NT2H_FN T acsc_t(T a0) { NT2H_STRICT return asin_t<P, S>(T(1) / a0); }
// asec: the pack form simd/common/asec.hpp:18-30 is Pio_2 - acsc(a0).  The
// scalar forms: double, scalar/asec.hpp:26-38, 0 at 1, else
// (Pio_2 - acsc(a0)) + 6.123e-17 (Pio_2's residue); float, :40-67, NaN below
// |a0| = 1, the series sqrt(2e)(1 - e(5/12 + e(43/160 - e(177/896 +
// e 2867/18432)))) in e = |a0| - 1 < 0.001 (pi minus it for a0 < 0), 0 at 1,
// else Pio_2 - acsc(a0).
template <class P, bool S, class T> NT2H_FN T asec_t(T a0) {
  NT2H_STRICT
  if constexpr (!S) {
    return c_pio2<T>() - acsc_t<P, S>(a0);
  } else if constexpr (sizeof(T) == 8) {
    if (a0 == 1.0) return 0.0;
    return (c_pio2<T>() - acsc_t<P, S>(a0)) + NT2H_D(0x3c91a62633145c07);
  } else {
    const float ax = __builtin_fabsf(a0);
    if (ax < 1.0f) return c_nan<float>();
    const float ax1 = ax - 1.0f;
    if (ax1 < 0.001f) {
      const float tmp = __builtin_sqrtf(2 * ax1) *
                        (1 - ax1 * ((5 / 12.0f) + ax1 * ((43 / 160.0f) - ax1 * ((177 / 896.0f) + ax1 * (2867 / 18432.0f)))));
      return a0 < 0.0f ? c_pi<float>() - tmp : tmp;
    }
    if (a0 == 1.0f) return 0.0f;
    return c_pio2<float>() - acsc_t<P, S>(a0);
  }
}
// atand: generic/atand.hpp:17-27, indeg(atan); asind, acosd: generic/asind.hpp:18-28,
// acosd.hpp:18-28, indeg(asin), indeg(acos); asecd, acscd: generic/asecd.hpp:18-28,
// acscd.hpp:17-27, indeg(asec), indeg(acsc)
template <class P, bool S, class T> NT2H_FN T atand_t(T a0) { return indeg(atan_t<P, S>(a0)); }
template <class P, bool S, class T> NT2H_FN T asind_t(T a0) { return indeg(asin_t<P, S>(a0)); }
template <class P, bool S, class T> NT2H_FN T acosd_t(T a0) { return indeg(acos_t<P, S>(a0)); }
template <class P, bool S, class T> NT2H_FN T asecd_t(T a0) { return indeg(asec_t<P, S>(a0)); }
template <class P, bool S, class T> NT2H_FN T acscd_t(T a0) { return indeg(acsc_t<P, S>(a0)); }
// acotd: generic/acotd.hpp:30-44, 90 - atand(|a0|) (0 subtracted at a0 = 0),
// 0 at +-Inf, a0's sign bit or-ed in
template <class P, bool S, class T> NT2H_FN T acotd_t(T a0) {
  NT2H_STRICT
  const T d = atand_t<P, S>(absv(a0));
  T z = T(90) - (a0 != T(0) ? d : T(0));
  z = __builtin_isinf(a0) ? T(0) : z;
  return or_sign(z, a0);
}
// atanpi, asinpi, acospi: generic/atanpi.hpp:18-28, asinpi.hpp:18-30,
// acospi.hpp:19-29, Invpi * f(a0)
template <class P, bool S, class T> NT2H_FN T atanpi_t(T a0) { NT2H_STRICT return c_invpi<T>() * atan_t<P, S>(a0); }
template <class P, bool S, class T> NT2H_FN T asinpi_t(T a0) { NT2H_STRICT return c_invpi<T>() * asin_t<P, S>(a0); }
template <class P, bool S, class T> NT2H_FN T acospi_t(T a0) { NT2H_STRICT return c_invpi<T>() * acos_t<P, S>(a0); }
// atan2d: generic/atan2d.hpp:16-27, indeg(atan2)
template <class P, bool S, class T> NT2H_FN T atan2d_t(T a0, T a1) { return indeg(atan2_tt<P, S>(a0, a1)); }

// ---- reciprocal trigonometric functions of degree and pi arguments ---------------
// secd: generic/secd.hpp:20-31, rec of cosd with its zeros made NaN
template <class P, class T> NT2H_FN T secd_t(T a0) {
  NT2H_STRICT
  const T r = trigu_t<P, u_deg>(k_cos, a0);
  return T(1) / (r == T(0) ? c_nan<T>() : r);
}
// cscpi: generic/cscpi.hpp:22-34, NaN at nonzero integers, else rec(sinpi)
template <class P, class T> NT2H_FN T cscpi_t(T a0) {
  NT2H_STRICT
  const T r = T(1) / trigu_t<P, u_pi>(k_sin, a0);
  return (is_flint(a0) && a0 != T(0)) ? c_nan<T>() : r;
}
// cscd: generic/cscd.hpp:18-28, cscpi(Oneo_180 * a0)
template <class P, class T> NT2H_FN T cscd_t(T a0) { NT2H_STRICT return cscpi_t<P>(c_oneo180<T>() * a0); }
// secpi: generic/secpi.hpp:22-32, NaN where 2 a0 is odd (is_odd of a float:
// boost/simd/predicates/functions/simd/common/is_odd.hpp:47, is_even(x - 1)
// = is_flint((x - 1)/2), is_even.hpp:43), else rec(cospi)
template <class P, class T> NT2H_FN T secpi_t(T a0) {
  NT2H_STRICT
  const T r = T(1) / trigu_t<P, u_pi>(k_cos, a0);
  return is_flint((T(2) * a0 - T(1)) * T(0.5)) ? c_nan<T>() : r;
}

// ---- sinc ----------------------------------------------------------------------
// sinc: simd/common/sinc.hpp:33-58, sin(a0)/a0 over the pack (sin's per-pack
// reduction), 1 where |a0| < Eps, 0 at +-Inf; the scalar form
// (scalar/sinc.hpp:28-49) gives the same per lane with the scalar sin.
template <class P, int G, class T> NT2H_FN void sinc_group(const T* a0, T* r) {
  NT2H_STRICT
  T s[G];
  if constexpr (sizeof(T) == 4) trig_group_f<P, G>(k_sin, a0, s, nullptr);
  else trig_group_d<P, G>(k_sin, a0, s, nullptr);
#pragma unroll
  for (int l = 0; l < G; ++l) {
    const T x = a0[l];
    T v = s[l] / x;
    v = absv(x) < c_eps<T>() ? T(1) : v;
    r[l] = __builtin_isinf(x) ? T(0) : v;
  }
}
// sincpi: the pack form simd/common/sincpi.hpp:35-60 is Invpi*(sinpi(a0)/a0),
// the scalar form scalar/sincpi.hpp:29-49 (Invpi*sinpi(a0))/a0; both 1 where
// |a0| < Eps and 0 at +-Inf
template <class P, bool S, class T> NT2H_FN T sincpi_t(T a0) {
  NT2H_STRICT
  const T sp = trigu_t<P, u_pi>(k_sin, a0);
  T v = S ? (c_invpi<T>() * sp) / a0 : c_invpi<T>() * (sp / a0);
  v = absv(a0) < c_eps<T>() ? T(1) : v;
  return __builtin_isinf(a0) ? T(0) : v;
}

// ---- exponential family ------------------------------------------------------------
// (modules/core/exponential/include/nt2/exponential/functions/)
// sqrt1pm1: generic/sqrt1pm1.hpp:23-36, t = sqrt(1 + a0); a0/(1 + t) for
// |a0| < 1/2, else t - 1
template <class T> NT2H_FN T sqrt1pm1(T a0) {
  NT2H_STRICT
  const T t = sqrtv(T(1) + a0);
  return absv(a0) < T(0.5) ? a0 / (T(1) + t) : t - T(1);
}
// exprecneg: generic/exprecneg.hpp:18-28, exp(-rec(a0));
// exprecnegc: generic/exprecnegc.hpp:18-28, -expm1(-rec(a0))
template <class P, class T> NT2H_FN T exprecneg_t(T a0) { NT2H_STRICT return exp_t<P>(-(T(1) / a0)); }
template <class P, class T> NT2H_FN T exprecnegc_t(T a0) { NT2H_STRICT return -expm1_t<P>(-(T(1) / a0)); }
// expx2: exp(a0^2) with a0 = m + f split (constants expx2c1.hpp:29-30 and
// expx2c2.hpp:29-30: m a multiple of 32 in float, 128 in double), Inf above
// Maxlog (maxlog.hpp:29-30) and at +-Inf.  The pack form (simd/common/
// expx2.hpp:37-61) forms m with fma(c2, x, 1/2) and u1 = fma(2m, f, f^2); the
// scalar form (scalar/expx2.hpp:32-54) with plain products, the same bits
// when fma is two roundings.
template <class P, bool S, class T> NT2H_FN T expx2_t(T a0) {
  NT2H_STRICT
  const T c1 = cst<T>(0x42000000u, 0x4060000000000000ull), c2 = cst<T>(0x3d000000u, 0x3f80000000000000ull);
  const T maxlog = cst<T>(0x42b0c0a5u, 0x40862e42fefa39efull);
  const T x = absv(a0);
  T m, f, u1;
  if constexpr (S) {
    m = c1 * floorv(c2 * x + T(0.5));
    f = x - m;
    u1 = T(2) * m * f + f * f;
  } else {
    m = c1 * floorv(fma<P>(c2, x, T(0.5)));
    f = x - m;
    u1 = fma<P>(m + m, f, f * f);
  }
  const T u = m * m;
  const T r = (u + u1 > maxlog) ? cst<T>(0x7f800000u, 0x7ff0000000000000ull) : exp_t<P>(u) * exp_t<P>(u1);
  return __builtin_isinf(x) ? x : r;
}
// logspace_add: scalar/logspace_add.hpp:25-42 (simd/common/logspace_add.hpp:30-46
// the same per lane), max(a0, a1) + log1p(exp(-|a0 - a1|)), a0 + a1 where the
// difference is NaN
template <class P, class T> NT2H_FN T logspace_add_t(T a0, T a1) {
  NT2H_STRICT
  const T tmp = -absv(a0 - a1);
  const T r = (a0 > a1 ? a0 : a1) + log1p_t<P>(exp_t<P>(tmp));
  return tmp != tmp ? a0 + a1 : r;
}
// logspace_sub: scalar/logspace_sub.hpp:20-32 (simd/common/logspace_sub.hpp:26-41
// the same per lane), x = a1 - a0; a0 + (x > -Log_2 ? log(-expm1(x)) :
// log1p(-exp(x))) (Log_2: constants/log_2.hpp:29-30)
template <class P, class T> NT2H_FN T logspace_sub_t(T a0, T a1) {
  NT2H_STRICT
  const T x = a1 - a0;
  const T tmp = x > -cst<T>(0x3f317218u, 0x3fe62e42fefa39efull) ? log_t<P>(-expm1_t<P>(x)) : log1p_t<P>(-exp_t<P>(x));
  return a0 + tmp;
}

// ---- inverse error functions -----------------------------------------------------------
// (modules/core/euler/include/nt2/euler/functions/): the polynomials of
// details/erfinv_kernel.hpp (float: erfinv1 :31-41, erfinv2 :44-54, the
// erfcinv2 rational :57-71, Giles' single-precision approximation; double:
// erfinv1 :80-104, erfinv2 :108-128, erfinv3 :132-150), coefficients by bit
// pattern, highest degree first.
template <class P> NT2H_FN float erfinv1_f(float w) {
  constexpr float c[9] = {NT2H_F(0x32f16588), NT2H_F(0x34b84b36), NT2H_F(0xb66c7357), NT2H_F(0xb6935ac1),
      NT2H_F(0x396532db), NT2H_F(0xbaa45408), NT2H_F(0xbb88e4ef), NT2H_F(0x3e7c8f63), NT2H_F(0x3fc02e2f)};
  return horner<P>(c, w);
}
template <class P> NT2H_FN float erfinv2_f(float w) {
  constexpr float c[9] = {NT2H_F(0xb951f09b), NT2H_F(0x38d3b56b), NT2H_F(0x3ab0dc72), NT2H_F(0xbb70bde7),
      NT2H_F(0x3bbc127b), NT2H_F(0xbbf9c5d7), NT2H_F(0x3c1aa57e), NT2H_F(0x3f8036db), NT2H_F(0x40354f7e)};
  return horner<P>(c, w);
}
template <class P> NT2H_FN float erfcinv2_f(float w) {
  NT2H_STRICT
  constexpr float p[5] = {NT2H_F(0x3b3a8aee), NT2H_F(0x3ba62c57), NT2H_F(0xbc8bcf63), NT2H_F(0x3c98e17c),
      NT2H_F(0xbbaf8bfe)};
  constexpr float q[5] = {NT2H_F(0x321631b2), NT2H_F(0x3b3a73c1), NT2H_F(0x3ba9025f), NT2H_F(0xbc5d80aa),
      NT2H_F(0x3c2f4022)};
  return horner<P>(p, w) / horner<P>(q, w);
}
template <class P> NT2H_FN double erfinv1_d(double w) {
  constexpr double c[23] = {NT2H_D(0xbbb135d2e746e627), NT2H_D(0xbc08ddf93324d327),
      NT2H_D(0x3c37b83eef0b7c9f), NT2H_D(0x3c69ba72cd589b91), NT2H_D(0xbca33689090a6b96),
      NT2H_D(0x3c782e11898132e0), NT2H_D(0x3cfde4acfd9e26ba), NT2H_D(0xbd26d33eed66c487),
      NT2H_D(0xbd36f2167040d8e2), NT2H_D(0x3d872a22c2d77e20), NT2H_D(0xbdac8859c4e5c0af),
      NT2H_D(0xbdcdc583d118a561), NT2H_D(0x3e120f47ccf46b3c), NT2H_D(0xbe31a9e38dc84d60),
      NT2H_D(0xbe5f36cd6d3d46a9), NT2H_D(0x3e9c6b4f5d03b787), NT2H_D(0xbeb6e8a5434ae8a2),
      NT2H_D(0xbeed1d1f7b8736f6), NT2H_D(0x3f2879c2a212f024), NT2H_D(0xbf4845769484fca8),
      NT2H_D(0xbf78b6c33114f909), NT2H_D(0x3fcebd80d9b13e28), NT2H_D(0x3ffa755e7c99ae86)};
  return horner<P>(c, w);
}
template <class P> NT2H_FN double erfinv2_d(double w) {
  constexpr double c[19] = {NT2H_D(0x3e23040f87dbd932), NT2H_D(0x3e785cbe52878635),
      NT2H_D(0xbe92777453dd3955), NT2H_D(0x3e5395abcd554c6c), NT2H_D(0x3eb936388a3790ad),
      NT2H_D(0xbed0d5db812b5083), NT2H_D(0x3ec8860cd5d652f6), NT2H_D(0x3eea29a0cacdfb23),
      NT2H_D(0xbf08cef1f80281f2), NT2H_D(0x3f11e684d0b9188a), NT2H_D(0x3ef932cd54c8a222),
      NT2H_D(0xbf37448a89ef8aa3), NT2H_D(0x3f4f3cc55ad40c25), NT2H_D(0xbf5ba924132f38b1),
      NT2H_D(0x3f6468eeca533cf8), NT2H_D(0xbf6ebadabb891bbd), NT2H_D(0x3f75ffcfe5b76afc),
      NT2H_D(0x3ff0158a6d641d39), NT2H_D(0x4008abcc380d5a48)};
  return horner<P>(c, w);
}
template <class P> NT2H_FN double erfinv3_d(double w) {
  constexpr double c[17] = {NT2H_D(0xbdbdcec3a7785389), NT2H_D(0xbdf18feec0e38727),
      NT2H_D(0x3e19e6bf2dda45e3), NT2H_D(0xbe30468fb24e2f5f), NT2H_D(0x3e405ac6a8fba182),
      NT2H_D(0xbe50102e495fb9c0), NT2H_D(0x3e5f4c20e1334af8), NT2H_D(0xbe722d220fdf9c3e),
      NT2H_D(0x3e8ebc8bb824cb54), NT2H_D(0xbeb0a8d40ea372cc), NT2H_D(0x3ed2fbd29d093d2b),
      NT2H_D(0xbef4a3497e1e0fac), NT2H_D(0x3f13ebf4eb00938f), NT2H_D(0xbf2c2f36a8fc5d53),
      NT2H_D(0xbf222ea5df04047c), NT2H_D(0x3ff02a30d1fba0dc), NT2H_D(0x4013664ddd1ad7fb)};
  return horner<P>(c, w);
}
template <class T> NT2H_FN T signnz(T a) { return __builtin_signbit(a) ? T(-1) : T(1); }
// erfinv: simd/common/erfinv.hpp:32-57 (float), :64-108 (double); the scalar
// forms (scalar/erfinv.hpp:22-52, :59-80) give the same value per lane.
// w = -log((1 - a0)(1 + a0)); float: a0 erfinv1(w - 2.5) below 5, else
// a0 erfinv2(sqrt(w) - 3); double: a0 erfinv1(w - 3.125) below 6.25,
// a0 erfinv2(sqrt(w) - 3.25) below 16, else a0 erfinv3(sqrt(w) - 5);
// signnz(a0) w where w is infinite (a0 = +-1)
template <class P, class T> NT2H_FN T erfinv_t(T a0) {
  NT2H_STRICT
  const T w = -log_t<P>((T(1) - a0) * (T(1) + a0));
  if (__builtin_isinf(w) && w > T(0)) return signnz(a0) * w;
  if constexpr (sizeof(T) == 4) {
    if (w < 5.0f) return a0 * erfinv1_f<P>(w - 2.5f);
    return a0 * erfinv2_f<P>(__builtin_sqrtf(w) - 3.0f);
  } else {
    if (w < 6.25) return a0 * erfinv1_d<P>(w - 3.125);
    if (w < 16.0) return a0 * erfinv2_d<P>(__builtin_sqrt(w) - 3.25);
    return a0 * erfinv3_d<P>(__builtin_sqrt(w) - 5.0);
  }
}
// erfcinv: simd/common/erfcinv.hpp:32-57 (float), :64-107 (double); the scalar
// forms (scalar/erfcinv.hpp:22-51, :58-80) give the same value per lane.
// a00 = 1 - a0, w = -log(a0 (2 - a0)); as erfinv with a00 for a0, except the
// float tail copysign(erfcinv2(sqrt(w)), a00)
template <class P, class T> NT2H_FN T erfcinv_t(T a0) {
  NT2H_STRICT
  const T a00 = T(1) - a0;
  const T w = -log_t<P>(a0 * (T(2) - a0));
  if (__builtin_isinf(w) && w > T(0)) return signnz(a00) * w;
  if constexpr (sizeof(T) == 4) {
    if (w < 5.0f) return a00 * erfinv1_f<P>(w - 2.5f);
    return __builtin_copysignf(erfcinv2_f<P>(__builtin_sqrtf(w)), a00);
  } else {
    if (w < 6.25) return a00 * erfinv1_d<P>(w - 3.125);
    if (w < 16.0) return a00 * erfinv2_d<P>(__builtin_sqrt(w) - 3.25);
    return a00 * erfinv3_d<P>(__builtin_sqrt(w) - 5.0);
  }
}

// sinhc = sinh(x)/x: hyperbolic/functions/simd/common/sinhc.hpp:41-80 (the
// scalar form scalar/sinhc.hpp:31-67 gives the same value per lane): below
// |x| = 1 sinhc_kernel (details/sinhc_kernel.hpp: a degree-3 polynomial in x^2
// for float, 1 + x^2 p(x^2)/q(x^2) for double); above, e = exp(x) and
// average(e, -1/e)/x, or with e = exp(x/2) beyond Maxlog - Log_2,
// ((e/2)/x) e; +Inf at +-Inf
template <class P, class T> NT2H_FN T sinhc_t(T a0) {
  NT2H_STRICT
  const T x = absv(a0);
  if (x < T(1)) {
    const T x2 = x * x;
    if constexpr (sizeof(T) == 4) {
      constexpr float c[4] = {NT2H_F(0x39559e2f), NT2H_F(0x3c087bbe), NT2H_F(0x3e2aaacc), NT2H_F(0x3f800000)};
      return horner<P>(c, x2);
    } else {
      constexpr double p[4] = {NT2H_D(0xbfe9435fe8bb3cd6), NT2H_D(0xc064773a398ff4fe), NT2H_D(0xc0c694b8c71d6182),
                               NT2H_D(0xc115782bdbf6ab05)};
      constexpr double q[4] = {NT2H_D(0x3ff0000000000000), NT2H_D(0xc0715b6096e96484), NT2H_D(0x40e1a7ba7ed72245),
                               NT2H_D(0xc1401a20e4f90044)};
      return 1.0 + (x2 * horner<P>(p, x2)) / horner<P>(q, x2);
    }
  }
  const T thr = cst<T>(0x42b0c0a5u, 0x40862e42fefa39efull) - cst<T>(0x3f317218u, 0x3fe62e42fefa39efull);
  const bool big = x > thr;
  const T e = exp_t<P>(x * (big ? T(0.5) : T(1)));
  const T r = big ? ((T(0.5) * e) / x) * e : fma<P>(e, T(0.5), -(T(1) / e) * T(0.5)) / x;
  return __builtin_isinf(x) ? x : r;
}

}}}  // namespace nt2::hip::dm
