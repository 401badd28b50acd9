// Accurate single-precision Black-Scholes for gfx950 (NT2H_ACCURATE policy).
//
// Same textbook formula as the NT2 evaluation (call and put from log, sqrt,
// exp and normcdf), evaluated with this library's own FP64 approximations:
// on MI355X a non-packed FP32 VALU op and an FP64 one issue at the same rate,
// so carrying the whole option in FP64 costs no more instructions than FP32,
// and it is what keeps deep out-of-the-money prices (where S*N(d1) and
// X*exp(-rT)*N(d2) cancel by up to ~170x over the benchmark distribution)
// within a few ulp of the exact price after the final rounding to float.
//
//   1/sqrt(T): v_rsq_f64 + one Newton step;  S/X: v_rcp_f64 + one Newton step
//   log(q):  q = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//            s = (m-1)/(m+1), series to s^11 (|s| <= 0.172: < 2e-11)
//   exp(w):  w = k ln2 + r, |r| <= ln2/2, degree-10 fit (1.6e-15), ldexp
//   N(d):    N(-|d|) = exp(-d^2/2) erfcx(|d|/sqrt2)/2 with erfcx a degree-12
//            polynomial in t = A*C/(C+z) + B (4.3e-11 relative on [0, 12]);
//            exp(-d2^2/2) = exp(-d1^2/2) * (S/X) * exp(rT) (exact identity)
//   exp(rT): Taylor degree 8 on [0, 0.2]
// Error budget of the f32 result: final rounding 0.5 ulp plus ~0.25 ulp from
// erfcx amplified by the cancellation, the rest below 1e-9 relative.
#pragma once

#include <hip/hip_runtime.h>

#include "detail/bs_accurate_coeffs.hpp"

namespace nt2 { namespace hip { namespace bsacc {

struct Params {
  double r, v, mu, inv_v;  // r, sigma, r + sigma^2/2, 1/sigma (from the f32 scalars)
};

__host__ __device__ inline Params make_params(float r, float v) {
  Params p;
  p.r = (double)r;
  p.v = (double)v;
  p.mu = p.r + 0.5 * p.v * p.v;
  p.inv_v = 1.0 / p.v;
  return p;
}

__device__ __forceinline__ double rcp_nr(double a) {
  double y = __builtin_amdgcn_rcp(a);
  const double e = __builtin_fma(-a, y, 1.0);
  return __builtin_fma(y, e, y);
}

template <int K>
__device__ __forceinline__ double horner_d(const double (&c)[K], double x) {
  double acc = c[0];
#pragma unroll
  for (int i = 1; i < K; ++i) acc = __builtin_fma(acc, x, c[i]);
  return acc;
}

// exp(w) for w in [-745, 0]
__device__ __forceinline__ double exp_neg(double w) {
  const double k = __builtin_rint(w * 1.4426950408889634);
  double r = __builtin_fma(-k, 6.93147180369123816490e-01, w);  // ln2 hi (32 bits)
  r = __builtin_fma(-k, 1.90821492927058770002e-10, r);         // ln2 lo
  const double p = horner_d(kExp, r);
  return __builtin_amdgcn_ldexp(p, (int)k);
}

// log(q) for normal positive q
__device__ __forceinline__ double log_pos(double q) {
  const uint64_t b = __builtin_bit_cast(uint64_t, q);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  const bool big = m > 1.4142135623730951;
  m = big ? 0.5 * m : m;
  e = big ? e + 1 : e;
  const double f = m - 1.0;
  const double s = f * rcp_nr(2.0 + f);
  const double s2 = s * s;
  // 2 atanh(s) = 2s + s*s2*(2/3 + s2*(2/5 + s2*(2/7 + s2*(2/9 + s2*2/11))))
  double p = __builtin_fma(s2, 2.0 / 11.0, 2.0 / 9.0);
  p = __builtin_fma(p, s2, 2.0 / 7.0);
  p = __builtin_fma(p, s2, 2.0 / 5.0);
  p = __builtin_fma(p, s2, 2.0 / 3.0);
  const double lm = __builtin_fma(s * s2, p, 2.0 * s);
  return __builtin_fma((double)e, 6.93147180559945309417e-01, lm);
}

__device__ __forceinline__ void bs_accurate_one(float Sf, float Kf, float Tf, const Params& P,
                                                float& call, float& put) {
  const double S = Sf, K = Kf, T = Tf;
  // 1/sqrt(T), sqrt(T)
  double y = __builtin_amdgcn_rsq(T);
  const double h = T * y;
  y = __builtin_fma(0.5 * y, __builtin_fma(-h, y, 1.0), y);
  const double tau = P.v * (T * y);
  const double inv_tau = y * P.inv_v;
  // q = S/K, log q
  const double q = S * rcp_nr(K);
  const double L = log_pos(q);
  const double d1 = __builtin_fma(P.mu, T, L) * inv_tau;
  const double d2 = d1 - tau;
  // E = exp(rT) (rT in [0, 0.2]), DF = 1/E, F = K*DF
  const double x = P.r * T;
  double E = __builtin_fma(x, 1.0 / 40320.0, 1.0 / 5040.0);
  E = __builtin_fma(E, x, 1.0 / 720.0);
  E = __builtin_fma(E, x, 1.0 / 120.0);
  E = __builtin_fma(E, x, 1.0 / 24.0);
  E = __builtin_fma(E, x, 1.0 / 6.0);
  E = __builtin_fma(E, x, 0.5);
  E = __builtin_fma(E, x, 1.0);
  E = __builtin_fma(E, x, 1.0);
  const double F = K * rcp_nr(E);
  // Gaussians
  const double e1 = exp_neg(-0.5 * d1 * d1);
  const double e2 = e1 * q * E;
  // erfcx(|d|/sqrt2) for both, one reciprocal for the two changes of variable
  const double z1 = __builtin_fmin(__builtin_fabs(d1) * 0.70710678118654752440, kErfcxZmax);
  const double z2 = __builtin_fmin(__builtin_fabs(d2) * 0.70710678118654752440, kErfcxZmax);
  const double a1 = kErfcxC + z1, a2 = kErfcxC + z2;
  const double ra = rcp_nr(a1 * a2);
  const double t1 = __builtin_fma(kErfcxA * kErfcxC, a2 * ra, kErfcxB);
  const double t2 = __builtin_fma(kErfcxA * kErfcxC, a1 * ra, kErfcxB);
  const double Q1 = 0.5 * e1 * horner_d(kErfcx, t1);  // N(-|d1|)
  const double Q2 = 0.5 * e2 * horner_d(kErfcx, t2);  // N(-|d2|)
  const double N1 = d1 < 0.0 ? Q1 : 1.0 - Q1, M1 = d1 < 0.0 ? 1.0 - Q1 : Q1;  // N(d1), N(-d1)
  const double N2 = d2 < 0.0 ? Q2 : 1.0 - Q2, M2 = d2 < 0.0 ? 1.0 - Q2 : Q2;
  const double C = S * N1 - F * N2;
  const double Pp = F * M2 - S * M1;
  call = (float)__builtin_fmax(C, 0.0);
  put = (float)__builtin_fmax(Pp, 0.0);
}

}}}  // namespace nt2::hip::bsacc
