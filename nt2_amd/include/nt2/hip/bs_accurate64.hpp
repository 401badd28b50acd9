// Accurate double-precision Black-Scholes for gfx950 (NT2H_ACCURATE policy of
// nt2h_blackscholes_f64).
//
// Same textbook formula as the NT2 evaluation (call and put from log, sqrt,
// exp and normcdf), evaluated so that every price stays within about one ulp
// of the exact price after the final rounding to double, including deep
// out-of-the-money options where S N(d1) and X exp(-rT) N(d2) cancel.
//
// Stationarity.  With d2 = d1 - tau formed exactly (tau = v sqrt(T) kept as a
// double-double), dC/dd1 = S phi(d1) - F phi(d1 - tau) vanishes at the true
// d1, so an error delta in d1 moves the prices by a relative ~d1^2 delta^2 / 2
// only.  d1 therefore comes from a plain double evaluation (log by the atanh
// series, error ~1e-16), while tau, F = X exp(-rT) and both normal tails are
// double-double:
//   exp(z)  for a double-double z <= 0: z = k ln2/32 + r with ln2/32 in
//           three parts (exact products for |k| < 2^17), |r| <= ln2/64,
//           e^r = 1 + r + r^2 P(r) (degree 7), times 2^(j/32) from a
//           double-double LDS table; the result is returned as a
//           double-double mantissa and a separate power of two, so tails far
//           below the double range keep their precision until the end;
//   Q(x) = N(-x) = exp(-x^2/2) * gt(y) * y for x >= 0 with y = C/(C + x)
//           in double-double and gt = R(x)(C + x)/(C sqrt(2 pi)) (R the Mills
//           ratio) a piecewise degree-9 polynomial in y (32 pieces, relative
//           error 2.4e-20) whose constant and linear coefficients are
//           double-double; x is clamped at 60, where Q < 2^-2590.
// Outputs.  Only the smaller price is evaluated from the tails (the call
// when S < F, else the put): D = A - B with A = S N(d1) and B = F N(d2) for
// the call, A = F N(-d2) and B = S N(-d1) for the put, each N(.) being
// Q(|.|) or 1 - Q(|.|).  The terms cancel by at most a factor of a few
// hundred over practical inputs, which the ~2^-62 relative accuracy of each
// term absorbs.  The other price is D + (F - S) (put-call parity) on the call
// side, or D - (F - S) on the put side: a non-negative term either way.
// Domain: S, X, T, v > 0; NaN inputs give NaN prices.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "detail/bs_accurate64_coeffs.hpp"

namespace nt2 { namespace hip { namespace bsacc64 {

constexpr double kRound52 = 0x1.8p52;  // adding it rounds |x| < 2^51 to an integer
constexpr int kTabDoubles = kGRows * kGN + 64;  // LDS per workgroup: gt table, then 2^(j/32) hi, lo
constexpr double kXMax = 60.0;                  // Q(60) < 2^-2590: clamped tails are zero

struct dd {
  double h, l;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b;
  const double bp = s - a;
  return {s, (a - (s - bp)) + (b - bp)};
}
// |a| >= |b| (or a == 0)
__device__ __forceinline__ dd fast_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, __builtin_fma(a, b, -p)};
}
// double-double product.  The pairs here need not be normalized (exp_dd and
// the tail polynomial leave up to ~1e-4 of the value in the low part), so the
// low-low product is kept: relative error ~2^-66 plus the inputs' own.
__device__ __forceinline__ dd mul_dd(dd a, dd b) {
  const double p = a.h * b.h;
  return {p, __builtin_fma(a.h, b.h, -p) + __builtin_fma(a.h, b.l, __builtin_fma(a.l, b.h, a.l * b.l))};
}

struct Params {
  double r, v, mu;  // rate, volatility, r + v^2/2
};

__host__ __device__ inline Params make_params(double r, double v) { return Params{r, v, r + 0.5 * v * v}; }

// exp(zh + zl) = (m.h + m.l) 2^k for -1800 <= zh <= 700, relative error
// below 2^-64 (m in [1, 2) up to the rounding of r)
__device__ __forceinline__ dd exp_dd(double zh, double zl, const double* __restrict__ ex2h,
                                     const double* __restrict__ ex2l, int& k) {
  const double sh = __builtin_fma(zh, k32oLn2, kRound52);
  const double kd = sh - kRound52;
  const int kk = (int)(uint32_t)__builtin_bit_cast(uint64_t, sh);
  const double r1 = __builtin_fma(-kd, kL1, zh);  // exact: kd*kL1 is exact and close to zh
  const dd r = two_sum(r1, -(kd * kL2));           // kd*kL2 exact
  const double rl = r.l + __builtin_fma(-kd, kL3, zl);
  const double rh = r.h;
  double t = __builtin_fma(rh, 1.0 / 5040.0, 1.0 / 720.0);
  t = __builtin_fma(rh, t, 1.0 / 120.0);
  t = __builtin_fma(rh, t, 1.0 / 24.0);
  t = __builtin_fma(rh, t, 1.0 / 6.0);
  t = __builtin_fma(rh, t, 0.5);
  const double q = (rh * rh) * t;                    // e^rh - 1 - rh, relative 2^-66
  const dd s = fast_two_sum(1.0, rh);
  // (1 + rh + q)(1 + rl) - s.h: rl reaches ~2^-44 through zl, so rl q counts
  const double lo = s.l + __builtin_fma(rl, s.h + q, q);
  const int j = kk & 31;
  const double th = ex2h[j], tl = ex2l[j];
  k = kk >> 5;
  const double mh = th * s.h;
  return {mh, __builtin_fma(th, s.h, -mh) + __builtin_fma(th, lo, tl * s.h)};
}

// Q(x) = N(-x) = (q.h + q.l) 2^k for a double-double x >= 0 (relative error
// ~2^-62)
__device__ __forceinline__ dd tail_dd(dd x, const double* __restrict__ gt, const double* __restrict__ ex2h,
                                      const double* __restrict__ ex2l, int& k) {
  double xh = x.h, xl = x.l;
  if (!(xh <= kXMax)) {  // beyond the clamp (NaN included: then the callers' S or F carries NaN)
    xh = kXMax;
    xl = 0.0;
  }
  // exp(-x^2/2)
  const dd p = two_prod(xh, xh);
  const double pl = __builtin_fma(2.0 * xh, xl, p.l);
  const dd m = exp_dd(-0.5 * p.h, -0.5 * pl, ex2h, ex2l, k);
  // y = C/(C + x) as a double-double
  const dd s = two_sum(kGC, xh);
  const double sl = s.l + xl;
  double inv = __builtin_amdgcn_rcp(s.h);
  inv = __builtin_fma(inv, __builtin_fma(-s.h, inv, 1.0), inv);
  const double yh = kGC * inv;
  const double yl = (__builtin_fma(-yh, s.h, kGC) - yh * sl) * inv;
  // piece j of y: u = 32 y - (j + 1/2), exact (y >= 3/63 makes j >= 1)
  const double t = yh * (double)kGN;
  int j = (int)t;
  j = j > kGN - 1 ? kGN - 1 : (j < 1 ? 1 : j);
  const double u = t - ((double)j + 0.5);
  const double ul = (double)kGN * yl;
  const double* c = gt + j;
  double P = c[11 * kGN];
#pragma unroll
  for (int row = 10; row >= 4; --row) P = __builtin_fma(P, u, c[row * kGN]);
  const double T2 = (u * u) * P;                       // c2 u^2 + ... + c9 u^9
  const double c0h = c[0], c0l = c[kGN], c1h = c[2 * kGN], c1l = c[3 * kGN];
  const dd p1 = two_prod(c1h, u);
  const dd g0 = fast_two_sum(c0h, p1.h);               // |c1 u| <= 0.05 c0
  // g'(u) ul = (c1 + 2 c2 u + ...) ul, to first order
  const double corr = __builtin_fma(2.0 * c[4 * kGN], u, c1h) * ul;
  const double gl = ((c0l + p1.l) + g0.l) + (__builtin_fma(c1l, u, T2) + corr);
  const dd a = mul_dd(m, dd{g0.h, gl});
  return mul_dd(a, dd{yh, yl});
}

// natural logarithm of q > 0 to ~1e-16 absolute: q = m 2^e, m in
// [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with s = (m - 1)/(m + 1), |s| < 0.1716
__device__ __forceinline__ double log_est(double q) {
  double m = __builtin_amdgcn_frexp_mant(q);  // [1/2, 1)
  int e = __builtin_amdgcn_frexp_exp(q);
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const double d = m + 1.0;
  double inv = __builtin_amdgcn_rcp(d);
  inv = __builtin_fma(inv, __builtin_fma(-d, inv, 1.0), inv);
  const double s = (m - 1.0) * inv;
  const double s2 = s * s;
  double P = __builtin_fma(s2, 2.0 / 17.0, 2.0 / 15.0);
  P = __builtin_fma(s2, P, 2.0 / 13.0);
  P = __builtin_fma(s2, P, 2.0 / 11.0);
  P = __builtin_fma(s2, P, 2.0 / 9.0);
  P = __builtin_fma(s2, P, 2.0 / 7.0);
  P = __builtin_fma(s2, P, 2.0 / 5.0);
  P = __builtin_fma(s2, P, 2.0 / 3.0);
  return __builtin_fma((double)e, 0.69314718055994530942, __builtin_fma(s * s2, P, 2.0 * s));
}

// x * 2^k for a double-double x
__device__ __forceinline__ dd ldexp_dd(dd x, int k) {
  return {__builtin_amdgcn_ldexp(x.h, k), __builtin_amdgcn_ldexp(x.l, k)};
}

// 1 - q 2^k for a tail q 2^k <= 1/2, as a double-double
__device__ __forceinline__ dd one_minus(dd q, int k) {
  const dd x = ldexp_dd(q, k);
  const dd w = fast_two_sum(1.0, -x.h);
  return {w.h, w.l - x.l};
}

// call and put of one option
__device__ __forceinline__ void bs_price(double S, double X, double T, const Params& P,
                                         const double* __restrict__ gt, const double* __restrict__ ex2h,
                                         const double* __restrict__ ex2l, double& call, double& put) {
  // tau = v sqrt(T): the hardware rsq refined by one Newton step, then
  // sqrt(T) = s0 + sl with s0 = T y and sl the first-order correction
  double y = __builtin_amdgcn_rsq(T);
  y = y * __builtin_fma(-0.5 * T * y, y, 1.5);
  const double s0 = T * y;
  const double sl = (0.5 * __builtin_fma(-s0, s0, T)) * y;
  const dd tau = {P.v * s0, __builtin_fma(P.v, s0, -(P.v * s0)) + P.v * sl};
  // F = X exp(-rT)
  const dd w = two_prod(P.r, T);
  int kE;
  const dd E0 = exp_dd(-w.h, -w.l, ex2h, ex2l, kE);
  const dd E = ldexp_dd(E0, kE);
  const dd F = {X * E.h, __builtin_fma(X, E.h, -(X * E.h)) + X * E.l};
  // d1 in double; d2 = d1 - tau in double-double
  double ix = __builtin_amdgcn_rcp(X);
  ix = __builtin_fma(ix, __builtin_fma(-X, ix, 1.0), ix);
  double it = __builtin_amdgcn_rcp(tau.h);
  it = __builtin_fma(it, __builtin_fma(-tau.h, it, 1.0), it);
  const double d1 = __builtin_fma(P.mu, T, log_est(S * ix)) * it;
  const dd d2e = two_sum(d1, -tau.h);
  const dd d2 = {d2e.h, d2e.l - tau.l};
  // the tails: Q1 = Q(|d1|), Q2 = Q(|d2|), one after the other (a loop kept
  // rolled, so that the registers of one evaluation serve both)
  const bool d2neg = d2.h < 0.0;
  dd Q1, Q2;
  int k1 = 0, k2 = 0;
#pragma unroll 1
  for (int q = 0; q < 2; ++q) {
    const dd x = q == 0 ? dd{__builtin_fabs(d1), 0.0} : (d2neg ? dd{-d2.h, -d2.l} : d2);
    int kq;
    const dd Qq = tail_dd(x, gt, ex2h, ex2l, kq);
    if (q == 0) {
      Q1 = Qq;
      k1 = kq;
    } else {
      Q2 = Qq;
      k2 = kq;
    }
  }
  // The smaller price is evaluated from the tails (the call when S < F,
  // else the put) and the other follows by put-call parity, which then adds
  // a non-negative term.  call = S N(d1) - F N(d2), put = F N(-d2) - S N(-d1),
  // each N(.) being Q or 1 - Q by the sign of its argument.
  const bool call_side = F.h > S || (F.h == S && F.l > 0.0);
  const bool u1 = call_side ? !(d1 < 0.0) : (d1 < 0.0);  // N of a positive argument: 1 - Q
  const bool u2 = call_side ? !d2neg : d2neg;
  dd P1 = Q1, P2 = Q2;
  int k1p = k1, k2p = k2;
  if (u1) {
    P1 = one_minus(Q1, k1);
    k1p = 0;
  }
  if (u2) {
    P2 = one_minus(Q2, k2);
    k2p = 0;
  }
  const dd SP1 = {S * P1.h, __builtin_fma(S, P1.h, -(S * P1.h)) + S * P1.l};
  const dd FP2 = mul_dd(F, P2);
  // D = A - B: the call (call side) or the put (put side)
  const dd A = call_side ? SP1 : FP2;
  const dd B = call_side ? FP2 : SP1;
  const int ka = call_side ? k1p : k2p;
  const int kb = call_side ? k2p : k1p;
  const dd Bs = ldexp_dd(B, kb - ka);
  const dd Dh = two_sum(A.h, -Bs.h);
  const dd D = ldexp_dd(dd{Dh.h, Dh.l + (A.l - Bs.l)}, ka);
  // the other side: D + (F - S) on the call side, D - (F - S) on the put side
  const dd fs = two_sum(F.h, -S);
  const double fsl = fs.l + F.l;
  const dd G = call_side ? two_sum(D.h, fs.h) : two_sum(D.h, -fs.h);
  const double Gl = G.l + (D.l + (call_side ? fsl : -fsl));
  const double dv = D.h + D.l, gv = G.h + Gl;
  call = call_side ? dv : gv;
  put = call_side ? gv : dv;
}

// workgroup-wide copy of the tables into LDS: gt (coefficient-major), then
// 2^(j/32) hi and lo
__device__ __forceinline__ void load_tables(double* tab) {
  for (int i = threadIdx.x; i < kGRows * kGN; i += blockDim.x) tab[i] = kG[i];
  for (int i = threadIdx.x; i < 32; i += blockDim.x) {
    tab[kGRows * kGN + i] = kExp2Hi[i];
    tab[kGRows * kGN + 32 + i] = kExp2Lo[i];
  }
  __syncthreads();
}

}}}  // namespace nt2::hip::bsacc64
