// Accurate single-precision Black-Scholes for gfx950 (NT2H_ACCURATE policy).
//
// Same textbook formula as the NT2 evaluation (call and put from log, sqrt,
// exp and normcdf), evaluated with this library's own approximations so that
// deep out-of-the-money prices, where S*N(d1) and X*exp(-rT)*N(d2) cancel by
// up to ~170x over the benchmark distribution, stay within about one ulp of
// the exact price after the final rounding to float.
//
// Stationarity.  With d2 = d1 - tau formed exactly (tau = v sqrt(T)), the
// derivative dC/dd1 = S phi(d1) - F phi(d1 - tau) vanishes at the true d1, so
// an error delta in d1 moves the call (and the put) only by
// -S phi(d1) tau delta^2 / 2: a relative ~a^2 delta^2 / 2 for a deep
// out-of-the-money option at a = |d1|.  d1 is therefore taken from the
// single-precision hardware rcp/log2/rsq (delta ~ 1e-6, below 1e-9 of the
// price), while tau, d2, F = X exp(-rT) and both normcdf are FP64:
//   exp(w)   = 2^(k/32) e^r, |r| <= ln2/64: 2^(j/32) from a 32-entry LDS
//              table, e^r of degree 3 (7e-11, evaluated in units of ln2/32
//              as a monic cubic), 2^(k>>5) by ldexp; the
//              discount factor exp(-rT) for |rT| <= 1/4 (every usual rate
//              and maturity) by one degree-7 polynomial instead (3e-12);
//   N(-|d|)  = exp(-d^2/2) * erfcx(|d|/sqrt2)/2 with the second factor a
//              piecewise polynomial (32 intervals, degree 4, 9e-11) in
//              y = C/(C+|d|); the two tails share one reciprocal.
// Both LDS tables hold 32 doubles per array, so a wave's gathers are free of
// bank conflicts.
// Outputs.  Only the out-of-the-money side is evaluated from the tails (the
// call when d1 < 0, else the put); the other price follows from put-call
// parity, C - P = S - F, which adds a term of the same sign (d1 < 0 implies
// F > S) or, for d1 >= 0 > d2, cancels: by at most a factor of ~2 near the
// money over the benchmark inputs, and by up to ~1e6 at v sqrt(T) ~ 10 where
// the call is tiny beside the put, which the FP64 intermediates still absorb
// for a float result.  On MI355X an FP64 FMA issues at the rate of an FP32
// one, which is what makes the FP64 evaluation affordable.
// Error budget of the f32 result: final rounding 0.5 ulp plus ~170 x 1.6e-10
// from the approximations at the worst cancellation (0.4 ulp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "detail/bs_accurate_coeffs.hpp"

namespace nt2 { namespace hip { namespace bsacc {

constexpr double kRound52 = 0x1.8p52;  // adding it rounds |x| < 2^51 to an integer
constexpr int kTabDoubles = (kTailDeg + 1) * kTailN + 32;  // LDS per workgroup (1536 B)

struct Params {
  double r, v;          // rate and volatility, promoted from the f32 scalars
  float mu_v, ln2_v;    // (r + v^2/2)/v and ln2/v in single precision, for the d1 estimate
};

__host__ __device__ inline Params make_params(float r, float v) {
  Params p;
  p.r = (double)r;
  p.v = (double)v;
  p.mu_v = (float)((p.r + 0.5 * p.v * p.v) / p.v);
  p.ln2_v = (float)(0.69314718055994530942 / p.v);
  return p;
}

// The table path works in units of ln2/32: with t = w*32/ln2 and k = rint(t),
// exp(w) = 2^(k>>5) * 2^((k&31)/32) * e^(r L), r = t - k in [-1/2, 1/2],
// L = ln2/32.  r comes from one fma (t - k with t's product unrounded), and
// e^(r L) = kE3 * (((r + kA2) r + kA1) r + kA0): kExpR rescaled to r and made
// monic, its leading coefficient kE3 folded into the LDS table of 2^(j/32).
// The first Horner step then reads one constant, and the reduction needs no
// separate multiply.
constexpr double kExpL = 0.693147180559945309417 / 32.0;
constexpr double kE1 = kExpR[1] * kExpL;
constexpr double kE2 = kExpR[2] * kExpL * kExpL;
constexpr double kE3 = kExpR[3] * kExpL * kExpL * kExpL;
constexpr double kA0 = kExpR[0] / kE3, kA1 = kE1 / kE3, kA2 = kE2 / kE3;
constexpr double k32oLn2 = 46.166241308446828;    // 32/ln2
constexpr double kM16oLn2 = -23.083120654223414;  // -16/ln2: exp(-q/2) = 2^(q * -16/ln2 / 32)

// 2^(k>>5) * ex2[k&31] * m(r), k the low word of sh = rint(t) + kRound52 and
// ex2[j] = 2^(j/32) kE3
__device__ __forceinline__ double exp_core(double sh, double r, const double* __restrict__ ex2) {
  const int k = (int)(uint32_t)__builtin_bit_cast(uint64_t, sh);
  double p = r + kA2;
  p = __builtin_fma(p, r, kA1);
  p = __builtin_fma(p, r, kA0);
  return __builtin_amdgcn_ldexp(ex2[k & 31] * p, k >> 5);
}

// exp(w) for w in [-700, 700]
__device__ __forceinline__ double exp_w(double w, const double* __restrict__ ex2) {
  const double sh = __builtin_fma(w, k32oLn2, kRound52);
  const double kd = sh - kRound52;
  return exp_core(sh, __builtin_fma(w, k32oLn2, -kd), ex2);
}

// exp(w) for |w| <= kExpSmall
__device__ __forceinline__ double exp_small(double w) {
  double p = __builtin_fma(kExpS[7], w, kExpS[6]);
#pragma unroll
  for (int i = 5; i >= 0; --i) p = __builtin_fma(p, w, kExpS[i]);
  return p;
}

// exp(w), by the small-argument polynomial or the table path: a function of
// w alone, so an option's price never depends on its neighbours.  A NaN w
// gives NaN.
__device__ __forceinline__ double discount(double w, const double* __restrict__ ex2) {
  if (__builtin_fabs(w) <= kExpSmall) return exp_small(w);
  return exp_w(w < -700.0 ? -700.0 : w, ex2);
}

// four discount factors, each equal to discount(w[k]); when every argument
// of the wave is small (the common case) the table path is skipped (a
// wave-uniform branch: it changes the work, not the bits)
__device__ __forceinline__ void discount4(const double (&w)[4], const double* __restrict__ ex2, double (&E)[4]) {
  bool small = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) small = small && __builtin_fabs(w[k]) <= kExpSmall;
  if (__all(small)) {
#pragma unroll
    for (int k = 0; k < 4; ++k) E[k] = exp_small(w[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) E[k] = discount(w[k], ex2);
  }
}

// exp(-q/2) for q in [0, 1400] (|r|'s error from the rounded constant: below
// 1400 * 23.1 * 2^-53 = 3.6e-12 units, 8e-14 relative)
__device__ __forceinline__ double exp_mhalf(double q, const double* __restrict__ ex2) {
  const double sh = __builtin_fma(q, kM16oLn2, kRound52);
  const double kd = sh - kRound52;
  return exp_core(sh, __builtin_fma(q, kM16oLn2, -kd), ex2);
}

// erfcx(.)/2 polynomial of interval j at u in [0, 1)
__device__ __forceinline__ double tail_poly(const double* __restrict__ c, double u) {
  double g = __builtin_fma(c[4 * kTailN], u, c[3 * kTailN]);
  g = __builtin_fma(g, u, c[2 * kTailN]);
  g = __builtin_fma(g, u, c[1 * kTailN]);
  return __builtin_fma(g, u, c[0]);
}

// Q1 = N(-|d1|), Q2 = N(-|d2|); |d| is clamped at kTailDmax (tails below 1e-63
// there).  A NaN d gives a finite Q: the callers carry NaN through S and F.
__device__ __forceinline__ void tails(double d1, double d2, const double* __restrict__ tab,
                                      const double* __restrict__ ex2, double& Q1, double& Q2) {
  static_assert(kTailDeg == 4, "tail_poly is written for degree 4");
  const double c1 = __builtin_fmin(__builtin_fabs(d1), kTailDmax);
  const double c2 = __builtin_fmin(__builtin_fabs(d2), kTailDmax);
  const double a1 = kTailC + c1, a2 = kTailC + c2;
  const double pr = a1 * a2;
  double ra = __builtin_amdgcn_rcp(pr);
  ra = __builtin_fma(ra, __builtin_fma(-pr, ra, 1.0), ra);  // one Newton step: 1/(a1 a2)
  const double cra = (kTailC * kTailA) * ra;
  const double x1 = __builtin_fma(a2, cra, kTailB);  // A C/(C + |d1|) + B in [0, 32)
  const double x2 = __builtin_fma(a1, cra, kTailB);
  const int j1 = (int)x1, j2 = (int)x2;
  Q1 = exp_mhalf(c1 * c1, ex2) * tail_poly(tab + j1, __builtin_amdgcn_fract(x1));
  Q2 = exp_mhalf(c2 * c2, ex2) * tail_poly(tab + j2, __builtin_amdgcn_fract(x2));
}

// the discount argument w = -rT of one option (exp(w) is its discount factor)
__device__ __forceinline__ double disc_arg(float Tf, const Params& P) { return -P.r * (double)Tf; }

// call and put of one option given its discount factor E = exp(-rT)
__device__ __forceinline__ void bs_price(float Sf, float Kf, float Tf, double E, const Params& P,
                                         const double* __restrict__ tab, const double* __restrict__ ex2,
                                         float& call, float& put) {
  // single-precision estimate of d1 = (ln(S/K) + (r + v^2/2) T) / (v sqrt(T))
  const float yf = __builtin_amdgcn_rsqf(Tf);
  const float qf = Sf * __builtin_amdgcn_rcpf(Kf);
  const float d1f = __builtin_fmaf(__builtin_amdgcn_logf(qf), P.ln2_v, P.mu_v * Tf) * yf;
  // FP64: tau = v sqrt(T) by one Newton step from the single-precision rsq
  const double T = Tf, y = yf;
  const double yT = T * y;
  const double h = __builtin_fma(-yT, yT, T);
  const double tau = __builtin_fma((0.5 * P.v) * y, h, P.v * yT);
  const double d1 = d1f;
  const double d2 = d1 - tau;
  const double F = (double)Kf * E;
  double Q1, Q2;
  tails(d1, d2, tab, ex2, Q1, Q2);
  // With Q = N(-|d|): d1 < 0 (so d2 < 0): call = S Q1 - F Q2;
  // d1 >= 0 > d2: put = F (1 - Q2) - S Q1; d2 >= 0: put = F Q2 - S Q1.
  // D = S Q1 - F W is the call, or minus the put; the other side is
  // D + (F - S) (put = call - S + F), or minus it.
  const double S = Sf;
  const bool call_otm = d1 < 0.0;
  const double W = (!call_otm && d2 < 0.0) ? 1.0 - Q2 : Q2;
  const double D = __builtin_fma(S, Q1, -(F * W));
  const double G = D + (F - S);
  const float fd = (float)D, fg = (float)G;
  call = call_otm ? fd : -fg;
  put = call_otm ? fg : -fd;
}

__device__ __forceinline__ void bs_accurate(float Sf, float Kf, float Tf, const Params& P,
                                            const double* __restrict__ tab,
                                            const double* __restrict__ ex2, float& call,
                                            float& put) {
  bs_price(Sf, Kf, Tf, discount(disc_arg(Tf, P), ex2), P, tab, ex2, call, put);
}

// four options (one float4 of each table)
template <class V4>
__device__ __forceinline__ void bs_accurate4(const V4& s4, const V4& x4, const V4& t4, const Params& P,
                                             const double* __restrict__ tab, const double* __restrict__ ex2,
                                             float (&call)[4], float (&put)[4]) {
  double w[4], E[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = disc_arg(t4[k], P);
  discount4(w, ex2, E);
#pragma unroll
  for (int k = 0; k < 4; ++k) bs_price(s4[k], x4[k], t4[k], E[k], P, tab, ex2, call[k], put[k]);
}

// workgroup-wide copy of the tables into LDS (coefficient-major tail table, then 2^(j/32) kE3)
__device__ __forceinline__ void load_tables(double* tab, double* ex2) {
  for (int i = threadIdx.x; i < (kTailDeg + 1) * kTailN; i += blockDim.x) tab[i] = kTail[i];
  for (int i = threadIdx.x; i < 32; i += blockDim.x) ex2[i] = kExp2j[i] * kE3;
  __syncthreads();
}

}}}  // namespace nt2::hip::bsacc
