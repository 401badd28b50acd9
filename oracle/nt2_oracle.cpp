// ============================================================================
// THE ORACLE — test infrastructure, not product code.
//
// A CPU restatement of NT2's own Boost.SIMD evaluation of the table-expression
// hot path: the math kernels (exp, log, sqrt, erfc, normcdf, fastnormcdf,
// cos/sin/sincos with NT2's full range-reduction cascade), the saturated
// uint8 ops, NT2's transform schedule (8-lane f32 / 4-lane f64 packs from
// index 0, scalar tail with the scalar kernels) and NT2's SIMD inner_fold
// order for sums.  Built with g++ -O3 -mavx2 -ffp-contract=off -fopenmp so
// that NT2's fma(a,b,c) is a*b+c with two roundings, as on x86 AVX
// (modules/boost/simd/base/include/boost/simd/arithmetic/functions/generic/
// fma.hpp:32).
//
// Only tests/, the smoke check and bench.py's cpu_baseline leg load this
// library, and only as the checker or as the timed CPU baseline.
//
// Every function cites the reference file:line it restates.  Paths below are
// relative to the reference tree root.  Parity pinning (golden vectors and
// unit-test KATs) lives in tests/test_oracle_*.py.
// ============================================================================
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <limits>
#include <omp.h>
#include <quadmath.h>
#include <unistd.h>

#include "pio2_tables.h"

#define F32(b) __builtin_bit_cast(float, (uint32_t)(b##u))
#define F64(b) __builtin_bit_cast(double, (uint64_t)(b##ull))

namespace {

inline uint32_t bits(float f) { return __builtin_bit_cast(uint32_t, f); }
inline uint64_t bits(double d) { return __builtin_bit_cast(uint64_t, d); }
inline float as_f(uint32_t b) { return __builtin_bit_cast(float, b); }
inline double as_d(uint64_t b) { return __builtin_bit_cast(double, b); }

// NT2 static Horner: coefficients in decreasing powers, each step
// fma(x, acc, c) = x*acc + c with two roundings.
// modules/core/polynomials/include/nt2/polynomials/functions/scalar/impl/horner.hpp:13-45
template <class T, size_t K>
inline T horner(const T (&c)[K], T x) {
  T acc = c[0];
  for (size_t i = 1; i < K; ++i) acc = x * acc + c[i];
  return acc;
}

// round2even: modules/boost/simd/base/include/boost/simd/arithmetic/functions/simd/common/round2even.hpp:31-40
inline float r2e(float x) { return __builtin_rintf(x); }
inline double r2e(double x) { return __builtin_rint(x); }

// ---------------------------------------------------------------------------
// single precision constants (bit patterns from the NT2 constant registry)
// ---------------------------------------------------------------------------
const float fInvlog_2 = F32(0x3fb8aa3b), fLog_2hi = F32(0x3f318000),
            fLog_2lo = F32(0xb95e8083), fMaxlog = F32(0x42b0c0a5),
            fMinlog = F32(0xc2b0c0a5);
// f_expo_reduction.hpp:62-70
const float fExpC[5] = {F32(0x3ab778cf), F32(0x3c098d8b), F32(0x3d2aa957),
                        F32(0x3e2aa9a5), F32(0x3f000000)};
const float fSqrt_2o_2 = F32(0x3f3504f3), fSmallest = F32(0x00800000),
            fTwotonmb = F32(0x4B000000), fMlogtwo2nmb = F32(0xc17f1402);
// f_kernel.hpp:62-72
const float fLogC[8] = {F32(0xbda5dff0), F32(0x3e0229f9), F32(0xbe04d6b7),
                        F32(0x3e116e80), F32(0xbe2a6aa0), F32(0x3e4cd0a3),
                        F32(0xbe800064), F32(0x3eaaaaa9)};
const float fTwothird = F32(0x3f2aaaab);
// erf_kernel.hpp:50-64 (erfc2, degree 10) and :72-84 (erfc3, degree 8)
const float fErfc2C[11] = {F32(0xc0283611), F32(0x40d67e3b), F32(0xc0cd1a85),
                           F32(0x4050b063), F32(0xbfaea865), F32(0x3e2006f0),
                           F32(0x3e1175c7), F32(0x3ec4ca6e), F32(0x3e243828),
                           F32(0xbf918a62), F32(0x3f0a0e8b)};
const float fErfc3C[9] = {F32(0x41aa8e55), F32(0x401b5680), F32(0xc010d956),
                          F32(0x3f2cff3b), F32(0xc016c985), F32(0xbffc9284),
                          F32(0xbfa11698), F32(0xbe036d7e), F32(0x3f7ffffe)};
// fastnormcdf.hpp:42-49
const float fFastC[5] = {F32(0x3faa466f), F32(0xbfe91eea), F32(0x3fe40778),
                         F32(0xbeb68f87), F32(0x3ea385fa)};
const float fInvsqrt_2pi = F32(0x3ecc422a);
// trigonometric constants
const float fPio_4 = F32(0x3f490fdb), fPio_2 = F32(0x3fc90fdb),
            f20pi = F32(0x427b53d1), fMedium_pi = F32(0x43490fdb),
            fDmpi = F32(0x49490fdb), fPio2_1 = F32(0x3fc90f80),
            fPio2_2 = F32(0x37354400), fPio2_3 = F32(0x2e85a300),
            fPio2_1t = F32(0x37354443), fPio2_2t = F32(0x2e85a308),
            fPio2_3t = F32(0x248d3132), fTwoopi = F32(0x3f22f983),
            fInvpio_2 = F32(0x3f22f984);
// f_trig_evaluation.hpp:47-64 (tan)
const float fTanC[6] = {F32(0x3c19c53b), F32(0x3b4c779c), F32(0x3cc821b5), F32(0x3d5ac5c9),
                        F32(0x3e0896dd), F32(0x3eaaaa6f)};
// f_trig_evaluation.hpp:35-46
const float fCosC[3] = {F32(0x37ccf5ce), F32(0xbab60619), F32(0x3d2aaaa5)};
const float fSinC[3] = {F32(0xb94ca1f9), F32(0x3c08839d), F32(0xbe2aaaa2)};

// ---------------------------------------------------------------------------
// double precision constants
// ---------------------------------------------------------------------------
const double dInvlog_2 = F64(0x3ff71547652b82fe),
             dLog_2hi = F64(0x3fe62e42fee00000),
             dLog_2lo = F64(0x3dea39ef35793c76),
             dMaxlog = F64(0x40862e42fefa39ef),
             dMinlog = F64(0xc086232bdd7abcd2);
// d_expo_reduction.hpp:66-73
const double dExpC[5] = {F64(0x3e66376972bea4d0), F64(0xbebbbd41c5d26bf1),
                         F64(0x3f11566aaf25de2c), F64(0xbf66c16c16bebd93),
                         F64(0x3fc555555555553e)};
const double dSqrt_2o_2 = F64(0x3FE6A09E667F3BCD);
// d_log.hpp:73-84
const double dLogT1[3] = {F64(0x3fc39a09d078c69f), F64(0x3fcc71c51d8e78af),
                          F64(0x3fd999999997fa04)};
const double dLogT2[4] = {F64(0x3fc2f112df3e5244), F64(0x3fc7466496cb03de),
                          F64(0x3fd2492494229359), F64(0x3fe5555555555593)};
// erf_kernel.hpp:96-166
const double dErf1P[5] = {F64(0x3f110512d5b20332), F64(0x3f53b7664358865a),
                          F64(0x3fa4a59a4f02579c), F64(0x3fc16500f106c0a5),
                          F64(0x3ff20dd750429b61)};
const double dErf1Q[5] = {F64(0x3f37ea4332348252), F64(0x3f8166f75999dbd1),
                          F64(0x3fb64536ca92ea2f), F64(0x3fdd0a84eb1ca867),
                          F64(0x3ff0000000000000)};
const double dErfc2P[7] = {0.0, F64(0x3f7cf4cfe0aacbb4), F64(0x3fb2488a6b5cb5e5),
                           F64(0x3fd53dd7a67c7e9f), F64(0x3fec1986509e687b),
                           F64(0x3ff54dfe9b258a60), F64(0x3feffffffbbb552b)};
const double dErfc2Q[7] = {F64(0x3f89a996639b0d00), F64(0x3fc033c113a7deee),
                           F64(0x3fe307622fcff772), F64(0x3ff9e677c2777c3c),
                           F64(0x40053b1052dca8bd), F64(0x4003adeae79b9708),
                           F64(0x3ff0000000000000)};
const double dErfc3P[7] = {0.0, F64(0x3f971d0907ea7a92), F64(0x3fc42210f88b9d43),
                           F64(0x3fe29be1cff90d94), F64(0x3ff44744306832ae),
                           F64(0x3ff9fa202deb88e5), F64(0x3fefff5a9e697ae2)};
const double dErfc3Q[7] = {F64(0x3fa47bd61bbb3843), F64(0x3fd1d7ab774bb837),
                           F64(0x3ff0cfd4cb6cde9f), F64(0x400315ffdfd5ce91),
                           F64(0x400afd487397568f), F64(0x400602f24bf3fdb6),
                           F64(0x3ff0000000000000)};
const double dPi = F64(0x400921fb54442d18);
const double dPio_4 = F64(0x3fe921fb54442d18), dPio_2 = F64(0x3ff921fb54442d18),
             d20pi = F64(0x404f6a7a2955385e), dMedium_pi = F64(0x412921fb54442d18),
             dPio2_1 = F64(0x3FF921FB54400000), dPio2_2 = F64(0x3DD0B4611A600000),
             dPio2_3 = F64(0x3BA3198A2E000000), dPio2_1t = F64(0x3DD0B4611A626331),
             dPio2_2t = F64(0x3BA3198A2E037073), dPio2_3t = F64(0x397B839A252049C1),
             dTwoopi = F64(0x3fe45f306dc9c883), dInvpio_2 = F64(0x3FE45F306DC9C883);
const double dDmpi = (double)F32(0x49490fdb);
// d_trig_evaluation.hpp:58-70 (tan)
const double dTanN[3] = {F64(0xc0c992d8d24f3f38), F64(0x413199eca5fc9ddd), F64(0xc1711fead3299176)};
const double dTanD[5] = {F64(0x3ff0000000000000), F64(0x40cab8a5eeb36572), F64(0xc13427bc582abc96),
                         F64(0x4177d98fc2ead8ef), F64(0xc189afe03cbe5a31)};
// d_trig_evaluation.hpp:33-50
const double dCosC[7] = {F64(0x3da8ff831ad9b219), F64(0xbe21eea7c1e514d4),
                         F64(0x3e927e4f8e06d9a5), F64(0xbefa01a019ddbcd9),
                         F64(0x3f56c16c16c15d47), F64(0xbfa5555555555551),
                         F64(0x3fe0000000000000)};
const double dSinC[6] = {F64(0x3de5d8fd1fcf0ec1), F64(0xbe5ae5e5a9291691),
                         F64(0x3ec71de3567d4896), F64(0xbf2a01a019bfdf03),
                         F64(0x3f8111111110f7d0), F64(0xbfc5555555555548)};

// ===========================================================================
// exp
// ===========================================================================
// f32: modules/core/exponential/include/nt2/exponential/functions/scalar/impl/expo/
//      f_expo_reduction.hpp:73-79 (reduce), :58-72 (approx), expo_scale.hpp:26-35
//      (2^k scaling), simd/common/impl/expo/expo_base.hpp:47-60 (special values).
//      The scalar variant (scalar/impl/expo/expo_base.hpp:33-45) returns the same
//      values, so one lane function serves both.
inline float exp_f(float a0) {
  float k = r2e(fInvlog_2 * a0);
  float x = a0 - k * fLog_2hi;
  x = x - k * fLog_2lo;
  float y = horner(fExpC, x);
  float r = (y * (x * x) + x) + 1.0f;
  float kc = std::fmin(std::fmax(k, -200.0f), 200.0f);
  uint32_t sb = (uint32_t)((int32_t)kc + 127) << 23;
  r = r * as_f(sb);
  r = (a0 <= fMinlog) ? 0.0f : r;
  r = (a0 >= fMaxlog) ? INFINITY : r;
  return (a0 != a0) ? a0 : r;
}

// f64: d_expo_reduction.hpp:57-84 (reduce/approx/finalize), expo_base.hpp:31-42
inline double exp_d(double a0) {
  double k = r2e(dInvlog_2 * a0);
  double hi = a0 - k * dLog_2hi;
  double lo = k * dLog_2lo;
  double x = hi - lo;
  double t = x * x;
  double c = x - t * horner(dExpC, t);
  c = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
  double kc = std::fmin(std::fmax(k, -2000.0), 2000.0);
  uint64_t sb = (uint64_t)((int64_t)kc + 1023) << 52;
  c = c * as_d(sb);
  c = (a0 <= dMinlog) ? 0.0 : c;
  c = (a0 >= dMaxlog) ? (double)INFINITY : c;
  return (a0 != a0) ? a0 : c;
}

// ===========================================================================
// log
// ===========================================================================
// fast_frexp: modules/boost/simd/base/include/boost/simd/ieee/functions/simd/common/fast_frexp.hpp:33-40
// f32 log: simd/common/impl/logs/f_log.hpp:101-125 + :173-186 (finalize),
//          scalar/impl/logs/f_kernel.hpp:47-78.  The scalar variant
//          (scalar/impl/logs/f_log.hpp) returns the same values.
inline float log_f(float a0) {
  float z = a0, t = 0.0f;
  if (std::fabs(z) < fSmallest) { z = z * fTwotonmb; t = fMlogtwo2nmb; }
  uint32_t b = bits(z);
  int32_t e = (int32_t)((b & 0x7f800000u) >> 23) - 126;
  float m = as_f((b & ~0x7f800000u) | 0x3f000000u);
  bool lt = m < fSqrt_2o_2;
  float fe = lt ? (float)e + (-1.0f) : (float)e;
  float x = (lt ? m + m : m) - 1.0f;
  float x2 = x * x;
  float y = (horner(fLogC, x) * x) * x2;
  y = fe * fLog_2lo + y;
  y = -0.5f * x2 + y;
  float r = fLog_2hi * fe + ((x + y) + t);
  if (a0 < 0.0f || a0 != a0) r = NAN;
  if (a0 == INFINITY) r = a0;
  if (a0 == 0.0f) r = -INFINITY;
  return r;
}

// f64 log: simd/common/impl/logs/d_log.hpp:60-97 (no denormal prescale in NT2's f64 path)
inline double log_d(double a0) {
  uint64_t b = bits(a0);
  int64_t k = (int64_t)((b & 0x7ff0000000000000ull) >> 52) - 1022;
  double x = as_d((b & ~0x7ff0000000000000ull) | 0x3fe0000000000000ull);
  bool lt = dSqrt_2o_2 > x;
  k = k + (lt ? -1 : 0);
  double f = (x + (lt ? x : 0.0)) - 1.0;
  double dk = (double)k;
  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  double t1 = w * horner(dLogT1, w);
  double t2 = z * horner(dLogT2, w);
  double R = t2 + t1;
  double hfsq = 0.5 * (f * f);
  double r = dk * dLog_2hi - ((hfsq - (s * (hfsq + R) + dk * dLog_2lo)) - f);
  if (a0 < 0.0 || a0 != a0) r = NAN;
  if (a0 == (double)INFINITY) r = a0;
  if (a0 == 0.0) r = -(double)INFINITY;
  return r;
}

// ===========================================================================
// exp2, exp10, log2, log10
// ===========================================================================
// NT2's expa with the two_tag and ten_tag reductions.  f32:
// scalar/impl/expo/f_expo_reduction.hpp:88-121 (two: k = round2even(a0),
// x = a0 - k, 1 + (y x^2 + x Log_2) with y a degree-4 Remez polynomial) and
// :123-165 (ten: k = round2even(Invlog10_2 a0), x = a0 - k Log10_2hi - k Log10_2lo,
// 1 + P(x) x); limits Minlog2/Maxlog2, Minlog10/Maxlog10 (constants/minlog2.hpp:29,
// maxlog2.hpp:29, minlog10.hpp:29, maxlog10.hpp:29).  Special values and scale as
// exp above (expo_base.hpp:47-60; scalar expo_base.hpp:56-68 gives the same values).
inline float pow2k_scale_f(float a0, float r, float k, float lo, float hi) {
  float kc = std::fmin(std::fmax(k, -200.0f), 200.0f);
  r = r * as_f((uint32_t)((int32_t)kc + 127) << 23);
  if (a0 <= lo) r = 0.0f;
  if (a0 >= hi) r = INFINITY;
  return (a0 != a0) ? a0 : r;
}
inline double pow2k_scale_d(double a0, double c, double k, double lo, double hi) {
  double kc = std::fmin(std::fmax(k, -2000.0), 2000.0);
  c = c * as_d((uint64_t)((int64_t)kc + 1023) << 52);
  if (a0 <= lo) c = 0.0;
  if (a0 >= hi) c = (double)INFINITY;
  return (a0 != a0) ? a0 : c;
}
const float fExp2C[5] = {F32(0x3922c8c4), F32(0x3ab01218), F32(0x3c1d9422),
                         F32(0x3d6356eb), F32(0x3e75fdf1)};
const float fExp10C[6] = {F32(0x3e54dff1), F32(0x3f0aacef), F32(0x3f95eb4c),
                          F32(0x400237da), F32(0x4029a926), F32(0x40135d8e)};
inline float exp2_f(float a0) {
  float k = r2e(a0);
  float x = a0 - k;
  float y = horner(fExp2C, x);
  float r = (y * (x * x) + x * F32(0x3f317218)) + 1.0f;
  return pow2k_scale_f(a0, r, k, F32(0xc2fe0000), F32(0x42fe0000));
}
inline float exp10_f(float a0) {
  float k = r2e(F32(0x40549a78) * a0);
  float x = a0 - k * F32(0x3e9a0000);
  x = x - k * F32(0x39826a14);
  float r = horner(fExp10C, x) * x + 1.0f;
  return pow2k_scale_f(a0, r, k, F32(0xc2179999), F32(0x4218ec59));
}
// f64: scalar/impl/expo/d_expo_reduction.hpp:85-123 (two: x = (a0 - k) Log_2, the
// exp approximation and finalize 1 - ((-(x c)/(2 - c)) - x)) and :125-170 (ten:
// the rational P/Q form, 1 + 2 x P/(Q - x P)); expo_base.hpp:31-42.
const double dExp10P[4] = {F64(0x3fa4fd75f3062dd4), F64(0x40277d9474c55934),
                           F64(0x40796b7a050349e4), F64(0x40a2b4798e134a01)};
const double dExp10Q[4] = {F64(0x3ff0000000000000), F64(0x405545fdce51ca08),
                           F64(0x4093e05eefd67782), F64(0x40a03f37650df6e2)};
inline double exp2_d(double a0) {
  double k = r2e(a0);
  double x = (a0 - k) * F64(0x3fe62e42fefa39ef);
  double t = x * x;
  double c = x - t * horner(dExpC, t);
  c = 1.0 - ((-(x * c) / (2.0 - c)) - x);
  return pow2k_scale_d(a0, c, k, F64(0xc08ff00000000000), F64(0x408ff80000000000));
}
inline double exp10_d(double a0) {
  double k = r2e(F64(0x400a934f0979a372) * a0);
  double x = a0 - k * F64(0x3fd3440000000000);
  x = x - k * F64(0x3ed3509f79fef312);
  double xx = x * x;
  double px = x * horner(dExp10P, xx);
  double x2 = px / (horner(dExp10Q, xx) - px);
  return pow2k_scale_d(a0, 1.0 + (x2 + x2), k, F64(0xc0734413509f79fe), F64(0x40734413509f79fe));
}
// f32 log2 / log10: simd/common/impl/logs/f_log.hpp:127-171 on the f_kernel.hpp
// kernel (the same lines as log_f up to y); denormals are scaled by 2^23 with
// Mlog2two2nmb = -23 / Mlog10two2nmb = -23 log10(2) added
// (constants/mlog2two2nmb.hpp:29, mlog10two2nmb.hpp:29).  The scalar variant
// (scalar/impl/logs/f_log.hpp:83-153) computes the same expressions.
inline void log_kernel_f(float a0, float tden, float& fe, float& x, float& x2, float& y, float& t) {
  float z = a0;
  t = 0.0f;
  if (std::fabs(z) < fSmallest) { z = z * fTwotonmb; t = tden; }
  uint32_t b = bits(z);
  int32_t e = (int32_t)((b & 0x7f800000u) >> 23) - 126;
  float m = as_f((b & ~0x7f800000u) | 0x3f000000u);
  bool lt = m < fSqrt_2o_2;
  fe = lt ? (float)e + (-1.0f) : (float)e;
  x = (lt ? m + m : m) - 1.0f;
  x2 = x * x;
  y = (horner(fLogC, x) * x) * x2;
}
inline float log_special_f(float a0, float r) {
  if (a0 < 0.0f || a0 != a0) r = NAN;
  if (a0 == INFINITY) r = a0;
  if (a0 == 0.0f) r = -INFINITY;
  return r;
}
inline float log2_f(float a0) {
  float fe, x, x2, y, t;
  log_kernel_f(a0, F32(0xc1b80000), fe, x, x2, y, t);
  y = -0.5f * x2 + y;
  const float l2em1 = F32(0x3ee2a8ed);  // log2(e) - 1
  float z = x * l2em1 + y * l2em1;
  return log_special_f(a0, (((z + y) + x) + fe) + t);
}
inline float log10_f(float a0) {
  float fe, x, x2, y, t;
  log_kernel_f(a0, F32(0xc0dd8ede), fe, x, x2, y, t);
  y = -0.5f * x2 + y;                                  // amul(y, Mhalf, x2)
  float z = (x + y) * F32(0x3a37b152);                 // Log10_elo
  z = y * F32(0x3ede0000) + z;                         // Log10_ehi
  z = x * F32(0x3ede0000) + z;
  z = fe * F32(0x3e9a0000) + z;                        // Log10_2hi
  return log_special_f(a0, fe * F32(0x39826a14) + (z + t));  // Log10_2lo
}
// f64 log2 / log10: simd/common/impl/logs/d_log.hpp:101-115 (and scalar
// d_log.hpp:107-137, the same expressions): the log kernel, then
// -((hfsq - s (hfsq + R)) - f) times Invlog_2 plus k, or times Invlog_10 plus
// k Log_2olog_10.
inline double log_frac_d(double a0, double& dk) {
  uint64_t b = bits(a0);
  int64_t k = (int64_t)((b & 0x7ff0000000000000ull) >> 52) - 1022;
  double x = as_d((b & ~0x7ff0000000000000ull) | 0x3fe0000000000000ull);
  bool lt = dSqrt_2o_2 > x;
  k = k + (lt ? -1 : 0);
  double f = (x + (lt ? x : 0.0)) - 1.0;
  dk = (double)k;
  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  double R = z * horner(dLogT2, w) + w * horner(dLogT1, w);
  double hfsq = 0.5 * (f * f);
  return -((hfsq - s * (hfsq + R)) - f);
}
inline double log_special_d(double a0, double r) {
  if (a0 < 0.0 || a0 != a0) r = NAN;
  if (a0 == (double)INFINITY) r = a0;
  if (a0 == 0.0) r = -(double)INFINITY;
  return r;
}
inline double log2_d(double a0) {
  double dk;
  double u = log_frac_d(a0, dk);
  return log_special_d(a0, u * F64(0x3ff71547652b82fe) + dk);
}
inline double log10_d(double a0) {
  double dk;
  double u = log_frac_d(a0, dk);
  return log_special_d(a0, u * F64(0x3fdbcb7b1526e50d) + dk * F64(0x3fd34413509f79ff));
}

// ===========================================================================
// expm1, log1p, tanh
// ===========================================================================
// expm1: modules/core/exponential/include/nt2/exponential/functions/details/
// expm1_kernel.hpp:26-50 (f32) and :53-85 (f64); simd/common/expm1.hpp:29-37
// (-1 below Logeps = constants/logeps.hpp:30, Inf above Maxlog).  2^-k and the
// final 2^k are the exponent-field constructions shl(Maxexponent - k, 23/52)
// and fast_ldexp (boost/simd/ieee/functions/simd/common/fast_ldexp.hpp).
inline float expm1_f(float a0) {
  const float cc[3] = {F32(0x3acf6db4), F32(0xbd08887f), F32(0x3f800000)};
  float k = r2e(fInvlog_2 * a0);
  float x = a0 - k * fLog_2hi;
  x = x - k * fLog_2lo;
  float hx = x * 0.5f;
  float hxs = x * hx;
  float r1 = horner(cc, hxs);
  float t = 3.0f - r1 * hx;
  float e = hxs * ((r1 - t) / (6.0f - x * t));
  e = x * e - hxs;
  int32_t ik = (k != k) ? 0 : (int32_t)std::fmin(std::fmax(k, -200.0f), 200.0f);
  float two2mk = as_f((uint32_t)(127 - ik) << 23);
  float y = (1.0f - two2mk) - (e - x);
  float r = y * as_f((uint32_t)(ik + 127) << 23);
  if (a0 > fMaxlog) r = INFINITY;
  if (a0 < F32(0xc17f1402)) r = -1.0f;
  return r;
}
inline double expm1_d(double a0) {
  const double cc[6] = {F64(0xBE8AFDB76E09C32D), F64(0x3ED0CFCA86E65239), F64(0xBF14CE199EAADBB7),
                        F64(0x3F5A01A019FE5585), F64(0xBFA11111111110F4), F64(0x3ff0000000000000)};
  double k = r2e(dInvlog_2 * a0);
  double hi = a0 - k * dLog_2hi;
  double lo = k * dLog_2lo;
  double x = hi - lo;
  double hxs = (x * x) * 0.5;
  double r1 = horner(cc, hxs);
  double t = 3.0 - (r1 * 0.5) * x;
  double e = hxs * ((r1 - t) / (6.0 - x * t));
  double c = (hi - x) - lo;
  e = (x * (e - c) - c) - hxs;
  int64_t ik = (k != k) ? 0 : (int64_t)std::fmin(std::fmax(k, -2000.0), 2000.0);
  double two2mk = as_d((uint64_t)(1023 - ik) << 52);
  double ct1 = (1.0 - two2mk) - (e - x);
  double ct2 = 1.0 + (x - (e + two2mk));
  double y = (k < 20.0) ? ct1 : ct2;
  double r = y * as_d((uint64_t)(ik + 1023) << 52);
  if (a0 > dMaxlog) r = (double)INFINITY;
  if (a0 < F64(0xc04205966f2b4f19)) r = -1.0;
  return r;
}
// log1p: simd/common/log1p.hpp:26-37 (scalar/log1p.hpp:27-37 gives the same values)
template <class T, class L> inline T log1p_t(T a0, L logf) {
  T u = T(1) + a0;
  T l = logf(u);
  T r = (u != T(0)) ? l + (a0 - (u - T(1))) / u : l;
  return (u == (T)INFINITY) ? u : r;
}
inline float log1p_f(float a0) { return log1p_t(a0, log_f); }
inline double log1p_d(double a0) { return log1p_t(a0, log_d); }
// tanh: modules/core/hyperbolic/include/nt2/hyperbolic/functions/simd/common/tanh.hpp:27-45,
// details/tanh_kernel.hpp:24-36 (f32) and :44-71 (f64)
inline float tanh_f(float a0) {
  const float cc[5] = {F32(0xbbbaf0ea), F32(0x3ca9134e), F32(0xbd5c1e2d), F32(0x3e088393), F32(0xbeaaaa99)};
  float x = std::fabs(a0), r;
  if (x < 0.625f) {
    float x2 = x * x;
    r = (horner(cc, x2) * x2) * x + x;
  } else {
    r = -2.0f * (1.0f / (1.0f + exp_f(x + x))) + 1.0f;
  }
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
inline double tanh_d(double a0) {
  const double p[3] = {F64(0xbfeedc5baafd6f4b), F64(0xc058d26a0e26682d), F64(0xc0993ac030580563)};
  const double q[4] = {F64(0x3FF0000000000000), F64(0x405c33f28a581b86), F64(0x40a176fa0e5535fa),
                       F64(0x40b2ec102442040c)};
  double x = std::fabs(a0), r;
  if (x < 0.625) {
    double x2 = x * x;
    r = ((x2 * horner(p, x2)) / horner(q, x2)) * x + x;
  } else {
    r = -2.0 * (1.0 / (1.0 + exp_d(x + x))) + 1.0;
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}

// sinh: modules/core/hyperbolic/include/nt2/hyperbolic/functions/simd/common/sinh.hpp:27-47,
// details/sinh_kernel.hpp:25-66; cosh: generic/cosh.hpp:27-45 with
// boost/simd average(a, b) = fma(a, 1/2, b/2) (arithmetic/functions/generic/average.hpp)
inline float sinh_f(float a0) {
  const float cc[4] = {F32(0x39559e2f), F32(0x3c087bbe), F32(0x3e2aaacc), F32(0x3f800000)};
  float x = std::fabs(a0), r;
  if (x < 1.0f) {
    r = horner(cc, x * x) * x;
  } else {
    bool big = x > fMaxlog - F32(0x3f317218);
    float e = exp_f(x * (big ? 0.5f : 1.0f));
    float h = 0.5f * e;
    r = big ? h * e : h - 0.5f * (1.0f / e);
  }
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
inline double sinh_d(double a0) {
  const double p[4] = {F64(0xbfe9435fe8bb3cd6), F64(0xc064773a398ff4fe), F64(0xc0c694b8c71d6182),
                       F64(0xc115782bdbf6ab05)};
  const double q[4] = {F64(0x3ff0000000000000), F64(0xc0715b6096e96484), F64(0x40e1a7ba7ed72245),
                       F64(0xc1401a20e4f90044)};
  double x = std::fabs(a0), r;
  if (x < 1.0) {
    double x2 = x * x;
    r = x * ((x2 * horner(p, x2)) / horner(q, x2)) + x;
  } else {
    bool big = x > dMaxlog - F64(0x3fe62e42fefa39ef);
    double e = exp_d(x * (big ? 0.5 : 1.0));
    double h = 0.5 * e;
    r = big ? h * e : h - 0.5 * (1.0 / e);
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
inline float cosh_f(float a0) {
  float x = std::fabs(a0);
  bool big = x > fMaxlog - F32(0x3f317218);
  float e = exp_f(x * (big ? 0.5f : 1.0f));
  return big ? (0.5f * e) * e : e * 0.5f + (1.0f / e) * 0.5f;
}
inline double cosh_d(double a0) {
  double x = std::fabs(a0);
  bool big = x > dMaxlog - F64(0x3fe62e42fefa39ef);
  double e = exp_d(x * (big ? 0.5 : 1.0));
  return big ? (0.5 * e) * e : e * 0.5 + (1.0 / e) * 0.5;
}

// atan: simd/common/impl/invtrig/f_invtrig.hpp:98-133, d_invtrig.hpp:128-166 (packs);
// scalar/impl/invtrig/f_invtrig.hpp:92-128, d_invtrig.hpp:112-159 (pack tail)
// (modules/core/trigonometric/include/nt2/trigonometric/functions/...)
inline float atan_f(float a0, bool scalar) {
  const float cc[4] = {F32(0x3da4f0d1), F32(0xbe0e1b85), F32(0x3e4c925f), F32(0xbeaaaa2a)};
  const float t3 = F32(0x401a827a), t1 = F32(0x3ed413cd), pio2 = F32(0x3fc90fdb), pio4 = F32(0x3f490fdb);
  const float pio2lo = F32(0xb33bbd2e), pio4lo = F32(0xb2bbbd2e);
  float x = std::fabs(a0), r;
  if (scalar) {
    if (a0 == 0.0f) r = 0.0f;
    else if (std::isinf(a0)) r = pio2;
    else {
      float y = 0.0f, more = 0.0f, xx = x;
      if (x > t3) { y = pio2; more = pio2lo; xx = -(1.0f / x); }
      else if (x > t1) { y = pio4; more = pio4lo; xx = (x - 1.0f) / (x + 1.0f); }
      float z = xx * xx;
      float z1 = xx * (horner(cc, z) * z) + xx;
      r = y + (z1 + more);
    }
  } else {
    bool f1 = x < t3, f2 = (x >= t1) && f1;
    float yy = f2 ? pio4 : (f1 ? 0.0f : pio2);
    float xx = f2 ? (x - 1.0f) / (x + 1.0f) : (f1 ? x : -(1.0f / x));
    float z = xx * xx;
    float z1 = xx * (horner(cc, z) * z) + xx;
    if (f2) z1 = z1 + pio4lo;
    if (!f1) z1 = z1 + pio2lo;
    r = yy + z1;
  }
  return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
}
inline double atan_d(double a0, bool scalar) {
  const double p[5] = {F64(0xbfec007fa1f72594), F64(0xc03028545b6b807a), F64(0xc052c08c36880273),
                       F64(0xc05eb8bf2d05ba25), F64(0xc0503669fd28ec8e)};
  const double q[6] = {F64(0x3ff0000000000000), F64(0x4038dbc45b14603c), F64(0x4064a0dd43b8fa25),
                       F64(0x407b0e18d2e2be3b), F64(0x407e563f13b049ea), F64(0x4068519efbbd62ec)};
  const double t3 = F64(0x4003504f333f9de6), t1 = F64(0x3fda827999fcef31);
  const double pio2 = F64(0x3ff921fb54442d18), pio4 = F64(0x3fe921fb54442d18);
  const double pio2lo = F64(0x3c91a62633145c07), pio4lo = F64(0x3c81a62633145c07);
  double x = std::fabs(a0), r;
  if (scalar) {
    if (a0 == 0.0) r = 0.0;
    else if (std::isinf(a0)) r = pio2;
    else {
      double y, flag = (x > t3) ? 1.0 : 0.0, xx = x;
      if (x > t3) { y = pio2; xx = -(1.0 / x); }
      else if (x <= F64(0x3fe51eb851eb851f)) y = 0.0;
      else { y = pio4; flag = 0.5; xx = (x - 1.0) / (x + 1.0); }
      double z = xx * xx;
      double zz = (z * horner(p, z)) / horner(q, z);
      zz = xx * zz + xx;
      zz = zz + flag * pio2lo;
      r = y + zz;
    }
  } else {
    bool f1 = x < t3, f2 = (x >= t1) && f1;
    double yy = f2 ? pio4 : (f1 ? 0.0 : pio2);
    double xx = f2 ? (x - 1.0) / (x + 1.0) : (f1 ? x : -(1.0 / x));
    double z = xx * xx;
    double zz = (z * horner(p, z)) / horner(q, z);
    zz = xx * zz + xx;
    if (f2) zz = zz + pio4lo;
    zz = zz + (f1 ? 0.0 : pio2lo);
    r = yy + zz;
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
// asin, acos (f32): simd/common/impl/invtrig/f_invtrig.hpp:55-97; the scalar
// forms of scalar/impl/invtrig/f_invtrig.hpp:49-90 give the same values (the
// early return below 1e-4 is what the polynomial rounds to; x > 1 is NaN both
// ways), so one restatement serves packs and tails.
inline float asin_f(float a0) {
  const float c[5] = {F32(0x3d2cb352), F32(0x3cc617e3), F32(0x3d3a3ec7), F32(0x3d9980f6), F32(0x3e2aaae4)};
  float x = std::fabs(a0);
  const bool big = x > 0.5f;
  const float z = big ? 0.5f * (1.0f - x) : x * x;
  if (big) x = std::sqrt(z);
  float z1 = horner(c, z);
  z1 = z1 * (z * x) + x;
  if (big) z1 = F32(0x3fc90fdb) - (z1 + z1);
  return as_f(bits(z1) ^ (bits(a0) & 0x80000000u));
}
inline float acos_f(float a0) {
  float x = std::fabs(a0);
  const bool big = x > 0.5f;
  x = big ? std::sqrt(-0.5f * x + 0.5f) : a0;
  x = asin_f(x);
  if (big) x = x + x;
  if (a0 < -0.5f) x = F32(0x40490fdb) - x;
  return big ? x : F32(0x3fc90fdb) - x;
}
// asin, acos (f64): simd/common/impl/invtrig/d_invtrig.hpp:57-127 (packs: both
// branches evaluated, the low part of pi/2 subtracted, acos through
// pi - 2 asin(sqrt((1-|x|)/2)) below -1/2); scalar/impl/invtrig/d_invtrig.hpp:46-111
// (pack tail: the low part of pi/2 added at :71, acos below 1/2 as
// (pi/4 - asin(x) + pi/4lo) + pi/4).
inline double asin_d(double a0, bool scalar) {
  const double rr[5] = {F64(0x3f684fc3988e9f08), F64(0xbfe2079259f9290f), F64(0x401bdff5baf33e6a),
                        F64(0xc03991aaac01ab68), F64(0x403c896240f3081d)};
  const double ss[5] = {F64(0x3ff0000000000000), F64(0xc035f2a2b6bf5d8c), F64(0x40626219af6a7f42),
                        F64(0xc077fe08959063ee), F64(0x40756709b0b644be)};
  const double pp[6] = {F64(0x3f716b9b0bd48ad3), F64(0xbfe34341333e5c16), F64(0x4015c74b178a2dd9),
                        F64(0xc0304331de27907b), F64(0x40339007da779259), F64(0xc020656c06ceafd5)};
  const double qq[6] = {F64(0x3ff0000000000000), F64(0xc02d7b590b5e0eab), F64(0x40519fc025fe9054),
                        F64(0xc06265bb6d3576d7), F64(0x4061705684ffbf9d), F64(0xc04898220a3607ac)};
  const double pio4 = F64(0x3fe921fb54442d18), pio2lo = F64(0x3c91a62633145c07);
  const double sqrteps = F64(0x3e50000000000000), ct1 = F64(0x3fe4000000000000);
  const double x = std::fabs(a0);
  double r;
  if (scalar) {
    if (x > 1.0) return std::numeric_limits<double>::quiet_NaN();
    if (x < sqrteps) return a0;
    if (x > ct1) {
      double zz = 1.0 - x;
      const double vp = (zz * horner(rr, zz)) / horner(ss, zz);
      zz = std::sqrt(zz + zz);
      double z = pio4 - zz;
      zz = zz * vp + pio2lo;
      z = z - zz;
      r = z + pio4;
    } else {
      const double zz = x * x;
      const double z = (zz * horner(pp, zz)) / horner(qq, zz);
      r = x * z + x;
    }
  } else {
    double zz1 = 1.0 - x;
    const double vp = (zz1 * horner(rr, zz1)) / horner(ss, zz1);
    zz1 = std::sqrt(zz1 + zz1);
    double z = pio4 - zz1;
    zz1 = zz1 * vp - pio2lo;
    z = z - zz1;
    zz1 = z + pio4;
    const double zz2a = a0 * a0;
    z = (zz2a * horner(pp, zz2a)) / horner(qq, zz2a);
    const double zz2 = x * z + x;
    if (x > 1.0) return std::numeric_limits<double>::quiet_NaN();
    r = x < sqrteps ? x : (x > ct1 ? zz1 : zz2);
  }
  return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
inline double acos_d(double a0, bool scalar) {
  const double pi = F64(0x400921fb54442d18), pio2 = F64(0x3ff921fb54442d18);
  const double pio4 = F64(0x3fe921fb54442d18), pio4lo = F64(0x3c81a62633145c07);
  if (scalar) {
    if (std::fabs(a0) > 1.0) return std::numeric_limits<double>::quiet_NaN();
    if (a0 > 0.5) return 2.0 * asin_d(std::sqrt(-0.5 * a0 + 0.5), true);
    double z = pio4 - asin_d(a0, true);
    z = z + pio4lo;
    return z + pio4;
  }
  double x = std::fabs(a0);
  const bool big = x > 0.5;
  x = big ? std::sqrt(-0.5 * x + 0.5) : a0;
  x = asin_d(x, false);
  if (big) x = x + x;
  if (a0 < -0.5) x = pi - x;
  return big ? x : pio2 - x;
}

// pow(a0, a1): modules/core/exponential/include/nt2/exponential/functions/
// simd/common/pow.hpp:38-46 and pow_abs.hpp:65-133 (packs), scalar/pow.hpp:30-36
// and pow_abs.hpp:51-111 (pack tail), kernels details/pow_kernel.hpp:30-228.
// Cephes' pow: log2 of the significand against a 17-entry table of 2^(-i/16),
// the product with a1 split in sixteenths, 2^(fraction) by a polynomial.
// Helpers restate boost.simd: frexp (ieee/functions/simd/common/frexp.hpp:52-75:
// Inf and NaN give exponent 0 and mantissa a0 + 0.5), ldexp
// (ieee/functions/simd/common/ldexp.hpp:57-71: two products, the second by
// 2^Minexponent or 2), toint (cvtt: INT_MIN when out of range or NaN),
// is_flint (x - trunc(x) == 0, false for Inf/NaN), is_odd(x) =
// is_flint((x - 1) / 2) (predicates/functions/simd/common/{is_odd,is_even,is_flint}.hpp).
template <class T> struct powk;
template <> struct powk<float> {
  using I = int32_t;
  using U = uint32_t;
  static constexpr int nbm = 23, maxe = 127, mine = -126, lime = 128;
  static float A(int i) {
    static const uint32_t t[17] = {0x3f800000, 0x3f75257d, 0x3f6ac0c7, 0x3f60ccdf, 0x3f5744fd, 0x3f4e248c,
                                   0x3f45672a, 0x3f3d08a4, 0x3f3504f3, 0x3f2d583f, 0x3f25fed7, 0x3f1ef532,
                                   0x3f1837f0, 0x3f11c3d3, 0x3f0b95c2, 0x3f05aac3, 0x3f000000};
    return as_f(t[i]);
  }
  static float B(int i) {
    static const uint32_t t[9] = {0x00000000, 0xb1c116de, 0xb254a58a, 0x318aa837, 0x324fe77a,
                                  0xb2ac9d5e, 0x32a31b71, 0xb1e0aba1, 0x00000000};
    return as_f(t[i]);
  }
  static float pow1(float x, float z) {
    const float c[4] = {F32(0xbe2a61b2), F32(0x3e4d2fa3), F32(0xbe800015), F32(0x3eaaaaa3)};
    return (z * x) * horner(c, x);
  }
  static float pow2(float x) {
    const float c[4] = {F32(0x3c1a49bc), F32(0x3d634d38), F32(0x3e75fde1), F32(0x3f317218)};
    return horner(c, x);
  }
  static float log2em1() { return F32(0x3ee2a8ed); }
  static float large() { return F32(0x44ffe000); }
  static float low() { return F32(0xc5160000); }
  static float tiny() { return F32(0x00800000); }
  static float bits2(U u) { return as_f(u); }
};
template <> struct powk<double> {
  using I = int64_t;
  using U = uint64_t;
  static constexpr int nbm = 52, maxe = 1023, mine = -1022, lime = 1024;
  static double A(int i) {
    static const uint64_t t[17] = {
        0x3ff0000000000000, 0x3feea4afa2a490da, 0x3fed5818dcfba487, 0x3fec199bdd85529c, 0x3feae89f995ad3ad,
        0x3fe9c49182a3f090, 0x3fe8ace5422aa0db, 0x3fe7a11473eb0187, 0x3fe6a09e667f3bcd, 0x3fe5ab07dd485429,
        0x3fe4bfdad5362a27, 0x3fe3dea64c123422, 0x3fe306fe0a31b715, 0x3fe2387a6e756238, 0x3fe172b83c7d517b,
        0x3fe0b5586cf9890f, 0x3fe0000000000000};
    return as_d(t[i]);
  }
  static double B(int i) {
    static const uint64_t t[9] = {0x0000000000000000, 0x3c72ed02d75b3707, 0x3c87a1cd345dcc81,
                                  0x3c86e9f156864b27, 0xbc8bdd3413b26456, 0x3c6d4397afec42e2,
                                  0x3c76f46ad23182e4, 0xbc719041b9d78a76, 0x0000000000000000};
    return as_d(t[i]);
  }
  static double pow1(double x, double z) {
    const double p[4] = {F64(0x3fdfdb997f5b5cf0), F64(0x400dddefea9edf15), F64(0x401eccbd7f78eb6f),
                         F64(0x4012aa83b65c9b74)};
    const double q[5] = {F64(0x3ff0000000000000), F64(0x4022aab49b20914e), F64(0x403bffff41c1c9f5),
                         F64(0x4040ccbc1b176402), F64(0x402bffc5918ae92e)};
    return x * ((z * horner(p, x)) / horner(q, x));
  }
  static double pow2(double x) {
    const double c[7] = {F64(0x3eef6307d7f2937f), F64(0x3f242fbe60fc9259), F64(0x3f55d87ec84aef1d),
                         F64(0x3f83b2ab6ef133b7), F64(0x3fac6b08d7041a92), F64(0x3fcebfbdff82c56d),
                         F64(0x3fe62e42fefa39ef)};
    return horner(c, x);
  }
  static double log2em1() { return F64(0x3fdc551d94ae0bf8); }
  static double large() { return F64(0x40cfff8000000000); }
  static double low() { return F64(0xc0d0c7c000000000); }
  static double tiny() { return F64(0x0010000000000000); }
  static double bits2(U u) { return as_d(u); }
};
template <class T> inline typename powk<T>::I pw_toint(T x) {
  using I = typename powk<T>::I;
  const T lim = (T)std::numeric_limits<I>::max();
  if (!(x > -lim - 1 && x < lim)) return std::numeric_limits<I>::min();
  return (I)x;
}
template <class T> inline T pw_frexp(T a, typename powk<T>::I& e) {
  if (a == 0) { e = 0; return a; }
  if (!std::isfinite(a)) { e = 0; return a + T(0.5); }
  int ee;
  const T m = std::frexp(a, &ee);
  e = ee;
  return m;
}
template <class T> inline T pw_ldexp(T a, typename powk<T>::I e) {
  using K = powk<T>;
  using U = typename K::U;
  T f = T(1);
  if (e < K::mine) { e -= K::mine; f = K::tiny(); }
  if (e == K::lime) { f = f + T(1); e -= 1; }
  const U b = (U)(e + K::maxe) << K::nbm;
  return (a * K::bits2(b)) * f;
}
template <class T> inline bool pw_flint(T x) { return std::isfinite(x) && (x - std::trunc(x)) == T(0); }
template <class T> inline bool pw_odd(T x) { return pw_flint((x - T(1)) * T(0.5)); }
template <class T> inline T pw_reduc(T x) { return T(0.0625) * std::floor(T(16) * x); }

// the arithmetic common to both forms: sets w (sixteenths of the exponent)
// and returns the unclamped result
template <class T> inline T pow_core(T ax, T a1, T& wout) {
  using K = powk<T>;
  using I = typename K::I;
  I e;
  T x = pw_frexp(ax, e);
  // select: the table entry just above the significand (pow_kernel.hpp:88-102)
  I i = 1;
  if (x <= K::A(9)) i = 9;
  if (x <= K::A(i + 4)) i += 4;
  if (x <= K::A(i + 2)) i += 2;
  if (x >= K::A(1)) i = -1;
  i = i + 1;
  const T tmp = K::A((int)i);
  x = x - tmp;
  x = x - K::B((int)(i >> 1));
  x = x / tmp;
  T z = x * x;
  T w = K::pow1(x, z);
  w = T(-0.5) * z + w;
  w = K::log2em1() * w + w;
  z = K::log2em1() * x + w;
  z = z + x;
  w = (T)(-i) * T(0.0625) + (T)e;
  const T ya = pw_reduc(a1);
  const T yb = a1 - ya;
  T W = z * a1 + w * yb;
  T Wa = pw_reduc(W);
  T Wb = W - Wa;
  W = w * ya + Wa;
  Wa = pw_reduc(W);
  const T u = W - Wa;
  W = Wb + u;
  Wb = pw_reduc(W);
  w = T(16) * (Wa + Wb);
  wout = w;
  e = pw_toint(w);
  Wb = W - Wb;
  if (Wb > T(0)) { e = e + 1; Wb = Wb - T(0.0625); }
  z = K::pow2(Wb) * Wb;
  // i = e/16 + (e > 0): the pack form (the tail's e >= 0 gives the same value)
  i = e / 16 + (e > 0 ? 1 : 0);
  e = i * 16 - e;
  const T wt = K::A((e >= 0 && e <= 16) ? (int)e : 0);  // other e only in lanes replaced below
  z = wt * z + wt;
  return pw_ldexp(z, i);
}
template <class T> inline T pow_abs_t(T a0, T a1, bool scalar) {
  using K = powk<T>;
  const T inf = std::numeric_limits<T>::infinity();
  const T x = std::fabs(a0);
  if (scalar) {
    if (x == T(1)) return x;
    if (x == T(0)) return a1 == T(0) ? T(1) : T(0);
    if (x == a1 && a1 == inf) return inf;
    if (x == inf && a1 == -inf) return T(0);
    if (a1 == inf) return (x < T(1)) ? T(0) : inf;
    if (a1 == -inf) return (x > T(1)) ? T(0) : inf;
    if (x == inf) return (a1 < T(0)) ? T(0) : ((a1 == T(0)) ? T(1) : inf);
    if (std::isnan(a0) || std::isnan(a1)) return std::numeric_limits<T>::quiet_NaN();
    T w;
    const T z = pow_core(x, a1, w);
    if (w > K::large()) return inf;
    if (w < K::low()) return T(0);
    return z;
  }
  T w;
  T z = pow_core(x, a1, w);
  const bool gtax1 = x > T(1);
  if (a1 == inf) z = gtax1 ? inf : T(0);
  if (a1 == -inf) z = gtax1 ? T(0) : inf;
  if (x == inf) z = (a1 > T(0)) ? inf : ((a1 < T(0)) ? T(0) : T(1));
  if (w < K::low()) z = T(0);
  if (w > K::large()) z = inf;
  if (x == T(1)) z = x;
  if (a0 == T(0)) z = (a1 == T(0)) ? T(1) : T(0);
  return z;
}
template <class T> inline T pow_t(T a0, T a1, bool scalar) {
  const bool ltz = a0 < T(0);
  if (scalar && ltz && !pw_flint(a1)) return std::numeric_limits<T>::quiet_NaN();
  T z = pow_abs_t(a0, a1, scalar);
  if (ltz && pw_odd(a1)) z = -z;
  if (ltz && !pw_flint(a1)) z = std::numeric_limits<T>::quiet_NaN();
  return z;
}

// atan2: simd/common/atan2.hpp:30-54, scalar/atan2.hpp:29-45
template <class T, class K> inline T atan2_t(T a0, T a1, T pi, K kernel) {
  if (a0 != a0 || a1 != a1) return a0 + a1;
  if (std::isinf(a0) && std::isinf(a1)) { a0 = std::copysign(T(1), a0); a1 = std::copysign(T(1), a1); }
  T z = kernel(std::fabs(a0 / a1));
  z = (a1 > T(0)) ? z : pi - z;
  if (std::signbit(a0)) z = -z;
  return (a0 == T(0)) ? ((a1 < T(0)) ? pi : T(0)) : z;
}
inline float atan2_f(float a0, float a1, bool scalar) {
  return atan2_t(a0, a1, F32(0x40490fdb), [&](float q) { return atan_f(q, scalar); });
}
inline double atan2_d(double a0, double a1, bool scalar) {
  return atan2_t(a0, a1, F64(0x400921fb54442d18), [&](double q) { return atan_d(q, scalar); });
}

// ===========================================================================
// rounding and sign
// ===========================================================================
// floor / ceil / trunc / round2even are the IEEE operations (SSE4.1/AVX
// roundps with the matching mode).  round: half away from zero,
// modules/boost/simd/base/include/boost/simd/arithmetic/functions/scalar/round.hpp:30-44
// (::roundf / ::round, or the ceil-based form, which agree for every input).
// sign: boost/simd/ieee/functions/simd/common/sign.hpp:55-66, (a > 0) - (a < 0),
// NaN for NaN.
template <class T> inline T sign_of(T a) {
  if (a != a) return a;
  return (T)(a > T(0)) - (T)(a < T(0));
}

// ===========================================================================
// erfc, normcdf, fastnormcdf
// ===========================================================================
// f32: modules/core/euler/include/nt2/euler/functions/simd/common/erfc.hpp:92-123
//      (scalar/erfc.hpp:72-100 returns the same values); kernels erf_kernel.hpp:46-85.
inline float erfc_f(float a0) {
  float x = std::fabs(a0);
  float z = x / (1.0f + x);
  float r;
  if (x < fTwothird) {
    r = (1.0f - z) * horner(fErfc3C, z);
  } else {
    float z2 = z - 0.4f;
    r = exp_f(-(x * x)) * horner(fErfc2C, z2);
  }
  if (x == INFINITY) r = 0.0f;
  return (a0 < 0.0f) ? 2.0f - r : r;
}

// f64 SIMD: simd/common/erfc.hpp:46-89.  f64 scalar (pack tail): scalar/erfc.hpp:38-70.
inline double erfc_d(double a0, bool scalar) {
  if (scalar) {
    if (a0 != a0) return a0;
    double x = std::fabs(a0), xx = x * x, z = 0.0;
    if (x <= 0.0000000001) z = 1.0 - (x * 2.0) / std::sqrt(dPi);
    else if (x < 0.65) z = 1.0 - x * (horner(dErf1P, xx) / horner(dErf1Q, xx));
    else if (x < 2.2) z = exp_d(-xx) * (horner(dErfc2P, x) / horner(dErfc2Q, x));
    else if (x < 6.0) z = exp_d(-xx) * (horner(dErfc3P, x) / horner(dErfc3Q, x));
    return (a0 < 0.0) ? 2.0 - z : z;
  }
  double x = std::fabs(a0), xx = x * x, r;
  if (x < 0.65) r = 1.0 - x * (horner(dErf1P, xx) / horner(dErf1Q, xx));
  else if (x < 2.2) r = exp_d(-xx) * (horner(dErfc2P, x) / horner(dErfc2Q, x));
  else r = exp_d(-xx) * (horner(dErfc3P, x) / horner(dErfc3Q, x));
  if (x == (double)INFINITY) r = 0.0;
  return (a0 < 0.0) ? 2.0 - r : r;
}

// normcdf(x) = Half*erfc(-Sqrt_2o_2*x):
// modules/core/statistics/include/nt2/statistics/functions/generic/normcdf.hpp:33
inline float normcdf_f(float a0) { return 0.5f * erfc_f((-fSqrt_2o_2) * a0); }
inline double normcdf_d(double a0, bool scalar) {
  return 0.5 * erfc_d((-dSqrt_2o_2) * a0, scalar);
}

// fastnormcdf (single only): generic/fastnormcdf.hpp:39-54
inline float fastnormcdf_f(float a0) {
  float l = std::fabs(a0);
  float k = 1.0f / (0.2316419f * l + 1.0f);
  float w = horner(fFastC, k);
  w = w * ((k * fInvsqrt_2pi) * exp_f((-(l * l)) * 0.5f));
  return (a0 > 0.0f) ? 1.0f - w : w;
}

// ===========================================================================
// erf, cbrt, asinh, acosh, atanh
// ===========================================================================
// erf f32 (one form): modules/core/euler/include/nt2/euler/functions/simd/
// common/erf.hpp:94-124; scalar/erf.hpp:76-104 gives the same values.
// erf1: erf_kernel.hpp:34-44.
const float fErf1C[6] = {F32(0xba1fc83b), F32(0x3ba4468d), F32(0xbcdb61f4),
                         F32(0x3de70f22), F32(0xbec0937e), F32(0x3f906eba)};
inline float erf_f(float a0) {
  float x = std::fabs(a0);
  if (std::isinf(a0)) return a0 > 0.0f ? 1.0f : -1.0f;
  if (x < fTwothird) return a0 * horner(fErf1C, x * x);
  float z = x / (1.0f + x) - 0.4f;
  float r = 1.0f - exp_f(-(x * x)) * horner(fErfc2C, z);
  return (a0 < 0.0f) ? -r : r;
}

// erf f64.  SIMD: simd/common/erf.hpp:47-92.  Scalar (a pack's tail):
// scalar/erf.hpp:42-74, whose last branch returns One for |x| > 6 whatever
// the sign of a0.
inline double erf_d(double a0, bool scalar) {
  double x = std::fabs(a0), xx = x * x;
  if (scalar) {
    if (a0 != a0) return a0;
    if (x <= 0.0000000001) return a0 * 2.0 / std::sqrt(dPi);
    if (x <= 0.65) return a0 * (horner(dErf1P, xx) / horner(dErf1Q, xx));
    if (x <= 2.2) {
      double z = 1.0 - exp_d(-xx) * (horner(dErfc2P, x) / horner(dErfc2Q, x));
      return (a0 < 0.0) ? -z : z;
    }
    if (x <= 6.0) {
      double z = 1.0 - exp_d(-xx) * (horner(dErfc3P, x) / horner(dErfc3Q, x));
      return (a0 < 0.0) ? -z : z;
    }
    return 1.0;
  }
  if (x < 0.65) return a0 * (horner(dErf1P, xx) / horner(dErf1Q, xx));
  double ex = exp_d(-xx), z;
  if (x < 2.2) z = 1.0 - ex * (horner(dErfc2P, x) / horner(dErfc2Q, x));
  else z = 1.0 - ex * (horner(dErfc3P, x) / horner(dErfc3Q, x));
  z = (a0 < 0.0) ? -z : z;
  if (std::isinf(a0)) z = (a0 > 0.0) ? 1.0 : -1.0;
  return z;
}

// cbrt: modules/core/exponential/include/nt2/exponential/functions/simd/
// common/cbrt.hpp:52-166 (scalar/cbrt.hpp:42-156: the same steps).
// fast_frexp: mantissa in [1/2, 1) and exponent from the bits; fast_ldexp:
// times 2^e built from bits (boost/simd/ieee/functions/scalar/fast_ldexp.hpp).
template <class T> struct cbrt_k;
template <> struct cbrt_k<float> {
  using U = uint32_t;
  static constexpr int nmb = 23, bias = 127, newton = 1;
  static constexpr U expmask = 0x7f800000u, half = 0x3f000000u, sign = 0x80000000u;
  static float smallest() { return F32(0x00800000); }
  static float twotonmb() { return F32(0x4B000000); }
  static float twotomnmbo3() { return F32(0x3ba14518); }
  static float third() { return F32(0x3EAAAAAB); }
  static float cb2(bool p) { return p ? F32(0x3fa14518) : F32(0x3f4b2ff5); }
  static float cb4(bool p) { return p ? F32(0x3fcb2ff5) : F32(0x3f214518); }
  static float poly(float x) {
    static const float c[5] = {F32(0xbe09e49a), F32(0x3f0bf0fe), F32(0xbf745265),
                               F32(0x3f91eb77), F32(0x3ece0609)};
    return horner(c, x);
  }
};
template <> struct cbrt_k<double> {
  using U = uint64_t;
  static constexpr int nmb = 52, bias = 1023, newton = 2;
  static constexpr U expmask = 0x7ff0000000000000ull, half = 0x3fe0000000000000ull,
                     sign = 0x8000000000000000ull;
  static double smallest() { return F64(0x0010000000000000); }
  static double twotonmb() { return F64(0x4330000000000000); }
  static double twotomnmbo3() { return F64(0x3ed428a2f98d7286); }
  static double third() { return F64(0x3FD5555555555555); }
  static double cb2(bool p) { return p ? F64(0x3ff428a2f98d728b) : F64(0x3fe965fea53d6e3d); }
  static double cb4(bool p) { return p ? F64(0x3ff965fea53d6e3d) : F64(0x3fe428a2f98d728b); }
  static double poly(double x) {
    static const double c[5] = {F64(0xbfc13c93386fdff6), F64(0x3fe17e1fc7e59d58), F64(0xbfee8a4ca3ba37b8),
                                F64(0x3ff23d6ee505873a), F64(0x3fd9c0c12122a4fe)};
    return horner(c, x);
  }
};
template <class T> inline T cbrt_t(T a0) {
  using K = cbrt_k<T>;
  using U = typename K::U;
  T z = std::fabs(a0);
  T f = T(1);
  if (z < K::smallest()) {
    z *= K::twotonmb();
    f = K::twotomnmbo3();
  }
  U b = __builtin_bit_cast(U, z);
  long e = (long)((b & K::expmask) >> K::nmb) - (K::bias - 1);
  T x = K::poly(__builtin_bit_cast(T, (U)((b & ~K::expmask) | K::half)));
  bool flag = e >= 0;
  long e1 = flag ? e : -e;
  long rem = e1;
  e1 /= 3;
  rem -= e1 * 3;
  e = (e > 0) ? e1 : ((e < 0) ? -e1 : 0);  // negate(e1, e) = e1 * sign(e)
  T fact = (rem == 1) ? K::cb2(flag) : T(1);
  fact = (rem == 2) ? K::cb4(flag) : fact;
  x = (x * fact) * __builtin_bit_cast(T, (U)((U)(e + K::bias) << K::nmb));
  for (int k = 0; k < K::newton; ++k) x -= (x - z / (x * x)) * K::third();
  x = __builtin_bit_cast(T, (U)(__builtin_bit_cast(U, x) | (__builtin_bit_cast(U, a0) & K::sign))) * f;
  if (a0 == T(0) || std::isinf(a0)) return a0;
  return x;
}

// hypot as NT2 writes it (boost/simd/arithmetic/functions/scalar/hypot.hpp:
// 24-40): both operands scaled by 2^-e, e = exponent(max) clamped to
// [Minexponent, Maxexponentm1], the root scaled back.
template <class T> inline T hypot_nt2(T a, T b) {
  if ((std::isnan(a) && std::isinf(b)) || (std::isinf(a) && std::isnan(b))) return INFINITY;
  T r = std::fabs(a), i = std::fabs(b);
  T m = std::max(i, r);
  int e = (m == T(0) || !std::isfinite(m)) ? 0 : std::ilogb(m);
  const int lo = sizeof(T) == 4 ? -126 : -1022, hi = sizeof(T) == 4 ? 126 : 1022;
  e = std::min(std::max(e, lo), hi);
  T sr = std::ldexp(r, -e), si = std::ldexp(i, -e);
  return std::ldexp(std::sqrt(sr * sr + si * si), e);
}

// asinh f32.  SIMD: modules/core/hyperbolic/include/nt2/hyperbolic/functions/
// simd/common/asinh.hpp:66-106; scalar (tail): scalar/asinh.hpp:68-111.
// average(a, b) = fma(a, 1/2, b/2) (boost/simd/arithmetic/functions/generic/average.hpp:40).
const float fAsinhC[5] = {F32(0x3ca4d6e6), F32(0xbd2ee581), F32(0x3d9949b1), F32(0xbe2aa9ad),
                          F32(0x3f800000)};
const float fLog_2 = F32(0x3f317218), fOneosqrteps = F32(0x453504f3), fOneotwoeps = F32(0x4A800000);
const double dLog_2 = F64(0x3fe62e42fefa39ef), dOneosqrteps = F64(0x4190000000000000),
             dOneotwoeps = F64(0x4320000000000000), dSqrteps = F64(0x3e50000000000000);
inline float asinh_f(float a0, bool scalar) {
  float x = std::fabs(a0), z;
  if (x < 0.5f) {
    z = horner(fAsinhC, x * x) * x;
  } else if (x > fOneosqrteps) {
    z = log_f(x) + fLog_2;
  } else if (scalar) {
    z = log_f(x + hypot_nt2(1.0f, x));
  } else {
    z = log_f(x * 0.5f + hypot_nt2(1.0f, x) * 0.5f) + fLog_2;
  }
  return as_f(bits(z) ^ (bits(a0) & 0x80000000u));
}
// asinh f64.  SIMD: simd/common/asinh.hpp:46-64; scalar: scalar/asinh.hpp:32-66.
inline double asinh_d(double a0, bool scalar) {
  double x = std::fabs(a0), z;
  if (scalar) {
    if (x < dSqrteps) return a0;
    if (x < 0.5) {
      double invx = 1.0 / x;
      z = log1p_d(x + x / (invx + std::sqrt(invx * invx + 1.0)));
    } else if (x > dOneosqrteps) {
      z = log_d(x) + dLog_2;
    } else {
      z = log_d(x + hypot_nt2(1.0, x));
    }
  } else {
    bool test = x > dOneosqrteps;
    z = test ? x - 1.0 : x + (x * x) / (1.0 + hypot_nt2(1.0, x));
    if (x == (double)INFINITY) z = x;
    z = test ? log1p_d(z) + dLog_2 : log1p_d(z);
  }
  return as_d(bits(z) ^ (bits(a0) & 0x8000000000000000ull));
}

// acosh: simd/common/acosh.hpp:26-42; scalar (tail): scalar/acosh.hpp:22-38.
template <class T, class L1, class L>
inline T acosh_t(T a0, bool scalar, T big, T ln2, L1 lg1p, L lg) {
  T t = a0 - T(1);
  if (scalar) {
    if (t <= big) return lg1p(t + std::sqrt(t + t + t * t));
    return lg(t) + ln2;
  }
  bool test = t > big;
  T z = test ? a0 : t + std::sqrt(t + t + t * t);
  return test ? lg1p(z) + ln2 : lg1p(z);
}
inline float acosh_f(float a0, bool scalar) { return acosh_t(a0, scalar, fOneotwoeps, fLog_2, log1p_f, log_f); }
inline double acosh_d(double a0, bool scalar) { return acosh_t(a0, scalar, dOneotwoeps, dLog_2, log1p_d, log_d); }

// atanh (one form): simd/common/atanh.hpp:28-46 = scalar/atanh.hpp:22-44.
template <class T, class L1> inline T atanh_t(T a0, L1 lg1p) {
  T ax = std::fabs(a0), t = ax + ax, z1 = T(1) - ax;
  bool test = ax < T(0.5);
  T tmp = (test ? ax : t) / z1;
  T r = T(0.5) * lg1p(test ? t * tmp + t : tmp);
  if constexpr (sizeof(T) == 4) return as_f(bits(r) ^ (bits(a0) & 0x80000000u));
  else return as_d(bits(r) ^ (bits(a0) & 0x8000000000000000ull));
}
inline float atanh_f(float a0) { return atanh_t(a0, log1p_f); }
inline double atanh_d(double a0) { return atanh_t(a0, log1p_d); }

// ===========================================================================
// trigonometric range reduction
// ===========================================================================
int g_ph_fix = 0;

// fdlibm __kernel_rem_pio2 / __kernel_rem_pio2f as embedded by NT2:
// modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:234-402
// (single, 8-bit chunks, jk = 9) and :690-905 (double, 24-bit chunks, jk = 4);
// both called with prec = 2.  One template restates both, parameterised by the
// chunk width B.
template <class T, int B>
int kernel_rem_pio2(const T* x, T* y, int e0, int nx, const int32_t* ipio2,
                    const T* PIo2) {
  const int jk = (B == 8) ? 9 : 4;
  const int jp = jk;
  const T two_b = (T)(1 << B);
  const T twon_b = (T)1 / two_b;
  const int32_t full = 1 << B;          // 0x100 / 0x1000000
  const int32_t maskall = full - 1;     // 0xff / 0xffffff
  int32_t jz, jx, jv, carry, n, iq[20], i, j, k, m, q0, ih;
  T z, fw, f[20], fq[20], q[20];

  jx = nx - 1;
  jv = (e0 - 3) / B;
  if (jv < 0) jv = 0;
  q0 = e0 - B * (jv + 1);
  j = jv - jx;
  m = jx + jk;
  for (i = 0; i <= m; i++, j++) f[i] = (j < 0) ? (T)0 : (T)ipio2[j];
  for (i = 0; i <= jk; i++) {
    for (j = 0, fw = 0; j <= jx; j++) fw += x[j] * f[jx + i - j];
    q[i] = fw;
  }
  jz = jk;
recompute:
  for (i = 0, j = jz, z = q[jz]; j > 0; i++, j--) {
    fw = (T)((int32_t)(twon_b * z));
    iq[i] = (int32_t)(z - two_b * fw);
    z = q[j - 1] + fw;
  }
  z = std::ldexp(z, q0);
  z -= (T)8 * std::floor(z * (T)0.125);
  n = (int32_t)z;
  z -= (T)n;
  ih = 0;
  if (q0 > 0) {
    i = (iq[jz - 1] >> (B - q0));
    n += i;
    iq[jz - 1] -= i << (B - q0);
    ih = iq[jz - 1] >> (B - 1 - q0);
  } else if (q0 == 0) {
    // NT2 single: iq>>8 (rem_pio2.hpp:290); double: iq>>23 (:791).  For
    // 8-bit chunks the top bit is bit 7, so NT2's shift by 8 always yields 0
    // and about 4% of |x| > 2^18*pi lose accuracy; g_ph_fix selects the
    // corrected shift so the tests can pin the rest of the restatement.
    ih = iq[jz - 1] >> ((B == 8) ? (g_ph_fix ? 7 : 8) : 23);
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
    if (q0 > 0) {
      switch (q0) {
        case 1: iq[jz - 1] &= (maskall >> 1); break;
        case 2: iq[jz - 1] &= (maskall >> 2); break;
      }
    }
    if (ih == 2) {
      z = (T)1 - z;
      if (carry != 0) z -= std::ldexp((T)1, q0);
    }
  }
  if (z == (T)0) {
    j = 0;
    for (i = jz - 1; i >= jk; i--) j |= iq[i];
    if (j == 0) {
      for (k = 1; iq[jk - k] == 0; k++) {
      }
      for (i = jz + 1; i <= jz + k; i++) {
        f[jx + i] = (T)ipio2[jv + i];
        for (j = 0, fw = 0; j <= jx; j++) fw += x[j] * f[jx + i - j];
        q[i] = fw;
      }
      jz += k;
      goto recompute;
    }
  }
  if (z == (T)0) {
    jz -= 1;
    q0 -= B;
    while (iq[jz] == 0) { jz--; q0 -= B; }
  } else {
    z = std::ldexp(z, -q0);
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
  fw = std::ldexp((T)1, q0);
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

float PIo2_f[11];
double PIo2_d[8];
struct TableInit {
  TableInit() {
    for (int i = 0; i < 11; ++i) PIo2_f[i] = as_f(oracle_tables::pio2_8_bits[i]);
    for (int i = 0; i < 8; ++i) PIo2_d[i] = as_d(oracle_tables::pio2_24_bits[i]);
  }
} table_init_;

// __ieee754_rem_pio2f: scalar/rem_pio2.hpp:114-229; NT2 wrapper :37-44 returns n&3.
int rem_pio2f_fdlibm(float x, float& xr) {
  const float two8 = 2.5600000000e+02f, invpio2 = F32(0x3f22f984),
              pio2_1 = F32(0x3fc90f80), pio2_1t = F32(0x37354443),
              pio2_2 = F32(0x37354400), pio2_2t = F32(0x2e85a308),
              pio2_3 = F32(0x2e85a300), pio2_3t = F32(0x248d3132);
  float y0, z, w, t, r, fn, tx[3];
  int32_t e0, i, j, nx, n, ix, hx;
  hx = (int32_t)bits(x);
  ix = hx & 0x7fffffff;
  if (ix <= 0x3f490fd8) { xr = x; return 0; }
  if (ix < 0x4016cbe4) {
    if (hx > 0) {
      z = x - pio2_1;
      if ((ix & 0xfffffff0) != 0x3fc90fd0) { y0 = z - pio2_1t; }
      else { z -= pio2_2; y0 = z - pio2_2t; }
      xr = y0; return 1;
    } else {
      z = x + pio2_1;
      if ((ix & 0xfffffff0) != 0x3fc90fd0) { y0 = z + pio2_1t; }
      else { z += pio2_2; y0 = z + pio2_2t; }
      xr = y0; return (-1) & 3;
    }
  }
  if (ix <= 0x43490f80) {
    t = std::fabs(x);
    n = (int32_t)(t * invpio2 + 0.5f);
    fn = (float)n;
    r = t - fn * pio2_1;
    w = fn * pio2_1t;
    if (n < 32 && (int32_t)(ix & 0xffffff00) != oracle_tables::npio2_hw_f[n - 1]) {
      y0 = r - w;
    } else {
      uint32_t high;
      j = ix >> 23;
      y0 = r - w;
      high = bits(y0);
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
    if (hx < 0) { xr = -y0; return (-n) & 3; }
    xr = y0; return n & 3;
  }
  if (ix >= 0x7f800000) { xr = x - x; return 0; }
  e0 = (ix >> 23) - 134;
  z = as_f((uint32_t)(ix - (int32_t)((uint32_t)e0 << 23)));
  for (i = 0; i < 2; i++) { tx[i] = (float)((int32_t)(z)); z = (z - tx[i]) * two8; }
  tx[2] = z;
  nx = 3;
  while (tx[nx - 1] == 0.0f) nx--;
  float y[2];
  n = kernel_rem_pio2<float, 8>(tx, y, e0, nx, oracle_tables::two_over_pi_8, PIo2_f);
  if (hx < 0) { xr = -y[0]; return (-n) & 3; }
  xr = y[0];
  return n & 3;
}

// __ieee754_rem_pio2 (double): scalar/rem_pio2.hpp:521-621; NT2 wrapper :414-427.
int rem_pio2d_fdlibm(double x, double& xr) {
  if (x == (double)INFINITY) { xr = NAN; return 0; }
  const double two24 = 1.67772160000000000000e+07,
               invpio2 = F64(0x3FE45F306DC9C883), pio2_1 = F64(0x3FF921FB54400000),
               pio2_1t = F64(0x3DD0B4611A626331), pio2_2 = F64(0x3DD0B4611A600000),
               pio2_2t = F64(0x3BA3198A2E037073), pio2_3 = F64(0x3BA3198A2E000000),
               pio2_3t = F64(0x397B839A252049C1);
  double z = 0.0, w, t, r, fn, y0, tx[3];
  int32_t e0, i, j, nx, n, ix, hx;
  uint64_t xb = bits(x);
  hx = (int32_t)(xb >> 32);
  ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) { xr = x; return 0; }
  if (ix < 0x4002d97c) {
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) { y0 = z - pio2_1t; }
      else { z -= pio2_2; y0 = z - pio2_2t; }
      xr = y0; return 1;
    } else {
      z = x + pio2_1;
      if (ix != 0x3ff921fb) { y0 = z + pio2_1t; }
      else { z += pio2_2; y0 = z + pio2_2t; }
      xr = y0; return (-1) & 3;
    }
  }
  if (ix <= 0x413921fb) {
    t = std::fabs(x);
    n = (int32_t)(t * invpio2 + 0.5);
    fn = (double)n;
    r = t - fn * pio2_1;
    w = fn * pio2_1t;
    if ((n < 32) && (n > 0) && (ix != oracle_tables::npio2_hw_d[n - 1])) {
      y0 = r - w;
    } else {
      uint32_t high;
      j = ix >> 20;
      y0 = r - w;
      high = (uint32_t)(bits(y0) >> 32);
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
    if (hx < 0) { xr = -y0; return (-n) & 3; }
    xr = y0; return n & 3;
  }
  if (ix >= 0x7ff00000) { xr = x - x; return 0; }
  uint32_t low = (uint32_t)xb;
  e0 = (ix >> 20) - 1046;
  z = as_d(((uint64_t)(uint32_t)(ix - (int32_t)((uint32_t)e0 << 20)) << 32) | low);
  for (i = 0; i < 2; i++) { tx[i] = (double)((int32_t)(z)); z = (z - tx[i]) * two24; }
  tx[2] = z;
  nx = 3;
  while (tx[nx - 1] == 0.0) nx--;
  double y[2];
  n = kernel_rem_pio2<double, 24>(tx, y, e0, nx, oracle_tables::two_over_pi_24, PIo2_d);
  if (hx < 0) { xr = -y[0]; return (-n) & 3; }
  xr = y[0];
  return n & 3;
}

// is_ngt(a, b) = !(a > b): NaN counts as reduced
inline bool ngt(float a, float b) { return !(a > b); }
inline bool ngt(double a, double b) { return !(a > b); }

// f64 range-reduction cascade on a group of G lanes (G = 4 for a SIMD pack,
// G = 1 for the scalar tail): scalar/impl/trigo/trig_reduction.hpp:94-215.
// For double, conversion to a wider type is not allowed, so lanes beyond
// Medium_pi (2^18 pi) go to fdlibm per lane (:212-215, :217-220).
template <int G>
void reduce_d(const double* x, double* xr, int64_t* n) {
  bool all4 = true, all2 = true, all20 = true, allm = true;
  for (int i = 0; i < G; ++i) {
    all4 = all4 && ngt(x[i], dPio_4);
    all2 = all2 && ngt(x[i], dPio_2);
    all20 = all20 && ngt(x[i], d20pi);
    allm = allm && ngt(x[i], dMedium_pi);
  }
  if (all4) {
    for (int i = 0; i < G; ++i) { xr[i] = x[i]; n[i] = 0; }
  } else if (all2) {  // rem_pio2_straight.hpp:58-63
    for (int i = 0; i < G; ++i) {
      double r = x[i] - dPio2_1;
      r -= dPio2_2;
      r -= dPio2_3;
      xr[i] = r; n[i] = 1;
    }
  } else if (all20) {  // rem_pio2_cephes.hpp:65-72
    for (int i = 0; i < G; ++i) {
      double xi = r2e(x[i] * dTwoopi);
      double r = x[i] - xi * dPio2_1;
      r -= xi * dPio2_2;
      r -= xi * dPio2_3;
      xr[i] = r; n[i] = ((int64_t)xi) & 3;
    }
  } else if (allm) {  // rem_pio2_medium.hpp:68-83
    for (int i = 0; i < G; ++i) {
      const double t = x[i];
      const double fn = r2e(t * dInvpio_2);
      double r = t - fn * dPio2_1;
      double w = fn * dPio2_1t;
      double t2 = r;
      w = fn * dPio2_2;
      r = t2 - w;
      w = fn * dPio2_2t - ((t2 - r) - w);
      t2 = r;
      w = fn * dPio2_3;
      r = t2 - w;
      w = fn * dPio2_3t - ((t2 - r) - w);
      xr[i] = r - w;
      n[i] = ((int64_t)fn) & 3;
    }
  } else {
    for (int i = 0; i < G; ++i) {
      double r;
      n[i] = rem_pio2d_fdlibm(x[i], r);
      xr[i] = r;
    }
  }
}

// f32 cascade on a group of G lanes (G = 8 SIMD, 1 scalar tail):
// trig_reduction.hpp:94-215 with use_conversion (:190-215): when every lane is
// within 2^18 pi, the pack is split into halves of 4 doubles and each half runs
// the f64 cascade; beyond that, fdlibm single per lane.
template <int G>
void reduce_f(const float* x, float* xr, int32_t* n) {
  bool all4 = true, all2 = true, all20 = true, allm = true, alld = true;
  for (int i = 0; i < G; ++i) {
    all4 = all4 && ngt(x[i], fPio_4);
    all2 = all2 && ngt(x[i], fPio_2);
    all20 = all20 && ngt(x[i], f20pi);
    allm = allm && ngt(x[i], fMedium_pi);
    alld = alld && ngt(x[i], fDmpi);
  }
  if (all4) {
    for (int i = 0; i < G; ++i) { xr[i] = x[i]; n[i] = 0; }
  } else if (all2) {
    for (int i = 0; i < G; ++i) {
      float r = x[i] - fPio2_1;
      r -= fPio2_2;
      r -= fPio2_3;
      xr[i] = r; n[i] = 1;
    }
  } else if (all20) {
    for (int i = 0; i < G; ++i) {
      float xi = r2e(x[i] * fTwoopi);
      float r = x[i] - xi * fPio2_1;
      r -= xi * fPio2_2;
      r -= xi * fPio2_3;
      xr[i] = r; n[i] = ((int32_t)xi) & 3;
    }
  } else if (allm) {
    for (int i = 0; i < G; ++i) {
      const float t = x[i];
      const float fn = r2e(t * fInvpio_2);
      float r = t - fn * fPio2_1;
      float w = fn * fPio2_1t;
      float t2 = r;
      w = fn * fPio2_2;
      r = t2 - w;
      w = fn * fPio2_2t - ((t2 - r) - w);
      t2 = r;
      w = fn * fPio2_3;
      r = t2 - w;
      w = fn * fPio2_3t - ((t2 - r) - w);
      xr[i] = r - w;
      n[i] = ((int32_t)fn) & 3;
    }
  } else if (alld) {
    constexpr int H = (G >= 4) ? 4 : G;
    for (int h = 0; h < G; h += H) {
      double ux[H], uxr[H];
      int64_t un[H];
      for (int i = 0; i < H; ++i) ux[i] = (double)x[h + i];
      reduce_d<H>(ux, uxr, un);
      for (int i = 0; i < H; ++i) { xr[h + i] = (float)uxr[i]; n[h + i] = (int32_t)un[i]; }
    }
  } else {
    for (int i = 0; i < G; ++i) {
      float r;
      n[i] = rem_pio2f_fdlibm(x[i], r);
      xr[i] = r;
    }
  }
}

// evaluation polynomials: f_trig_evaluation.hpp:35-46, d_trig_evaluation.hpp:33-50
inline float cos_eval_f(float z) {
  float y = horner(fCosC, z);
  return 1.0f + (z * -0.5f + y * (z * z));
}
inline float sin_eval_f(float z, float x) {
  float y1 = horner(fSinC, z);
  return (y1 * z) * x + x;
}
inline double cos_eval_d(double z) { return 1.0 - horner(dCosC, z) * z; }
inline double sin_eval_d(double z, double x) {
  double y1 = horner(dSinC, z);
  return (y1 * z) * x + x;
}

// kind: 0 = cos, 1 = sin, 2 = sincos (s -> out0, c -> out1)
// SIMD trig_base: simd/common/impl/trigo/trig_base.hpp:59-71 (cos), :80-91 (sin),
// :140-156 (sincos); scalar trig_base: scalar/impl/trigo/trig_base.hpp:126-140,
// :142-156, :183-207 (invalid inputs return NaN first).
template <int G>
void trig_group_f(int kind, const float* a0, float* out0, float* out1) {
  float x[G], xr[G];
  int32_t n[G];
  for (int i = 0; i < G; ++i) x[i] = std::fabs(a0[i]);
  reduce_f<G>(x, xr, n);
  for (int i = 0; i < G; ++i) {
    const uint32_t swap = (uint32_t)n[i] & 1u;
    const float z = xr[i] * xr[i];
    const float se = sin_eval_f(z, xr[i]);
    const float ce = cos_eval_f(z);
    const uint32_t cos_sign = (swap ^ (((uint32_t)n[i] & 2u) >> 1)) << 31;
    const uint32_t sin_sign = (((uint32_t)n[i] & 2u) << 30) ^ (bits(a0[i]) & 0x80000000u);
    float c = as_f(bits(swap ? se : ce) ^ cos_sign);
    float s = as_f(bits(swap ? ce : se) ^ sin_sign);
    if (kind == 3) {  // tan: f_trig_evaluation.hpp:47-64, trig_base.hpp:115-124 (scalar :165-176)
      const float zz = xr[i] * xr[i];
      float y = ((horner(fTanC, zz) * zz) * xr[i]) + xr[i];
      if (n[i] & 1) y = -(1.0f / y);
      float t = as_f(bits(y) ^ (bits(a0[i]) & 0x80000000u));
      if (G == 1 && !std::isfinite(a0[i])) t = NAN;
      else if (G == 1 && a0[i] == 0.0f) t = a0[i];
      out0[i] = t;
      continue;
    }
    if (kind == 4) {  // cot: cot_eval (1/y even, -y odd), trig_base.hpp:126-138 (scalar :178-188)
      const float zz = xr[i] * xr[i];
      float y = ((horner(fTanC, zz) * zz) * xr[i]) + xr[i];
      y = (n[i] & 1) ? -y : 1.0f / y;
      float t = as_f(bits(y) ^ (bits(a0[i]) & 0x80000000u));
      if (G == 1 && !std::isfinite(a0[i])) t = NAN;
      else if (a0[i] == 0.0f) t = 1.0f / a0[i];
      out0[i] = t;
      continue;
    }
    if (G == 1 && !std::isfinite(a0[i])) { c = NAN; s = NAN; }
    if (kind == 0) out0[i] = c;
    else if (kind == 1) out0[i] = s;
    else if (kind == 5) out0[i] = 1.0f / c;  // sec: generic/sec.hpp, rec(cos(a0))
    else if (kind == 6) out0[i] = 1.0f / s;  // csc: generic/csc.hpp, rec(sin(a0))
    else { out0[i] = s; out1[i] = c; }
  }
}

template <int G>
void trig_group_d(int kind, const double* a0, double* out0, double* out1) {
  double x[G], xr[G];
  int64_t n[G];
  for (int i = 0; i < G; ++i) x[i] = std::fabs(a0[i]);
  reduce_d<G>(x, xr, n);
  for (int i = 0; i < G; ++i) {
    const uint64_t swap = (uint64_t)n[i] & 1u;
    const double z = xr[i] * xr[i];
    const double se = sin_eval_d(z, xr[i]);
    const double ce = cos_eval_d(z);
    const uint64_t cos_sign = (swap ^ (((uint64_t)n[i] & 2u) >> 1)) << 63;
    const uint64_t sin_sign = (((uint64_t)n[i] & 2u) << 62) ^ (bits(a0[i]) & 0x8000000000000000ull);
    double c = as_d(bits(swap ? se : ce) ^ cos_sign);
    double s = as_d(bits(swap ? ce : se) ^ sin_sign);
    if (kind == 3) {  // tan: d_trig_evaluation.hpp:58-78
      const double zz = xr[i] * xr[i];
      double y = xr[i] + xr[i] * (zz * (horner(dTanN, zz) / horner(dTanD, zz)));
      if (n[i] & 1) y = -(1.0 / y);
      double t = as_d(bits(y) ^ (bits(a0[i]) & 0x8000000000000000ull));
      if (G == 1 && !std::isfinite(a0[i])) t = NAN;
      else if (G == 1 && a0[i] == 0.0) t = a0[i];
      out0[i] = t;
      continue;
    }
    if (kind == 4) {  // cot
      const double zz = xr[i] * xr[i];
      double y = xr[i] + xr[i] * (zz * (horner(dTanN, zz) / horner(dTanD, zz)));
      y = (n[i] & 1) ? -y : 1.0 / y;
      double t = as_d(bits(y) ^ (bits(a0[i]) & 0x8000000000000000ull));
      if (G == 1 && !std::isfinite(a0[i])) t = NAN;
      else if (a0[i] == 0.0) t = 1.0 / a0[i];
      out0[i] = t;
      continue;
    }
    if (G == 1 && !std::isfinite(a0[i])) { c = NAN; s = NAN; }
    if (kind == 0) out0[i] = c;
    else if (kind == 1) out0[i] = s;
    else if (kind == 5) out0[i] = 1.0 / c;
    else if (kind == 6) out0[i] = 1.0 / s;
    else { out0[i] = s; out1[i] = c; }
  }
}

// sinpi ... cotd: trig_base<A0, pi_tag | degree_tag, ..., big_>
// (simd/common/impl/trigo/trig_base.hpp:53-138, scalar/impl/trigo/trig_base.hpp:
// 126-188) with the lane-wise reductions of scalar/impl/trigo/trig_reduction.hpp:
// 262-318: pi: xi = round2even(x*2), xr = (x - xi*Half)*Pi; degree:
// xi = round2even(x*Oneo_90), xr = inrad(x - xi*_90) = (...)*Deginrad
// (generic/inrad.hpp); n = fast_toint(xi) with x86's truncating conversion
// (the integer indefinite value out of range).  tan_invalid / cot_invalid:
// trig_reduction.hpp:268-269 (degree, times Oneo_180), :308-309 (pi).
// kind: 0 cos, 1 sin, 3 tan, 4 cot.
inline bool is_flint_v(float v) { return v - std::trunc(v) == 0.0f; }
inline bool is_flint_v(double v) { return v - std::trunc(v) == 0.0; }
template <class T> struct unit_c;
template <> struct unit_c<float> {
  using I = int32_t;
  using U = uint32_t;
  static constexpr U sign = 0x80000000u;
  static float pi() { return F32(0x40490fdb); }
  static float oneo90() { return F32(0x3c360b61); }
  static float oneo180() { return F32(0x3bb60b61); }
  static float deginrad() { return F32(0x3c8efa35); }
  static I toint(float v) { return (v < 2147483648.0f) ? (I)v : (I)INT32_MIN; }
  static float tan_base(float z) {
    const float zz = z * z;
    return ((horner(fTanC, zz) * zz) * z) + z;
  }
  static float sin_eval(float z, float x) { return sin_eval_f(z, x); }
  static float cos_eval(float z) { return cos_eval_f(z); }
};
template <> struct unit_c<double> {
  using I = int64_t;
  using U = uint64_t;
  static constexpr U sign = 0x8000000000000000ull;
  static double pi() { return F64(0x400921fb54442d18); }
  static double oneo90() { return F64(0x3f86c16c16c16c17); }
  static double oneo180() { return F64(0x3f76c16c16c16c17); }
  static double deginrad() { return F64(0x3f91df46a2529d39); }
  static I toint(double v) { return (v < 9223372036854775808.0) ? (I)v : (I)INT64_MIN; }
  static double tan_base(double z) {
    const double zz = z * z;
    return z + z * (zz * (horner(dTanN, zz) / horner(dTanD, zz)));
  }
  static double sin_eval(double z, double x) { return sin_eval_d(z, x); }
  static double cos_eval(double z) { return cos_eval_d(z); }
};
template <class T, bool PI> T trig_unit(int kind, T a0) {
  using C = unit_c<T>;
  using U = typename C::U;
  const T x = std::fabs(a0);
  T xi, xr;
  if (PI) {
    xi = std::nearbyint(x * T(2));
    xr = (x - xi * T(0.5)) * C::pi();
  } else {
    xi = std::nearbyint(x * C::oneo90());
    xr = (x - xi * T(90)) * C::deginrad();
  }
  const typename C::I n = C::toint(xi);
  const U sa = __builtin_bit_cast(U, a0) & C::sign;
  auto flip = [](T v, U m) { return __builtin_bit_cast(T, (U)(__builtin_bit_cast(U, v) ^ m)); };
  if (kind == 3 || kind == 4) {
    const T y = C::tan_base(xr);
    T r = (kind == 3) ? ((n & 1) ? -(T(1) / y) : y) : ((n & 1) ? -y : T(1) / y);
    r = flip(r, sa);
    bool bad;
    if (kind == 3) bad = is_flint_v(PI ? a0 - T(0.5) : (a0 - T(90)) * C::oneo180());
    else bad = a0 != T(0) && is_flint_v(PI ? a0 : a0 * C::oneo180());
    if (bad) r = NAN;
    if (kind == 4 && a0 == T(0)) r = T(1) / a0;
    return r;
  }
  const U swap = (U)n & 1u;
  const T z = xr * xr;
  const T se = C::sin_eval(z, xr), ce = C::cos_eval(z);
  const int sh = sizeof(T) * 8 - 1;
  if (kind == 0) return flip(swap ? se : ce, (swap ^ (((U)n & 2u) >> 1)) << sh);
  return flip(swap ? ce : se, (((U)n & 2u) << (sh - 1)) ^ sa);
}

// ===========================================================================
// Compositions over the kernels above, each NT2's own definition (its
// operation order fixes the bits); `scalar` selects NT2's scalar form (the
// tail of a table) where the inner function or the definition has one.
// coth: modules/core/hyperbolic/include/nt2/hyperbolic/functions/simd/common/
// coth.hpp:35-63 (scalar/coth.hpp:29-55), tanh_kernel's coth
// (details/tanh_kernel.hpp:40-43, :55-59); sech, csch, acoth, asech, acsch:
// hyperbolic/functions/generic/{sech.hpp:19-29,csch.hpp:18-30,acoth.hpp:18-28,
// asech.hpp:18-28,acsch.hpp:18-28}; acot, acsc, atand, asind, acosd, asecd,
// acscd, acotd, atanpi, asinpi, acospi, atan2d, indeg, inrad, secd, cscd,
// secpi, cscpi: modules/core/trigonometric/include/nt2/trigonometric/functions/
// generic/*.hpp; asec: simd/common/asec.hpp:18-30 and scalar/asec.hpp:26-67;
// sinc, sincpi: simd/common/{sinc.hpp:33-58,sincpi.hpp:35-60} and
// scalar/{sinc.hpp:28-49,sincpi.hpp:29-49}.
// ===========================================================================
namespace ov {
inline float atan(float x, bool s) { return atan_f(x, s); }
inline double atan(double x, bool s) { return atan_d(x, s); }
inline float asin(float x, bool) { return asin_f(x); }
inline double asin(double x, bool s) { return asin_d(x, s); }
inline float acos(float x, bool) { return acos_f(x); }
inline double acos(double x, bool s) { return acos_d(x, s); }
inline float asinh(float x, bool s) { return asinh_f(x, s); }
inline double asinh(double x, bool s) { return asinh_d(x, s); }
inline float acosh(float x, bool s) { return acosh_f(x, s); }
inline double acosh(double x, bool s) { return acosh_d(x, s); }
inline float atanh(float x) { return atanh_f(x); }
inline double atanh(double x) { return atanh_d(x); }
inline float sinh(float x) { return sinh_f(x); }
inline double sinh(double x) { return sinh_d(x); }
inline float cosh(float x) { return cosh_f(x); }
inline double cosh(double x) { return cosh_d(x); }
inline float exp(float x) { return exp_f(x); }
inline double exp(double x) { return exp_d(x); }
inline float expm1(float x) { return expm1_f(x); }
inline double expm1(double x) { return expm1_d(x); }
inline float log(float x) { return log_f(x); }
inline double log(double x) { return log_d(x); }
inline float log1p(float x) { return log1p_f(x); }
inline double log1p(double x) { return log1p_d(x); }
inline float atan2(float a, float b, bool s) { return atan2_f(a, b, s); }
inline double atan2(double a, double b, bool s) { return atan2_d(a, b, s); }
template <class T> T cst(uint32_t f, uint64_t d) {
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(float, f);
  else return __builtin_bit_cast(double, d);
}
template <class T> T pio2() { return cst<T>(0x3fc90fdbu, 0x3ff921fb54442d18ull); }
template <class T> T invpi() { return cst<T>(0x3ea2f983u, 0x3fd45f306dc9c883ull); }
template <class T> T eps() { return std::numeric_limits<T>::epsilon(); }
template <class T> T signbit_of(T a) {
  using U = typename unit_c<T>::U;
  return __builtin_bit_cast(T, (U)(__builtin_bit_cast(U, a) & unit_c<T>::sign));
}
template <class T> T bxor(T a, T b) {
  using U = typename unit_c<T>::U;
  return __builtin_bit_cast(T, (U)(__builtin_bit_cast(U, a) ^ __builtin_bit_cast(U, b)));
}
template <class T> T bor(T a, T b) {
  using U = typename unit_c<T>::U;
  return __builtin_bit_cast(T, (U)(__builtin_bit_cast(U, a) | __builtin_bit_cast(U, b)));
}
}  // namespace ov

// indeg: a0*Radindeg - a0*Radindegr (constants/radindeg.hpp:29-30,
// radindegr.hpp:28-29); inrad: a0*Deginrad (constants/deginrad.hpp:29-30)
template <class T> T indeg_o(T a0) {
  return a0 * ov::cst<T>(0x42652ee1u, 0x404ca5dc1a63c1f8ull) - a0 * ov::cst<T>(0xb53387c1u, 0xbce1e7ab456405f8ull);
}
template <class T> T inrad_o(T a0) { return a0 * ov::cst<T>(0x3c8efa35u, 0x3f91df46a2529d39ull); }
inline float coth_o(float a0) {
  const float cc[5] = {F32(0xbbbaf0ea), F32(0x3ca9134e), F32(0xbd5c1e2d), F32(0x3e088393), F32(0xbeaaaa99)};
  const float x = std::fabs(a0);
  float r;
  if (x < 0.625f) {
    const float x2 = x * x;
    r = 1.0f / ((horner(cc, x2) * x2) * x + x);
  } else {
    r = 2.0f * (1.0f / (exp_f(x + x) - 1.0f)) + 1.0f;
  }
  return ov::bxor(r, ov::signbit_of(a0));
}
inline double coth_o(double a0) {
  const double p[3] = {F64(0xbfeedc5baafd6f4b), F64(0xc058d26a0e26682d), F64(0xc0993ac030580563)};
  const double q[4] = {F64(0x3FF0000000000000), F64(0x405c33f28a581b86), F64(0x40a176fa0e5535fa),
                       F64(0x40b2ec102442040c)};
  const double x = std::fabs(a0);
  double r;
  if (x < 0.625) {
    const double x2 = x * x;
    const double qv = horner(q, x2);
    r = qv / (x * (horner(p, x2) * x2 + qv));
  } else {
    r = 2.0 * (1.0 / (exp_d(x + x) - 1.0)) + 1.0;
  }
  return ov::bxor(r, ov::signbit_of(a0));
}
template <class T> T asec_o(T a0, bool scalar) {
  if (!scalar) return ov::pio2<T>() - ov::asin(T(1) / a0, false);
  if constexpr (sizeof(T) == 8) {
    if (a0 == 1.0) return 0.0;
    return (ov::pio2<T>() - ov::asin(T(1) / a0, true)) + F64(0x3c91a62633145c07);
  } else {
    const float ax = std::fabs(a0);
    if (ax < 1.0f) return NAN;
    const float ax1 = ax - 1.0f;
    if (ax1 < 0.001f) {
      const float tmp = std::sqrt(2 * ax1) *
                        (1 - ax1 * ((5 / 12.0f) + ax1 * ((43 / 160.0f) - ax1 * ((177 / 896.0f) + ax1 * (2867 / 18432.0f)))));
      return a0 < 0.0f ? F32(0x40490fdb) - tmp : tmp;
    }
    if (a0 == 1.0f) return 0.0f;
    return ov::pio2<T>() - ov::asin(T(1) / a0, true);
  }
}
template <class T> T cscpi_o(T a0) {
  const T r = T(1) / trig_unit<T, true>(1, a0);
  return (is_flint_v(a0) && a0 != T(0)) ? T(NAN) : r;
}
template <class T> T erfinv_o(T a0);
template <class T> T sinhc_o(T a0);
template <class T> T erfcinv_o(T a0);
// one lane of a composed op; kinds as in the C ABI (44..69, 68 = sinc is by pack)
template <class T> T compose_o(int op, T x, bool sc) {
  switch (op) {
    case 44: return coth_o(x);
    case 45: return T(1) / ov::cosh(x);
    case 46: return T(1) / ov::sinh(x);
    case 47: return ov::atanh(T(1) / x);
    case 48: return ov::acosh(T(1) / x, sc);
    case 49: return ov::asinh(T(1) / x, sc);
    case 50: return ov::atan(T(1) / x, sc);
    case 51: return asec_o(x, sc);
    case 52: return ov::asin(T(1) / x, sc);
    case 53: {
      const T d = indeg_o(ov::atan(std::fabs(x), sc));
      T z = T(90) - (x != T(0) ? d : T(0));
      if (std::isinf(x)) z = T(0);
      return ov::bor(z, ov::signbit_of(x));
    }
    case 54: return indeg_o(asec_o(x, sc));
    case 55: return indeg_o(ov::asin(T(1) / x, sc));
    case 56: return indeg_o(ov::acos(x, sc));
    case 57: return ov::invpi<T>() * ov::acos(x, sc);
    case 58: return indeg_o(ov::asin(x, sc));
    case 59: return ov::invpi<T>() * ov::asin(x, sc);
    case 60: return indeg_o(ov::atan(x, sc));
    case 61: return ov::invpi<T>() * ov::atan(x, sc);
    case 62: return indeg_o(x);
    case 63: return inrad_o(x);
    case 64: {
      const T r = trig_unit<T, false>(0, x);
      return T(1) / (r == T(0) ? T(NAN) : r);
    }
    case 65: return cscpi_o(ov::cst<T>(0x3bb60b61u, 0x3f76c16c16c16c17ull) * x);
    case 66: {
      const T r = T(1) / trig_unit<T, true>(0, x);
      return is_flint_v((T(2) * x - T(1)) * T(0.5)) ? T(NAN) : r;
    }
    case 67: return cscpi_o(x);
    // the exponential family: modules/core/exponential/include/nt2/exponential/
    // functions/{generic/sqrt1pm1.hpp:23-36,generic/exprecneg.hpp:18-28,
    // generic/exprecnegc.hpp:18-28,generic/reallog.hpp:20-32,
    // generic/realsqrt.hpp:20-32,simd/common/expx2.hpp:37-61,scalar/expx2.hpp:32-54}
    case 70: {
      const T t = std::sqrt(T(1) + x);
      return std::fabs(x) < T(0.5) ? x / (T(1) + t) : t - T(1);
    }
    case 71: return ov::exp(-(T(1) / x));
    case 72: return -ov::expm1(-(T(1) / x));
    case 73: return ov::log(x);
    case 74: return std::sqrt(x);
    case 75: {  // fma(c2, x, 1/2) and fma(2m, f, f^2) with two roundings: the scalar form's bits
      const T c1 = ov::cst<T>(0x42000000u, 0x4060000000000000ull), c2 = ov::cst<T>(0x3d000000u, 0x3f80000000000000ull);
      const T maxlog = ov::cst<T>(0x42b0c0a5u, 0x40862e42fefa39efull);
      const T a = std::fabs(x);
      const T m = c1 * std::floor(c2 * a + T(0.5));
      const T f = a - m;
      const T u = m * m, u1 = (m + m) * f + f * f;
      const T r = (u + u1 > maxlog) ? T(INFINITY) : ov::exp(u) * ov::exp(u1);
      return std::isinf(a) ? a : r;
    }
    case 76: return erfinv_o(x);
    case 77: return erfcinv_o(x);
    case 78: return sinhc_o(x);
    case 69: {
      const T sp = trig_unit<T, true>(1, x);
      if (std::isinf(x)) return T(0);
      if (std::fabs(x) < ov::eps<T>()) return T(1);
      return sc ? (ov::invpi<T>() * sp) / x : ov::invpi<T>() * (sp / x);
    }
  }
  return T(NAN);
}
// ---- inverse error functions -----------------------------------------------------------
// (modules/core/euler/include/nt2/euler/functions/): the polynomials of
// details/erfinv_kernel.hpp (float: erfinv1 :31-41, erfinv2 :44-54, the
// erfcinv2 rational :57-71, Giles' single-precision approximation; double:
// erfinv1 :80-104, erfinv2 :108-128, erfinv3 :132-150), coefficients by bit
// pattern, highest degree first.
inline float erfinv1_f(float w) {
  const float c[9] = {F32(0x32f16588), F32(0x34b84b36), F32(0xb66c7357), F32(0xb6935ac1),
      F32(0x396532db), F32(0xbaa45408), F32(0xbb88e4ef), F32(0x3e7c8f63), F32(0x3fc02e2f)};
  return horner(c, w);
}
inline float erfinv2_f(float w) {
  const float c[9] = {F32(0xb951f09b), F32(0x38d3b56b), F32(0x3ab0dc72), F32(0xbb70bde7),
      F32(0x3bbc127b), F32(0xbbf9c5d7), F32(0x3c1aa57e), F32(0x3f8036db), F32(0x40354f7e)};
  return horner(c, w);
}
inline float erfcinv2_f(float w) {
  const float p[5] = {F32(0x3b3a8aee), F32(0x3ba62c57), F32(0xbc8bcf63), F32(0x3c98e17c),
      F32(0xbbaf8bfe)};
  const float q[5] = {F32(0x321631b2), F32(0x3b3a73c1), F32(0x3ba9025f), F32(0xbc5d80aa),
      F32(0x3c2f4022)};
  return horner(p, w) / horner(q, w);
}
inline double erfinv1_d(double w) {
  const double c[23] = {F64(0xbbb135d2e746e627), F64(0xbc08ddf93324d327),
      F64(0x3c37b83eef0b7c9f), F64(0x3c69ba72cd589b91), F64(0xbca33689090a6b96),
      F64(0x3c782e11898132e0), F64(0x3cfde4acfd9e26ba), F64(0xbd26d33eed66c487),
      F64(0xbd36f2167040d8e2), F64(0x3d872a22c2d77e20), F64(0xbdac8859c4e5c0af),
      F64(0xbdcdc583d118a561), F64(0x3e120f47ccf46b3c), F64(0xbe31a9e38dc84d60),
      F64(0xbe5f36cd6d3d46a9), F64(0x3e9c6b4f5d03b787), F64(0xbeb6e8a5434ae8a2),
      F64(0xbeed1d1f7b8736f6), F64(0x3f2879c2a212f024), F64(0xbf4845769484fca8),
      F64(0xbf78b6c33114f909), F64(0x3fcebd80d9b13e28), F64(0x3ffa755e7c99ae86)};
  return horner(c, w);
}
inline double erfinv2_d(double w) {
  const double c[19] = {F64(0x3e23040f87dbd932), F64(0x3e785cbe52878635),
      F64(0xbe92777453dd3955), F64(0x3e5395abcd554c6c), F64(0x3eb936388a3790ad),
      F64(0xbed0d5db812b5083), F64(0x3ec8860cd5d652f6), F64(0x3eea29a0cacdfb23),
      F64(0xbf08cef1f80281f2), F64(0x3f11e684d0b9188a), F64(0x3ef932cd54c8a222),
      F64(0xbf37448a89ef8aa3), F64(0x3f4f3cc55ad40c25), F64(0xbf5ba924132f38b1),
      F64(0x3f6468eeca533cf8), F64(0xbf6ebadabb891bbd), F64(0x3f75ffcfe5b76afc),
      F64(0x3ff0158a6d641d39), F64(0x4008abcc380d5a48)};
  return horner(c, w);
}
inline double erfinv3_d(double w) {
  const double c[17] = {F64(0xbdbdcec3a7785389), F64(0xbdf18feec0e38727),
      F64(0x3e19e6bf2dda45e3), F64(0xbe30468fb24e2f5f), F64(0x3e405ac6a8fba182),
      F64(0xbe50102e495fb9c0), F64(0x3e5f4c20e1334af8), F64(0xbe722d220fdf9c3e),
      F64(0x3e8ebc8bb824cb54), F64(0xbeb0a8d40ea372cc), F64(0x3ed2fbd29d093d2b),
      F64(0xbef4a3497e1e0fac), F64(0x3f13ebf4eb00938f), F64(0xbf2c2f36a8fc5d53),
      F64(0xbf222ea5df04047c), F64(0x3ff02a30d1fba0dc), F64(0x4013664ddd1ad7fb)};
  return horner(c, w);
}
template <class T> T signnz_o(T a) { return std::signbit(a) ? T(-1) : T(1); }
// erfinv: simd/common/erfinv.hpp:32-57 (float), :64-108 (double); the scalar
// forms (scalar/erfinv.hpp:22-52, :59-80) give the same value per lane.
// w = -log((1 - a0)(1 + a0)); float: a0 erfinv1(w - 2.5) below 5, else
// a0 erfinv2(sqrt(w) - 3); double: a0 erfinv1(w - 3.125) below 6.25,
// a0 erfinv2(sqrt(w) - 3.25) below 16, else a0 erfinv3(sqrt(w) - 5);
// signnz_o(a0) w where w is infinite (a0 = +-1)
template <class T> T erfinv_o(T a0) {
  const T w = -ov::log((T(1) - a0) * (T(1) + a0));
  if (std::isinf(w) && w > T(0)) return signnz_o(a0) * w;
  if constexpr (sizeof(T) == 4) {
    if (w < 5.0f) return a0 * erfinv1_f(w - 2.5f);
    return a0 * erfinv2_f(std::sqrt(w) - 3.0f);
  } else {
    if (w < 6.25) return a0 * erfinv1_d(w - 3.125);
    if (w < 16.0) return a0 * erfinv2_d(std::sqrt(w) - 3.25);
    return a0 * erfinv3_d(std::sqrt(w) - 5.0);
  }
}
// erfcinv: simd/common/erfcinv.hpp:32-57 (float), :64-107 (double); the scalar
// forms (scalar/erfcinv.hpp:22-51, :58-80) give the same value per lane.
// a00 = 1 - a0, w = -log(a0 (2 - a0)); as erfinv with a00 for a0, except the
// float tail copysign(erfcinv2(sqrt(w)), a00)
template <class T> T erfcinv_o(T a0) {
  const T a00 = T(1) - a0;
  const T w = -ov::log(a0 * (T(2) - a0));
  if (std::isinf(w) && w > T(0)) return signnz_o(a00) * w;
  if constexpr (sizeof(T) == 4) {
    if (w < 5.0f) return a00 * erfinv1_f(w - 2.5f);
    return std::copysign(erfcinv2_f(std::sqrt(w)), a00);
  } else {
    if (w < 6.25) return a00 * erfinv1_d(w - 3.125);
    if (w < 16.0) return a00 * erfinv2_d(std::sqrt(w) - 3.25);
    return a00 * erfinv3_d(std::sqrt(w) - 5.0);
  }
}

// sinhc: hyperbolic/functions/simd/common/sinhc.hpp:41-80, details/sinhc_kernel.hpp
template <class T> T sinhc_o(T a0) {
  const T x = std::fabs(a0);
  if (x < T(1)) {
    const T x2 = x * x;
    if constexpr (sizeof(T) == 4) {
      const float c[4] = {F32(0x39559e2f), F32(0x3c087bbe), F32(0x3e2aaacc), F32(0x3f800000)};
      return horner(c, x2);
    } else {
      const double p[4] = {F64(0xbfe9435fe8bb3cd6), F64(0xc064773a398ff4fe), F64(0xc0c694b8c71d6182), F64(0xc115782bdbf6ab05)};
      const double q[4] = {F64(0x3ff0000000000000), F64(0xc0715b6096e96484), F64(0x40e1a7ba7ed72245), F64(0xc1401a20e4f90044)};
      return 1.0 + (x2 * horner(p, x2)) / horner(q, x2);
    }
  }
  const T thr = ov::cst<T>(0x42b0c0a5u, 0x40862e42fefa39efull) - ov::cst<T>(0x3f317218u, 0x3fe62e42fefa39efull);
  const bool big = x > thr;
  const T e = ov::exp(x * (big ? T(0.5) : T(1)));
  const T r = big ? ((T(0.5) * e) / x) * e : (e * T(0.5) + (-(T(1) / e)) * T(0.5)) / x;
  return std::isinf(x) ? x : r;
}
// logspace_add: exponential/functions/scalar/logspace_add.hpp:25-42;
// logspace_sub: scalar/logspace_sub.hpp:20-32 (Log_2: constants/log_2.hpp:29-30)
template <class T> T logspace_add_o(T a0, T a1) {
  const T tmp = -std::fabs(a0 - a1);
  const T r = (a0 > a1 ? a0 : a1) + ov::log1p(ov::exp(tmp));
  return std::isnan(tmp) ? a0 + a1 : r;
}
template <class T> T logspace_sub_o(T a0, T a1) {
  const T x = a1 - a0;
  const T tmp = x > -ov::cst<T>(0x3f317218u, 0x3fe62e42fefa39efull) ? ov::log(-ov::expm1(x)) : ov::log1p(-ov::exp(x));
  return a0 + tmp;
}
// sinc over a pack of G lanes (sin with the pack's reduction)
template <class T, int G> void sinc_o(const T* x, T* y) {
  T s[G];
  if constexpr (sizeof(T) == 4) trig_group_f<G>(1, x, s, nullptr);
  else trig_group_d<G>(1, x, s, nullptr);
  for (int l = 0; l < G; ++l) {
    if (std::isinf(x[l])) y[l] = T(0);
    else if (std::fabs(x[l]) < ov::eps<T>()) y[l] = T(1);
    else y[l] = s[l] / x[l];
  }
}

// ===========================================================================
// NT2 transform schedule: modules/core/base/include/nt2/core/functions/simd/
// transform.hpp:40-57 (packs of N from the range start, scalar tail), run in
// parallel blocks of grain = L2/sizeof(T) when size > 8*grain
// (modules/arch/shared_memory/include/nt2/core/functions/shared_memory/
// transform.hpp:36-49, modules/arch/openmp/include/nt2/sdk/openmp/spawner/
// transform.hpp:44-72).  Grain is a multiple of the pack width, so packs start
// at multiples of N from index 0 and the last n%N elements use the scalar path.
// ===========================================================================
size_t l2_grain(size_t elem) {
  long l2 = sysconf(_SC_LEVEL2_CACHE_SIZE);
  if (l2 <= 0) l2 = 256 * 1024;
  size_t g = (size_t)l2 / elem;
  return std::max<size_t>(64, g & ~(size_t)63);
}

template <class Body>
void nt2_transform(size_t n, size_t N, size_t elem, Body body) {
  const size_t grain = l2_grain(elem);
  const size_t packs = n / N;
  if (n > 8 * grain) {
    const size_t ppb = grain / N;
    const size_t nblk = (packs + ppb - 1) / ppb;
#pragma omp parallel for schedule(static)
    for (size_t b = 0; b < nblk; ++b) {
      const size_t p1 = std::min(packs, (b + 1) * ppb);
      for (size_t p = b * ppb; p < p1; ++p) body(p * N, true);
    }
  } else {
    for (size_t p = 0; p < packs; ++p) body(p * N, true);
  }
  for (size_t i = packs * N; i < n; ++i) body(i, false);
}

int g_threads = 0;

// NT2 SIMD inner_fold of plus over a contiguous table (single thread):
// modules/core/reduction/include/nt2/core/functions/simd/inner_fold.hpp:44-67,
// details/fold_step.hpp:31-43, and the AVX butterfly
// modules/boost/simd/base/include/boost/simd/swar/functions/simd/sse/avx/all_reduce.hpp:52-66
// which sums 8 lanes as ((l1+l5)+(l3+l7)) + ((l2+l6)+(l0+l4)).
float hsum8(const float* a) {
  return ((a[1] + a[5]) + (a[3] + a[7])) + ((a[2] + a[6]) + (a[0] + a[4]));
}
double hsum4(const double* a) {  // all_reduce.hpp:29-36
  return (a[1] + a[3]) + (a[0] + a[2]);
}

template <class Elem>
float fold_sum_f32(size_t n, Elem elem) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const size_t packs = n / 8;
  for (size_t p = 0; p < packs; ++p) {
    float v[8];
    elem(p * 8, v, 8);
    for (int l = 0; l < 8; ++l) acc[l] = acc[l] + v[l];
  }
  float s = hsum8(acc);
  for (size_t i = packs * 8; i < n; ++i) {
    float v;
    elem(i, &v, 1);
    s = s + v;
  }
  return s;
}

// sum(a, dim) and cumsum(a, dim) in NT2's order.  The reduction dispatch
// (modules/core/reduction/include/nt2/dsl/functions/reduction/run.hpp:50-90):
// with lo_dim the first non-singleton dimension, red <= lo_dim folds the
// table reshaped to (size of red, product of the dimensions after red) with
// inner_fold (SIMD: pack accumulators from each column's start, the AVX
// horizontal sum, scalar tail; simd/inner_fold.hpp:44-67); otherwise it folds
// (product before red, size of red, product after red) with outer_fold, a
// sequential sum over the reduced dimension per output
// (details/outer_fold_step.hpp:33-60).  Scans are sequential per line
// (modules/core/swar/include/nt2/core/functions/scalar/inner_scan.hpp:36-55,
// outer_scan.hpp:34-55, details/scan_step.hpp:30-44).
struct Lines {
  size_t inner, len, outer;
  bool inner_fold;
};
Lines lines_of(const size_t* ext, int dim) {
  Lines L{1, 1, 1, false};
  size_t lo_dim = 1;
  for (int k = 0; k < 4; ++k)
    if (ext[k] != 1) { lo_dim = (size_t)k + 1; break; }
  for (int k = 0; k < 4; ++k) {
    if (k < dim - 1) L.inner *= ext[k];
    else if (k == dim - 1) L.len = ext[k];
    else L.outer *= ext[k];
  }
  L.inner_fold = (size_t)dim <= lo_dim;
  return L;
}
template <class T, int N>
void sum_dim(const T* x, const size_t* ext, int dim, T* out) {
  const Lines L = lines_of(ext, dim);
  if (L.inner_fold) {
    for (size_t o = 0; o < L.outer; ++o) {
      const T* c = x + o * L.len;
      T acc[N];
      for (int l = 0; l < N; ++l) acc[l] = T(0);
      const size_t nb = L.len / N;
      for (size_t p = 0; p < nb; ++p)
        for (int l = 0; l < N; ++l) acc[l] = acc[l] + c[p * N + l];
      T r;
      if constexpr (N == 8) r = hsum8(acc);
      else r = hsum4(acc);
      for (size_t i = nb * N; i < L.len; ++i) r = r + c[i];
      out[o] = r;
    }
  } else {
    // NT2 runs this loop nest with m outside and i inside (details/outer_fold_step.hpp:51-59):
    // each output is still summed in order over m
    for (size_t o = 0; o < L.outer; ++o) {
      T* acc = out + L.inner * o;
      for (size_t i = 0; i < L.inner; ++i) acc[i] = T(0);
      for (size_t m = 0; m < L.len; ++m) {
        const T* row = x + L.inner * (m + L.len * o);
        for (size_t i = 0; i < L.inner; ++i) acc[i] = acc[i] + row[i];
      }
    }
  }
}
template <class T>
void cumsum_dim(const T* x, const size_t* ext, int dim, T* out) {
  const Lines L = lines_of(ext, dim);
  for (size_t o = 0; o < L.outer; ++o) {
    const size_t b = L.inner * L.len * o;
    for (size_t m = 0; m < L.len; ++m)  // each line summed in order over m
      for (size_t i = 0; i < L.inner; ++i) {
        const size_t idx = b + i + L.inner * m;
        out[idx] = (m ? out[idx - L.inner] : T(0)) + x[idx];
      }
  }
}
// cumprod, cummin, cummax: the cumulative_ functors
// (modules/boost/simd/base/include/boost/simd/swar/functions/cumprod.hpp:27
// multiplies_ from One, cummin.hpp:26 min_ from Inf, cummax.hpp:27 max_ from
// Minf), each step summary = bop(summary, x) in order (details/scan_step.hpp:39);
// scalar min / max: boost/simd/arithmetic/functions/scalar/min.hpp:27, max.hpp:27.
template <class T>
void cumop_dim(const T* x, const size_t* ext, int dim, int op, T* out) {
  const Lines L = lines_of(ext, dim);
  const T neutral = op == 1 ? T(1) : (op == 2 ? T(INFINITY) : T(-INFINITY));
  for (size_t o = 0; o < L.outer; ++o) {
    const size_t b = L.inner * L.len * o;
    for (size_t i = 0; i < L.inner; ++i) {
      T s = neutral;
      for (size_t m = 0; m < L.len; ++m) {
        const size_t idx = b + i + L.inner * m;
        const T v = x[idx];
        if (op == 1) s = s * v;
        else if (op == 2) s = (s < v) ? s : v;
        else s = (s > v) ? s : v;
        out[idx] = s;
      }
    }
  }
}

}  // namespace

// ===========================================================================
// C ABI used by the tests and by bench.py's CPU baseline
// ===========================================================================
extern "C" {

enum {
  OR_EXP = 0, OR_LOG = 1, OR_SQRT = 2, OR_ERFC = 3, OR_NORMCDF = 4,
  OR_FASTNORMCDF = 5, OR_COS = 6, OR_SIN = 7, OR_ABS = 8, OR_EXP2 = 9, OR_EXP10 = 10,
  OR_LOG2 = 11, OR_LOG10 = 12, OR_FLOOR = 13, OR_CEIL = 14, OR_TRUNC = 15, OR_ROUND = 16,
  OR_ROUND2EVEN = 17, OR_SIGN = 18, OR_EXPM1 = 19, OR_LOG1P = 20, OR_TANH = 21, OR_SINH = 22,
  OR_COSH = 23, OR_ATAN = 24, OR_TAN = 25, OR_ASIN = 26, OR_ACOS = 27, OR_ERF = 28,
  OR_CBRT = 29, OR_ASINH = 30, OR_ACOSH = 31, OR_ATANH = 32, OR_COT = 33, OR_SEC = 34, OR_CSC = 35,
  OR_SINPI = 36, OR_COSPI = 37, OR_TANPI = 38, OR_COTPI = 39, OR_SIND = 40, OR_COSD = 41, OR_TAND = 42,
  OR_COTD = 43,
  // 44..69: the compositions (compose_o; 68 = sinc by pack)
  OR_SINC = 68, OR_SINCPI = 69, OR_EXPX2 = 75, OR_SINHC = 78,
  OR_LAST = OR_SINHC
};

int oracle_abi_version(void) { return 1; }

// 0 (default): NT2's single-precision Payne-Hanek as written; 1: with the
// corrected top-bit shift (test-only switch, used to pin the restatement).
void oracle_set_ph_fix(int on) { g_ph_fix = on ? 1 : 0; }

void oracle_set_threads(int n) {
  g_threads = n;
  if (n > 0) omp_set_num_threads(n);
}
int oracle_get_threads(void) { return omp_get_max_threads(); }

// y = f(x) over a table of n floats, NT2 transform semantics
int oracle_map_f32(int op, const float* x, float* y, size_t n) {
  if (op < 0 || op > OR_LAST) return -1;
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    const size_t w = pack ? 8 : 1;
    switch (op) {
      case OR_EXP: for (size_t l = 0; l < w; ++l) y[i + l] = exp_f(x[i + l]); break;
      case OR_LOG: for (size_t l = 0; l < w; ++l) y[i + l] = log_f(x[i + l]); break;
      case OR_SQRT: for (size_t l = 0; l < w; ++l) y[i + l] = std::sqrt(x[i + l]); break;
      case OR_ERFC: for (size_t l = 0; l < w; ++l) y[i + l] = erfc_f(x[i + l]); break;
      case OR_NORMCDF: for (size_t l = 0; l < w; ++l) y[i + l] = normcdf_f(x[i + l]); break;
      case OR_FASTNORMCDF: for (size_t l = 0; l < w; ++l) y[i + l] = fastnormcdf_f(x[i + l]); break;
      case OR_ABS: for (size_t l = 0; l < w; ++l) y[i + l] = std::fabs(x[i + l]); break;
      case OR_EXP2: for (size_t l = 0; l < w; ++l) y[i + l] = exp2_f(x[i + l]); break;
      case OR_EXP10: for (size_t l = 0; l < w; ++l) y[i + l] = exp10_f(x[i + l]); break;
      case OR_LOG2: for (size_t l = 0; l < w; ++l) y[i + l] = log2_f(x[i + l]); break;
      case OR_LOG10: for (size_t l = 0; l < w; ++l) y[i + l] = log10_f(x[i + l]); break;
      case OR_FLOOR: for (size_t l = 0; l < w; ++l) y[i + l] = std::floor(x[i + l]); break;
      case OR_CEIL: for (size_t l = 0; l < w; ++l) y[i + l] = std::ceil(x[i + l]); break;
      case OR_TRUNC: for (size_t l = 0; l < w; ++l) y[i + l] = std::trunc(x[i + l]); break;
      case OR_ROUND: for (size_t l = 0; l < w; ++l) y[i + l] = std::round(x[i + l]); break;
      case OR_ROUND2EVEN: for (size_t l = 0; l < w; ++l) y[i + l] = std::nearbyint(x[i + l]); break;
      case OR_SIGN: for (size_t l = 0; l < w; ++l) y[i + l] = sign_of(x[i + l]); break;
      case OR_EXPM1: for (size_t l = 0; l < w; ++l) y[i + l] = expm1_f(x[i + l]); break;
      case OR_LOG1P: for (size_t l = 0; l < w; ++l) y[i + l] = log1p_f(x[i + l]); break;
      case OR_TANH: for (size_t l = 0; l < w; ++l) y[i + l] = tanh_f(x[i + l]); break;
      case OR_SINH: for (size_t l = 0; l < w; ++l) y[i + l] = sinh_f(x[i + l]); break;
      case OR_COSH: for (size_t l = 0; l < w; ++l) y[i + l] = cosh_f(x[i + l]); break;
      case OR_ATAN: for (size_t l = 0; l < w; ++l) y[i + l] = atan_f(x[i + l], !pack); break;
      case OR_ASIN: for (size_t l = 0; l < w; ++l) y[i + l] = asin_f(x[i + l]); break;
      case OR_ACOS: for (size_t l = 0; l < w; ++l) y[i + l] = acos_f(x[i + l]); break;
      case OR_ERF: for (size_t l = 0; l < w; ++l) y[i + l] = erf_f(x[i + l]); break;
      case OR_CBRT: for (size_t l = 0; l < w; ++l) y[i + l] = cbrt_t(x[i + l]); break;
      case OR_ASINH: for (size_t l = 0; l < w; ++l) y[i + l] = asinh_f(x[i + l], !pack); break;
      case OR_ACOSH: for (size_t l = 0; l < w; ++l) y[i + l] = acosh_f(x[i + l], !pack); break;
      case OR_ATANH: for (size_t l = 0; l < w; ++l) y[i + l] = atanh_f(x[i + l]); break;
      case OR_COS:
      case OR_SIN:
      case OR_TAN:
      case OR_COT:
      case OR_SEC:
      case OR_CSC: {
        const int kind = op == OR_COS ? 0 : op == OR_SIN ? 1 : op == OR_TAN ? 3 : op == OR_COT ? 4 : op == OR_SEC ? 5 : 6;
        if (pack) trig_group_f<8>(kind, x + i, y + i, nullptr);
        else trig_group_f<1>(kind, x + i, y + i, nullptr);
        break;
      }
      case OR_SINPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, true>(1, x[i + l]); break;
      case OR_COSPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, true>(0, x[i + l]); break;
      case OR_TANPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, true>(3, x[i + l]); break;
      case OR_COTPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, true>(4, x[i + l]); break;
      case OR_SIND: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, false>(1, x[i + l]); break;
      case OR_COSD: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, false>(0, x[i + l]); break;
      case OR_TAND: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, false>(3, x[i + l]); break;
      case OR_COTD: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<float, false>(4, x[i + l]); break;
      case OR_SINC:
        if (pack) sinc_o<float, 8>(x + i, y + i);
        else sinc_o<float, 1>(x + i, y + i);
        break;
      default: for (size_t l = 0; l < w; ++l) y[i + l] = compose_o<float>(op, x[i + l], !pack); break;
    }
  });
  return 0;
}

int oracle_map_f64(int op, const double* x, double* y, size_t n) {
  if (op < 0 || op > OR_LAST || op == OR_FASTNORMCDF) return -1;
  nt2_transform(n, 4, sizeof(double), [&](size_t i, bool pack) {
    const size_t w = pack ? 4 : 1;
    switch (op) {
      case OR_EXP: for (size_t l = 0; l < w; ++l) y[i + l] = exp_d(x[i + l]); break;
      case OR_LOG: for (size_t l = 0; l < w; ++l) y[i + l] = log_d(x[i + l]); break;
      case OR_SQRT: for (size_t l = 0; l < w; ++l) y[i + l] = std::sqrt(x[i + l]); break;
      case OR_ERFC: for (size_t l = 0; l < w; ++l) y[i + l] = erfc_d(x[i + l], !pack); break;
      case OR_NORMCDF: for (size_t l = 0; l < w; ++l) y[i + l] = normcdf_d(x[i + l], !pack); break;
      case OR_ABS: for (size_t l = 0; l < w; ++l) y[i + l] = std::fabs(x[i + l]); break;
      case OR_EXP2: for (size_t l = 0; l < w; ++l) y[i + l] = exp2_d(x[i + l]); break;
      case OR_EXP10: for (size_t l = 0; l < w; ++l) y[i + l] = exp10_d(x[i + l]); break;
      case OR_LOG2: for (size_t l = 0; l < w; ++l) y[i + l] = log2_d(x[i + l]); break;
      case OR_LOG10: for (size_t l = 0; l < w; ++l) y[i + l] = log10_d(x[i + l]); break;
      case OR_FLOOR: for (size_t l = 0; l < w; ++l) y[i + l] = std::floor(x[i + l]); break;
      case OR_CEIL: for (size_t l = 0; l < w; ++l) y[i + l] = std::ceil(x[i + l]); break;
      case OR_TRUNC: for (size_t l = 0; l < w; ++l) y[i + l] = std::trunc(x[i + l]); break;
      case OR_ROUND: for (size_t l = 0; l < w; ++l) y[i + l] = std::round(x[i + l]); break;
      case OR_ROUND2EVEN: for (size_t l = 0; l < w; ++l) y[i + l] = std::nearbyint(x[i + l]); break;
      case OR_SIGN: for (size_t l = 0; l < w; ++l) y[i + l] = sign_of(x[i + l]); break;
      case OR_EXPM1: for (size_t l = 0; l < w; ++l) y[i + l] = expm1_d(x[i + l]); break;
      case OR_LOG1P: for (size_t l = 0; l < w; ++l) y[i + l] = log1p_d(x[i + l]); break;
      case OR_TANH: for (size_t l = 0; l < w; ++l) y[i + l] = tanh_d(x[i + l]); break;
      case OR_SINH: for (size_t l = 0; l < w; ++l) y[i + l] = sinh_d(x[i + l]); break;
      case OR_COSH: for (size_t l = 0; l < w; ++l) y[i + l] = cosh_d(x[i + l]); break;
      case OR_ATAN: for (size_t l = 0; l < w; ++l) y[i + l] = atan_d(x[i + l], !pack); break;
      case OR_ASIN: for (size_t l = 0; l < w; ++l) y[i + l] = asin_d(x[i + l], !pack); break;
      case OR_ACOS: for (size_t l = 0; l < w; ++l) y[i + l] = acos_d(x[i + l], !pack); break;
      case OR_ERF: for (size_t l = 0; l < w; ++l) y[i + l] = erf_d(x[i + l], !pack); break;
      case OR_CBRT: for (size_t l = 0; l < w; ++l) y[i + l] = cbrt_t(x[i + l]); break;
      case OR_ASINH: for (size_t l = 0; l < w; ++l) y[i + l] = asinh_d(x[i + l], !pack); break;
      case OR_ACOSH: for (size_t l = 0; l < w; ++l) y[i + l] = acosh_d(x[i + l], !pack); break;
      case OR_ATANH: for (size_t l = 0; l < w; ++l) y[i + l] = atanh_d(x[i + l]); break;
      case OR_COS:
      case OR_SIN:
      case OR_TAN:
      case OR_COT:
      case OR_SEC:
      case OR_CSC: {
        const int kind = op == OR_COS ? 0 : op == OR_SIN ? 1 : op == OR_TAN ? 3 : op == OR_COT ? 4 : op == OR_SEC ? 5 : 6;
        if (pack) trig_group_d<4>(kind, x + i, y + i, nullptr);
        else trig_group_d<1>(kind, x + i, y + i, nullptr);
        break;
      }
      case OR_SINPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, true>(1, x[i + l]); break;
      case OR_COSPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, true>(0, x[i + l]); break;
      case OR_TANPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, true>(3, x[i + l]); break;
      case OR_COTPI: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, true>(4, x[i + l]); break;
      case OR_SIND: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, false>(1, x[i + l]); break;
      case OR_COSD: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, false>(0, x[i + l]); break;
      case OR_TAND: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, false>(3, x[i + l]); break;
      case OR_COTD: for (size_t l = 0; l < w; ++l) y[i + l] = trig_unit<double, false>(4, x[i + l]); break;
      case OR_SINC:
        if (pack) sinc_o<double, 4>(x + i, y + i);
        else sinc_o<double, 1>(x + i, y + i);
        break;
      default: for (size_t l = 0; l < w; ++l) y[i + l] = compose_o<double>(op, x[i + l], !pack); break;
    }
  });
  return 0;
}

// out = atan2(x, y) over tables of n elements, NT2 transform semantics
int oracle_map2_f32(int op, const float* x, const float* y, float* out, size_t n) {
  if (op < 0 || op > 4) return -1;
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    const size_t w = pack ? 8 : 1;
    for (size_t l = 0; l < w; ++l)
      out[i + l] = op == 0   ? atan2_f(x[i + l], y[i + l], !pack)
                   : op == 1 ? pow_t<float>(x[i + l], y[i + l], !pack)
                   : op == 2 ? indeg_o(atan2_f(x[i + l], y[i + l], !pack))
                   : op == 3 ? logspace_add_o(x[i + l], y[i + l])
                             : logspace_sub_o(x[i + l], y[i + l]);
  });
  return 0;
}
int oracle_map2_f64(int op, const double* x, const double* y, double* out, size_t n) {
  if (op < 0 || op > 4) return -1;
  nt2_transform(n, 4, sizeof(double), [&](size_t i, bool pack) {
    const size_t w = pack ? 4 : 1;
    for (size_t l = 0; l < w; ++l)
      out[i + l] = op == 0   ? atan2_d(x[i + l], y[i + l], !pack)
                   : op == 1 ? pow_t<double>(x[i + l], y[i + l], !pack)
                   : op == 2 ? indeg_o(atan2_d(x[i + l], y[i + l], !pack))
                   : op == 3 ? logspace_add_o(x[i + l], y[i + l])
                             : logspace_sub_o(x[i + l], y[i + l]);
  });
  return 0;
}

// pow(x, k) with an integral exponent:
// modules/core/exponential/include/nt2/exponential/functions/generic/pow.hpp:110-132
// (|k| by repeated squaring in the base's precision) and :154-168 (1/p for k < 0);
// elementwise, so the pack schedule does not matter
extern "C++" template <class T> T powi_t(T base, int64_t k) {
  uint64_t e = k < 0 ? (uint64_t)0 - (uint64_t)k : (uint64_t)k;
  T p = T(1);
  while (e) {
    if (e & 1u) p = p * base;
    e >>= 1;
    base = base * base;
  }
  return k < 0 ? T(1) / p : p;
}
int oracle_powi_f32(const float* x, const int64_t* k, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = powi_t(x[i], k[i]);
  return 0;
}
int oracle_powi_f64(const double* x, const int64_t* k, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = powi_t(x[i], k[i]);
  return 0;
}

// [s, c] = sincos(x): generic/sincos.hpp:50-64
int oracle_sincos_f32(const float* x, float* s, float* c, size_t n) {
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    if (pack) trig_group_f<8>(2, x + i, s + i, c + i);
    else trig_group_f<1>(2, x + i, s + i, c + i);
  });
  return 0;
}
int oracle_sincos_f64(const double* x, double* s, double* c, size_t n) {
  nt2_transform(n, 4, sizeof(double), [&](size_t i, bool pack) {
    if (pack) trig_group_d<4>(2, x + i, s + i, c + i);
    else trig_group_d<1>(2, x + i, s + i, c + i);
  });
  return 0;
}

// Y = alpha*X + Y: demo/saxpy/nt2/saxpy_nt2.cpp:39-42
int oracle_saxpy_f32(size_t n, float a, const float* x, float* y) {
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    const size_t w = pack ? 8 : 1;
    for (size_t l = 0; l < w; ++l) y[i + l] = a * x[i + l] + y[i + l];
  });
  return 0;
}

float oracle_sum_f32(const float* x, size_t n) {
  return fold_sum_f32(n, [&](size_t i, float* v, int w) {
    for (int l = 0; l < w; ++l) v[l] = x[i + l];
  });
}

double oracle_sum_f64(const double* x, size_t n) {
  double acc[4] = {0, 0, 0, 0};
  const size_t packs = n / 4;
  for (size_t p = 0; p < packs; ++p)
    for (int l = 0; l < 4; ++l) acc[l] = acc[l] + x[p * 4 + l];
  double s = hsum4(acc);
  for (size_t i = packs * 4; i < n; ++i) s = s + x[i];
  return s;
}

// res = globalsum(X*Y): demo/sdot/nt2/sdot_nt2.cpp:35-38;
// dot(a,b) = sum(a*conj(b)): modules/core/reduction/include/nt2/reduction/functions/expr/dot.hpp:22-29
float oracle_dot_f32(const float* x, const float* y, size_t n) {
  return fold_sum_f32(n, [&](size_t i, float* v, int w) {
    for (int l = 0; l < w; ++l) v[l] = x[i + l] * y[i + l];
  });
}

// NT2's OpenMP inner_fold for the CPU baseline timing: per-thread pack
// partials combined under a critical section in arrival order
// (modules/arch/openmp/include/nt2/sdk/openmp/spawner/fold.hpp:49-66).
float oracle_dot_f32_omp(const float* x, const float* y, size_t n) {
  const size_t grain = l2_grain(sizeof(float));
  if (n <= 8 * grain) return oracle_dot_f32(x, y, n);
  const size_t packs = n / 8, ppb = grain / 8, nblk = (packs + ppb - 1) / ppb;
  float total[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp parallel
  {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static)
    for (size_t b = 0; b < nblk; ++b) {
      const size_t p1 = std::min(packs, (b + 1) * ppb);
      for (size_t p = b * ppb; p < p1; ++p)
        for (int l = 0; l < 8; ++l) acc[l] = acc[l] + x[p * 8 + l] * y[p * 8 + l];
    }
#pragma omp critical
    for (int l = 0; l < 8; ++l) total[l] = total[l] + acc[l];
  }
  float s = hsum8(total);
  for (size_t i = packs * 8; i < n; ++i) s = s + x[i] * y[i];
  return s;
}

// Exact reference for float sums: every float product/value is summed in
// double-double (Dekker/Knuth TwoSum), returned as the rounded double.
double oracle_dot_f32_exact(const float* x, const float* y, size_t n) {
  double hi = 0.0, lo = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double p = (double)x[i] * (double)y[i];  // exact: 24+24 bits
    double s = hi + p, bb = s - hi, e = (hi - (s - bb)) + (p - bb);
    hi = s;
    lo += e;
  }
  return hi + lo;
}
double oracle_sum_f32_exact(const float* x, size_t n) {
  double hi = 0.0, lo = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double p = (double)x[i];
    double s = hi + p, bb = s - hi, e = (hi - (s - bb)) + (p - bb);
    hi = s;
    lo += e;
  }
  return hi + lo;
}

// ---------------------------------------------------------------------------
// Black-Scholes.  mode 0 = textbook call/put with erfc-based normcdf:
//   da = sqrt(T); vsT = v*da; d1 = (log(S/X) + fma(sqr(v), 0.5, r)*T)/vsT;
//   d2 = d1 - vsT; e = X*exp(-r*T);
//   call = S*normcdf(d1) - e*normcdf(d2); put = e*normcdf(-d2) - S*normcdf(-d1)
// mode 1 = the NT2 demo formula (demo/blackscholes/nt2/blackandscholes_nt2.cpp:55-59):
//   d1 = log(S/X) + (fma(sqr(v),0.5,r)*T)/(v*da); d2 = fnms(v,da,d1);
//   call = fnms(X*exp(-r*T), fastnormcdf(d2), S*fastnormcdf(d1)); put untouched.
// fma/fnms are NT2's two-rounding forms (generic/fma.hpp:32, generic/fnms.hpp:32).
// ---------------------------------------------------------------------------
int oracle_bs_f32(size_t n, const float* S, const float* X, const float* T,
                  float r, float v, float* call, float* put, int mode) {
  if (mode != 0 && mode != 1) return -1;
  const float drift = (v * v) * 0.5f + r;
  const float mr = -r;
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    const size_t w = pack ? 8 : 1;
    for (size_t l = 0; l < w; ++l) {
      const size_t j = i + l;
      const float da = std::sqrt(T[j]);
      const float vsT = v * da;
      const float e = X[j] * exp_f(mr * T[j]);
      if (mode == 0) {
        const float d1 = (log_f(S[j] / X[j]) + drift * T[j]) / vsT;
        const float d2 = d1 - vsT;
        call[j] = S[j] * normcdf_f(d1) - e * normcdf_f(d2);
        put[j] = e * normcdf_f(-d2) - S[j] * normcdf_f(-d1);
      } else {
        const float d1 = log_f(S[j] / X[j]) + (drift * T[j]) / vsT;
        const float d2 = d1 - v * da;
        call[j] = S[j] * fastnormcdf_f(d1) - e * fastnormcdf_f(d2);
      }
    }
  });
  return 0;
}

int oracle_bs_f64(size_t n, const double* S, const double* X, const double* T,
                  double r, double v, double* call, double* put) {
  const double drift = (v * v) * 0.5 + r;
  const double mr = -r;
  nt2_transform(n, 4, sizeof(double), [&](size_t i, bool pack) {
    const size_t w = pack ? 4 : 1;
    for (size_t l = 0; l < w; ++l) {
      const size_t j = i + l;
      const double da = std::sqrt(T[j]);
      const double vsT = v * da;
      const double e = X[j] * exp_d(mr * T[j]);
      const double d1 = (log_d(S[j] / X[j]) + drift * T[j]) / vsT;
      const double d2 = d1 - vsT;
      call[j] = S[j] * normcdf_d(d1, !pack) - e * normcdf_d(d2, !pack);
      put[j] = e * normcdf_d(-d2, !pack) - S[j] * normcdf_d(-d1, !pack);
    }
  });
  return 0;
}

// FP64 accuracy reference for Black-Scholes: the textbook formula evaluated in
// double with the C library's log/exp/erfc (independent of NT2's kernels).
// Inputs are the f32 table values promoted exactly; outputs stay in double.
int oracle_bs_fp64ref_f32in(size_t n, const float* S, const float* X, const float* T,
                            double r, double v, double* call, double* put) {
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < n; ++j) {
    const double s = S[j], x = X[j], t = T[j];
    const double vsT = v * std::sqrt(t);
    const double d1 = (std::log(s / x) + (r + 0.5 * v * v) * t) / vsT;
    const double d2 = d1 - vsT;
    const double e = x * std::exp(-r * t);
    const double n1 = 0.5 * std::erfc(-d1 / std::sqrt(2.0));
    const double n2 = 0.5 * std::erfc(-d2 / std::sqrt(2.0));
    const double m1 = 0.5 * std::erfc(d1 / std::sqrt(2.0));
    const double m2 = 0.5 * std::erfc(d2 / std::sqrt(2.0));
    call[j] = s * n1 - e * n2;
    put[j] = e * m2 - s * m1;
  }
  return 0;
}
int oracle_bs_fp64ref_f64in(size_t n, const double* S, const double* X, const double* T,
                            double r, double v, double* call, double* put) {
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < n; ++j) {
    const double s = S[j], x = X[j], t = T[j];
    const double vsT = v * std::sqrt(t);
    const double d1 = (std::log(s / x) + (r + 0.5 * v * v) * t) / vsT;
    const double d2 = d1 - vsT;
    const double e = x * std::exp(-r * t);
    call[j] = s * (0.5 * std::erfc(-d1 / std::sqrt(2.0))) - e * (0.5 * std::erfc(-d2 / std::sqrt(2.0)));
    put[j] = e * (0.5 * std::erfc(d2 / std::sqrt(2.0))) - s * (0.5 * std::erfc(d1 / std::sqrt(2.0)));
  }
  return 0;
}

// Quad-precision (113-bit, libquadmath) Black-Scholes reference for f64
// prices, returned as the double-double (hi, lo) of the quad value.  The
// textbook formula (the same as oracle_bs_fp64ref_*) with the smaller price
// (the call when S < F = X exp(-rT), else the put) evaluated from
// N(x) = erfc(-x/sqrt2)/2 and the other by put-call parity, which then adds a
// non-negative term; the direct terms cancel by at most a few hundred times
// over the benchmark inputs against 113 bits: the result is good to ~2^-100
// relative there.
static inline __float128 qncdf(__float128 x) {
  return (__float128)0.5 * erfcq(-x / sqrtq((__float128)2));
}
int oracle_bs_quad_f64in(size_t n, const double* S, const double* X, const double* T, double r,
                         double v, double* call_hi, double* call_lo, double* put_hi, double* put_lo) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (size_t j = 0; j < n; ++j) {
    const __float128 s = S[j], x = X[j], t = T[j], rr = r, vv = v;
    const __float128 tau = vv * sqrtq(t);
    const __float128 d1 = (logq(s / x) + (rr + vv * vv / 2) * t) / tau, d2 = d1 - tau;
    const __float128 F = x * expq(-rr * t);
    __float128 c, p;
    if (F > s) {  // the call is the smaller price
      c = s * qncdf(d1) - F * qncdf(d2);
      p = c + (F - s);
    } else {
      p = F * qncdf(-d2) - s * qncdf(-d1);
      c = p + (s - F);
    }
    call_hi[j] = (double)c;
    call_lo[j] = (double)(c - (__float128)call_hi[j]);
    put_hi[j] = (double)p;
    put_lo[j] = (double)(p - (__float128)put_hi[j]);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// uint8 ops and sigma-delta
// adds (u8) = min(255, a+b): arithmetic/functions/scalar/adds.hpp:54-67
// selinc/seldec = c ? a+1 : a / c ? a-1 : a: boolean/functions/scalar/selinc.hpp:18-29, seldec.hpp:18-29
// if_zero_else_one(c) = c ? 0 : 1: boolean/functions/scalar/if_zero_else_one.hpp:21-52
// frame step: demo/sigmadelta/nt2/sigmadelta_nt2.cpp:37-58, driver :100-106
// mul = adds applied N-1 times (N = 3 reproduces the demo's adds(adds(d,d),d)).
// ---------------------------------------------------------------------------
static inline uint8_t adds_u8(uint8_t a, uint8_t b) {
  unsigned s = (unsigned)a + b;
  return (uint8_t)(s > 255 ? 255 : s);
}
static inline uint8_t subs_u8(uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : 0); }

int oracle_u8_binop(int op, const uint8_t* a, const uint8_t* b, uint8_t* y, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    switch (op) {
      case 0: y[i] = adds_u8(a[i], b[i]); break;
      case 1: y[i] = subs_u8(a[i], b[i]); break;
      case 2: y[i] = (uint8_t)(a[i] < b[i] ? a[i] + 1 : a[i]); break;  // selinc(a<b, a)
      case 3: y[i] = (uint8_t)(a[i] > b[i] ? a[i] - 1 : a[i]); break;  // seldec(a>b, a)
      case 4: y[i] = (uint8_t)(a[i] < b[i] ? 0 : 1); break;            // if_zero_else_one(a<b)
      case 5: y[i] = (uint8_t)(std::max(a[i], b[i]) - std::min(a[i], b[i])); break;
      default: return -1;
    }
  }
  return 0;
}

int oracle_sigmadelta_u8(size_t npix, int nframes, const uint8_t* frames,
                         size_t frame_stride, uint8_t* bkg, uint8_t* var,
                         uint8_t* label, int N, uint32_t* frame_counts) {
  if (N < 1 || nframes < 1) return -1;
  if (frame_counts)
    for (int k = 0; k < nframes; ++k) frame_counts[k] = 0;
  for (int k = 1; k < nframes; ++k) {
    const uint8_t* fr = frames + (size_t)k * frame_stride;
    uint64_t cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
    for (size_t i = 0; i < npix; ++i) {
      uint8_t b = bkg[i];
      const uint8_t f = fr[i];
      b = (b < f) ? (uint8_t)(b + 1) : ((b > f) ? (uint8_t)(b - 1) : b);
      const uint8_t d = (uint8_t)(std::max(b, f) - std::min(b, f));
      uint8_t m = d;
      for (int t = 1; t < N; ++t) m = adds_u8(m, d);
      uint8_t vv = var[i];
      if (d != 0) vv = (vv < m) ? (uint8_t)(vv + 1) : ((vv > m) ? (uint8_t)(vv - 1) : vv);
      const uint8_t lab = (d < vv) ? 0 : 1;
      bkg[i] = b;
      var[i] = vv;
      label[i] = lab;
      cnt += lab;
    }
    if (frame_counts) frame_counts[k] = (uint32_t)cnt;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Batched expression z = exp(-x.*x/2).*cos(y) + sqrt(abs(x)), s = sum(z).
// Parsed as NT2 would: ((-x).*x)/2, exp, cos, times, plus sqrt(abs(x)).
// ---------------------------------------------------------------------------
static inline void expr4_pack(const float* x, const float* y, float* z, int w) {
  float c[8];
  if (w == 8) trig_group_f<8>(0, y, c, nullptr);
  else for (int l = 0; l < w; ++l) trig_group_f<1>(0, y + l, c + l, nullptr);
  for (int l = 0; l < w; ++l) {
    const float t = ((-x[l]) * x[l]) / 2.0f;
    z[l] = exp_f(t) * c[l] + std::sqrt(std::fabs(x[l]));
  }
}

int oracle_expr4_f32(const float* x, const float* y, float* z, size_t n) {
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    expr4_pack(x + i, y + i, z + i, pack ? 8 : 1);
  });
  return 0;
}

float oracle_expr4_sum_f32(const float* x, const float* y, size_t n) {
  return fold_sum_f32(n, [&](size_t i, float* v, int w) { expr4_pack(x + i, y + i, v, w); });
}

// FP64 reference of the batched expression: z computed in double with the C
// library from the f32 inputs, summed in double-double.
double oracle_expr4_sum_fp64(const float* x, const float* y, size_t n) {
  double hi = 0.0, lo = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double xd = x[i], yd = y[i];
    const double p = std::exp(-xd * xd / 2.0) * std::cos(yd) + std::sqrt(std::fabs(xd));
    double s = hi + p, bb = s - hi, e = (hi - (s - bb)) + (p - bb);
    hi = s;
    lo += e;
  }
  return hi + lo;
}

// sum(a, dim) / cumsum(a, dim) of a column-major table of extent ext[0..3]
int oracle_sum_dim_f32(const float* x, const size_t* ext, int dim, float* out) {
  if (dim < 1) return -1;
  sum_dim<float, 8>(x, ext, dim, out);
  return 0;
}
int oracle_sum_dim_f64(const double* x, const size_t* ext, int dim, double* out) {
  if (dim < 1) return -1;
  sum_dim<double, 4>(x, ext, dim, out);
  return 0;
}
int oracle_cumsum_dim_f32(const float* x, const size_t* ext, int dim, float* out) {
  if (dim < 1) return -1;
  cumsum_dim<float>(x, ext, dim, out);
  return 0;
}
int oracle_cumsum_dim_f64(const double* x, const size_t* ext, int dim, double* out) {
  if (dim < 1) return -1;
  cumsum_dim<double>(x, ext, dim, out);
  return 0;
}
// op: 1 cumprod, 2 cummin, 3 cummax
int oracle_cumop_dim_f32(const float* x, const size_t* ext, int dim, int op, float* out) {
  if (dim < 1 || op < 1 || op > 3) return -1;
  cumop_dim<float>(x, ext, dim, op, out);
  return 0;
}
int oracle_cumop_dim_f64(const double* x, const size_t* ext, int dim, int op, double* out) {
  if (dim < 1 || op < 1 || op > 3) return -1;
  cumop_dim<double>(x, ext, dim, op, out);
  return 0;
}


// ---------------------------------------------------------------------------
// The image demos.  rgb2grey: gr = 0.3 r + 0.59 g + 0.11 b in float
// (demo/rgb2grey/nt2/rgb2grey_nt2.cpp:42-45); rgb2yuv: y, then u and v from the
// stored y (demo/rgb2yuv/nt2/rgb2yuv_nt2.cpp:30-35).  Each product and sum
// rounds once (AVX without FMA).
// ---------------------------------------------------------------------------
int oracle_rgb2grey_f32(size_t n, const float* r, const float* g, const float* b, float* gr) {
  const float cr = 0.3f, cg = 0.59f, cb = 0.11f;
  for (size_t i = 0; i < n; ++i) gr[i] = (cr * r[i] + cg * g[i]) + cb * b[i];
  return 0;
}
// uint8 variant: the weights in 8.8 fixed point, rounded, saturated
int oracle_rgb2grey_u8(size_t n, const uint8_t* r, const uint8_t* g, const uint8_t* b, uint8_t* gr) {
  for (size_t i = 0; i < n; ++i) {
    const unsigned v = (77u * r[i] + 151u * g[i] + 28u * b[i] + 128u) >> 8;
    gr[i] = (uint8_t)(v > 255u ? 255u : v);
  }
  return 0;
}
int oracle_rgb2yuv_f32(size_t n, const float* r, const float* g, const float* b, float* y, float* u,
                       float* v) {
  for (size_t i = 0; i < n; ++i) {
    const float yy = (0.299f * r[i] + 0.587f * g[i]) + 0.114f * b[i];
    y[i] = yy;
    u[i] = 0.492f * (b[i] - yy);
    v[i] = 0.877f * (r[i] - yy);
  }
  return 0;
}

// linspace(l, u, n): element p (0-based) is u for p = n - 1, else p*step + l
// with step = (u - l)/(n - 1), two roundings
// (modules/core/generative/include/nt2/core/functions/details/linspace.hpp:22-48).
extern "C++" template <class T> void linspace_t(size_t n, T l, T u, T* out) {
  const T step = (u - l) / T(n - 1);
  for (size_t p = 0; p < n; ++p) out[p] = (p == n - 1) ? u : T(p) * step + l;
}
int oracle_linspace_f32(size_t n, float l, float u, float* out) { linspace_t(n, l, u, out); return 0; }
int oracle_linspace_f64(size_t n, double l, double u, double* out) { linspace_t(n, l, u, out); return 0; }

// mandelbrot: C = arrayfun(step, A, B) (demo/mandelbrot/nt2/mandelbrot_nt2.cpp:35-70,
// :100-103).  The step runs on a whole pack: z = z^2 + c for every lane, a lane
// counts an iteration while its |z|^2 before the update is below 4, and the
// pack stops when no lane counted or after max_iter iterations (256 in the
// demo).  Packs of 8 from index 0 (8 floats, 8 int32 on AVX), then one lane
// at a time for the tail.
extern "C++" template <int W>
void mandel_pack(const float* a, const float* b, int32_t* c, int max_iter) {
  float x[W], y[W];
  int32_t it[W];
  for (int l = 0; l < W; ++l) { x[l] = 0.0f; y[l] = 0.0f; it[l] = 0; }
  int i = 0;
  bool any;
  do {
    any = false;
    for (int l = 0; l < W; ++l) {
      const float x2 = x[l] * x[l], y2 = y[l] * y[l], xy = (2.0f * x[l]) * y[l];
      x[l] = (x2 - y2) + a[l];
      y[l] = xy + b[l];
      const bool in = (x2 + y2) < 4.0f;
      it[l] += in ? 1 : 0;
      any = any || in;
    }
    ++i;
  } while (any && i < max_iter);
  for (int l = 0; l < W; ++l) c[l] = it[l];
}
int oracle_mandelbrot_f32(size_t n, const float* a, const float* b, int32_t* c, int max_iter) {
  if (max_iter < 1) return -1;
  nt2_transform(n, 8, sizeof(float), [&](size_t i, bool pack) {
    if (pack) mandel_pack<8>(a + i, b + i, c + i, max_iter);
    else mandel_pack<1>(a + i, b + i, c + i, max_iter);
  });
  return 0;
}

}  // extern "C"
