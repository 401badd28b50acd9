// Per-option Black-Scholes evaluation shared by the fused kernel and the
// serving engine, so that a request gives the same bits alone or batched.
#pragma once

#include <hip/hip_runtime.h>

#include "nt2/hip/math.hpp"
#include "nt2h.h"

namespace nt2h {
namespace dm = nt2::hip::dm;

// textbook: da = sqrt(T); vsT = v*da; d1 = (log(S/X) + drift*T)/vsT with
// drift = fma(sqr(v), 0.5, r); d2 = d1 - vsT; e = X*exp(-r*T);
// call = S*N(d1) - e*N(d2); put = e*N(-d2) - S*N(-d1), N = normcdf.
// N(d) and N(-d) share one erfc evaluation: erfc(-c*d) and erfc(c*d) have the
// same |argument|, NT2's erfc folds the sign afterwards (erfc.hpp:118-121).
// demo: demo/blackscholes/nt2/blackandscholes_nt2.cpp:55-59.
template <class P, int MODE>
__device__ __forceinline__ void bs_one_f32(float S, float X, float T, float mr, float v,
                                           float drift, float& call, float& put) {
  _Pragma("clang fp contract(off)")
  const float da = __builtin_sqrtf(T);
  const float vsT = v * da;
  const float e = X * dm::exp_f<P>(mr * T);
  if constexpr (MODE == NT2H_BS_TEXTBOOK) {
    const float d1 = (dm::log_f<P>(S / X) + drift * T) / vsT;
    const float d2 = d1 - vsT;
    const float c = NT2H_F(0x3f3504f3);
    const float a1 = (-c) * d1, a2 = (-c) * d2;  // erfc arguments of N(d1), N(d2)
    const float r1 = dm::erfc_abs_f<P>(__builtin_fabsf(a1));
    const float r2 = dm::erfc_abs_f<P>(__builtin_fabsf(a2));
    const float n1 = 0.5f * dm::erfc_fold_f(a1, r1), m1 = 0.5f * dm::erfc_fold_f(-a1, r1);
    const float n2 = 0.5f * dm::erfc_fold_f(a2, r2), m2 = 0.5f * dm::erfc_fold_f(-a2, r2);
    call = S * n1 - e * n2;
    put = e * m2 - S * m1;
  } else {
    const float d1 = dm::log_f<P>(S / X) + (drift * T) / vsT;
    const float d2 = d1 - v * da;
    call = S * dm::fastnormcdf_f<P>(d1) - e * dm::fastnormcdf_f<P>(d2);
    put = 0.0f;
  }
}

template <class P, bool Scalar>
__device__ __forceinline__ void bs_one_f64(double S, double X, double T, double mr, double v,
                                           double drift, double& call, double& put) {
  _Pragma("clang fp contract(off)")
  const double da = __builtin_sqrt(T);
  const double vsT = v * da;
  const double e = X * dm::exp_d<P>(mr * T);
  const double d1 = (dm::log_d<P>(S / X) + drift * T) / vsT;
  const double d2 = d1 - vsT;
  call = S * dm::normcdf_d<P, Scalar>(d1) - e * dm::normcdf_d<P, Scalar>(d2);
  put = e * dm::normcdf_d<P, Scalar>(-d2) - S * dm::normcdf_d<P, Scalar>(-d1);
}

}  // namespace nt2h
