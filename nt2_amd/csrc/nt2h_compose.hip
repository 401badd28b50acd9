// Elementwise maps of NT2's composed functions (C ABI op codes NT2H_OP_COTH
// to NT2H_OP_SINHC): the reciprocal hyperbolic functions and their inverses,
// the inverse trigonometric functions of reciprocals and in degree or pi
// units, the reciprocal trigonometric functions of degree and pi arguments,
// sinc, and the exponential family's sqrt1pm1, exprecneg(c), reallog,
// realsqrt and expx2, the inverse error functions, and sinhc.  The schedule is the other maps' (packs of 8 floats / 4 doubles
// from index 0, NT2's scalar form for the tail); the functions are in
// nt2/hip/math_compose.hpp.  A translation unit of its own so that the
// library's sources compile in parallel.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.hpp"
#include "map_packs.hpp"
#include "nt2/hip/math_compose.hpp"
#include "nt2h.h"

#pragma clang fp contract(off)

using nt2::hip::exact_policy;
using nt2::hip::fast_policy;
namespace dm = nt2::hip::dm;

namespace nt2h {

template <class P, int OP> struct compose_op {
  template <int G, class T> __device__ __forceinline__ static void run(const T* a, T* r) {
    constexpr bool S = (G == 1);
    if constexpr (OP == NT2H_OP_SINC) {
      dm::sinc_group<P, G>(a, r);
    } else {
#pragma unroll
      for (int l = 0; l < G; ++l) {
        const T x = a[l];
        T v;
        if constexpr (OP == NT2H_OP_COTH) v = dm::coth_t<P>(x);
        else if constexpr (OP == NT2H_OP_SECH) v = dm::sech_t<P>(x);
        else if constexpr (OP == NT2H_OP_CSCH) v = dm::csch_t<P>(x);
        else if constexpr (OP == NT2H_OP_ACOTH) v = dm::acoth_t<P>(x);
        else if constexpr (OP == NT2H_OP_ASECH) v = dm::asech_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACSCH) v = dm::acsch_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACOT) v = dm::acot_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ASEC) v = dm::asec_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACSC) v = dm::acsc_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACOTD) v = dm::acotd_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ASECD) v = dm::asecd_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACSCD) v = dm::acscd_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACOSD) v = dm::acosd_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ACOSPI) v = dm::acospi_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ASIND) v = dm::asind_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ASINPI) v = dm::asinpi_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ATAND) v = dm::atand_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ATANPI) v = dm::atanpi_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_INDEG) v = dm::indeg(x);
        else if constexpr (OP == NT2H_OP_INRAD) v = dm::inrad(x);
        else if constexpr (OP == NT2H_OP_SECD) v = dm::secd_t<P>(x);
        else if constexpr (OP == NT2H_OP_CSCD) v = dm::cscd_t<P>(x);
        else if constexpr (OP == NT2H_OP_SECPI) v = dm::secpi_t<P>(x);
        else if constexpr (OP == NT2H_OP_CSCPI) v = dm::cscpi_t<P>(x);
        else if constexpr (OP == NT2H_OP_SINCPI) v = dm::sincpi_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_SQRT1PM1) v = dm::sqrt1pm1(x);
        else if constexpr (OP == NT2H_OP_EXPRECNEG) v = dm::exprecneg_t<P>(x);
        else if constexpr (OP == NT2H_OP_EXPRECNEGC) v = dm::exprecnegc_t<P>(x);
        else if constexpr (OP == NT2H_OP_REALLOG) v = dm::log_t<P>(x);
        else if constexpr (OP == NT2H_OP_REALSQRT) v = dm::sqrtv(x);
        else if constexpr (OP == NT2H_OP_EXPX2) v = dm::expx2_t<P, S>(x);
        else if constexpr (OP == NT2H_OP_ERFINV) v = dm::erfinv_t<P>(x);
        else if constexpr (OP == NT2H_OP_ERFCINV) v = dm::erfcinv_t<P>(x);
        else v = dm::sinhc_t<P>(x);
        r[l] = v;
      }
    }
  }
};

template <class P, int OP, class T, int W>
__global__ __launch_bounds__(kBlock) void compose_k(const T* __restrict__ x, T* __restrict__ y, size_t n, int vec) {
  map_packs<compose_op<P, OP>, T, W>(x, y, n, vec);
}

template <class P, class T, int W>
static int launch_t(const char* name, int op, size_t n, const T* x, T* y, hipStream_t s) {
  const int vec = aligned16(x) && aligned16(y);
  const dim3 g(grid_for(n / W + n % W)), b(kBlock);
  switch (op) {
#define NT2H_COMPOSE(OPC) case OPC: hipLaunchKernelGGL((compose_k<P, OPC, T, W>), g, b, 0, s, x, y, n, vec); break;
    NT2H_COMPOSE(NT2H_OP_COTH) NT2H_COMPOSE(NT2H_OP_SECH) NT2H_COMPOSE(NT2H_OP_CSCH)
    NT2H_COMPOSE(NT2H_OP_ACOTH) NT2H_COMPOSE(NT2H_OP_ASECH) NT2H_COMPOSE(NT2H_OP_ACSCH)
    NT2H_COMPOSE(NT2H_OP_ACOT) NT2H_COMPOSE(NT2H_OP_ASEC) NT2H_COMPOSE(NT2H_OP_ACSC)
    NT2H_COMPOSE(NT2H_OP_ACOTD) NT2H_COMPOSE(NT2H_OP_ASECD) NT2H_COMPOSE(NT2H_OP_ACSCD)
    NT2H_COMPOSE(NT2H_OP_ACOSD) NT2H_COMPOSE(NT2H_OP_ACOSPI) NT2H_COMPOSE(NT2H_OP_ASIND)
    NT2H_COMPOSE(NT2H_OP_ASINPI) NT2H_COMPOSE(NT2H_OP_ATAND) NT2H_COMPOSE(NT2H_OP_ATANPI)
    NT2H_COMPOSE(NT2H_OP_INDEG) NT2H_COMPOSE(NT2H_OP_INRAD) NT2H_COMPOSE(NT2H_OP_SECD)
    NT2H_COMPOSE(NT2H_OP_CSCD) NT2H_COMPOSE(NT2H_OP_SECPI) NT2H_COMPOSE(NT2H_OP_CSCPI)
    NT2H_COMPOSE(NT2H_OP_SINC) NT2H_COMPOSE(NT2H_OP_SINCPI)
    NT2H_COMPOSE(NT2H_OP_SQRT1PM1) NT2H_COMPOSE(NT2H_OP_EXPRECNEG) NT2H_COMPOSE(NT2H_OP_EXPRECNEGC)
    NT2H_COMPOSE(NT2H_OP_REALLOG) NT2H_COMPOSE(NT2H_OP_REALSQRT) NT2H_COMPOSE(NT2H_OP_EXPX2)
    NT2H_COMPOSE(NT2H_OP_ERFINV) NT2H_COMPOSE(NT2H_OP_ERFCINV) NT2H_COMPOSE(NT2H_OP_SINHC)
#undef NT2H_COMPOSE
    default: return fail(NT2H_EINVAL, name);
  }
  NT2H_CHECK_LAUNCH(name);
  return NT2H_OK;
}

template <class T, int W>
static int launch_policy(const char* name, int op, size_t n, const T* x, T* y, int policy, hipStream_t s) {
  if (n == 0) return NT2H_OK;
  if (policy == NT2H_EXACT) return launch_t<exact_policy, T, W>(name, op, n, x, y, s);
  if (policy == NT2H_FAST) return launch_t<fast_policy, T, W>(name, op, n, x, y, s);
  return fail(NT2H_EINVAL, name);
}

int launch_compose(int op, size_t n, const float* x, float* y, int policy, hipStream_t s) {
  return launch_policy<float, 8>("nt2h_map_f32: unknown op or bad policy", op, n, x, y, policy, s);
}
int launch_compose(int op, size_t n, const double* x, double* y, int policy, hipStream_t s) {
  return launch_policy<double, 4>("nt2h_map_f64: unknown op or bad policy", op, n, x, y, policy, s);
}

}  // namespace nt2h
