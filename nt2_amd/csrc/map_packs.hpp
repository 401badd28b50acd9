// The elementwise map schedule of NT2's transform_
// (modules/core/base/include/nt2/core/functions/simd/transform.hpp:27-58):
// work item t < n/W evaluates the pack [W t, W t + W) through the pack form
// F::run<W>, and the items beyond take the n % W tail elements one each
// through NT2's scalar form F::run<1>.  Packs move as two 16-byte accesses
// when both tables are 16-byte aligned (vec != 0).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace nt2h {

template <class F, class T, int W>
__device__ __forceinline__ void map_packs(const T* __restrict__ x, T* __restrict__ y, size_t n, int vec) {
  constexpr int H = 16 / (int)sizeof(T);  // lanes per 16-byte access
  static_assert(W == 2 * H, "a pack is two 16-byte accesses");
  typedef T v16 __attribute__((ext_vector_type(H)));
  const size_t npacks = n / W;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npacks) {
    T a[W], r[W];
    const size_t i = t * W;
    if (vec) {
      const v16 v0 = *reinterpret_cast<const v16*>(x + i);
      const v16 v1 = *reinterpret_cast<const v16*>(x + i + H);
#pragma unroll
      for (int l = 0; l < H; ++l) {
        a[l] = v0[l];
        a[H + l] = v1[l];
      }
    } else {
#pragma unroll
      for (int l = 0; l < W; ++l) a[l] = x[i + l];
    }
    F::template run<W>(a, r);
    if (vec) {
      v16 w0, w1;
#pragma unroll
      for (int l = 0; l < H; ++l) {
        w0[l] = r[l];
        w1[l] = r[H + l];
      }
      *reinterpret_cast<v16*>(y + i) = w0;
      *reinterpret_cast<v16*>(y + i + H) = w1;
    } else {
#pragma unroll
      for (int l = 0; l < W; ++l) y[i + l] = r[l];
    }
  } else if (t < npacks + (n - npacks * W)) {
    const size_t i = npacks * W + (t - npacks);
    T a = x[i], r;
    F::template run<1>(&a, &r);
    y[i] = r;
  }
}

// maps of the composed functions (op codes NT2H_OP_COTH and above), in
// nt2h_compose.hip; the policy is NT2H_EXACT or NT2H_FAST
int launch_compose(int op, size_t n, const float* x, float* y, int policy, hipStream_t s);
int launch_compose(int op, size_t n, const double* x, double* y, int policy, hipStream_t s);

}  // namespace nt2h
