// The image demos as single fused kernels: rgb2grey and rgb2yuv (HBM-bound
// elementwise passes), linspace (NT2's generator) and the mandelbrot escape
// count with NT2's pack-uniform iteration.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"
#include "nt2/hip/math.hpp"
#include "nt2h.h"

#pragma clang fp contract(off)

using nt2::hip::exact_policy;
using nt2::hip::fast_policy;
namespace dm = nt2::hip::dm;

namespace nt2h {

typedef float v4f __attribute__((ext_vector_type(4)));

// gr = 0.3 r + 0.59 g + 0.11 b, left to right: demo/rgb2grey/nt2/rgb2grey_nt2.cpp:42-45
template <class P>
__device__ __forceinline__ float grey_of(float r, float g, float b) {
  return dm::fma<P>(0.11f, b, dm::fma<P>(0.59f, g, 0.3f * r));
}

// y, then u and v from y: demo/rgb2yuv/nt2/rgb2yuv_nt2.cpp:30-35
template <class P>
__device__ __forceinline__ void yuv_of(float r, float g, float b, float& y, float& u, float& v) {
  y = dm::fma<P>(0.114f, b, dm::fma<P>(0.587f, g, 0.299f * r));
  u = 0.492f * (b - y);
  v = 0.877f * (r - y);
}

// One float4 of pixels per work item when the planes are 16-byte aligned,
// then one pixel per work item for the rest.
template <class P>
__global__ __launch_bounds__(kBlock) void rgb2grey_k(size_t n, const float* __restrict__ r,
                                                     const float* __restrict__ g,
                                                     const float* __restrict__ b,
                                                     float* __restrict__ gr, int vec) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n4 = vec ? n / 4 : 0;
  if (t < n4) {
    const v4f rv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(r) + t);
    const v4f gv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(g) + t);
    const v4f bv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(b) + t);
    v4f o;
    o.x = grey_of<P>(rv.x, gv.x, bv.x);
    o.y = grey_of<P>(rv.y, gv.y, bv.y);
    o.z = grey_of<P>(rv.z, gv.z, bv.z);
    o.w = grey_of<P>(rv.w, gv.w, bv.w);
    __builtin_nontemporal_store(o, reinterpret_cast<v4f*>(gr) + t);
  } else {
    const size_t i = n4 * 4 + (t - n4);
    if (i < n) gr[i] = grey_of<P>(r[i], g[i], b[i]);
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void rgb2yuv_k(size_t n, const float* __restrict__ r,
                                                    const float* __restrict__ g,
                                                    const float* __restrict__ b, float* __restrict__ y,
                                                    float* __restrict__ u, float* __restrict__ v, int vec) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n4 = vec ? n / 4 : 0;
  if (t < n4) {
    const v4f rv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(r) + t);
    const v4f gv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(g) + t);
    const v4f bv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(b) + t);
    v4f yv, uv, vv;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      float yy, uu, vvl;
      yuv_of<P>(rv[l], gv[l], bv[l], yy, uu, vvl);
      yv[l] = yy;
      uv[l] = uu;
      vv[l] = vvl;
    }
    __builtin_nontemporal_store(yv, reinterpret_cast<v4f*>(y) + t);
    __builtin_nontemporal_store(uv, reinterpret_cast<v4f*>(u) + t);
    __builtin_nontemporal_store(vv, reinterpret_cast<v4f*>(v) + t);
  } else {
    const size_t i = n4 * 4 + (t - n4);
    if (i < n) yuv_of<P>(r[i], g[i], b[i], y[i], u[i], v[i]);
  }
}

// uint8 grey: the demo's weights 0.3, 0.59, 0.11 in 8.8 fixed point (77, 151,
// 28; they sum to 256), rounded and saturated:
// grey = min(255, (77 r + 151 g + 28 b + 128) >> 8).  Four pixels per 32-bit
// word: the even and the odd bytes are spread into 16-bit lanes, where every
// product and sum stays below 2^16.
__device__ __forceinline__ uint32_t grey4_u8(uint32_t r, uint32_t g, uint32_t b) {
  const uint32_t m = 0x00ff00ffu;
  const uint32_t ev = ((r & m) * 77u + (g & m) * 151u + (b & m) * 28u + 0x00800080u) >> 8;
  const uint32_t od = (((r >> 8) & m) * 77u + ((g >> 8) & m) * 151u + ((b >> 8) & m) * 28u + 0x00800080u) >> 8;
  return (ev & m) | ((od & m) << 8);
}
__device__ __forceinline__ uint8_t grey1_u8(uint32_t r, uint32_t g, uint32_t b) {
  const uint32_t v = (77u * r + 151u * g + 28u * b + 128u) >> 8;
  return (uint8_t)(v > 255u ? 255u : v);
}
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void rgb2grey_u8_k(size_t n, const uint8_t* __restrict__ r,
                                                        const uint8_t* __restrict__ g,
                                                        const uint8_t* __restrict__ b,
                                                        uint8_t* __restrict__ gr, int vec) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n16 = vec ? n / 16 : 0;
  if (t < n16) {
    const v4u rv = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(r) + t);
    const v4u gv = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(g) + t);
    const v4u bv = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(b) + t);
    v4u o;
#pragma unroll
    for (int l = 0; l < 4; ++l) o[l] = grey4_u8(rv[l], gv[l], bv[l]);
    __builtin_nontemporal_store(o, reinterpret_cast<v4u*>(gr) + t);
  } else {
    const size_t i = n16 * 16 + (t - n16);
    if (i < n) gr[i] = grey1_u8(r[i], g[i], b[i]);
  }
}

// linspace(l, u, n): u at the last index, p*step + l elsewhere
// (modules/core/generative/include/nt2/core/functions/details/linspace.hpp:33-45)
template <class P, class T>
__global__ __launch_bounds__(kBlock) void linspace_k(size_t n, T l, T step, T u, T* __restrict__ out) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = (p == n - 1) ? u : dm::fma<P>((T)p, step, l);
}

// mandelbrot: C = arrayfun(step, A, B), demo/mandelbrot/nt2/mandelbrot_nt2.cpp:35-70.
// One pixel per thread.  NT2 iterates a pack of 8 pixels until none of its
// lanes is inside the radius-2 disc (or max_iter iterations); here the 8
// lanes of a pack are 8 adjacent threads of a wave and the pack's decision
// is their bits of a wave ballot.  Pixels of the scalar tail (past the last
// whole pack) decide alone.  Every lane of a pack takes the same decision at
// the same iteration, so each ballot is executed by all eight.
template <class P>
__global__ __launch_bounds__(kBlock) void mandelbrot_k(size_t n, const float* __restrict__ a,
                                                       const float* __restrict__ b, int32_t* __restrict__ c,
                                                       int max_iter) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t npacked = n & ~(size_t)7;
  const unsigned lane = threadIdx.x & 63u;
  const uint64_t group = t < npacked ? (0xffull << (lane & 56u)) : (1ull << lane);
  const bool live = t < n;
  const float ca = live ? a[t] : 0.0f;
  const float cb = live ? b[t] : 0.0f;
  float x = 0.0f, y = 0.0f;
  int32_t it = 0;
  int i = 0;
  bool go = live;
  while (go) {
    const float y2 = y * y;
    const float m2 = dm::fma<P>(x, x, y2);             // |z|^2 before the update
    const float xn = dm::fma<P>(x, x, -y2) + ca;       // (x^2 - y^2) + a
    y = dm::fma<P>(2.0f * x, y, cb);                   // (2x) y + b
    x = xn;
    const bool in = m2 < 4.0f;
    it += in ? 1 : 0;
    ++i;
    const uint64_t inside = __ballot(in);
    go = (inside & group) != 0 && i < max_iter;
  }
  if (live) c[t] = it;
}

template <class P>
static int launch_rgb2grey(size_t n, const float* r, const float* g, const float* b, float* gr, hipStream_t s) {
  const int vec = aligned16(r) && aligned16(g) && aligned16(b) && aligned16(gr);
  const size_t items = vec ? n / 4 + n % 4 : n;
  hipLaunchKernelGGL(rgb2grey_k<P>, dim3(grid_for(items)), dim3(kBlock), 0, s, n, r, g, b, gr, vec);
  NT2H_CHECK_LAUNCH("rgb2grey_f32");
  return NT2H_OK;
}

template <class P>
static int launch_rgb2yuv(size_t n, const float* r, const float* g, const float* b, float* y, float* u, float* v,
                          hipStream_t s) {
  const int vec = aligned16(r) && aligned16(g) && aligned16(b) && aligned16(y) && aligned16(u) && aligned16(v);
  const size_t items = vec ? n / 4 + n % 4 : n;
  hipLaunchKernelGGL(rgb2yuv_k<P>, dim3(grid_for(items)), dim3(kBlock), 0, s, n, r, g, b, y, u, v, vec);
  NT2H_CHECK_LAUNCH("rgb2yuv_f32");
  return NT2H_OK;
}

template <class T>
static int linspace_t(const char* name, size_t n, T lo, T hi, T* out, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!out) return fail(NT2H_EINVAL, name);
  const T step = (hi - lo) / (T)(n - 1);
  const dim3 g(grid_for(n)), bl(kBlock);
  if (policy == NT2H_EXACT)
    hipLaunchKernelGGL((linspace_k<exact_policy, T>), g, bl, 0, as_stream(stream), n, lo, step, hi, out);
  else if (policy == NT2H_FAST)
    hipLaunchKernelGGL((linspace_k<fast_policy, T>), g, bl, 0, as_stream(stream), n, lo, step, hi, out);
  else
    return fail(NT2H_EINVAL, name);
  NT2H_CHECK_LAUNCH(name);
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_rgb2grey_f32(size_t n, const float* r, const float* g, const float* b, float* grey, int policy,
                      void* stream) {
  if (n == 0) return NT2H_OK;
  if (!r || !g || !b || !grey) return fail(NT2H_EINVAL, "nt2h_rgb2grey_f32: null pointer");
  if (policy == NT2H_EXACT) return launch_rgb2grey<exact_policy>(n, r, g, b, grey, as_stream(stream));
  if (policy == NT2H_FAST) return launch_rgb2grey<fast_policy>(n, r, g, b, grey, as_stream(stream));
  return fail(NT2H_EINVAL, "nt2h_rgb2grey_f32: bad policy");
}

int nt2h_rgb2grey_u8(size_t n, const uint8_t* r, const uint8_t* g, const uint8_t* b, uint8_t* grey,
                     void* stream) {
  if (n == 0) return NT2H_OK;
  if (!r || !g || !b || !grey) return fail(NT2H_EINVAL, "nt2h_rgb2grey_u8: null pointer");
  const int vec = aligned16(r) && aligned16(g) && aligned16(b) && aligned16(grey);
  const size_t items = vec ? n / 16 + n % 16 : n;
  hipLaunchKernelGGL(rgb2grey_u8_k, dim3(grid_for(items)), dim3(kBlock), 0, as_stream(stream), n, r, g, b, grey, vec);
  NT2H_CHECK_LAUNCH("rgb2grey_u8");
  return NT2H_OK;
}

int nt2h_rgb2yuv_f32(size_t n, const float* r, const float* g, const float* b, float* y, float* u, float* v,
                     int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!r || !g || !b || !y || !u || !v) return fail(NT2H_EINVAL, "nt2h_rgb2yuv_f32: null pointer");
  if (policy == NT2H_EXACT) return launch_rgb2yuv<exact_policy>(n, r, g, b, y, u, v, as_stream(stream));
  if (policy == NT2H_FAST) return launch_rgb2yuv<fast_policy>(n, r, g, b, y, u, v, as_stream(stream));
  return fail(NT2H_EINVAL, "nt2h_rgb2yuv_f32: bad policy");
}

int nt2h_linspace_f32(size_t n, float lo, float hi, float* out, int policy, void* stream) {
  return linspace_t<float>("nt2h_linspace_f32", n, lo, hi, out, policy, stream);
}

int nt2h_linspace_f64(size_t n, double lo, double hi, double* out, int policy, void* stream) {
  return linspace_t<double>("nt2h_linspace_f64", n, lo, hi, out, policy, stream);
}

int nt2h_mandelbrot_f32(size_t n, const float* a, const float* b, int32_t* out, int max_iter, int policy,
                        void* stream) {
  if (n == 0) return NT2H_OK;
  if (!a || !b || !out) return fail(NT2H_EINVAL, "nt2h_mandelbrot_f32: null pointer");
  if (max_iter < 1) return fail(NT2H_EINVAL, "nt2h_mandelbrot_f32: max_iter must be at least 1");
  const dim3 g(grid_for(n)), bl(kBlock);
  if (policy == NT2H_EXACT)
    hipLaunchKernelGGL(mandelbrot_k<exact_policy>, g, bl, 0, as_stream(stream), n, a, b, out, max_iter);
  else if (policy == NT2H_FAST)
    hipLaunchKernelGGL(mandelbrot_k<fast_policy>, g, bl, 0, as_stream(stream), n, a, b, out, max_iter);
  else
    return fail(NT2H_EINVAL, "nt2h_mandelbrot_f32: bad policy");
  NT2H_CHECK_LAUNCH("mandelbrot_f32");
  return NT2H_OK;
}

}  // extern "C"
