// libnt2h core: the HIP architecture site's kernels for NT2's transform_ and
// inner_fold_ over tables, the fused Black-Scholes and sigma-delta workloads,
// the seeded synthetic-table generators and the runtime services of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>
#include <mutex>
#include <string>

#include "bs_device.hpp"
#include "common.hpp"
#include "map_packs.hpp"
#include "nt2/hip/math_compose.hpp"
#include "nt2/hip/bs_accurate.hpp"
#include "nt2/hip/bs_accurate64.hpp"
#include "nt2/hip/math.hpp"
#include "nt2h.h"

#pragma clang fp contract(off)

using nt2::hip::exact_policy;
using nt2::hip::fast_policy;
namespace dm = nt2::hip::dm;

namespace nt2h {

// ============================================================================
// errors
// ============================================================================
static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const char* what) {
  set_error(what);
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? NT2H_ENOMEM : NT2H_EHIP;
}

// ============================================================================
// reduction scratch, one per (device, stream).  A buffer that has to grow is
// replaced by one at least twice its size and the old one is kept until the
// process ends (not freed): work already queued on the stream, or a graph
// captured from it, may still hold its address.  Growth is rare and the
// retained total stays below the largest buffer; no stream synchronisation.
// ============================================================================
static std::mutex g_scratch_mu;
static std::map<std::pair<int, void*>, Scratch> g_scratch;
static std::vector<double*> g_scratch_retired;

int get_scratch(size_t nblocks, Scratch* out, hipStream_t s) {
  int dev = 0;
  NT2H_CHECK(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  Scratch& sc = g_scratch[{dev, (void*)s}];
  if (sc.cap < nblocks) {
    size_t cap = nblocks < kMaxFoldBlocks ? kMaxFoldBlocks : nblocks;
    if (sc.partial && cap < 2 * sc.cap) cap = 2 * sc.cap;
    double* p = nullptr;
    NT2H_CHECK(hipMalloc(&p, cap * sizeof(double)), "scratch alloc");
    if (sc.partial) g_scratch_retired.push_back(sc.partial);
    sc.partial = p;
    sc.cap = cap;
  }
  *out = sc;
  return NT2H_OK;
}

// ============================================================================
// generators: SplitMix64 stream out_i = mix(seed + (i+1)*golden)
// ============================================================================
__host__ __device__ inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline double unit_at(uint64_t seed, uint64_t i) {
  return (double)(splitmix64_at(seed, i) >> 11) * 0x1.0p-53;
}

template <class T>
__global__ void fill_uniform_k(T* out, size_t n, double lo, double hi, uint64_t seed) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (T)(lo + (hi - lo) * unit_at(seed, i));
}

__global__ void fill_normal_k(float* out, size_t n, uint64_t seed) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t p = i >> 1;
    const double u1 = unit_at(seed, 2 * p), u2 = unit_at(seed, 2 * p + 1);
    const double r = sqrt(-2.0 * log(1.0 - u1));
    const double a = 6.283185307179586 * u2;
    out[i] = (float)((i & 1) ? r * sin(a) : r * cos(a));
  }
}

struct Rect {
  int x0, y0, w, h, vx, vy, val;
};
struct Rects {
  Rect r[8];
  int n;
};

__global__ void fill_frames_k(uint8_t* frames, int h, int w, int nframes, Rects rects, int noise,
                              uint64_t seed) {
  const size_t npix = (size_t)h * w;
  const size_t total = npix * (size_t)nframes;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const int k = (int)(idx / npix);
    const size_t p = idx - (size_t)k * npix;
    const int i = (int)(p % h), j = (int)(p / h);  // column-major (row i, column j)
    int v = 16 + (int)(splitmix64_at(seed, p) % 185u);
    for (int q = 0; q < rects.n; ++q) {
      const Rect& R = rects.r[q];
      const int cx = (((R.x0 + R.vx * k) % w) + w) % w;
      const int cy = (((R.y0 + R.vy * k) % h) + h) % h;
      const int dx = ((j - cx) % w + w) % w, dy = ((i - cy) % h + h) % h;
      if (dx < R.w && dy < R.h) v = R.val > v ? R.val : v;
    }
    if (noise > 0) {
      const int nz = (int)(splitmix64_at(seed ^ 0xA5A5A5A5A5A5A5A5ull, idx) % (uint64_t)(2 * noise + 1)) - noise;
      v += nz;
    }
    frames[idx] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
  }
}

// ============================================================================
// transform_: elementwise maps with NT2 pack semantics.  Work item t < npacks
// is the pack [W*t, W*t+W); items beyond are the scalar tail elements.
// ============================================================================
// the radian trigonometric ops, whose range reduction is chosen per pack
template <int OP> constexpr bool radian_trig() {
  return OP == NT2H_OP_COS || OP == NT2H_OP_SIN || OP == NT2H_OP_TAN || OP == NT2H_OP_COT || OP == NT2H_OP_SEC ||
         OP == NT2H_OP_CSC;
}
template <int OP> constexpr int radian_kind() {
  return OP == NT2H_OP_COS ? dm::k_cos
         : OP == NT2H_OP_SIN ? dm::k_sin
         : OP == NT2H_OP_TAN ? dm::k_tan
         : OP == NT2H_OP_COT ? dm::k_cot
         : OP == NT2H_OP_SEC ? dm::k_sec
                             : dm::k_csc;
}

template <class P, int OP, int G>
__device__ __forceinline__ void apply_f32(const float* x, float* y) {
  constexpr bool scalar = (G == 1);
  if constexpr (radian_trig<OP>()) {
    dm::trig_group_f<P, G>(radian_kind<OP>(), x, y, nullptr);
  } else {
#pragma unroll
    for (int l = 0; l < G; ++l) {
      const float a = x[l];
      float r;
      if constexpr (OP == NT2H_OP_EXP) r = dm::exp_f<P>(a);
      else if constexpr (OP == NT2H_OP_LOG) r = dm::log_f<P>(a);
      else if constexpr (OP == NT2H_OP_SQRT) r = __builtin_sqrtf(a);
      else if constexpr (OP == NT2H_OP_ERFC) r = dm::erfc_f<P>(a);
      else if constexpr (OP == NT2H_OP_NORMCDF) r = dm::normcdf_f<P>(a);
      else if constexpr (OP == NT2H_OP_FASTNORMCDF) r = dm::fastnormcdf_f<P>(a);
      else if constexpr (OP == NT2H_OP_EXP2) r = dm::exp2_f<P>(a);
      else if constexpr (OP == NT2H_OP_EXP10) r = dm::exp10_f<P>(a);
      else if constexpr (OP == NT2H_OP_LOG2) r = dm::log2_f<P>(a);
      else if constexpr (OP == NT2H_OP_LOG10) r = dm::log10_f<P>(a);
      else if constexpr (OP == NT2H_OP_FLOOR) r = __builtin_floorf(a);
      else if constexpr (OP == NT2H_OP_CEIL) r = __builtin_ceilf(a);
      else if constexpr (OP == NT2H_OP_TRUNC) r = __builtin_truncf(a);
      else if constexpr (OP == NT2H_OP_ROUND) r = dm::round_nt2(a);
      else if constexpr (OP == NT2H_OP_ROUND2EVEN) r = __builtin_rintf(a);
      else if constexpr (OP == NT2H_OP_SIGN) r = dm::sign_nt2(a);
      else if constexpr (OP == NT2H_OP_EXPM1) r = dm::expm1_f<P>(a);
      else if constexpr (OP == NT2H_OP_LOG1P) r = dm::log1p_f<P>(a);
      else if constexpr (OP == NT2H_OP_TANH) r = dm::tanh_f<P>(a);
      else if constexpr (OP == NT2H_OP_SINH) r = dm::sinh_f<P>(a);
      else if constexpr (OP == NT2H_OP_COSH) r = dm::cosh_f<P>(a);
      else if constexpr (OP == NT2H_OP_ATAN) r = dm::atan_f<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ASIN) r = dm::asin_f<P>(a);
      else if constexpr (OP == NT2H_OP_ACOS) r = dm::acos_f<P>(a);
      else if constexpr (OP == NT2H_OP_ERF) r = dm::erf_f<P>(a);
      else if constexpr (OP == NT2H_OP_CBRT) r = dm::cbrt_f<P>(a);
      else if constexpr (OP == NT2H_OP_ASINH) r = dm::asinh_f<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ACOSH) r = dm::acosh_f<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ATANH) r = dm::atanh_f<P>(a);
      else if constexpr (OP == NT2H_OP_SINPI) r = dm::trigu_f<P, dm::u_pi>(dm::k_sin, a);
      else if constexpr (OP == NT2H_OP_COSPI) r = dm::trigu_f<P, dm::u_pi>(dm::k_cos, a);
      else if constexpr (OP == NT2H_OP_TANPI) r = dm::trigu_f<P, dm::u_pi>(dm::k_tan, a);
      else if constexpr (OP == NT2H_OP_COTPI) r = dm::trigu_f<P, dm::u_pi>(dm::k_cot, a);
      else if constexpr (OP == NT2H_OP_SIND) r = dm::trigu_f<P, dm::u_deg>(dm::k_sin, a);
      else if constexpr (OP == NT2H_OP_COSD) r = dm::trigu_f<P, dm::u_deg>(dm::k_cos, a);
      else if constexpr (OP == NT2H_OP_TAND) r = dm::trigu_f<P, dm::u_deg>(dm::k_tan, a);
      else if constexpr (OP == NT2H_OP_COTD) r = dm::trigu_f<P, dm::u_deg>(dm::k_cot, a);
      else r = __builtin_fabsf(a);
      y[l] = r;
    }
  }
}

template <class P, int OP, int G>
__device__ __forceinline__ void apply_f64(const double* x, double* y) {
  constexpr bool scalar = (G == 1);
  if constexpr (radian_trig<OP>()) {
    dm::trig_group_d<P, G>(radian_kind<OP>(), x, y, nullptr);
  } else {
#pragma unroll
    for (int l = 0; l < G; ++l) {
      const double a = x[l];
      double r;
      if constexpr (OP == NT2H_OP_EXP) r = dm::exp_d<P>(a);
      else if constexpr (OP == NT2H_OP_LOG) r = dm::log_d<P>(a);
      else if constexpr (OP == NT2H_OP_SQRT) r = __builtin_sqrt(a);
      else if constexpr (OP == NT2H_OP_ERFC) r = dm::erfc_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_NORMCDF) r = dm::normcdf_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_EXP2) r = dm::exp2_d<P>(a);
      else if constexpr (OP == NT2H_OP_EXP10) r = dm::exp10_d<P>(a);
      else if constexpr (OP == NT2H_OP_LOG2) r = dm::log2_d<P>(a);
      else if constexpr (OP == NT2H_OP_LOG10) r = dm::log10_d<P>(a);
      else if constexpr (OP == NT2H_OP_FLOOR) r = __builtin_floor(a);
      else if constexpr (OP == NT2H_OP_CEIL) r = __builtin_ceil(a);
      else if constexpr (OP == NT2H_OP_TRUNC) r = __builtin_trunc(a);
      else if constexpr (OP == NT2H_OP_ROUND) r = dm::round_nt2(a);
      else if constexpr (OP == NT2H_OP_ROUND2EVEN) r = __builtin_rint(a);
      else if constexpr (OP == NT2H_OP_SIGN) r = dm::sign_nt2(a);
      else if constexpr (OP == NT2H_OP_EXPM1) r = dm::expm1_d<P>(a);
      else if constexpr (OP == NT2H_OP_LOG1P) r = dm::log1p_d<P>(a);
      else if constexpr (OP == NT2H_OP_TANH) r = dm::tanh_d<P>(a);
      else if constexpr (OP == NT2H_OP_SINH) r = dm::sinh_d<P>(a);
      else if constexpr (OP == NT2H_OP_COSH) r = dm::cosh_d<P>(a);
      else if constexpr (OP == NT2H_OP_ATAN) r = dm::atan_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ASIN) r = dm::asin_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ACOS) r = dm::acos_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ERF) r = dm::erf_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_CBRT) r = dm::cbrt_d<P>(a);
      else if constexpr (OP == NT2H_OP_ASINH) r = dm::asinh_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ACOSH) r = dm::acosh_d<P, scalar>(a);
      else if constexpr (OP == NT2H_OP_ATANH) r = dm::atanh_d<P>(a);
      else if constexpr (OP == NT2H_OP_SINPI) r = dm::trigu_d<P, dm::u_pi>(dm::k_sin, a);
      else if constexpr (OP == NT2H_OP_COSPI) r = dm::trigu_d<P, dm::u_pi>(dm::k_cos, a);
      else if constexpr (OP == NT2H_OP_TANPI) r = dm::trigu_d<P, dm::u_pi>(dm::k_tan, a);
      else if constexpr (OP == NT2H_OP_COTPI) r = dm::trigu_d<P, dm::u_pi>(dm::k_cot, a);
      else if constexpr (OP == NT2H_OP_SIND) r = dm::trigu_d<P, dm::u_deg>(dm::k_sin, a);
      else if constexpr (OP == NT2H_OP_COSD) r = dm::trigu_d<P, dm::u_deg>(dm::k_cos, a);
      else if constexpr (OP == NT2H_OP_TAND) r = dm::trigu_d<P, dm::u_deg>(dm::k_tan, a);
      else if constexpr (OP == NT2H_OP_COTD) r = dm::trigu_d<P, dm::u_deg>(dm::k_cot, a);
      else r = __builtin_fabs(a);
      y[l] = r;
    }
  }
}

template <class P, int OP> struct apply32 {
  template <int G> __device__ __forceinline__ static void run(const float* a, float* r) { apply_f32<P, OP, G>(a, r); }
};
template <class P, int OP> struct apply64 {
  template <int G> __device__ __forceinline__ static void run(const double* a, double* r) { apply_f64<P, OP, G>(a, r); }
};

template <class P, int OP>
__global__ __launch_bounds__(kBlock) void map_f32_k(const float* __restrict__ x,
                                                    float* __restrict__ y, size_t n, int vec) {
  map_packs<apply32<P, OP>, float, 8>(x, y, n, vec);
}

template <class P, int OP>
__global__ __launch_bounds__(kBlock) void map_f64_k(const double* __restrict__ x,
                                                    double* __restrict__ y, size_t n, int vec) {
  map_packs<apply64<P, OP>, double, 4>(x, y, n, vec);
}

// binary maps: out = f(x, y), packs of W from index 0 (f(pack) form), then
// the scalar tail (NT2's scalar form)
template <class P, int OP, class T, int G>
__device__ __forceinline__ T apply2_one(T a, T b) {
  constexpr bool scalar = (G == 1);
  if constexpr (OP == NT2H_OP2_POW) return dm::pow_t<P, scalar, T>(a, b);
  else if constexpr (OP == NT2H_OP2_ATAN2D) return dm::atan2d_t<P, scalar>(a, b);
  else if constexpr (OP == NT2H_OP2_LOGSPACE_ADD) return dm::logspace_add_t<P>(a, b);
  else if constexpr (OP == NT2H_OP2_LOGSPACE_SUB) return dm::logspace_sub_t<P>(a, b);
  else if constexpr (sizeof(T) == 4) return dm::atan2_f<P, scalar>(a, b);
  else return dm::atan2_d<P, scalar>(a, b);
}
template <class P, int OP, class T, int W>
__global__ __launch_bounds__(kBlock) void map2_k(const T* __restrict__ x, const T* __restrict__ y,
                                                 T* __restrict__ out, size_t n) {
  const size_t npacks = n / W;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npacks) {
    const size_t i = t * W;
    T r[W];
#pragma unroll
    for (int l = 0; l < W; ++l) r[l] = apply2_one<P, OP, T, W>(x[i + l], y[i + l]);
#pragma unroll
    for (int l = 0; l < W; ++l) out[i + l] = r[l];
  } else if (t < npacks + (n - npacks * W)) {
    const size_t i = npacks * W + (t - npacks);
    out[i] = apply2_one<P, OP, T, 1>(x[i], y[i]);
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void sincos_f32_k(const float* __restrict__ x,
                                                       float* __restrict__ s,
                                                       float* __restrict__ c, size_t n) {
  const size_t npacks = n / 8;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npacks) {
    float a[8], rs[8], rc[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) a[l] = x[t * 8 + l];
    dm::trig_group_f<P, 8>(dm::k_sincos, a, rs, rc);
#pragma unroll
    for (int l = 0; l < 8; ++l) { s[t * 8 + l] = rs[l]; c[t * 8 + l] = rc[l]; }
  } else if (t < npacks + (n - npacks * 8)) {
    const size_t i = npacks * 8 + (t - npacks);
    float a = x[i], rs, rc;
    dm::trig_group_f<P, 1>(dm::k_sincos, &a, &rs, &rc);
    s[i] = rs;
    c[i] = rc;
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void sincos_f64_k(const double* __restrict__ x,
                                                       double* __restrict__ s,
                                                       double* __restrict__ c, size_t n) {
  const size_t npacks = n / 4;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npacks) {
    double a[4], rs[4], rc[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) a[l] = x[t * 4 + l];
    dm::trig_group_d<P, 4>(dm::k_sincos, a, rs, rc);
#pragma unroll
    for (int l = 0; l < 4; ++l) { s[t * 4 + l] = rs[l]; c[t * 4 + l] = rc[l]; }
  } else if (t < npacks + (n - npacks * 4)) {
    const size_t i = npacks * 4 + (t - npacks);
    double a = x[i], rs, rc;
    dm::trig_group_d<P, 1>(dm::k_sincos, &a, &rs, &rc);
    s[i] = rs;
    c[i] = rc;
  }
}

// Y = a*X + Y, one float4 per work item, scalar tail
__global__ __launch_bounds__(kBlock) void saxpy_k(size_t n, float a, const float* __restrict__ x,
                                                  float* __restrict__ y, int vec) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n4 = vec ? n / 4 : 0;
  typedef float f4 __attribute__((ext_vector_type(4)));
  if (t < n4) {  // streamed once: nontemporal 16-byte loads and stores
    const f4 xv = __builtin_nontemporal_load(reinterpret_cast<const f4*>(x) + t);
    f4 yv = __builtin_nontemporal_load(reinterpret_cast<const f4*>(y) + t);
#pragma unroll
    for (int l = 0; l < 4; ++l) yv[l] = a * xv[l] + yv[l];
    __builtin_nontemporal_store(yv, reinterpret_cast<f4*>(y) + t);
  } else {
    const size_t i = n4 * 4 + (t - n4);
    if (i < n) y[i] = a * x[i] + y[i];
  }
}

// ============================================================================
// HBM reference kernel: y = x (16-byte vectors, grid-stride, streaming
// stores).  Used to measure the achievable HBM bandwidth on the box.
// ============================================================================
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void stream_copy_k(const v4f* __restrict__ x,
                                                        v4f* __restrict__ y, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const v4f v = __builtin_nontemporal_load(x + i);
    __builtin_nontemporal_store(v, y + i);
  }
}

// ============================================================================
// inner_fold_: sum(x) and sum(x.*y), compensated FP32 per lane
// ============================================================================
// Block b folds the contiguous range [b*per, min((b+1)*per, units)) of units
// (float4s when every table is 16-byte aligned, else floats); each thread
// strides through it by the block size with kFoldUnroll units of every table
// loaded ahead of the dependent TwoSum chain, and the last block adds the
// n % 4 trailing elements of the float4 form.  The block's compensated sum is
// stored as partial[b]; fold_final_k adds the partials in a fixed order.
constexpr int kFoldUnroll = 4;
constexpr unsigned kFoldBlocks = 256;  // partials of one fold (at most)
typedef float f32x4v __attribute__((ext_vector_type(4)));

template <bool Dot, bool Vec>
__global__ __launch_bounds__(kBlock) void fold_part_k(const float* __restrict__ x, const float* __restrict__ y,
                                                      size_t n, size_t per, double* __restrict__ partial) {
  float s = 0.0f, c = 0.0f;
  constexpr int U = Vec ? kFoldUnroll : 4 * kFoldUnroll;
  const size_t units = Vec ? n / 4 : n;
  const size_t lo = (size_t)blockIdx.x * per;
  const size_t hi = lo + per < units ? lo + per : units;
  size_t i = lo + threadIdx.x;
  if constexpr (Vec) {
    const f32x4v* x4 = reinterpret_cast<const f32x4v*>(x);
    const f32x4v* y4 = reinterpret_cast<const f32x4v*>(y);
    for (; i + (size_t)(U - 1) * kBlock < hi; i += (size_t)U * kBlock) {
      f32x4v a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        a[u] = __builtin_nontemporal_load(x4 + i + u * kBlock);
        if (Dot) b[u] = __builtin_nontemporal_load(y4 + i + u * kBlock);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int l = 0; l < 4; ++l) comp_add(s, c, Dot ? a[u][l] * b[u][l] : a[u][l]);
    }
    for (; i < hi; i += kBlock) {
      const f32x4v a = x4[i];
      if (Dot) {
        const f32x4v b = y4[i];
#pragma unroll
        for (int l = 0; l < 4; ++l) comp_add(s, c, a[l] * b[l]);
      } else {
#pragma unroll
        for (int l = 0; l < 4; ++l) comp_add(s, c, a[l]);
      }
    }
    if (blockIdx.x == gridDim.x - 1) {
      const size_t t = units * 4 + threadIdx.x;
      if (t < n) comp_add(s, c, Dot ? x[t] * y[t] : x[t]);
    }
  } else {
    for (; i + (size_t)(U - 1) * kBlock < hi; i += (size_t)U * kBlock) {
      float a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        a[u] = __builtin_nontemporal_load(x + i + u * kBlock);
        if (Dot) b[u] = __builtin_nontemporal_load(y + i + u * kBlock);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) comp_add(s, c, Dot ? a[u] * b[u] : a[u]);
    }
    for (; i < hi; i += kBlock) comp_add(s, c, Dot ? x[i] * y[i] : x[i]);
  }
  const double b = block_sum_comp(s, c);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// ============================================================================
// fused Black-Scholes
// ============================================================================
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(f32x4* p, float4 v) {
  __builtin_nontemporal_store(f32x4{v.x, v.y, v.z, v.w}, p);
}
__device__ __forceinline__ void store_nt(f64x2* p, f64x2 v) { __builtin_nontemporal_store(v, p); }
// accurate policy: grid-stride over float4s of options (the LDS tables are
// loaded once per workgroup), FP64 inside.  I is the float4 index type: 32-bit
// when the tables hold fewer than 2^33 options, which lets the loads and
// stores address from a scalar base plus one 32-bit offset.
template <class I>
__global__ __launch_bounds__(kBlock) void bs_acc_f32_k(I n4, size_t n, const float* __restrict__ S,
                                                       const float* __restrict__ X,
                                                       const float* __restrict__ T,
                                                       nt2::hip::bsacc::Params prm,
                                                       float* __restrict__ call,
                                                       float* __restrict__ put) {
  namespace ba = nt2::hip::bsacc;
  __shared__ double tab[(ba::kTailDeg + 1) * ba::kTailN];
  __shared__ double ex2[32];
  ba::load_tables(tab, ex2);
  const I stride = (I)gridDim.x * blockDim.x;
  for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += stride) {
    const f32x4 s4 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(S) + t);
    const f32x4 x4 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(X) + t);
    const f32x4 t4 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(T) + t);
    float c[4], p[4];
    ba::bs_accurate4(s4, x4, t4, prm, tab, ex2, c, p);
    __builtin_nontemporal_store(f32x4{c[0], c[1], c[2], c[3]}, reinterpret_cast<f32x4*>(call) + t);
    __builtin_nontemporal_store(f32x4{p[0], p[1], p[2], p[3]}, reinterpret_cast<f32x4*>(put) + t);
  }
  const size_t sstride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += sstride)
    ba::bs_accurate(S[i], X[i], T[i], prm, tab, ex2, call[i], put[i]);
}

// accurate f64 policy: grid-stride over options, one per thread and
// iteration (a wave reads 512 contiguous bytes of each table; two options per
// thread made the compiler interleave them at 242 VGPRs, 2 waves per SIMD),
// tables in LDS
template <class I>
__global__ __launch_bounds__(kBlock) void bs_acc_f64_k(I n, const double* __restrict__ S,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ T,
                                                       nt2::hip::bsacc64::Params prm,
                                                       double* __restrict__ call,
                                                       double* __restrict__ put) {
  namespace bb = nt2::hip::bsacc64;
  __shared__ double tab[bb::kTabDoubles];
  bb::load_tables(tab);
  const double* gt = tab;
  const double* ex2h = tab + bb::kGRows * bb::kGN;
  const double* ex2l = ex2h + 32;
  const I stride = (I)gridDim.x * blockDim.x;
  for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double c, p;
    bb::bs_price(__builtin_nontemporal_load(S + i), __builtin_nontemporal_load(X + i),
                 __builtin_nontemporal_load(T + i), prm, gt, ex2h, ex2l, c, p);
    __builtin_nontemporal_store(c, call + i);
    __builtin_nontemporal_store(p, put + i);
  }
}

template <class P, int MODE>
__global__ __launch_bounds__(kBlock) void bs_f32_k(size_t n, const float* __restrict__ S,
                                                   const float* __restrict__ X,
                                                   const float* __restrict__ T, float r, float v,
                                                   float* __restrict__ call,
                                                   float* __restrict__ put, int vec) {
  _Pragma("clang fp contract(off)")
  const float drift = (v * v) * 0.5f + r;
  const float mr = -r;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n4 = vec ? n / 4 : 0;
  if (t < n4) {
    const float4 s4 = reinterpret_cast<const float4*>(S)[t];
    const float4 x4 = reinterpret_cast<const float4*>(X)[t];
    const float4 t4 = reinterpret_cast<const float4*>(T)[t];
    float4 c4, p4;
    bs_one_f32<P, MODE>(s4.x, x4.x, t4.x, mr, v, drift, c4.x, p4.x);
    bs_one_f32<P, MODE>(s4.y, x4.y, t4.y, mr, v, drift, c4.y, p4.y);
    bs_one_f32<P, MODE>(s4.z, x4.z, t4.z, mr, v, drift, c4.z, p4.z);
    bs_one_f32<P, MODE>(s4.w, x4.w, t4.w, mr, v, drift, c4.w, p4.w);
    store_nt(reinterpret_cast<f32x4*>(call) + t, c4);
    if (MODE == NT2H_BS_TEXTBOOK) store_nt(reinterpret_cast<f32x4*>(put) + t, p4);
  } else {
    const size_t i = n4 * 4 + (t - n4);
    if (i < n) {
      float c, p;
      bs_one_f32<P, MODE>(S[i], X[i], T[i], mr, v, drift, c, p);
      call[i] = c;
      if (MODE == NT2H_BS_TEXTBOOK) put[i] = p;
    }
  }
}

// f64: one NT2 pack (4 options) per work item; the n%4 tail uses NT2's scalar erfc
template <class P>
__global__ __launch_bounds__(kBlock) void bs_f64_k(size_t n, const double* __restrict__ S,
                                                   const double* __restrict__ X,
                                                   const double* __restrict__ T, double r,
                                                   double v, double* __restrict__ call,
                                                   double* __restrict__ put, int vec) {
  _Pragma("clang fp contract(off)")
  const double drift = (v * v) * 0.5 + r;
  const double mr = -r;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t npacks = n / 4;
  if (t < npacks) {
    const size_t i = t * 4;
    double s[4], x[4], tt[4], c[4], p[4];
    if (vec) {
      const double2 a0 = reinterpret_cast<const double2*>(S + i)[0], a1 = reinterpret_cast<const double2*>(S + i)[1];
      const double2 b0 = reinterpret_cast<const double2*>(X + i)[0], b1 = reinterpret_cast<const double2*>(X + i)[1];
      const double2 c0 = reinterpret_cast<const double2*>(T + i)[0], c1 = reinterpret_cast<const double2*>(T + i)[1];
      s[0] = a0.x; s[1] = a0.y; s[2] = a1.x; s[3] = a1.y;
      x[0] = b0.x; x[1] = b0.y; x[2] = b1.x; x[3] = b1.y;
      tt[0] = c0.x; tt[1] = c0.y; tt[2] = c1.x; tt[3] = c1.y;
    } else {
#pragma unroll
      for (int l = 0; l < 4; ++l) { s[l] = S[i + l]; x[l] = X[i + l]; tt[l] = T[i + l]; }
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) bs_one_f64<P, false>(s[l], x[l], tt[l], mr, v, drift, c[l], p[l]);
    if (vec) {
      store_nt(reinterpret_cast<f64x2*>(call + i), f64x2{c[0], c[1]});
      store_nt(reinterpret_cast<f64x2*>(call + i) + 1, f64x2{c[2], c[3]});
      store_nt(reinterpret_cast<f64x2*>(put + i), f64x2{p[0], p[1]});
      store_nt(reinterpret_cast<f64x2*>(put + i) + 1, f64x2{p[2], p[3]});
    } else {
#pragma unroll
      for (int l = 0; l < 4; ++l) { call[i + l] = c[l]; put[i + l] = p[l]; }
    }
  } else if (t < npacks + (n - npacks * 4)) {
    const size_t i = npacks * 4 + (t - npacks);
    double c, p;
    bs_one_f64<P, true>(S[i], X[i], T[i], mr, v, drift, c, p);
    call[i] = c;
    put[i] = p;
  }
}

// ============================================================================
// sigma-delta (uint8): one pixel per byte lane; state kept in registers
// across frames.  Frame rule: demo/sigmadelta/nt2/sigmadelta_nt2.cpp:37-58.
// ============================================================================
__device__ __forceinline__ void sd_pixel(uint32_t f, uint32_t& b, uint32_t& v, uint32_t& lab,
                                         int N) {
  b = (b < f) ? b + 1u : ((b > f) ? b - 1u : b);            // selinc / seldec
  const uint32_t d = (b > f ? b : f) - (b < f ? b : f);     // max - min
  uint32_t m = d;
  for (int t = 1; t < N; ++t) m = dm::adds_u8(m, d);        // adds(...adds(d,d)...)
  if (d != 0u) v = (v < m) ? v + 1u : ((v > m) ? v - 1u : v);
  lab = (d < v) ? 0u : 1u;                                  // if_zero_else_one(d < v)
}

__device__ __forceinline__ uint32_t wave_count(uint32_t c) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
  return c;
}

// 16 pixels per thread (one 128-bit load per frame), two pixels per 32-bit
// register as the halves of a packed f16 pair, so that every step is one
// packed op per pixel pair.  Values 0..255 and every intermediate below are
// integers of magnitude < 2048 (exact in f16), except N d - 1, which only
// feeds a maximum whose result is then capped at 254 (above 2048 it rounds to
// a value >= 2048); f16 gives the absolute value (a sign mask) and the
// [0, 1] clamp as free modifiers, and gfx950 has three-input packed
// minimum/maximum.  Per pair, with B = b + 1024 (the f16 whose low byte is b:
// unpacking and packing are byte permutes) and W = v - 1:
//   B = B + clamp(F - B) - clamp(B - F)  selinc(b<f, seldec(b>f, b))
//   d = |F - B|                          max(b,f) - min(b,f)
//   W = min(W + nz, 254, max(W - nz, N d - 1)), nz = min(d, 1)
//                                        v toward m = adds(...adds(d,d)...),
//                                        no change when d == 0
//   l = clamp(d - W) = (d >= v)          if_zero_else_one(d < v)
// With W in [-1, 254] the cap at 254 on W + nz gives what capping m at 255
// gives.  The frame loop keeps 4 frames in flight (the load of frame k + 4 is
// issued as frame k is consumed).  Per-frame label counts: each lane's count
// (at most 16) goes to one byte of a word per 4 frames in LDS; at the end the
// workgroup sums them per frame into partials[k][block], and sd_counts_k
// adds the partials of each frame in block order.
typedef _Float16 sd_h2 __attribute__((ext_vector_type(2)));
constexpr int kSdChunk = 64;   // frames per launch (LDS holds their per-lane counts)
constexpr int kSdAhead = 4;    // frames in flight per thread

__device__ __forceinline__ sd_h2 sd_lo(uint32_t w) {  // bytes 0,1 -> 1024 + byte
  return __builtin_bit_cast(sd_h2, __builtin_amdgcn_perm(0x64646464u, w, 0x04010400u));
}
__device__ __forceinline__ sd_h2 sd_hi(uint32_t w) {  // bytes 2,3 -> 1024 + byte
  return __builtin_bit_cast(sd_h2, __builtin_amdgcn_perm(0x64646464u, w, 0x04030402u));
}
__device__ __forceinline__ uint32_t sd_pack(sd_h2 lo, sd_h2 hi) {  // low bytes of 4 halves (each 1024 + byte)
  return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, hi), __builtin_bit_cast(uint32_t, lo), 0x06040200u);
}
__device__ __forceinline__ sd_h2 sd_clamp01(sd_h2 x) {
  return __builtin_elementwise_min(__builtin_elementwise_max(x, (sd_h2){0, 0}), (sd_h2){1, 1});
}
// one frame step of P pixel pairs, stage by stage over the pairs (each
// stage of one pair is independent of the other pairs, so the packed ops of
// consecutive pairs fill the wait states a dependent chain would need)
template <int P>
__device__ __forceinline__ void sd_pairs(const sd_h2 (&F)[P], sd_h2 (&B)[P], sd_h2 (&W)[P], sd_h2 (&L)[P], sd_h2 n) {
  const sd_h2 one = {1, 1};
  sd_h2 c1[P], c2[P], d[P], nz[P], m1[P];
#pragma unroll
  for (int q = 0; q < P; ++q) c1[q] = sd_clamp01(F[q] - B[q]);
#pragma unroll
  for (int q = 0; q < P; ++q) c2[q] = sd_clamp01(B[q] - F[q]);
#pragma unroll
  for (int q = 0; q < P; ++q) B[q] = (B[q] + c1[q]) - c2[q];
#pragma unroll
  for (int q = 0; q < P; ++q) { const sd_h2 e = F[q] - B[q]; d[q] = __builtin_elementwise_max(e, -e); }
#pragma unroll
  for (int q = 0; q < P; ++q) nz[q] = __builtin_elementwise_min(d[q], one);
#pragma unroll
  for (int q = 0; q < P; ++q) m1[q] = __builtin_elementwise_fma(d[q], n, -one);
#pragma unroll
  for (int q = 0; q < P; ++q)
    W[q] = __builtin_elementwise_minimum(__builtin_elementwise_minimum(W[q] + nz[q], (sd_h2){254, 254}),
                                         __builtin_elementwise_maximum(W[q] - nz[q], m1[q]));
#pragma unroll
  for (int q = 0; q < P; ++q) L[q] = sd_clamp01(d[q] - W[q]);
}

// NW 32-bit words (4 pixels each) per thread: 16 pixels (NW = 4) keeps the
// most work per load, 8 or 4 pixels run more waves per SIMD.
template <int NW> struct sd_vec;
template <> struct sd_vec<4> { typedef uint4 type; static __device__ void words(const uint4& a, uint32_t (&w)[4]) { w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; } static __device__ uint4 make(const uint32_t (&w)[4]) { return make_uint4(w[0], w[1], w[2], w[3]); } };
template <> struct sd_vec<2> { typedef uint2 type; static __device__ void words(const uint2& a, uint32_t (&w)[2]) { w[0] = a.x; w[1] = a.y; } static __device__ uint2 make(const uint32_t (&w)[2]) { return make_uint2(w[0], w[1]); } };
template <> struct sd_vec<1> { typedef uint32_t type; static __device__ void words(const uint32_t& a, uint32_t (&w)[1]) { w[0] = a; } static __device__ uint32_t make(const uint32_t (&w)[1]) { return w[0]; } };

template <int NW>
__global__ __launch_bounds__(kBlock) void sigmadelta_k(const uint8_t* __restrict__ frames,
                                                       size_t stride, int k0, int k1, size_t n16,
                                                       uint8_t* __restrict__ bkg,
                                                       uint8_t* __restrict__ var,
                                                       uint8_t* __restrict__ label, int N,
                                                       uint32_t* __restrict__ partials) {
  typedef typename sd_vec<NW>::type V;
  static_assert(kSdChunk % kSdAhead == 0 && kSdAhead == 4, "one count word per group of frames in flight");
  __shared__ uint32_t wc[kSdChunk / 4][kBlock];
  // XCD-aware tiles: workgroup b runs on XCD b % 8, and each XCD takes a
  // contiguous eighth of the image, so that its L2 sees contiguous spans of
  // every frame (the grid is a multiple of 8)
  const unsigned per = gridDim.x / 8;
  const size_t tile = (size_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  const size_t t = tile * blockDim.x + threadIdx.x;
  const bool active = t < n16;
  const size_t ta = active ? t : 0;
  const _Float16 nn = (_Float16)(N < 255 ? N : 255);
  const sd_h2 n = {nn, nn};
  // frames k0 .. k0+3 in flight (clamped indices: every load in bounds and
  // issued unconditionally)
  V ff[kSdAhead];
#pragma unroll
  for (int j = 0; j < kSdAhead; ++j)
    ff[j] = reinterpret_cast<const V*>(frames + (size_t)(k0 + j < k1 ? k0 + j : k1 - 1) * stride)[ta];
  sd_h2 B[2 * NW], W[2 * NW], L[2 * NW];
  {
    uint32_t bw[NW], vw[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) bw[q] = vw[q] = 0u;
    if (active) {
      sd_vec<NW>::words(reinterpret_cast<const V*>(bkg)[t], bw);
      sd_vec<NW>::words(reinterpret_cast<const V*>(var)[t], vw);
    }
    const sd_h2 off = {1025, 1025};
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      B[2 * q] = sd_lo(bw[q]);
      B[2 * q + 1] = sd_hi(bw[q]);
      W[2 * q] = sd_lo(vw[q]) - off;
      W[2 * q + 1] = sd_hi(vw[q]) - off;
      L[2 * q] = L[2 * q + 1] = (sd_h2){0, 0};
    }
  }
  uint32_t packed = 0u;
  for (int k = k0; k < k1; k += kSdAhead) {
#pragma unroll
    for (int j = 0; j < kSdAhead; ++j) {
      if (k + j >= k1) break;
      uint32_t fw[NW];
      sd_vec<NW>::words(ff[j], fw);
      const int kn = k + j + kSdAhead;
      ff[j] = reinterpret_cast<const V*>(frames + (size_t)(kn < k1 ? kn : k1 - 1) * stride)[ta];
      sd_h2 F[2 * NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) { F[2 * q] = sd_lo(fw[q]); F[2 * q + 1] = sd_hi(fw[q]); }
      sd_pairs<2 * NW>(F, B, W, L, n);
      sd_h2 acc = {1024, 0};  // 1024 + count: its low byte is the count
#pragma unroll
      for (int q = 0; q < 2 * NW; ++q) acc += L[q];
      const uint32_t c = __builtin_bit_cast(unsigned short, (_Float16)(acc.x + acc.y));
      // byte j of the group's word <- the count's low byte
      packed = __builtin_amdgcn_perm(c, packed, j == 0 ? 0x03020104u : j == 1 ? 0x03020400u
                                                : j == 2 ? 0x03040100u : 0x04020100u);
      if (j == kSdAhead - 1) wc[(k + j - k0) >> 2][threadIdx.x] = active ? packed : 0u;
    }
  }
  const int nf = k1 - k0;
  if (nf & 3) wc[(nf - 1) >> 2][threadIdx.x] = active ? packed : 0u;
  if (active) {
    uint32_t bw[NW], vw[NW], lw[NW];
    const sd_h2 off = {1025, 1025}, loff = {1024, 1024};
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      bw[q] = sd_pack(B[2 * q], B[2 * q + 1]);
      vw[q] = sd_pack(W[2 * q] + off, W[2 * q + 1] + off);
      lw[q] = sd_pack(L[2 * q] + loff, L[2 * q + 1] + loff);
    }
    reinterpret_cast<V*>(bkg)[t] = sd_vec<NW>::make(bw);
    reinterpret_cast<V*>(var)[t] = sd_vec<NW>::make(vw);
    reinterpret_cast<V*>(label)[t] = sd_vec<NW>::make(lw);
  }
  if (partials) {
    __syncthreads();
    // 16 threads per group of 4 frames: each adds 16 words (the lanes
    // s, s + 16, ...) with the bytes split over 16-bit fields (at most
    // 16 x 16 per field), then the 16 threads combine (at most 4096)
    const int g = threadIdx.x >> 4, s = threadIdx.x & 15;
    uint32_t e = 0u, o = 0u;
    if (g < (nf + 3) >> 2) {
#pragma unroll
      for (int u = 0; u < kBlock / 16; ++u) {
        const uint32_t w = wc[g][u * 16 + s];
        e += w & 0x00ff00ffu;
        o += (w >> 8) & 0x00ff00ffu;
      }
    }
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) {
      e += __shfl_xor(e, off, 64);
      o += __shfl_xor(o, off, 64);
    }
    if (s == 0) {
      const uint32_t c4[4] = {e & 0xffffu, o & 0xffffu, e >> 16, o >> 16};
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (4 * g + r < nf) partials[(size_t)(4 * g + r) * gridDim.x + blockIdx.x] = c4[r];
    }
  }
}

// counts[k0 + i] += sum over blocks of partials[i][block]: one workgroup per
// frame (integer sums, so the result does not depend on the order)
__global__ __launch_bounds__(kBlock) void sd_counts_k(const uint32_t* __restrict__ partials,
                                                      int nblocks, uint32_t* __restrict__ counts) {
  __shared__ uint32_t ws[kBlock / 64];
  const uint32_t* row = partials + (size_t)blockIdx.x * nblocks;
  uint32_t c = 0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) c += row[b];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += ws[w];
    counts[blockIdx.x] += t;
  }
}

__global__ __launch_bounds__(kBlock) void zero_u32_k(uint32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// one pixel per thread: pixels [p0, npix) (tails and unaligned buffers)
__global__ __launch_bounds__(kBlock) void sigmadelta1_k(const uint8_t* __restrict__ frames,
                                                        size_t stride, int nframes, size_t p0,
                                                        size_t npix, uint8_t* __restrict__ bkg,
                                                        uint8_t* __restrict__ var,
                                                        uint8_t* __restrict__ label, int N,
                                                        uint32_t* __restrict__ counts) {
  const size_t p = p0 + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < npix;
  uint32_t b = 0, v = 0, l = 0;
  if (active) { b = bkg[p]; v = var[p]; l = label[p]; }
  for (int k = 1; k < nframes; ++k) {
    uint32_t cnt = 0;
    if (active) {
      sd_pixel(frames[(size_t)k * stride + p], b, v, l, N);
      cnt = l;
    }
    if (counts) {
      cnt = wave_count(cnt);
      if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&counts[k], cnt);
    }
  }
  if (active) { bkg[p] = (uint8_t)b; var[p] = (uint8_t)v; label[p] = (uint8_t)l; }
}

// ============================================================================
// batched expression z = exp(-x.*x/2).*cos(y) + sqrt(abs(x)) (one table)
// ============================================================================
template <class P, int G>
__device__ __forceinline__ void expr4_group(const float* x, const float* y, float* z) {
  _Pragma("clang fp contract(off)")
  float c[G];
  dm::trig_group_f<P, G>(dm::k_cos, y, c, nullptr);
#pragma unroll
  for (int l = 0; l < G; ++l) {
    const float t = ((-x[l]) * x[l]) / 2.0f;
    z[l] = dm::exp_f<P>(t) * c[l] + __builtin_sqrtf(__builtin_fabsf(x[l]));
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void expr4_k(const float* __restrict__ x,
                                                  const float* __restrict__ y,
                                                  float* __restrict__ z, size_t n) {
  const size_t npacks = n / 8;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npacks) {
    float a[8], b[8], r[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) { a[l] = x[t * 8 + l]; b[l] = y[t * 8 + l]; }
    expr4_group<P, 8>(a, b, r);
#pragma unroll
    for (int l = 0; l < 8; ++l) z[t * 8 + l] = r[l];
  } else if (t < npacks + (n - npacks * 8)) {
    const size_t i = npacks * 8 + (t - npacks);
    float a = x[i], b = y[i], r;
    expr4_group<P, 1>(&a, &b, &r);
    z[i] = r;
  }
}

// per-block compensated partials of sum(z), added by fold_final_k
template <class P>
__global__ __launch_bounds__(kBlock) void expr4_sum_k(const float* __restrict__ x,
                                                      const float* __restrict__ y, size_t n,
                                                      double* __restrict__ partial) {
  float s = 0.0f, c = 0.0f;
  const size_t npacks = n / 8;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < npacks + (n - npacks * 8);
       t += stride) {
    if (t < npacks) {
      float a[8], b[8], r[8];
#pragma unroll
      for (int l = 0; l < 8; ++l) { a[l] = x[t * 8 + l]; b[l] = y[t * 8 + l]; }
      expr4_group<P, 8>(a, b, r);
#pragma unroll
      for (int l = 0; l < 8; ++l) comp_add(s, c, r[l]);
    } else {
      const size_t i = npacks * 8 + (t - npacks);
      float a = x[i], b = y[i], r;
      expr4_group<P, 1>(&a, &b, &r);
      comp_add(s, c, r);
    }
  }
  const double bsum = block_sum_comp(s, c);
  if (threadIdx.x == 0) partial[blockIdx.x] = bsum;
}

// ---- transform_ ---------------------------------------------------------
template <class P>
static int launch_map_f32(int op, size_t n, const float* x, float* y, hipStream_t s) {
  const size_t items = n / 8 + n % 8;
  const int vec = aligned16(x) && aligned16(y);
  const dim3 g(grid_for(items)), b(kBlock);
  switch (op) {
    case NT2H_OP_EXP: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_EXP>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_LOG: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_LOG>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_SQRT: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_SQRT>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_ERFC: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_ERFC>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_NORMCDF: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_NORMCDF>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_FASTNORMCDF: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_FASTNORMCDF>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_COS: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_COS>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_SIN: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_SIN>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_ABS: hipLaunchKernelGGL((map_f32_k<P, NT2H_OP_ABS>), g, b, 0, s, x, y, n, vec); break;
#define NT2H_MAP32(OPC) case OPC: hipLaunchKernelGGL((map_f32_k<P, OPC>), g, b, 0, s, x, y, n, vec); break;
    NT2H_MAP32(NT2H_OP_EXP2) NT2H_MAP32(NT2H_OP_EXP10) NT2H_MAP32(NT2H_OP_LOG2) NT2H_MAP32(NT2H_OP_LOG10)
    NT2H_MAP32(NT2H_OP_FLOOR) NT2H_MAP32(NT2H_OP_CEIL) NT2H_MAP32(NT2H_OP_TRUNC) NT2H_MAP32(NT2H_OP_ROUND)
    NT2H_MAP32(NT2H_OP_ROUND2EVEN) NT2H_MAP32(NT2H_OP_SIGN)
    NT2H_MAP32(NT2H_OP_EXPM1) NT2H_MAP32(NT2H_OP_LOG1P) NT2H_MAP32(NT2H_OP_TANH)
    NT2H_MAP32(NT2H_OP_SINH) NT2H_MAP32(NT2H_OP_COSH) NT2H_MAP32(NT2H_OP_ATAN) NT2H_MAP32(NT2H_OP_TAN)
    NT2H_MAP32(NT2H_OP_ASIN) NT2H_MAP32(NT2H_OP_ACOS)
    NT2H_MAP32(NT2H_OP_ERF) NT2H_MAP32(NT2H_OP_CBRT) NT2H_MAP32(NT2H_OP_ASINH) NT2H_MAP32(NT2H_OP_ACOSH)
    NT2H_MAP32(NT2H_OP_ATANH) NT2H_MAP32(NT2H_OP_COT) NT2H_MAP32(NT2H_OP_SEC) NT2H_MAP32(NT2H_OP_CSC)
    NT2H_MAP32(NT2H_OP_SINPI) NT2H_MAP32(NT2H_OP_COSPI) NT2H_MAP32(NT2H_OP_TANPI) NT2H_MAP32(NT2H_OP_COTPI)
    NT2H_MAP32(NT2H_OP_SIND) NT2H_MAP32(NT2H_OP_COSD) NT2H_MAP32(NT2H_OP_TAND) NT2H_MAP32(NT2H_OP_COTD)
#undef NT2H_MAP32
    default: return fail(NT2H_EINVAL, "nt2h_map_f32: unknown op");
  }
  NT2H_CHECK_LAUNCH("map_f32");
  return NT2H_OK;
}

template <class P>
static int launch_map_f64(int op, size_t n, const double* x, double* y, hipStream_t s) {
  const size_t items = n / 4 + n % 4;
  const int vec = aligned16(x) && aligned16(y);
  const dim3 g(grid_for(items)), b(kBlock);
  switch (op) {
    case NT2H_OP_EXP: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_EXP>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_LOG: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_LOG>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_SQRT: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_SQRT>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_ERFC: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_ERFC>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_NORMCDF: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_NORMCDF>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_COS: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_COS>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_SIN: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_SIN>), g, b, 0, s, x, y, n, vec); break;
    case NT2H_OP_ABS: hipLaunchKernelGGL((map_f64_k<P, NT2H_OP_ABS>), g, b, 0, s, x, y, n, vec); break;
#define NT2H_MAP64(OPC) case OPC: hipLaunchKernelGGL((map_f64_k<P, OPC>), g, b, 0, s, x, y, n, vec); break;
    NT2H_MAP64(NT2H_OP_EXP2) NT2H_MAP64(NT2H_OP_EXP10) NT2H_MAP64(NT2H_OP_LOG2) NT2H_MAP64(NT2H_OP_LOG10)
    NT2H_MAP64(NT2H_OP_FLOOR) NT2H_MAP64(NT2H_OP_CEIL) NT2H_MAP64(NT2H_OP_TRUNC) NT2H_MAP64(NT2H_OP_ROUND)
    NT2H_MAP64(NT2H_OP_ROUND2EVEN) NT2H_MAP64(NT2H_OP_SIGN)
    NT2H_MAP64(NT2H_OP_EXPM1) NT2H_MAP64(NT2H_OP_LOG1P) NT2H_MAP64(NT2H_OP_TANH)
    NT2H_MAP64(NT2H_OP_SINH) NT2H_MAP64(NT2H_OP_COSH) NT2H_MAP64(NT2H_OP_ATAN) NT2H_MAP64(NT2H_OP_TAN)
    NT2H_MAP64(NT2H_OP_ASIN) NT2H_MAP64(NT2H_OP_ACOS)
    NT2H_MAP64(NT2H_OP_ERF) NT2H_MAP64(NT2H_OP_CBRT) NT2H_MAP64(NT2H_OP_ASINH) NT2H_MAP64(NT2H_OP_ACOSH)
    NT2H_MAP64(NT2H_OP_ATANH) NT2H_MAP64(NT2H_OP_COT) NT2H_MAP64(NT2H_OP_SEC) NT2H_MAP64(NT2H_OP_CSC)
    NT2H_MAP64(NT2H_OP_SINPI) NT2H_MAP64(NT2H_OP_COSPI) NT2H_MAP64(NT2H_OP_TANPI) NT2H_MAP64(NT2H_OP_COTPI)
    NT2H_MAP64(NT2H_OP_SIND) NT2H_MAP64(NT2H_OP_COSD) NT2H_MAP64(NT2H_OP_TAND) NT2H_MAP64(NT2H_OP_COTD)
#undef NT2H_MAP64
    default: return fail(NT2H_EINVAL, "nt2h_map_f64: unknown op (fastnormcdf is single precision only)");
  }
  NT2H_CHECK_LAUNCH("map_f64");
  return NT2H_OK;
}

// ---- static scratch of nt2h_dot_f32_noalloc ---------------------------------
// 256 FP64 block partials (2 KiB) per device, held until the process ends.
constexpr unsigned kScratchDoubles = 256;
static_assert(kScratchDoubles >= kFoldBlocks, "the fold's partials must fit the static scratch");
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_sym_mu;
static double* g_scratch_sym[kMaxDevices];

// the current device's address of `scratch`, looked up on its first use
static int scratch_ptr(double** out) {
  int dev = 0;
  NT2H_CHECK(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(NT2H_EINVAL, "nt2h_dot_f32_noalloc: device index beyond 64");
  std::lock_guard<std::mutex> lk(g_scratch_sym_mu);
  if (!g_scratch_sym[dev]) {
    void* p = nullptr;
    NT2H_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(NT2H_EHIP, "hipGetSymbolAddress(scratch): null address");
    g_scratch_sym[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_sym[dev];
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int nt2h_abi_version(void) { return 2; }

const char* nt2h_last_error(void) { return g_last_error.c_str(); }

int nt2h_device_info(int* cus, size_t* hbm_bytes, char* name, size_t name_len) {
  int dev = 0;
  NT2H_CHECK(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t p;
  NT2H_CHECK(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  if (cus) *cus = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
  if (name && name_len) {
    std::strncpy(name, p.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  return NT2H_OK;
}

int nt2h_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(NT2H_EINVAL, "nt2h_malloc: null out pointer");
  *dptr = nullptr;
  if (bytes == 0) return NT2H_OK;
  NT2H_CHECK(hipMalloc(dptr, bytes), "hipMalloc");
  return NT2H_OK;
}

int nt2h_free(void* dptr) {
  if (!dptr) return NT2H_OK;
  NT2H_CHECK(hipFree(dptr), "hipFree");
  return NT2H_OK;
}

int nt2h_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream) {
  if (bytes == 0) return NT2H_OK;
  if (!dst || !src) return fail(NT2H_EINVAL, "nt2h_memcpy: null pointer");
  hipMemcpyKind k;
  switch (kind) {
    case NT2H_H2D: k = hipMemcpyHostToDevice; break;
    case NT2H_D2H: k = hipMemcpyDeviceToHost; break;
    case NT2H_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return fail(NT2H_EINVAL, "nt2h_memcpy: bad kind");
  }
  NT2H_CHECK(hipMemcpyAsync(dst, src, bytes, k, as_stream(stream)), "hipMemcpyAsync");
  return NT2H_OK;
}

int nt2h_stream_create(void** stream) {
  if (!stream) return fail(NT2H_EINVAL, "nt2h_stream_create: null out pointer");
  hipStream_t s;
  NT2H_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *stream = s;
  return NT2H_OK;
}
int nt2h_stream_destroy(void* stream) {
  if (stream) NT2H_CHECK(hipStreamDestroy(as_stream(stream)), "hipStreamDestroy");
  return NT2H_OK;
}
int nt2h_stream_sync(void* stream) {
  NT2H_CHECK(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
  return NT2H_OK;
}

static unsigned gen_grid(size_t n) {
  size_t g = (n + kBlock - 1) / kBlock;
  return (unsigned)(g > 65536 ? 65536 : (g ? g : 1));
}

int nt2h_fill_uniform_f32(float* out, size_t n, double lo, double hi, uint64_t seed, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!out) return fail(NT2H_EINVAL, "nt2h_fill_uniform_f32: null pointer");
  hipLaunchKernelGGL(fill_uniform_k<float>, dim3(gen_grid(n)), dim3(kBlock), 0, as_stream(stream), out, n, lo, hi, seed);
  NT2H_CHECK_LAUNCH("fill_uniform_f32");
  return NT2H_OK;
}
int nt2h_fill_uniform_f64(double* out, size_t n, double lo, double hi, uint64_t seed, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!out) return fail(NT2H_EINVAL, "nt2h_fill_uniform_f64: null pointer");
  hipLaunchKernelGGL(fill_uniform_k<double>, dim3(gen_grid(n)), dim3(kBlock), 0, as_stream(stream), out, n, lo, hi, seed);
  NT2H_CHECK_LAUNCH("fill_uniform_f64");
  return NT2H_OK;
}
int nt2h_fill_normal_f32(float* out, size_t n, uint64_t seed, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!out) return fail(NT2H_EINVAL, "nt2h_fill_normal_f32: null pointer");
  hipLaunchKernelGGL(fill_normal_k, dim3(gen_grid(n)), dim3(kBlock), 0, as_stream(stream), out, n, seed);
  NT2H_CHECK_LAUNCH("fill_normal_f32");
  return NT2H_OK;
}

int nt2h_fill_frames_u8(uint8_t* frames, int h, int w, int nframes, int nrect, int noise,
                        uint64_t seed, void* stream) {
  if (!frames || h <= 0 || w <= 0 || nframes <= 0 || nrect < 0 || nrect > 8 || noise < 0 || noise > 64)
    return fail(NT2H_EINVAL, "nt2h_fill_frames_u8: bad argument");
  Rects R;
  R.n = nrect;
  for (int q = 0; q < nrect; ++q) {
    auto u = [&](int j) { return splitmix64_at(seed ^ 0x5bd1e995ull, (uint64_t)(q * 16 + j)); };
    Rect& r = R.r[q];
    r.w = 48 + (int)(u(0) % 113u);
    r.h = 48 + (int)(u(1) % 113u);
    r.x0 = (int)(u(2) % (uint64_t)w);
    r.y0 = (int)(u(3) % (uint64_t)h);
    r.vx = (int)(u(4) % 17u) - 8;
    r.vy = (int)(u(5) % 17u) - 8;
    r.val = 230 + (int)(u(6) % 26u);
  }
  const size_t total = (size_t)h * w * nframes;
  hipLaunchKernelGGL(fill_frames_k, dim3(gen_grid(total)), dim3(kBlock), 0, as_stream(stream), frames, h, w, nframes, R, noise, seed);
  NT2H_CHECK_LAUNCH("fill_frames_u8");
  return NT2H_OK;
}

int nt2h_map_f32(int op, size_t n, const float* x, float* y, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !y) return fail(NT2H_EINVAL, "nt2h_map_f32: null pointer");
  if (op >= NT2H_OP_COTH) return launch_compose(op, n, x, y, policy, as_stream(stream));
  if (policy == NT2H_EXACT) return launch_map_f32<exact_policy>(op, n, x, y, as_stream(stream));
  if (policy == NT2H_FAST) return launch_map_f32<fast_policy>(op, n, x, y, as_stream(stream));
  return fail(NT2H_EINVAL, "nt2h_map_f32: bad policy");
}

int nt2h_map_f64(int op, size_t n, const double* x, double* y, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !y) return fail(NT2H_EINVAL, "nt2h_map_f64: null pointer");
  if (op >= NT2H_OP_COTH) return launch_compose(op, n, x, y, policy, as_stream(stream));
  if (policy == NT2H_EXACT) return launch_map_f64<exact_policy>(op, n, x, y, as_stream(stream));
  if (policy == NT2H_FAST) return launch_map_f64<fast_policy>(op, n, x, y, as_stream(stream));
  return fail(NT2H_EINVAL, "nt2h_map_f64: bad policy");
}

extern "C++" template <class T, int W>
int launch_map2(const char* name, int op, size_t n, const T* x, const T* y, T* out, int policy,
                       void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !y || !out) return fail(NT2H_EINVAL, name);
  if (op < NT2H_OP2_ATAN2 || op > NT2H_OP2_LOGSPACE_SUB) return fail(NT2H_EINVAL, name);
  if (policy != NT2H_EXACT && policy != NT2H_FAST) return fail(NT2H_EINVAL, name);
  const dim3 g(grid_for(n / W + n % W)), b(kBlock);
  hipStream_t s = as_stream(stream);
  if (op == NT2H_OP2_ATAN2) {
    if (policy == NT2H_EXACT) hipLaunchKernelGGL((map2_k<exact_policy, NT2H_OP2_ATAN2, T, W>), g, b, 0, s, x, y, out, n);
    else hipLaunchKernelGGL((map2_k<fast_policy, NT2H_OP2_ATAN2, T, W>), g, b, 0, s, x, y, out, n);
  } else if (op == NT2H_OP2_LOGSPACE_ADD) {
    if (policy == NT2H_EXACT) hipLaunchKernelGGL((map2_k<exact_policy, NT2H_OP2_LOGSPACE_ADD, T, W>), g, b, 0, s, x, y, out, n);
    else hipLaunchKernelGGL((map2_k<fast_policy, NT2H_OP2_LOGSPACE_ADD, T, W>), g, b, 0, s, x, y, out, n);
  } else if (op == NT2H_OP2_LOGSPACE_SUB) {
    if (policy == NT2H_EXACT) hipLaunchKernelGGL((map2_k<exact_policy, NT2H_OP2_LOGSPACE_SUB, T, W>), g, b, 0, s, x, y, out, n);
    else hipLaunchKernelGGL((map2_k<fast_policy, NT2H_OP2_LOGSPACE_SUB, T, W>), g, b, 0, s, x, y, out, n);
  } else if (op == NT2H_OP2_ATAN2D) {
    if (policy == NT2H_EXACT) hipLaunchKernelGGL((map2_k<exact_policy, NT2H_OP2_ATAN2D, T, W>), g, b, 0, s, x, y, out, n);
    else hipLaunchKernelGGL((map2_k<fast_policy, NT2H_OP2_ATAN2D, T, W>), g, b, 0, s, x, y, out, n);
  } else {
    if (policy == NT2H_EXACT) hipLaunchKernelGGL((map2_k<exact_policy, NT2H_OP2_POW, T, W>), g, b, 0, s, x, y, out, n);
    else hipLaunchKernelGGL((map2_k<fast_policy, NT2H_OP2_POW, T, W>), g, b, 0, s, x, y, out, n);
  }
  NT2H_CHECK_LAUNCH(name);
  return NT2H_OK;
}

int nt2h_map2_f32(int op, size_t n, const float* x, const float* y, float* out, int policy, void* stream) {
  return launch_map2<float, 8>("nt2h_map2_f32", op, n, x, y, out, policy, stream);
}
int nt2h_map2_f64(int op, size_t n, const double* x, const double* y, double* out, int policy, void* stream) {
  return launch_map2<double, 4>("nt2h_map2_f64", op, n, x, y, out, policy, stream);
}

int nt2h_sincos_f32(size_t n, const float* x, float* s, float* c, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !s || !c) return fail(NT2H_EINVAL, "nt2h_sincos_f32: null pointer");
  const dim3 g(grid_for(n / 8 + n % 8)), b(kBlock);
  if (policy == NT2H_EXACT) hipLaunchKernelGGL(sincos_f32_k<exact_policy>, g, b, 0, as_stream(stream), x, s, c, n);
  else if (policy == NT2H_FAST) hipLaunchKernelGGL(sincos_f32_k<fast_policy>, g, b, 0, as_stream(stream), x, s, c, n);
  else return fail(NT2H_EINVAL, "nt2h_sincos_f32: bad policy");
  NT2H_CHECK_LAUNCH("sincos_f32");
  return NT2H_OK;
}

int nt2h_sincos_f64(size_t n, const double* x, double* s, double* c, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !s || !c) return fail(NT2H_EINVAL, "nt2h_sincos_f64: null pointer");
  const dim3 g(grid_for(n / 4 + n % 4)), b(kBlock);
  if (policy == NT2H_EXACT) hipLaunchKernelGGL(sincos_f64_k<exact_policy>, g, b, 0, as_stream(stream), x, s, c, n);
  else if (policy == NT2H_FAST) hipLaunchKernelGGL(sincos_f64_k<fast_policy>, g, b, 0, as_stream(stream), x, s, c, n);
  else return fail(NT2H_EINVAL, "nt2h_sincos_f64: bad policy");
  NT2H_CHECK_LAUNCH("sincos_f64");
  return NT2H_OK;
}

int nt2h_stream_copy(const void* src, void* dst, size_t bytes, int grid, void* stream) {
  if (bytes == 0) return NT2H_OK;
  if (!src || !dst || !aligned16(src) || !aligned16(dst) || bytes % 16)
    return fail(NT2H_EINVAL, "nt2h_stream_copy: needs 16-byte aligned buffers and size");
  const size_t n4 = bytes / 16;
  const unsigned g = grid > 0 ? (unsigned)grid : grid_for(n4);
  hipLaunchKernelGGL(stream_copy_k, dim3(g), dim3(kBlock), 0, as_stream(stream),
                     (const v4f*)src, (v4f*)dst, n4);
  NT2H_CHECK_LAUNCH("stream_copy");
  return NT2H_OK;
}

int nt2h_saxpy_f32(size_t n, float a, const float* x, float* y, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !y) return fail(NT2H_EINVAL, "nt2h_saxpy_f32: null pointer");
  const int vec = aligned16(x) && aligned16(y);
  const size_t items = vec ? n / 4 + n % 4 : n;
  hipLaunchKernelGGL(saxpy_k, dim3(grid_for(items)), dim3(kBlock), 0, as_stream(stream), n, a, x, y, vec);
  NT2H_CHECK_LAUNCH("saxpy_f32");
  return NT2H_OK;
}

// ---- inner_fold_ ----------------------------------------------------------
static unsigned fold_grid(size_t work_items) {
  size_t g = (work_items + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > kMaxFoldBlocks) g = kMaxFoldBlocks;
  return (unsigned)g;
}

// fold geometry: at most kFoldBlocks blocks, each a contiguous range of a
// multiple of kBlock units, at least kBlock*kFoldUnroll units per block when
// there are that many
static void fold_geometry(size_t units, unsigned* g, size_t* per) {
  const size_t chunk = (size_t)kBlock * kFoldUnroll;
  size_t gg = (units + chunk - 1) / chunk;
  if (gg < 1) gg = 1;
  if (gg > kFoldBlocks) gg = kFoldBlocks;
  size_t p = (units + gg - 1) / gg;
  p = (p + kBlock - 1) / kBlock * kBlock;
  if (p == 0) p = kBlock;
  gg = (units + p - 1) / p;
  if (gg < 1) gg = 1;
  *g = (unsigned)gg;
  *per = p;
}

// the two launches of a fold into the partials buffer (kFoldBlocks doubles)
static int enqueue_fold(size_t n, const float* x, const float* y, double* partial, double* out, hipStream_t s) {
  const bool vec = aligned16(x) && (!y || aligned16(y));
  unsigned g;
  size_t per;
  fold_geometry(vec ? n / 4 : n, &g, &per);
  const dim3 grid(g), block(kBlock);
  if (y) {
    if (vec) hipLaunchKernelGGL((fold_part_k<true, true>), grid, block, 0, s, x, y, n, per, partial);
    else hipLaunchKernelGGL((fold_part_k<true, false>), grid, block, 0, s, x, y, n, per, partial);
  } else {
    if (vec) hipLaunchKernelGGL((fold_part_k<false, true>), grid, block, 0, s, x, y, n, per, partial);
    else hipLaunchKernelGGL((fold_part_k<false, false>), grid, block, 0, s, x, y, n, per, partial);
  }
  NT2H_CHECK_LAUNCH("fold_f32");
  hipLaunchKernelGGL(fold_final_k, dim3(1), block, 0, s, partial, g, out);
  NT2H_CHECK_LAUNCH("fold_f32 final");
  return NT2H_OK;
}

static int launch_fold(size_t n, const float* x, const float* y, double* out, hipStream_t s) {
  Scratch sc;
  int rc = get_scratch(kFoldBlocks, &sc, s);
  if (rc) return rc;
  return enqueue_fold(n, x, y, sc.partial, out, s);
}

int nt2h_sum_f32(size_t n, const float* x, double* out, void* stream) {
  if (!out || (n && !x)) return fail(NT2H_EINVAL, "nt2h_sum_f32: null pointer");
  if (n == 0) return hipMemsetAsync(out, 0, sizeof(double), as_stream(stream)) == hipSuccess ? NT2H_OK : NT2H_EHIP;
  return launch_fold(n, x, nullptr, out, as_stream(stream));
}

int nt2h_dot_f32(size_t n, const float* x, const float* y, double* out, void* stream) {
  if (!out || (n && (!x || !y))) return fail(NT2H_EINVAL, "nt2h_dot_f32: null pointer");
  if (n == 0) return hipMemsetAsync(out, 0, sizeof(double), as_stream(stream)) == hipSuccess ? NT2H_OK : NT2H_EHIP;
  return launch_fold(n, x, y, out, as_stream(stream));
}

// nt2h_dot_f32 without allocation: the fold's block partials live in the
// file-scope array `scratch` (one copy per device), whose device address is
// looked up once per device with hipGetSymbolAddress (the array's name is
// only a host-side stand-in and must never be passed to a kernel).  The fold
// writes every partial it reads, so nothing has to be cleared between calls;
// calls that overlap on one device share the array and race.
int nt2h_dot_f32_noalloc(size_t n, const float* x, const float* y, double* out, void* stream) {
  if (!out || (n && (!x || !y))) return fail(NT2H_EINVAL, "nt2h_dot_f32_noalloc: null pointer");
  unsigned g;
  size_t per;
  fold_geometry(aligned16(x) && aligned16(y) ? n / 4 : n, &g, &per);
  if (g > kScratchDoubles) return fail(NT2H_EINVAL, "nt2h_dot_f32_noalloc: more block partials than the static scratch holds");
  double* partial = nullptr;
  if (int rc = scratch_ptr(&partial)) return rc;
  // n == 0 runs the fold too (one block, no element read): it writes 0 without a memset node
  return enqueue_fold(n, x, y, partial, out, as_stream(stream));
}

// ---- Black-Scholes ----------------------------------------------------------
static unsigned g_acc_grid = 256 * 32;  // workgroups of the grid-stride accurate kernel (32 per CU)

/* tuning hook: workgroups launched by the accurate Black-Scholes kernel */
int nt2h_bs_set_grid(int workgroups) {
  if (workgroups < 1) return fail(NT2H_EINVAL, "nt2h_bs_set_grid: bad count");
  g_acc_grid = (unsigned)workgroups;
  return NT2H_OK;
}

int nt2h_blackscholes_f32(size_t n, const float* S, const float* X, const float* T, float r,
                          float v, float* call, float* put, int mode, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!S || !X || !T || !call || (mode == NT2H_BS_TEXTBOOK && !put))
    return fail(NT2H_EINVAL, "nt2h_blackscholes_f32: null pointer");
  if (mode != NT2H_BS_TEXTBOOK && mode != NT2H_BS_DEMO)
    return fail(NT2H_EINVAL, "nt2h_blackscholes_f32: bad mode");
  if (policy != NT2H_EXACT && policy != NT2H_FAST && policy != NT2H_ACCURATE)
    return fail(NT2H_EINVAL, "nt2h_blackscholes_f32: bad policy");
  if (policy == NT2H_ACCURATE && mode != NT2H_BS_TEXTBOOK)
    return fail(NT2H_EINVAL, "nt2h_blackscholes_f32: the accurate policy evaluates the textbook formula");
  const int vec = aligned16(S) && aligned16(X) && aligned16(T) && aligned16(call) &&
                  (mode == NT2H_BS_DEMO || aligned16(put));
  const size_t items = vec ? n / 4 + n % 4 : n;
  const dim3 g(grid_for(items)), b(kBlock);
  hipStream_t s = as_stream(stream);
  if (policy == NT2H_ACCURATE) {
    const size_t n4 = vec ? n / 4 : 0;
    const unsigned ga = std::min<unsigned>(grid_for(vec ? n4 + n % 4 : n), g_acc_grid);
    const auto prm = nt2::hip::bsacc::make_params(r, v);
    if (n4 < (size_t(1) << 31))
      hipLaunchKernelGGL(bs_acc_f32_k<uint32_t>, dim3(ga), b, 0, s, (uint32_t)n4, n, S, X, T, prm, call, put);
    else
      hipLaunchKernelGGL(bs_acc_f32_k<size_t>, dim3(ga), b, 0, s, n4, n, S, X, T, prm, call, put);
  } else if (policy == NT2H_EXACT) {
    if (mode == NT2H_BS_TEXTBOOK) hipLaunchKernelGGL((bs_f32_k<exact_policy, NT2H_BS_TEXTBOOK>), g, b, 0, s, n, S, X, T, r, v, call, put, vec);
    else hipLaunchKernelGGL((bs_f32_k<exact_policy, NT2H_BS_DEMO>), g, b, 0, s, n, S, X, T, r, v, call, put, vec);
  } else {
    if (mode == NT2H_BS_TEXTBOOK) hipLaunchKernelGGL((bs_f32_k<fast_policy, NT2H_BS_TEXTBOOK>), g, b, 0, s, n, S, X, T, r, v, call, put, vec);
    else hipLaunchKernelGGL((bs_f32_k<fast_policy, NT2H_BS_DEMO>), g, b, 0, s, n, S, X, T, r, v, call, put, vec);
  }
  NT2H_CHECK_LAUNCH("blackscholes_f32");
  return NT2H_OK;
}

int nt2h_blackscholes_f64(size_t n, const double* S, const double* X, const double* T,
                          double r, double v, double* call, double* put, int policy,
                          void* stream) {
  if (n == 0) return NT2H_OK;
  if (!S || !X || !T || !call || !put) return fail(NT2H_EINVAL, "nt2h_blackscholes_f64: null pointer");
  const int vec = aligned16(S) && aligned16(X) && aligned16(T) && aligned16(call) && aligned16(put);
  const dim3 g(grid_for(n / 4 + n % 4)), b(kBlock);
  hipStream_t s = as_stream(stream);
  if (policy == NT2H_ACCURATE) {
    const unsigned ga = std::min<unsigned>(grid_for(n), g_acc_grid);
    const auto prm = nt2::hip::bsacc64::make_params(r, v);
    if (n < (size_t(1) << 31))
      hipLaunchKernelGGL(bs_acc_f64_k<uint32_t>, dim3(ga), b, 0, s, (uint32_t)n, S, X, T, prm, call, put);
    else
      hipLaunchKernelGGL(bs_acc_f64_k<size_t>, dim3(ga), b, 0, s, n, S, X, T, prm, call, put);
  } else if (policy == NT2H_EXACT) hipLaunchKernelGGL(bs_f64_k<exact_policy>, g, b, 0, s, n, S, X, T, r, v, call, put, vec);
  else if (policy == NT2H_FAST) hipLaunchKernelGGL(bs_f64_k<fast_policy>, g, b, 0, s, n, S, X, T, r, v, call, put, vec);
  else return fail(NT2H_EINVAL, "nt2h_blackscholes_f64: bad policy");
  NT2H_CHECK_LAUNCH("blackscholes_f64");
  return NT2H_OK;
}

// ---- sigma-delta --------------------------------------------------------------
static int g_sd_pixels = 16;
int nt2h_sigmadelta_set_config(int pixels_per_thread) {
  if (pixels_per_thread != 16 && pixels_per_thread != 8 && pixels_per_thread != 4)
    return fail(NT2H_EINVAL, "nt2h_sigmadelta_set_config: 16, 8 or 4 pixels per thread");
  g_sd_pixels = pixels_per_thread;
  return NT2H_OK;
}

int nt2h_sigmadelta_u8(size_t npix, int nframes, const uint8_t* frames, size_t frame_stride,
                       uint8_t* bkg, uint8_t* var, uint8_t* label, int N,
                       uint32_t* frame_counts, void* stream) {
  if (npix == 0) return NT2H_OK;
  if (!frames || !bkg || !var || !label || nframes < 1 || N < 1 || frame_stride < npix)
    return fail(NT2H_EINVAL, "nt2h_sigmadelta_u8: bad argument");
  hipStream_t s = as_stream(stream);
  // counts are accumulated by the kernels below: start from zero (a fill
  // kernel, so that the sequence can be captured in a graph)
  if (frame_counts) {
    hipLaunchKernelGGL(zero_u32_k, dim3((nframes + kBlock - 1) / kBlock), dim3(kBlock), 0, s, frame_counts, nframes);
    NT2H_CHECK_LAUNCH("sigmadelta zero counts");
  }
  const int px = g_sd_pixels;  // pixels per thread: 16, 8 or 4
  const bool vec = ((reinterpret_cast<uintptr_t>(frames) | reinterpret_cast<uintptr_t>(bkg) |
                     reinterpret_cast<uintptr_t>(var) | reinterpret_cast<uintptr_t>(label) | frame_stride) %
                    (size_t)px) == 0;
  size_t p0 = 0;
  if (vec) {
    const size_t n16 = npix / px;
    if (n16) {
      const unsigned grid = (grid_for(n16) + 7) / 8 * 8;  // a multiple of the 8 XCDs
      Scratch sc;
      if (frame_counts) {
        const size_t need = (size_t)kSdChunk * grid;  // uint32 partials of one chunk
        if (int e = get_scratch((need + 1) / 2, &sc, s)) return e;
      }
      uint32_t* part = frame_counts ? reinterpret_cast<uint32_t*>(sc.partial) : nullptr;
      for (int k0 = 1; k0 < nframes; k0 += kSdChunk) {
        const int k1 = std::min(nframes, k0 + kSdChunk);
        if (px == 16)
          hipLaunchKernelGGL(sigmadelta_k<4>, dim3(grid), dim3(kBlock), 0, s, frames, frame_stride, k0, k1, n16,
                             bkg, var, label, N, part);
        else if (px == 8)
          hipLaunchKernelGGL(sigmadelta_k<2>, dim3(grid), dim3(kBlock), 0, s, frames, frame_stride, k0, k1, n16,
                             bkg, var, label, N, part);
        else
          hipLaunchKernelGGL(sigmadelta_k<1>, dim3(grid), dim3(kBlock), 0, s, frames, frame_stride, k0, k1, n16,
                             bkg, var, label, N, part);
        NT2H_CHECK_LAUNCH("sigmadelta");
        if (part) {
          hipLaunchKernelGGL(sd_counts_k, dim3(k1 - k0), dim3(kBlock), 0, s, part, (int)grid,
                             frame_counts + k0);
          NT2H_CHECK_LAUNCH("sigmadelta counts");
        }
      }
    }
    p0 = n16 * px;
  }
  if (p0 < npix) {
    hipLaunchKernelGGL(sigmadelta1_k, dim3(grid_for(npix - p0)), dim3(kBlock), 0, s, frames, frame_stride, nframes, p0, npix, bkg, var, label, N, frame_counts);
    NT2H_CHECK_LAUNCH("sigmadelta1");
  }
  return NT2H_OK;
}

int nt2h_sigmadelta_step_u8(size_t npix, const uint8_t* frame, uint8_t* bkg, uint8_t* var,
                            uint8_t* label, int N, void* stream) {
  if (npix == 0) return NT2H_OK;
  if (!frame || !bkg || !var || !label || N < 1) return fail(NT2H_EINVAL, "nt2h_sigmadelta_step_u8: bad argument");
  // a two-frame run whose frame 1 is `frame`: base pointer one stride before it
  const uint8_t* base = frame - npix;
  return nt2h_sigmadelta_u8(npix, 2, base, npix, bkg, var, label, N, nullptr, stream);
}

// ---- batched expression, single table ---------------------------------------
int nt2h_expr4_f32(size_t n, const float* x, const float* y, float* z, int policy, void* stream) {
  if (n == 0) return NT2H_OK;
  if (!x || !y || !z) return fail(NT2H_EINVAL, "nt2h_expr4_f32: null pointer");
  const dim3 g(grid_for(n / 8 + n % 8)), b(kBlock);
  if (policy == NT2H_EXACT) hipLaunchKernelGGL(expr4_k<exact_policy>, g, b, 0, as_stream(stream), x, y, z, n);
  else if (policy == NT2H_FAST) hipLaunchKernelGGL(expr4_k<fast_policy>, g, b, 0, as_stream(stream), x, y, z, n);
  else return fail(NT2H_EINVAL, "nt2h_expr4_f32: bad policy");
  NT2H_CHECK_LAUNCH("expr4_f32");
  return NT2H_OK;
}

int nt2h_expr4_sum_f32(size_t n, const float* x, const float* y, double* out, int policy,
                       void* stream) {
  if (!out || (n && (!x || !y))) return fail(NT2H_EINVAL, "nt2h_expr4_sum_f32: null pointer");
  hipStream_t s = as_stream(stream);
  if (n == 0) return hipMemsetAsync(out, 0, sizeof(double), s) == hipSuccess ? NT2H_OK : NT2H_EHIP;
  const unsigned g = fold_grid(n / 8 + n % 8);
  Scratch sc;
  int rc = get_scratch(g, &sc, s);
  if (rc) return rc;
  if (policy == NT2H_EXACT) hipLaunchKernelGGL(expr4_sum_k<exact_policy>, dim3(g), dim3(kBlock), 0, s, x, y, n, sc.partial);
  else if (policy == NT2H_FAST) hipLaunchKernelGGL(expr4_sum_k<fast_policy>, dim3(g), dim3(kBlock), 0, s, x, y, n, sc.partial);
  else return fail(NT2H_EINVAL, "nt2h_expr4_sum_f32: bad policy");
  NT2H_CHECK_LAUNCH("expr4_sum_f32");
  hipLaunchKernelGGL(fold_final_k, dim3(1), dim3(kBlock), 0, s, sc.partial, g, out);
  NT2H_CHECK_LAUNCH("expr4_sum_f32 final");
  return NT2H_OK;
}

}  // extern "C"
