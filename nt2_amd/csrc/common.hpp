// Shared host/device helpers of libnt2h: error reporting, launch geometry,
// wave64 reductions and the deterministic cross-block FP64 combine.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "nt2h.h"

namespace nt2h {

// ---- per-thread last error ----------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const char* what);
int hip_fail(hipError_t e, const char* what);

#define NT2H_CHECK_LAUNCH(what)                         \
  do {                                                  \
    hipError_t e_ = hipGetLastError();                  \
    if (e_ != hipSuccess) return nt2h::hip_fail(e_, what); \
  } while (0)

#define NT2H_CHECK(call, what)                          \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return nt2h::hip_fail(e_, what); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

constexpr int kBlock = 256;  // 4 wave64 per workgroup

inline unsigned grid_for(size_t items, size_t per_block = kBlock) {
  size_t g = (items + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > 0x7fffffffu) g = 0x7fffffffu;
  return (unsigned)g;
}

// ---- reduction scratch ----------------------------------------------------
// Per-(device, stream) buffer of FP64 block partials, grown on demand and
// kept until exit.
struct Scratch {
  double* partial = nullptr;
  size_t cap = 0;
};
int get_scratch(size_t nblocks, Scratch* out, hipStream_t s);

constexpr unsigned kMaxFoldBlocks = 2048;

// ---- device-side helpers --------------------------------------------------
// Compensated FP32 accumulation: the running value is s + c, where every
// addition's rounding error is recovered exactly with Knuth's TwoSum and
// accumulated in c (Kahan-Babuska style).  Error ~ n*eps^2*sum|x|.
__device__ __forceinline__ void comp_add(float& s, float& c, float v) {
  _Pragma("clang fp contract(off)")
  const float t = s + v;
  const float bp = t - s;
  const float e = (s - (t - bp)) + (v - bp);
  s = t;
  c = c + e;
}

// Error-free merge of two compensated partials (s + c) and (s2 + c2).
__device__ __forceinline__ void comp_merge(float& s, float& c, float s2, float c2) {
  _Pragma("clang fp contract(off)")
  const float t = s + s2;
  const float bp = t - s;
  const float e = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = (c + c2) + e;
}

// Block-wide combine of per-thread (s, c): pairwise wave64 butterfly with
// the error-free merge, then the 4 wave results summed in double in lane
// order.  Returns the block sum (valid in thread 0).
__device__ __forceinline__ double block_sum_comp(float s, float c) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float s2 = __shfl_xor(s, off, 64);
    const float c2 = __shfl_xor(c, off, 64);
    comp_merge(s, c, s2, c2);
  }
  __shared__ double wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = (double)s + (double)c;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
  }
  return t;
}

// Fixed-order sum of g block partials by one workgroup of kBlock threads:
// thread t adds partials t, t + kBlock, ... in order, the wave adds its lanes
// by a butterfly, and thread 0 adds the kBlock/64 wave sums in order.  The
// partials come from the previous kernel on the stream, so no fence or atomic
// is needed (a per-block release fence with a last-block-done counter measured
// ~50 ns per block, serialized, on gfx950: 4096 blocks cost 0.2 ms).
__device__ __forceinline__ double block_final_sum(const double* __restrict__ partial, unsigned g) {
  double acc = 0.0;
  for (unsigned i = threadIdx.x; i < g; i += kBlock) acc += partial[i];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ double w[kBlock / 64];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) t += w[k];
  }
  return t;
}

// *out = the fixed-order sum of partial[0..g); launched as one workgroup
static __global__ __launch_bounds__(kBlock) void fold_final_k(const double* __restrict__ partial, unsigned g,
                                                              double* __restrict__ out) {
  const double t = block_final_sum(partial, g);
  if (threadIdx.x == 0) *out = t;
}

}  // namespace nt2h
