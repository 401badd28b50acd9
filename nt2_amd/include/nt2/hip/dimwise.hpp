// Reductions and scans along one dimension of an N-D table (sum(a, dim),
// cumsum(a, dim)) on gfx950.
//
// A table of extent (d0, d1, d2, d3), column-major, is seen from dimension
// `dim` as lines: inner = d0 ... d(dim-2) (the stride of the dimension), len =
// d(dim-1), outer = d(dim) ... d3; line (i, o) holds the elements
// i + inner * (m + len * o), m = 0 .. len-1.
//
// Exact order (NT2's CPU evaluation on the x86 AVX build, paths relative to
// the reference tree):
//  * inner fold (dim 1): per column, W-lane pack accumulators over the packs
//    from the column start, NT2's horizontal sum of the pack, then the scalar
//    tail (modules/core/reduction/include/nt2/core/functions/simd/inner_fold.hpp:44-67,
//    details/fold_step.hpp:31-43; the AVX butterfly
//    modules/boost/simd/base/include/boost/simd/swar/functions/simd/sse/avx/all_reduce.hpp:29-66);
//  * outer fold (dim > 1): per output, the sum over m in order from 0, with
//    the expression evaluated as W-lane packs along the inner dimension up to
//    the largest multiple of 2W (NT2's 64-byte cache-line bound) and as
//    scalars beyond (modules/core/reduction/include/nt2/core/functions/simd/outer_fold.hpp:46-76,
//    details/outer_fold_step.hpp:33-60);
//  * scans: per line, the running sum in order, each element evaluated as a
//    scalar (modules/core/swar/include/nt2/core/functions/scalar/inner_scan.hpp:36-55,
//    scalar/outer_scan.hpp:34-55, details/scan_step.hpp:30-44).
// In T arithmetic with no contraction, these give NT2's bits.
//
// Fast order (the parallel forms): a contiguous line (inner == 1) is cut in
// chunks that workgroups reduce with compensated (error-free) accumulation
// and combine in FP64 in chunk order; scans accumulate in FP64 (chunk
// totals, then per-chunk workgroup scans from their offsets) and round each
// output once.  Strided lines keep one thread per output, accumulating in
// FP64 (compensated for double).  Results are deterministic for a shape.
//
// A source is any type with
//   template <int G> __device__ void load(size_t i, T (&r)[G]) const;
// giving the G elements from linear index i (a pointer, or a fused expression),
// and a constant `pack_free`: true when loading elements one at a time gives
// the values a pack load gives (false for expressions whose functions decide
// per pack, such as NT2's trigonometric range reduction).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace nt2 { namespace hip { namespace dw {

template <class T> struct pw;
template <> struct pw<float> { static constexpr int value = 8; };
template <> struct pw<double> { static constexpr int value = 4; };

template <class T> struct ptr_src {
  static constexpr bool pack_free = true;
  const T* p;
  template <int G> __device__ __forceinline__ void load(size_t i, T (&r)[G]) const {
#pragma unroll
    for (int l = 0; l < G; ++l) r[l] = p[i + l];
  }
};

// NT2's horizontal sum of an AVX pack: ((l1+l5)+(l3+l7)) + ((l2+l6)+(l0+l4))
// for 8 floats, (l1+l3) + (l0+l2) for 4 doubles
__device__ __forceinline__ float hsum(const float (&a)[8]) {
  _Pragma("clang fp contract(off)")
  return ((a[1] + a[5]) + (a[3] + a[7])) + ((a[2] + a[6]) + (a[0] + a[4]));
}
__device__ __forceinline__ double hsum(const double (&a)[4]) {
  _Pragma("clang fp contract(off)")
  return (a[1] + a[3]) + (a[0] + a[2]);
}

// error-free accumulation: s + c carries the running sum
template <class T>
__device__ __forceinline__ void two_sum_add(T& s, T& c, T v) {
  _Pragma("clang fp contract(off)")
  const T t = s + v;
  const T bp = t - s;
  c = c + ((s - (t - bp)) + (v - bp));
  s = t;
}

// ---- exact order ---------------------------------------------------------------

// inner fold: one thread per column (len elements from col * len)
template <class T, class Src>
__global__ __launch_bounds__(256) void inner_fold_exact_k(Src src, size_t len, size_t ncol, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int W = pw<T>::value;
  const size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const size_t k = col * len;
  const size_t nb = len / W;
  T acc[W];
#pragma unroll
  for (int l = 0; l < W; ++l) acc[l] = T(0);
  for (size_t p = 0; p < nb; ++p) {
    T v[W];
    src.template load<W>(k + p * W, v);
#pragma unroll
    for (int l = 0; l < W; ++l) acc[l] = acc[l] + v[l];
  }
  T s = hsum(acc);
  for (size_t i = nb * W; i < len; ++i) {
    T v[1];
    src.template load<1>(k + i, v);
    s = s + v[0];
  }
  out[col] = s;
}

// inner fold of pack-free sources: W consecutive threads per column, thread l
// accumulating lane l of every pack (the same chain as the pack accumulator's
// lane), then NT2's horizontal sum and the tail in the column's first thread
template <class T, class Src>
__global__ __launch_bounds__(256) void inner_fold_lanes_k(Src src, size_t len, size_t ncol, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int W = pw<T>::value;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t col = g / W;
  const int lane = (int)(g % W);
  const bool live = col < ncol;
  T acc = T(0);
  if (live) {
    const size_t k = col * len;
    const size_t nb = len / W;
    for (size_t p = 0; p < nb; ++p) {
      T v[1];
      src.template load<1>(k + p * W + lane, v);
      acc = acc + v[0];
    }
  }
  T a[W];
  const int first = (int)(threadIdx.x & 63) - lane;  // the column's first lane in the wave
#pragma unroll
  for (int l = 0; l < W; ++l) a[l] = __shfl(acc, first + l, 64);
  if (live && lane == 0) {
    const size_t k = col * len;
    T s = hsum(a);
    for (size_t i = len / W * W; i < len; ++i) {
      T v[1];
      src.template load<1>(k + i, v);
      s = s + v[0];
    }
    out[col] = s;
  }
}

// outer fold: per outer index o, items are the W-packs of the inner positions
// below ipack (a multiple of 2W), then the single positions from ipack
template <class T, class Src>
__global__ __launch_bounds__(256) void outer_fold_exact_k(Src src, size_t inner, size_t len, size_t outer,
                                                          size_t ipack, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int W = pw<T>::value;
  const size_t per_o = ipack / W + (inner - ipack);
  const size_t item = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= per_o * outer) return;
  const size_t o = item / per_o;
  const size_t q = item - o * per_o;
  const size_t base = o * inner * len;
  if (q < ipack / W) {
    const size_t i0 = q * W;
    T acc[W];
#pragma unroll
    for (int l = 0; l < W; ++l) acc[l] = T(0);
    for (size_t m = 0; m < len; ++m) {
      T v[W];
      src.template load<W>(base + i0 + inner * m, v);
#pragma unroll
      for (int l = 0; l < W; ++l) acc[l] = acc[l] + v[l];
    }
#pragma unroll
    for (int l = 0; l < W; ++l) out[o * inner + i0 + l] = acc[l];
  } else {
    const size_t i = ipack + (q - ipack / W);
    T acc = T(0);
    for (size_t m = 0; m < len; ++m) {
      T v[1];
      src.template load<1>(base + i + inner * m, v);
      acc = acc + v[0];
    }
    out[o * inner + i] = acc;
  }
}

// scan: one thread per line (i, o), running sum in order
template <class T, class Src>
__global__ __launch_bounds__(256) void scan_exact_k(Src src, size_t inner, size_t len, size_t outer,
                                                    T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  size_t idx = i + o * inner * len;
  T acc = T(0);
  for (size_t m = 0; m < len; ++m, idx += inner) {
    T v[1];
    src.template load<1>(idx, v);
    acc = acc + v[0];
    out[idx] = acc;
  }
}

// ---- fast order ------------------------------------------------------------------

constexpr int kChunkThreads = 256;
constexpr size_t kChunkElems = (size_t)kChunkThreads * 32;  // elements per chunk of a contiguous line

// FP64 block sum of one value per thread, in thread order (threads of a wave
// combined by a fixed butterfly, waves in order)
__device__ __forceinline__ double block_sum_d(double v, double* lds) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
  __syncthreads();
  return s;  // meaningful in thread 0
}

// partial[line * nchunk + c] = FP64 sum of chunk c of contiguous line `line`
// (compensated in T within each thread)
template <class T, class Src>
__global__ __launch_bounds__(kChunkThreads) void chunk_sum_k(Src src, size_t len, size_t nchunk,
                                                             double* __restrict__ partial) {
  __shared__ double lds[kChunkThreads / 64];
  const size_t line = blockIdx.y;
  const size_t c = blockIdx.x;
  const size_t b = line * len + c * kChunkElems;
  const size_t e = line * len + ((c + 1) * kChunkElems < len ? (c + 1) * kChunkElems : len);
  T s = T(0), cc = T(0);
  for (size_t i = b + threadIdx.x; i < e; i += kChunkThreads) {
    T v[1];
    src.template load<1>(i, v);
    two_sum_add(s, cc, v[0]);
  }
  const double t = block_sum_d((double)s + (double)cc, lds);
  if (threadIdx.x == 0) partial[line * nchunk + c] = t;
}

// out[line] = sum of the line's chunk partials in chunk order
template <class T>
__global__ __launch_bounds__(256) void chunk_final_k(const double* __restrict__ partial, size_t nchunk,
                                                     size_t nline, T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nline) return;
  double s = 0.0;
  for (size_t c = 0; c < nchunk; ++c) s += partial[line * nchunk + c];
  out[line] = (T)s;
}

// strided lines: one thread per output, compensated over m
template <class T, class Src>
__global__ __launch_bounds__(256) void strided_sum_k(Src src, size_t inner, size_t len, size_t outer,
                                                     T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  size_t idx = i + o * inner * len;
  T s = T(0), c = T(0);
  for (size_t m = 0; m < len; ++m, idx += inner) {
    T v[1];
    src.template load<1>(idx, v);
    two_sum_add(s, c, v[0]);
  }
  out[line] = (T)((double)s + (double)c);
}

// exclusive prefix of the chunk totals of each line (in place), FP64
__global__ __launch_bounds__(256) inline void chunk_offsets_k(double* __restrict__ partial, size_t nchunk,
                                                              size_t nline) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nline) return;
  double run = 0.0;
  for (size_t c = 0; c < nchunk; ++c) {
    const double t = partial[line * nchunk + c];
    partial[line * nchunk + c] = run;
    run += t;
  }
}

// scan of chunk c of contiguous line `line` from its FP64 offset: each thread
// takes 32 consecutive elements, the threads' totals are scanned through LDS
template <class T, class Src>
__global__ __launch_bounds__(kChunkThreads) void chunk_scan_k(Src src, size_t len, size_t nchunk,
                                                              const double* __restrict__ offsets,
                                                              T* __restrict__ out) {
  constexpr int E = (int)(kChunkElems / kChunkThreads);
  __shared__ double tot[kChunkThreads];
  const size_t line = blockIdx.y;
  const size_t c = blockIdx.x;
  const size_t lb = line * len;
  const size_t b = c * kChunkElems + (size_t)threadIdx.x * E;  // index within the line
  double v[E];
  double run = 0.0;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    T x[1] = {T(0)};
    if (b + k < len) src.template load<1>(lb + b + k, x);
    run += (double)x[0];
    v[k] = run;
  }
  tot[threadIdx.x] = run;
  __syncthreads();
  // inclusive Hillis-Steele scan of the thread totals (FP64, fixed order)
  for (int off = 1; off < kChunkThreads; off <<= 1) {
    const double a = threadIdx.x >= (unsigned)off ? tot[threadIdx.x - off] : 0.0;
    __syncthreads();
    tot[threadIdx.x] += a;
    __syncthreads();
  }
  const double before = offsets[line * nchunk + c] + (threadIdx.x ? tot[threadIdx.x - 1] : 0.0);
#pragma unroll
  for (int k = 0; k < E; ++k)
    if (b + k < len) out[lb + b + k] = (T)(before + v[k]);
}

// strided lines: one thread per line, FP64 running sum
template <class T, class Src>
__global__ __launch_bounds__(256) void strided_scan_k(Src src, size_t inner, size_t len, size_t outer,
                                                      T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  size_t idx = i + o * inner * len;
  double run = 0.0;
  for (size_t m = 0; m < len; ++m, idx += inner) {
    T v[1];
    src.template load<1>(idx, v);
    run += (double)v[0];
    out[idx] = (T)run;
  }
}

// ---- host-side drivers -------------------------------------------------------------

struct lines {
  size_t inner, len, outer;
};
// dim is 1-based; dimensions beyond the extent have size 1
inline lines lines_of(const size_t (&ext)[4], int dim) {
  lines r{1, 1, 1};
  for (int k = 0; k < 4; ++k) {
    if (k < dim - 1) r.inner *= ext[k];
    else if (k == dim - 1) r.len = ext[k];
    else r.outer *= ext[k];
  }
  return r;
}
inline unsigned blocks_for(size_t items, unsigned per = 256) {
  const size_t g = (items + per - 1) / per;
  return (unsigned)(g ? g : 1);
}
inline size_t chunks_of(size_t len) { return (len + kChunkElems - 1) / kChunkElems; }
// the fast order's chunked form: contiguous lines, at least one full chunk
// long, few enough to index by blockIdx.y; other shapes take one thread per line
inline bool chunked(const lines& L) { return L.inner == 1 && L.len >= kChunkElems && L.outer <= 65535; }
// FP64 scratch a fast-order operation needs (0 when none)
inline size_t scratch_doubles(const lines& L, bool fast) {
  return fast && chunked(L) ? chunks_of(L.len) * L.outer : 0;
}

// out has inner * outer elements.  `scratch` holds scratch_doubles() doubles.
template <class T, class Src>
void launch_sum(const Src& src, const lines& L, bool fast, double* scratch, T* out, hipStream_t s) {
  const size_t nout = L.inner * L.outer;
  if (nout == 0) return;
  if (!fast) {
    if (L.inner == 1 && Src::pack_free) {
      constexpr int W = pw<T>::value;
      hipLaunchKernelGGL((inner_fold_lanes_k<T, Src>), dim3(blocks_for(L.outer * W)), dim3(256), 0, s, src, L.len,
                         L.outer, out);
    } else if (L.inner == 1) {
      hipLaunchKernelGGL((inner_fold_exact_k<T, Src>), dim3(blocks_for(L.outer)), dim3(256), 0, s, src, L.len,
                         L.outer, out);
    } else {
      constexpr int W = pw<T>::value;
      const size_t ipack = L.inner / (2 * W) * (2 * W);
      const size_t items = (ipack / W + (L.inner - ipack)) * L.outer;
      hipLaunchKernelGGL((outer_fold_exact_k<T, Src>), dim3(blocks_for(items)), dim3(256), 0, s, src, L.inner,
                         L.len, L.outer, ipack, out);
    }
  } else if (chunked(L)) {
    const size_t nch = chunks_of(L.len);
    hipLaunchKernelGGL((chunk_sum_k<T, Src>), dim3((unsigned)nch, (unsigned)L.outer), dim3(kChunkThreads), 0, s,
                       src, L.len, nch, scratch);
    hipLaunchKernelGGL((chunk_final_k<T>), dim3(blocks_for(L.outer)), dim3(256), 0, s, scratch, nch, L.outer, out);
  } else {
    hipLaunchKernelGGL((strided_sum_k<T, Src>), dim3(blocks_for(nout)), dim3(256), 0, s, src, L.inner, L.len,
                       L.outer, out);
  }
}

// out has the source's extent
template <class T, class Src>
void launch_scan(const Src& src, const lines& L, bool fast, double* scratch, T* out, hipStream_t s) {
  if (L.inner * L.outer == 0 || L.len == 0) return;
  if (!fast) {
    hipLaunchKernelGGL((scan_exact_k<T, Src>), dim3(blocks_for(L.inner * L.outer)), dim3(256), 0, s, src, L.inner,
                       L.len, L.outer, out);
  } else if (chunked(L)) {
    const size_t nch = chunks_of(L.len);
    hipLaunchKernelGGL((chunk_sum_k<T, Src>), dim3((unsigned)nch, (unsigned)L.outer), dim3(kChunkThreads), 0, s,
                       src, L.len, nch, scratch);
    hipLaunchKernelGGL(chunk_offsets_k, dim3(blocks_for(L.outer)), dim3(256), 0, s, scratch, nch, L.outer);
    hipLaunchKernelGGL((chunk_scan_k<T, Src>), dim3((unsigned)nch, (unsigned)L.outer), dim3(kChunkThreads), 0, s,
                       src, L.len, nch, (const double*)scratch, out);
  } else {
    hipLaunchKernelGGL((strided_scan_k<T, Src>), dim3(blocks_for(L.inner * L.outer)), dim3(256), 0, s, src,
                       L.inner, L.len, L.outer, out);
  }
}

}}}  // namespace nt2::hip::dw
