// Reductions and scans along one dimension of an N-D table (sum(a, dim),
// cumsum(a, dim), and cumprod, cummin, cummax in NT2's order under every
// policy) on gfx950.
//
// A table of extent (d0, d1, d2, d3), column-major, is seen from dimension
// `dim` as lines: inner = d0 ... d(dim-2) (the stride of the dimension), len =
// d(dim-1), outer = d(dim) ... d3; line (i, o) holds the elements
// i + inner * (m + len * o), m = 0 .. len-1.  A line is contiguous when
// inner == 1.
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
// In T arithmetic with no contraction, these give NT2's bits.  The chains are
// sequential by definition, so the GPU forms give each chain one thread (or,
// for a column's pack lanes, one thread per lane) and spend the parallelism on
// the number of chains: loads are issued ahead of the dependent additions,
// contiguous lines are staged through LDS so that memory accesses coalesce.
//
// Fast order (the parallel forms, fast policy): chains are cut in chunks whose
// sums are combined in FP64 in chunk order.  Sums accumulate with
// compensated (error-free) additions in T; scans accumulate in FP64 (chunk
// totals, exclusive offsets, then each chunk rescanned from its offset) and
// round each output once.  Results are deterministic for a shape.
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

constexpr int kUnroll = 16;  // loads issued ahead of a chain's additions
constexpr int kDeep = 64;    // the same for the few long chains of strided folds and scans

// The scans' operations: NT2's cumulative_<tag, bop, neutral> functors
// (modules/boost/simd/base/include/boost/simd/swar/functions/cumsum.hpp:27,
// cumprod.hpp:27, cummin.hpp:26, cummax.hpp:27), each step bop(summary, x)
// (details/scan_step.hpp:39).  min and max are NT2's scalar forms
// (boost/simd/arithmetic/functions/scalar/min.hpp:27, max.hpp:27): a NaN
// element gives NaN at its own position and the next element restarts the
// chain; of two equal values (+0 and -0) the element wins.
struct op_sum {
  template <class A> __device__ __forceinline__ static A neutral() { return A(0); }
  template <class A> __device__ __forceinline__ static A step(A s, A x) { return s + x; }
};
struct op_prod {
  template <class A> __device__ __forceinline__ static A neutral() { return A(1); }
  template <class A> __device__ __forceinline__ static A step(A s, A x) { return s * x; }
};
struct op_min {
  template <class A> __device__ __forceinline__ static A neutral() { return A(__builtin_inf()); }
  template <class A> __device__ __forceinline__ static A step(A s, A x) { return (s < x) ? s : x; }
};
struct op_max {
  template <class A> __device__ __forceinline__ static A neutral() { return A(-__builtin_inf()); }
  template <class A> __device__ __forceinline__ static A step(A s, A x) { return (s > x) ? s : x; }
};

// running Op over n elements from idx with stride `step`, applied to acc in
// order (A: accumulator type); f(m, acc) after each step.  Loads go kUnroll
// ahead of the dependent steps: a group of kUnroll loads, then their steps,
// so the memory latency is paid once per group (kDeep for the strided folds
// and scans whose chains, one per output, are few and each stream a line).
template <class A, class T, class Op = op_sum, int kUnroll = 16, class Src, class F>
__device__ __forceinline__ void chain(const Src& src, size_t idx, size_t step, size_t n, A& acc, F f) {
  _Pragma("clang fp contract(off)")
  size_t m = 0;
  for (; m + kUnroll <= n; m += kUnroll) {
    T v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      T x[1];
      src.template load<1>(idx + (m + u) * step, x);
      v[u] = x[0];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      acc = Op::step(acc, (A)v[u]);
      f(m + u, acc);
    }
  }
  for (; m < n; ++m) {
    T x[1];
    src.template load<1>(idx + m * step, x);
    acc = Op::step(acc, (A)x[0]);
    f(m, acc);
  }
}

// compensated sum (s + c) over n elements from idx with stride `step`, loads
// issued kUnroll ahead
template <class T, class Src>
__device__ __forceinline__ void chain_comp(const Src& src, size_t idx, size_t step, size_t n, T& s, T& c) {
  size_t m = 0;
  for (; m + kUnroll <= n; m += kUnroll) {
    T v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      T x[1];
      src.template load<1>(idx + (m + u) * step, x);
      v[u] = x[0];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) two_sum_add(s, c, v[u]);
  }
  for (; m < n; ++m) {
    T x[1];
    src.template load<1>(idx + m * step, x);
    two_sum_add(s, c, x[0]);
  }
}

// ---- exact order: folds ------------------------------------------------------------

// inner fold, any source: one thread per column (len elements from col * len)
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
  size_t p = 0;
  for (; p + 2 <= nb; p += 2) {
    T v0[W], v1[W];
    src.template load<W>(k + p * W, v0);
    src.template load<W>(k + (p + 1) * W, v1);
#pragma unroll
    for (int l = 0; l < W; ++l) acc[l] = acc[l] + v0[l];
#pragma unroll
    for (int l = 0; l < W; ++l) acc[l] = acc[l] + v1[l];
  }
  for (; p < nb; ++p) {
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
  if (live) chain<T, T>(src, col * len + lane, (size_t)W, len / W, acc, [](size_t, T) {});
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

// Long contiguous lines, few of them: one workgroup per line stages tiles of
// kStage elements through LDS with coalesced loads (the next tile's loads are
// in flight while the current one is consumed) and runs the line's chains on
// the tile from LDS.
constexpr int kStageThreads = 256;
template <class T> struct stage_elems { static constexpr int value = 16384 / (int)sizeof(T); };  // 16 KB per tile

// inner fold of a pack-free source: threads 0..W-1 run the pack lanes' chains
template <class T, class Src>
__global__ __launch_bounds__(kStageThreads) void inner_fold_staged_k(Src src, size_t len, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int W = pw<T>::value;
  constexpr int TE = stage_elems<T>::value;
  constexpr int PER = TE / kStageThreads;
  __shared__ T buf[2][TE];
  const size_t col = blockIdx.x;
  const size_t k = col * len;
  const size_t npk = len / W * W;  // the pack part
  const int t = threadIdx.x;
  T acc = T(0);
  T v[PER];
  auto load_tile = [&](size_t tb) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const size_t i = tb + (size_t)e * kStageThreads + t;
      T x[1] = {T(0)};
      if (i < npk) src.template load<1>(k + i, x);
      v[e] = x[0];
    }
  };
  auto store_tile = [&](int b) {
#pragma unroll
    for (int e = 0; e < PER; ++e) buf[b][e * kStageThreads + t] = v[e];
  };
  if (npk) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  for (size_t tb = 0; tb < npk; tb += TE) {
    const int b = (int)((tb / TE) & 1);
    const bool more = tb + TE < npk;
    if (more) load_tile(tb + TE);  // in flight during the chains below
    if (t < W) {
      const int n = (int)(npk - tb < (size_t)TE ? npk - tb : (size_t)TE);
      int i = t;
      for (; i + (kUnroll - 1) * W < n; i += kUnroll * W) {  // LDS reads issued ahead of the additions
        T r[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) r[u] = buf[b][i + u * W];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) acc = acc + r[u];
      }
      for (; i < n; i += W) acc = acc + buf[b][i];
    }
    if (more) store_tile(b ^ 1);
    __syncthreads();
  }
  T a[W];
#pragma unroll
  for (int l = 0; l < W; ++l) a[l] = __shfl(acc, l, 64);
  if (t == 0) {
    T s = hsum(a);
    for (size_t i = npk; i < len; ++i) {
      T x[1];
      src.template load<1>(k + i, x);
      s = s + x[0];
    }
    out[col] = s;
  }
}

// scan of contiguous lines: thread 0 runs the line's chain on each tile in LDS,
// the workgroup (one wave) loads and stores the tiles (A: accumulator type).
// Small workgroups and 4 KB tiles keep up to 20 chains per CU (LDS-bound).
constexpr int kScanThreads = 64;
constexpr int kScanAhead = 32;  // LDS reads issued ahead of the chain's steps
template <class T> struct scan_elems { static constexpr int value = 4096 / (int)sizeof(T); };  // 4 KB per tile
template <class T, class A, class Src, class Op = op_sum>
__global__ __launch_bounds__(kScanThreads) void scan_staged_k(Src src, size_t len, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int kStageThreads = kScanThreads;
  constexpr int kUnroll = kScanAhead;
  constexpr int TE = scan_elems<T>::value;
  constexpr int PER = TE / kStageThreads;
  __shared__ __attribute__((aligned(16))) T buf[2][TE];
  const size_t base = (size_t)blockIdx.x * len;
  const int t = threadIdx.x;
  A acc = Op::template neutral<A>();
  T v[PER];
  auto load_tile = [&](size_t tb) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const size_t i = tb + (size_t)e * kStageThreads + t;
      T x[1] = {T(0)};
      if (i < len) src.template load<1>(base + i, x);
      v[e] = x[0];
    }
  };
  load_tile(0);
#pragma unroll
  for (int e = 0; e < PER; ++e) buf[0][e * kStageThreads + t] = v[e];
  __syncthreads();
  for (size_t tb = 0; tb < len; tb += TE) {
    const int b = (int)((tb / TE) & 1);
    const bool more = tb + TE < len;
    if (more) load_tile(tb + TE);  // in flight during the chain below
    const int n = (int)(len - tb < (size_t)TE ? len - tb : (size_t)TE);
    if (t == 0) {
      // 16-byte LDS reads and writes (4 floats or 2 doubles per access), all
      // reads of a group issued ahead of its steps
      constexpr int VE = 16 / (int)sizeof(T);
      typedef T vec_t __attribute__((ext_vector_type(VE)));
      vec_t* bv = reinterpret_cast<vec_t*>(&buf[b][0]);
      int i = 0;
      for (; i + kUnroll <= n; i += kUnroll) {
        vec_t r[kUnroll / VE];
#pragma unroll
        for (int u = 0; u < kUnroll / VE; ++u) r[u] = bv[i / VE + u];
#pragma unroll
        for (int u = 0; u < kUnroll / VE; ++u)
#pragma unroll
          for (int l = 0; l < VE; ++l) {
            acc = Op::step(acc, (A)r[u][l]);
            r[u][l] = (T)acc;
          }
#pragma unroll
        for (int u = 0; u < kUnroll / VE; ++u) bv[i / VE + u] = r[u];
      }
      for (; i < n; ++i) {
        acc = Op::step(acc, (A)buf[b][i]);
        buf[b][i] = (T)acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {  // results out, next tile in
      const int i = e * kStageThreads + t;
      if ((size_t)i < (size_t)n) out[base + tb + i] = buf[b][i];
      if (more) buf[b ^ 1][i] = v[e];
    }
    __syncthreads();
  }
}

// outer fold of a pack-free source: one thread per output, in order over m
template <class T, class Src>
__global__ __launch_bounds__(256) void outer_fold_scalar_k(Src src, size_t inner, size_t len, size_t outer,
                                                           T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  T acc = T(0);
  chain<T, T, op_sum, kDeep>(src, i + o * inner * len, inner, len, acc, [](size_t, T) {});
  out[line] = acc;
}

// outer fold, any source: per outer index o, items are the W-packs of the inner
// positions below ipack (a multiple of 2W), then the single positions from ipack
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
    size_t m = 0;
    for (; m + 4 <= len; m += 4) {
      T v[4][W];
#pragma unroll
      for (int u = 0; u < 4; ++u) src.template load<W>(base + i0 + inner * (m + u), v[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int l = 0; l < W; ++l) acc[l] = acc[l] + v[u][l];
    }
    for (; m < len; ++m) {
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
    chain<T, T, op_sum, kDeep>(src, base + i, inner, len, acc, [](size_t, T) {});
    out[o * inner + i] = acc;
  }
}

// ---- scans (A: the accumulator type: T in the exact order, double in the fast) ----

// strided lines: one thread per line, in order over m (coalesced across threads)
template <class T, class A, class Src, class Op = op_sum>
__global__ __launch_bounds__(256) void scan_strided_k(Src src, size_t inner, size_t len, size_t outer,
                                                      T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  const size_t idx = i + o * inner * len;
  A acc = Op::template neutral<A>();
  chain<A, T, Op, kDeep>(src, idx, inner, len, acc, [&](size_t m, A a) { out[idx + m * inner] = (T)a; });
}

// contiguous lines: one wave scans 64 lines, tile by tile of K elements staged
// through LDS (coalesced loads and stores, one sequential chain per thread)
constexpr int kRowLines = 64;
constexpr size_t kRowsMin = 16384;  // fewer contiguous lines: a workgroup per line (scan_staged_k)
constexpr size_t kStagedLines = 4096;  // inner folds of fewer, long columns: a workgroup per column
constexpr size_t kStagedLen = 65536;
template <class T> struct row_tile { static constexpr int value = 256 / (int)sizeof(T); };  // 64 f32, 32 f64

template <class T, class A, class Src, class Op = op_sum>
__global__ __launch_bounds__(kRowLines) void scan_rows_k(Src src, size_t len, size_t nlines, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int K = row_tile<T>::value;
  __shared__ T tile[kRowLines][K + 1];
  const size_t line0 = (size_t)blockIdx.x * kRowLines;
  const int t = threadIdx.x;
  A acc = Op::template neutral<A>();
  for (size_t c0 = 0; c0 < len; c0 += K) {
    // lane t loads column t % K of rows t / K + e * (64 / K): 256-byte runs of one line
    T xs[K];
#pragma unroll
    for (int e = 0; e < K; ++e) {  // every load issued before the first LDS write
      const int f = e * kRowLines + t;
      const int r = f / K, c = f % K;
      const size_t line = line0 + r;
      T x[1] = {T(0)};
      if (line < nlines && c0 + c < len) src.template load<1>(line * len + c0 + c, x);
      xs[e] = x[0];
    }
#pragma unroll
    for (int e = 0; e < K; ++e) {
      const int f = e * kRowLines + t;
      tile[f / K][f % K] = xs[e];
    }
    __syncthreads();
    const int kmax = (int)(len - c0 < (size_t)K ? len - c0 : (size_t)K);
    for (int c = 0; c < kmax; ++c) {
      acc = Op::step(acc, (A)tile[t][c]);
      tile[t][c] = (T)acc;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < K; ++e) {
      const int f = e * kRowLines + t;
      const int r = f / K, c = f % K;
      const size_t line = line0 + r;
      if (line < nlines && c0 + c < len) out[line * len + c0 + c] = tile[r][c];
    }
    __syncthreads();
  }
}

// scan of contiguous lines, kLinesPerWave lines per wave (one wave per
// workgroup): the wave loads a tile of K elements of each of its lines
// (64 consecutive elements of one line per load), the next tile is in flight
// during the chains; lanes 0 .. R-1 each run one line's chain over the tile in
// LDS with 16-byte reads and writes (rows padded by 16 bytes, so the R lanes'
// accesses fall in distinct banks); then the wave stores the tile.  One
// instruction advances R chains, against one for the workgroup-per-line form.
constexpr int kLinesPerWave = 8;
constexpr size_t kLinesMin = 2048;  // contiguous lines from which the R-lines kernel is used (below kRowsMin)
template <class T> struct lines_tile { static constexpr int value = 2048 / (int)sizeof(T); };  // 2 KB of a line
template <class T, class A, class Src, class Op = op_sum>
__global__ __launch_bounds__(64) void scan_lines_k(Src src, size_t len, size_t nlines, T* __restrict__ out) {
  _Pragma("clang fp contract(off)")
  constexpr int R = kLinesPerWave;
  constexpr int K = lines_tile<T>::value;
  constexpr int VE = 16 / (int)sizeof(T);
  constexpr int KP = K + VE;          // padded row
  constexpr int PER = R * K / 64;     // elements per thread and tile
  constexpr int G = 32;               // elements per group of LDS reads
  typedef T vec_t __attribute__((ext_vector_type(VE)));
  __shared__ __attribute__((aligned(16))) T buf[R * KP];
  const size_t line0 = (size_t)blockIdx.x * R;
  const int t = threadIdx.x;
  A acc = Op::template neutral<A>();
  T xs[PER];
  auto load_tile = [&](size_t c0) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int f = e * 64 + t;
      const int r = f / K, c = f % K;
      const size_t line = line0 + r;
      T x[1] = {T(0)};
      if (line < nlines && c0 + c < len) src.template load<1>(line * len + c0 + c, x);
      xs[e] = x[0];
    }
  };
  load_tile(0);
  for (size_t c0 = 0; c0 < len; c0 += K) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int f = e * 64 + t;
      buf[(f / K) * KP + f % K] = xs[e];
    }
    __syncthreads();
    if (c0 + K < len) load_tile(c0 + K);  // in flight during the chains and the stores
    const int kmax = (int)(len - c0 < (size_t)K ? len - c0 : (size_t)K);
    if (t < R) {
      vec_t* row = reinterpret_cast<vec_t*>(buf + t * KP);
      int c = 0;
      for (; c + G <= kmax; c += G) {
        vec_t rr[G / VE];
#pragma unroll
        for (int u = 0; u < G / VE; ++u) rr[u] = row[c / VE + u];
#pragma unroll
        for (int u = 0; u < G / VE; ++u)
#pragma unroll
          for (int l = 0; l < VE; ++l) {
            acc = Op::step(acc, (A)rr[u][l]);
            rr[u][l] = (T)acc;
          }
#pragma unroll
        for (int u = 0; u < G / VE; ++u) row[c / VE + u] = rr[u];
      }
      T* rs = buf + t * KP;
      for (; c < kmax; ++c) {
        acc = Op::step(acc, (A)rs[c]);
        rs[c] = (T)acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int f = e * 64 + t;
      const int r = f / K, c = f % K;
      const size_t line = line0 + r;
      if (line < nlines && c0 + c < len) out[line * len + c0 + c] = buf[r * KP + c];
    }
    __syncthreads();
  }
}

// ---- fast order: chunked chains ------------------------------------------------------

constexpr int kChunkThreads = 256;
constexpr size_t kChunkElems = (size_t)kChunkThreads * 32;  // elements per chunk of a contiguous line

// FP64 block sum of one value per thread, in a fixed order (meaningful in thread 0)
__device__ __forceinline__ double block_sum_d(double v, double* lds) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
  __syncthreads();
  return s;
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
  constexpr int E = (int)(kChunkElems / kChunkThreads);
  T v[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {  // all loads first
    const size_t i = b + threadIdx.x + (size_t)k * kChunkThreads;
    T x[1] = {T(0)};
    if (i < e) src.template load<1>(i, x);
    v[k] = x[0];
  }
  T s = T(0), cc = T(0);
#pragma unroll
  for (int k = 0; k < E; ++k) two_sum_add(s, cc, v[k]);
  const double t = block_sum_d((double)s + (double)cc, lds);
  if (threadIdx.x == 0) partial[line * nchunk + c] = t;
}

// out[line] = FP64 sum of the line's chunk partials: one wave per line, lane l
// adding chunks l, l + 64, ... in order, then a fixed butterfly
template <class T>
__global__ __launch_bounds__(256) void chunk_final_k(const double* __restrict__ partial, size_t nchunk,
                                                     size_t nline, T* __restrict__ out) {
  const size_t line = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  if (line < nline)
    for (size_t c = lane; c < nchunk; c += 64) s += partial[line * nchunk + c];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (line < nline && lane == 0) out[line] = (T)s;
}

// exclusive prefix of the chunk totals of each line (in place), FP64: one
// wave per line, 64 chunks at a time (inclusive shuffle scan, then the base)
static __global__ __launch_bounds__(256) void chunk_offsets_k(double* __restrict__ partial, size_t nchunk,
                                                              size_t nline) {
  const size_t line = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (line >= nline) return;  // uniform across the wave
  double base = 0.0;
  for (size_t c0 = 0; c0 < nchunk; c0 += 64) {
    const size_t c = c0 + lane;
    const double t = c < nchunk ? partial[line * nchunk + c] : 0.0;
    double inc = t;
    for (int off = 1; off < 64; off <<= 1) {
      const double u = __shfl_up(inc, off, 64);
      if (lane >= off) inc += u;
    }
    if (c < nchunk) partial[line * nchunk + c] = base + (inc - t);
    base += __shfl(inc, 63, 64);
  }
}

// scan of chunk c of contiguous line `line` from its FP64 offset: the chunk is
// loaded coalesced into LDS (one padding slot per 32 elements, so that a
// thread's run of 32 is conflict-free), each thread scans 32 consecutive
// elements in FP64, the threads' totals are scanned through LDS, and the
// results leave through LDS with coalesced stores
template <class T, class Src>
__global__ __launch_bounds__(kChunkThreads) void chunk_scan_k(Src src, size_t len, size_t nchunk,
                                                              const double* __restrict__ offsets,
                                                              T* __restrict__ out) {
  constexpr int E = (int)(kChunkElems / kChunkThreads);
  constexpr int N = (int)kChunkElems;
  __shared__ T tile[N + N / 32];
  __shared__ double tot[kChunkThreads];
  const size_t line = blockIdx.y;
  const size_t c = blockIdx.x;
  const size_t lb = line * len + c * kChunkElems;  // the chunk's first element
  const size_t n = len - c * kChunkElems < kChunkElems ? len - c * kChunkElems : kChunkElems;
  const int t = threadIdx.x;
  T x[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {  // coalesced: element t + 256 k
    const int i = t + k * kChunkThreads;
    T y[1] = {T(0)};
    if ((size_t)i < n) src.template load<1>(lb + i, y);
    x[k] = y[0];
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = t + k * kChunkThreads;
    tile[i + i / 32] = x[k];
  }
  __syncthreads();
  double v[E];
  double run = 0.0;
#pragma unroll
  for (int k = 0; k < E; ++k) {  // this thread's run: elements 32 t .. 32 t + 31
    run += (double)tile[t * E + k + t];
    v[k] = run;
  }
  tot[t] = run;
  __syncthreads();
  // inclusive Hillis-Steele scan of the thread totals (FP64, fixed order)
  for (int off = 1; off < kChunkThreads; off <<= 1) {
    const double a = t >= off ? tot[t - off] : 0.0;
    __syncthreads();
    tot[t] += a;
    __syncthreads();
  }
  const double off = offsets ? offsets[line * nchunk + c] : 0.0;  // null: one chunk per line
  const double before = off + (t ? tot[t - 1] : 0.0);
#pragma unroll
  for (int k = 0; k < E; ++k) tile[t * E + k + t] = (T)(before + v[k]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = t + k * kChunkThreads;
    if ((size_t)i < n) out[lb + i] = tile[i + i / 32];
  }
}

// contiguous lines too short (or too many) to chunk: 8 threads per line, each
// compensated over every 8th element, combined in FP64 in lane order
template <class T, class Src>
__global__ __launch_bounds__(256) void lanes_sum_fast_k(Src src, size_t len, size_t nline, T* __restrict__ out) {
  constexpr int W = 8;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t line = g / W;
  const int lane = (int)(g % W);
  T s = T(0), c = T(0);
  if (line < nline && (size_t)lane < len) chain_comp(src, line * len + lane, (size_t)W, (len - lane + W - 1) / W, s, c);
  const double d = (double)s + (double)c;
  const int first = (int)(threadIdx.x & 63) - lane;
  double a[W];
#pragma unroll
  for (int l = 0; l < W; ++l) a[l] = __shfl(d, first + l, 64);
  if (line < nline && lane == 0) {
    double t = 0.0;
#pragma unroll
    for (int l = 0; l < W; ++l) t += a[l];
    out[line] = (T)t;
  }
}

// strided lines cut in nch chunks of mlen along m: chunk (line, ch) summed
// (compensated in T for sums, FP64 for scan totals) into partial[line * nch + ch],
// or, for a sum in one chunk, straight into the output
template <class T, bool SCAN, class Src>
__global__ __launch_bounds__(256) void strided_chunk_k(Src src, size_t inner, size_t len, size_t outer, size_t mlen,
                                                       size_t nch, double* __restrict__ partial, T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t ch = blockIdx.y;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  const size_t m0 = ch * mlen < len ? ch * mlen : len;
  const size_t m1 = m0 + mlen < len ? m0 + mlen : len;
  const size_t idx = i + o * inner * len + m0 * inner;
  double r;
  if constexpr (SCAN) {
    double acc = 0.0;
    chain<double, T>(src, idx, inner, m1 - m0, acc, [](size_t, double) {});
    r = acc;
  } else {
    T s = T(0), c = T(0);
    chain_comp(src, idx, inner, m1 - m0, s, c);
    r = (double)s + (double)c;
  }
  if (nch == 1 && !SCAN) out[line] = (T)r;
  else partial[line * nch + ch] = r;
}

// strided scans, second pass: chunk (line, ch) rescanned from its offset
template <class T, class Src>
__global__ __launch_bounds__(256) void strided_rescan_k(Src src, size_t inner, size_t len, size_t outer, size_t mlen,
                                                        size_t nch, const double* __restrict__ offsets,
                                                        T* __restrict__ out) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t ch = blockIdx.y;
  if (line >= inner * outer) return;
  const size_t o = line / inner;
  const size_t i = line - o * inner;
  const size_t m0 = ch * mlen < len ? ch * mlen : len;
  const size_t m1 = m0 + mlen < len ? m0 + mlen : len;
  const size_t idx = i + o * inner * len + m0 * inner;
  double acc = offsets[line * nch + ch];
  chain<double, T>(src, idx, inner, m1 - m0, acc, [&](size_t m, double a) { out[idx + m * inner] = (T)a; });
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
// the fast order's chunked form of contiguous lines: at least one full chunk
// long, few enough to index by blockIdx.y
inline bool chunked(const lines& L) { return L.inner == 1 && L.len >= kChunkElems && L.outer <= 65535; }
// chunks along m of strided lines: enough chains to fill the GPU (~64k),
// chunks of at least 256 elements, at most 65535 chunks
inline size_t strided_chunks(const lines& L) {
  const size_t chains = L.inner * L.outer;
  size_t n = chains >= 65536 ? 1 : (65536 + chains - 1) / chains;
  const size_t most = L.len / 256 ? L.len / 256 : 1;
  if (n > most) n = most;
  if (n > 65535) n = 65535;
  return n;
}
// FP64 scratch a fast-order operation needs (0 when none)
inline size_t scratch_doubles(const lines& L, bool fast, bool scan) {
  if (!fast || L.inner * L.outer == 0) return 0;
  if (chunked(L)) return chunks_of(L.len) * L.outer;
  if (L.inner == 1) return 0;
  const size_t nch = strided_chunks(L);
  return nch > 1 ? nch * L.inner * L.outer : 0;
}

// out has inner * outer elements; `scratch` holds scratch_doubles() doubles
template <class T, class Src>
void launch_sum(const Src& src, const lines& L, bool fast, double* scratch, T* out, hipStream_t s) {
  const size_t nout = L.inner * L.outer;
  if (nout == 0) return;
  constexpr int W = pw<T>::value;
  if (!fast) {
    if (L.inner == 1 && Src::pack_free && L.outer < kStagedLines && L.len >= kStagedLen) {
      hipLaunchKernelGGL((inner_fold_staged_k<T, Src>), dim3((unsigned)L.outer), dim3(kStageThreads), 0, s, src,
                         L.len, out);
    } else if (L.inner == 1 && Src::pack_free) {
      hipLaunchKernelGGL((inner_fold_lanes_k<T, Src>), dim3(blocks_for(L.outer * W)), dim3(256), 0, s, src, L.len,
                         L.outer, out);
    } else if (L.inner == 1) {
      hipLaunchKernelGGL((inner_fold_exact_k<T, Src>), dim3(blocks_for(L.outer)), dim3(256), 0, s, src, L.len,
                         L.outer, out);
    } else if (Src::pack_free) {
      hipLaunchKernelGGL((outer_fold_scalar_k<T, Src>), dim3(blocks_for(nout)), dim3(256), 0, s, src, L.inner, L.len,
                         L.outer, out);
    } else {
      const size_t ipack = L.inner / (2 * W) * (2 * W);
      const size_t items = (ipack / W + (L.inner - ipack)) * L.outer;
      hipLaunchKernelGGL((outer_fold_exact_k<T, Src>), dim3(blocks_for(items)), dim3(256), 0, s, src, L.inner,
                         L.len, L.outer, ipack, out);
    }
  } else if (chunked(L)) {
    const size_t nch = chunks_of(L.len);
    hipLaunchKernelGGL((chunk_sum_k<T, Src>), dim3((unsigned)nch, (unsigned)L.outer), dim3(kChunkThreads), 0, s,
                       src, L.len, nch, scratch);
    hipLaunchKernelGGL((chunk_final_k<T>), dim3(blocks_for(L.outer * 64)), dim3(256), 0, s, scratch, nch, L.outer, out);
  } else if (L.inner == 1) {
    hipLaunchKernelGGL((lanes_sum_fast_k<T, Src>), dim3(blocks_for(L.outer * 8)), dim3(256), 0, s, src, L.len,
                       L.outer, out);
  } else {
    const size_t nch = strided_chunks(L);
    const size_t mlen = (L.len + nch - 1) / nch;
    hipLaunchKernelGGL((strided_chunk_k<T, false, Src>), dim3(blocks_for(nout), (unsigned)nch), dim3(256), 0, s,
                       src, L.inner, L.len, L.outer, mlen, nch, scratch, out);
    if (nch > 1)
      hipLaunchKernelGGL((chunk_final_k<T>), dim3(blocks_for(nout * 64)), dim3(256), 0, s, scratch, nch, nout, out);
  }
}

// NT2's order for any scan operation: one chain per line, in T
template <class T, class Op, class Src>
void launch_scan_ordered(const Src& src, const lines& L, T* out, hipStream_t s) {
  const size_t nl = L.inner * L.outer;
  if (nl == 0 || L.len == 0) return;
  if (L.inner == 1 && nl >= kRowsMin)
    hipLaunchKernelGGL((scan_rows_k<T, T, Src, Op>), dim3(blocks_for(nl, kRowLines)), dim3(kRowLines), 0, s, src,
                       L.len, nl, out);
  else if (L.inner == 1 && nl >= kLinesMin)
    hipLaunchKernelGGL((scan_lines_k<T, T, Src, Op>), dim3(blocks_for(nl, kLinesPerWave)), dim3(64), 0, s, src,
                       L.len, nl, out);
  else if (L.inner == 1 && nl <= 65535)
    hipLaunchKernelGGL((scan_staged_k<T, T, Src, Op>), dim3((unsigned)nl), dim3(kScanThreads), 0, s, src, L.len,
                       out);
  else
    hipLaunchKernelGGL((scan_strided_k<T, T, Src, Op>), dim3(blocks_for(nl)), dim3(256), 0, s, src, L.inner, L.len,
                       L.outer, out);
}

// out has the source's extent (it may be the source's own table)
template <class T, class Src>
void launch_scan(const Src& src, const lines& L, bool fast, double* scratch, T* out, hipStream_t s) {
  const size_t nl = L.inner * L.outer;
  if (nl == 0 || L.len == 0) return;
  if (!fast) {
    launch_scan_ordered<T, op_sum>(src, L, out, s);
  } else if (chunked(L)) {
    const size_t nch = chunks_of(L.len);
    if (nch == 1) {  // every line is one chunk: its offset is 0, a single pass
      hipLaunchKernelGGL((chunk_scan_k<T, Src>), dim3(1u, (unsigned)L.outer), dim3(kChunkThreads), 0, s, src,
                         L.len, nch, (const double*)nullptr, out);
      return;
    }
    hipLaunchKernelGGL((chunk_sum_k<T, Src>), dim3((unsigned)nch, (unsigned)L.outer), dim3(kChunkThreads), 0, s,
                       src, L.len, nch, scratch);
    hipLaunchKernelGGL(chunk_offsets_k, dim3(blocks_for(L.outer * 64)), dim3(256), 0, s, scratch, nch, L.outer);
    hipLaunchKernelGGL((chunk_scan_k<T, Src>), dim3((unsigned)nch, (unsigned)L.outer), dim3(kChunkThreads), 0, s,
                       src, L.len, nch, (const double*)scratch, out);
  } else if (L.inner == 1 && nl >= kRowsMin) {
    hipLaunchKernelGGL((scan_rows_k<T, double, Src>), dim3(blocks_for(nl, kRowLines)), dim3(kRowLines), 0, s, src,
                       L.len, nl, out);
  } else if (L.inner == 1 && nl >= kLinesMin) {
    hipLaunchKernelGGL((scan_lines_k<T, double, Src>), dim3(blocks_for(nl, kLinesPerWave)), dim3(64), 0, s, src,
                       L.len, nl, out);
  } else if (L.inner == 1) {
    hipLaunchKernelGGL((scan_staged_k<T, double, Src>), dim3((unsigned)nl), dim3(kScanThreads), 0, s, src, L.len,
                       out);
  } else {
    const size_t nch = strided_chunks(L);
    if (nch == 1) {
      hipLaunchKernelGGL((scan_strided_k<T, double, Src>), dim3(blocks_for(nl)), dim3(256), 0, s, src, L.inner,
                         L.len, L.outer, out);
      return;
    }
    const size_t mlen = (L.len + nch - 1) / nch;
    hipLaunchKernelGGL((strided_chunk_k<T, true, Src>), dim3(blocks_for(nl), (unsigned)nch), dim3(256), 0, s, src,
                       L.inner, L.len, L.outer, mlen, nch, scratch, out);
    hipLaunchKernelGGL(chunk_offsets_k, dim3(blocks_for(nl * 64)), dim3(256), 0, s, scratch, nch, nl);
    hipLaunchKernelGGL((strided_rescan_k<T, Src>), dim3(blocks_for(nl), (unsigned)nch), dim3(256), 0, s, src,
                       L.inner, L.len, L.outer, mlen, nch, (const double*)scratch, out);
  }
}

}}}  // namespace nt2::hip::dw
