// Batched evaluation of many uneven tables in one persistent launch.
//
// Each table i holds x_i, y_i of len_i floats and the batch computes
// s_i = sum(exp(-x.*x/2).*cos(y) + sqrt(abs(x))) for every table.  The
// tables are cut into tiles of `tile` elements (a multiple of the 8-float
// pack, so NT2's pack grouping restarts at each table start); a persistent
// grid pulls tile indices from a device work queue and reduces each tile to a
// compensated FP32 partial stored as a double; a last kernel sums each
// table's partials in a fixed order (deterministic; a table of length 0 gets
// 0).  Tiles with arguments that need the out-of-line trig reduction are
// listed and redone by a second kernel, so that the persistent one holds no
// calls.  No tile publishes anything with a fence: the per-table sums are
// read by the next kernel on the stream.  The reset, persistent, slow-tile
// and sum kernels are captured once in a hipGraph and replayed.  An autotuner picks block size, elements per thread and
// workgroups per CU by timing the candidates.
//
// No counterpart in the reference: NT2 evaluates one expression per
// statement; this is the batched form of the same
// inner_fold_ over the same transform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "nt2/hip/math.hpp"
#include "nt2h.h"

#pragma clang fp contract(off)

using nt2::hip::exact_policy;
using nt2::hip::fast_policy;
namespace dm = nt2::hip::dm;

namespace nt2h {

struct TableDesc {
  const float* x;
  const float* y;
  uint64_t len;
  uint32_t first_tile;
  uint32_t ntiles;
  uint32_t vec;  // both pointers 16-byte aligned
  uint32_t pad;
};

template <class P, int G, bool BAIL = false>
__device__ __forceinline__ void expr4_g(const float* x, const float* y, float* z, bool* bail = nullptr) {
  _Pragma("clang fp contract(off)")
  float c[G];
  dm::trig_group_f<P, G, BAIL>(dm::k_cos, y, c, nullptr, bail);
#pragma unroll
  for (int l = 0; l < G; ++l) {
    const float t = ((-x[l]) * x[l]) / 2.0f;
    z[l] = dm::exp_f<P>(t) * c[l] + __builtin_sqrtf(__builtin_fabsf(x[l]));
  }
}

__global__ void batch_reset_k(unsigned* state) {
  if (threadIdx.x < 2) state[threadIdx.x] = 0u;
}

// one thread's compensated sum over its elements of a tile: 8-lane packs,
// then the table's last n%8 elements through NT2's scalar path
template <class P, int BLOCK, bool BAIL>
__device__ __forceinline__ void tile_sum(const TableDesc& d, uint64_t start, uint64_t full_end, uint64_t end,
                                         float& s, float& c, bool* bail) {
  for (uint64_t i = start + 8ull * threadIdx.x; i < full_end; i += 8ull * BLOCK) {
    float a[8], b[8], z[8];
    if (d.vec) {
      const float4 a0 = *reinterpret_cast<const float4*>(d.x + i);
      const float4 a1 = *reinterpret_cast<const float4*>(d.x + i + 4);
      const float4 b0 = *reinterpret_cast<const float4*>(d.y + i);
      const float4 b1 = *reinterpret_cast<const float4*>(d.y + i + 4);
      a[0] = a0.x; a[1] = a0.y; a[2] = a0.z; a[3] = a0.w;
      a[4] = a1.x; a[5] = a1.y; a[6] = a1.z; a[7] = a1.w;
      b[0] = b0.x; b[1] = b0.y; b[2] = b0.z; b[3] = b0.w;
      b[4] = b1.x; b[5] = b1.y; b[6] = b1.z; b[7] = b1.w;
    } else {
#pragma unroll
      for (int l = 0; l < 8; ++l) { a[l] = d.x[i + l]; b[l] = d.y[i + l]; }
    }
    expr4_g<P, 8, BAIL>(a, b, z, bail);
#pragma unroll
    for (int l = 0; l < 8; ++l) comp_add(s, c, z[l]);
  }
  for (uint64_t i = full_end + threadIdx.x; i < end; i += BLOCK) {
    float a = d.x[i], b = d.y[i], z;
    expr4_g<P, 1, BAIL>(&a, &b, &z, bail);
    comp_add(s, c, z);
  }
}

// the tile's FP64 partial: error-free wave merge, FP64 across waves in order
template <int BLOCK>
__device__ __forceinline__ void finish_tile(float s, float c, unsigned t, double* __restrict__ tile_partial,
                                            double* wsum) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float s2 = __shfl_xor(s, off, 64);
    const float c2 = __shfl_xor(c, off, 64);
    comp_merge(s, c, s2, c2);
  }
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (double)s + (double)c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) v += wsum[w];
    tile_partial[t] = v;
  }
  __syncthreads();
}

// sums[i] = the fixed-order sum of table i's tile partials (one wave per
// table: lane l adds partials l, l + 64, ... in order, then a butterfly)
__global__ __launch_bounds__(256) void batch_sums_k(const TableDesc* __restrict__ tables, int count,
                                                    const double* __restrict__ tile_partial,
                                                    double* __restrict__ sums) {
  const int ti = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (ti >= count) return;  // uniform per wave; no barrier below
  const uint32_t first = tables[ti].first_tile, nt = tables[ti].ntiles;
  double v = 0.0;
  for (uint32_t k = lane; k < nt; k += 64) v += tile_partial[first + k];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane == 0) sums[ti] = v;
}

// state[0]: the work queue head, state[1]: the number of tiles handed to the
// slow kernel.  The persistent kernel holds no calls: a tile with an argument
// that needs the out-of-line fdlibm reduction (beyond 2^18*pi) is listed in
// `slow` and left to batch_slow_k, which runs after it.
template <class P, int BLOCK>
__global__ __launch_bounds__(BLOCK) void batch_persistent_k(
    const TableDesc* __restrict__ tables, const uint32_t* __restrict__ tile_table,
    uint32_t ntiles, uint32_t tile, unsigned* state, uint32_t* __restrict__ slow,
    double* __restrict__ tile_partial) {
  __shared__ unsigned cur;
  __shared__ double wsum[BLOCK / 64];
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(&state[0], 1u);
    __syncthreads();
    const unsigned t = cur;
    __syncthreads();
    if (t >= ntiles) break;  // uniform across the block: every wave exits
    const uint32_t ti = tile_table[t];
    const TableDesc d = tables[ti];
    const uint64_t start = (uint64_t)(t - d.first_tile) * tile;
    const uint64_t end = (start + tile < d.len) ? start + tile : d.len;
    const uint64_t full_end = start + ((end - start) / 8) * 8;
    float s = 0.0f, c = 0.0f;
    bool bail = false;
    tile_sum<P, BLOCK, true>(d, start, full_end, end, s, c, &bail);
    if (__syncthreads_or(bail)) {  // uniform: the whole tile goes to the slow kernel
      if (threadIdx.x == 0) slow[atomicAdd(&state[1], 1u)] = t;
      // A barrier between this thread-0 block and the one at the loop head:
      // without it the optimizer threads the two `threadIdx.x == 0` tests
      // into one path and makes the loop head's barrier an inner loop that
      // lanes 1..63 of wave 0 re-enter while lane 0 waits outside it.
      __syncthreads();
      continue;
    }
    finish_tile<BLOCK>(s, c, t, tile_partial, wsum);
  }
}

// the tiles listed by batch_persistent_k, with the out-of-line reduction
template <class P, int BLOCK>
__global__ __launch_bounds__(BLOCK) void batch_slow_k(
    const TableDesc* __restrict__ tables, const uint32_t* __restrict__ tile_table,
    uint32_t tile, const unsigned* state, const uint32_t* __restrict__ slow,
    double* __restrict__ tile_partial) {
  __shared__ double wsum[BLOCK / 64];
  const unsigned nslow = state[1];
  for (unsigned u = blockIdx.x; u < nslow; u += gridDim.x) {  // uniform across the block
    const unsigned t = slow[u];
    const uint32_t ti = tile_table[t];
    const TableDesc d = tables[ti];
    const uint64_t start = (uint64_t)(t - d.first_tile) * tile;
    const uint64_t end = (start + tile < d.len) ? start + tile : d.len;
    const uint64_t full_end = start + ((end - start) / 8) * 8;
    float s = 0.0f, c = 0.0f;
    tile_sum<P, BLOCK, false>(d, start, full_end, end, s, c, nullptr);
    finish_tile<BLOCK>(s, c, t, tile_partial, wsum);
  }
}

struct BatchConfig {
  int block;       // threads per workgroup
  int ept;         // elements per thread per tile (multiple of 8)
  int wg_per_cu;   // persistent workgroups per compute unit
};

struct Batch {
  int count = 0;
  int policy = NT2H_FAST;
  std::vector<const float*> hx, hy;
  std::vector<uint64_t> hlen;
  BatchConfig cfg{256, 16, 8};
  TableDesc* tables = nullptr;
  uint32_t* tile_table = nullptr;
  double* tile_partial = nullptr;
  uint32_t* slow = nullptr;     // tiles left to the slow kernel
  unsigned* state = nullptr;    // queue head, slow tile count
  uint32_t ntiles = 0;
  size_t tile_cap = 0;
  int cus = 256;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double* graph_sums = nullptr;
  hipStream_t graph_stream = nullptr;
};

static int layout(Batch* b) {
  const uint32_t tile = (uint32_t)(b->cfg.block * b->cfg.ept);
  std::vector<TableDesc> td(b->count);
  std::vector<uint32_t> tt;
  uint64_t total_tiles = 0;
  for (int i = 0; i < b->count; ++i) total_tiles += (b->hlen[i] + tile - 1) / tile;
  if (total_tiles >= 0xffffffffull) return fail(NT2H_EINVAL, "nt2h_batch: too many tiles");
  tt.reserve(total_tiles);
  for (int i = 0; i < b->count; ++i) {
    const uint32_t nt = (uint32_t)((b->hlen[i] + tile - 1) / tile);
    td[i] = TableDesc{b->hx[i], b->hy[i], b->hlen[i], (uint32_t)tt.size(), nt,
                      (uint32_t)(aligned16(b->hx[i]) && aligned16(b->hy[i])), 0u};
    for (uint32_t k = 0; k < nt; ++k) tt.push_back((uint32_t)i);
  }
  b->ntiles = (uint32_t)tt.size();
  if (!b->tables) NT2H_CHECK(hipMalloc(&b->tables, sizeof(TableDesc) * b->count), "batch alloc");
  if (!b->state) NT2H_CHECK(hipMalloc(&b->state, 2 * sizeof(unsigned)), "batch alloc");
  if (b->tile_cap < tt.size()) {
    if (b->tile_table) { NT2H_CHECK(hipFree(b->tile_table), "batch free"); b->tile_table = nullptr; }
    if (b->tile_partial) { NT2H_CHECK(hipFree(b->tile_partial), "batch free"); b->tile_partial = nullptr; }
    if (b->slow) { NT2H_CHECK(hipFree(b->slow), "batch free"); b->slow = nullptr; }
    const size_t cap = std::max<size_t>(tt.size(), 1);
    NT2H_CHECK(hipMalloc(&b->tile_table, sizeof(uint32_t) * cap), "batch alloc");
    NT2H_CHECK(hipMalloc(&b->tile_partial, sizeof(double) * cap), "batch alloc");
    NT2H_CHECK(hipMalloc(&b->slow, sizeof(uint32_t) * cap), "batch alloc");
    b->tile_cap = cap;
  }
  NT2H_CHECK(hipMemcpy(b->tables, td.data(), sizeof(TableDesc) * b->count, hipMemcpyHostToDevice), "batch upload");
  if (!tt.empty())
    NT2H_CHECK(hipMemcpy(b->tile_table, tt.data(), sizeof(uint32_t) * tt.size(), hipMemcpyHostToDevice), "batch upload");
  return NT2H_OK;
}

template <class P, int BLOCK>
static void launch_persistent(Batch* b, double* sums, hipStream_t s) {
  const unsigned grid = (unsigned)std::max(1, std::min<int>(b->cus * b->cfg.wg_per_cu, (int)b->ntiles));
  const uint32_t tile = (uint32_t)(BLOCK * b->cfg.ept);
  hipLaunchKernelGGL((batch_persistent_k<P, BLOCK>), dim3(grid), dim3(BLOCK), 0, s, b->tables,
                     b->tile_table, b->ntiles, tile, b->state, b->slow, b->tile_partial);
  const unsigned sgrid = (unsigned)std::max(1, std::min<int>(b->cus, (int)b->ntiles));
  hipLaunchKernelGGL((batch_slow_k<P, BLOCK>), dim3(sgrid), dim3(BLOCK), 0, s, b->tables, b->tile_table,
                     tile, b->state, b->slow, b->tile_partial);
  hipLaunchKernelGGL(batch_sums_k, dim3((b->count + 3) / 4), dim3(256), 0, s, b->tables, b->count,
                     b->tile_partial, sums);
}

static int enqueue(Batch* b, double* sums, hipStream_t s) {
  hipLaunchKernelGGL(batch_reset_k, dim3(1), dim3(64), 0, s, b->state);
  const bool fast = b->policy == NT2H_FAST;
  switch (b->cfg.block) {
    case 128: fast ? launch_persistent<fast_policy, 128>(b, sums, s) : launch_persistent<exact_policy, 128>(b, sums, s); break;
    case 256: fast ? launch_persistent<fast_policy, 256>(b, sums, s) : launch_persistent<exact_policy, 256>(b, sums, s); break;
    case 512: fast ? launch_persistent<fast_policy, 512>(b, sums, s) : launch_persistent<exact_policy, 512>(b, sums, s); break;
    default: return fail(NT2H_EINVAL, "nt2h_batch: unsupported block size");
  }
  NT2H_CHECK_LAUNCH("batch_persistent");
  return NT2H_OK;
}

static int drop_graph(Batch* b) {
  if (b->exec) { NT2H_CHECK(hipGraphExecDestroy(b->exec), "graph destroy"); b->exec = nullptr; }
  if (b->graph) { NT2H_CHECK(hipGraphDestroy(b->graph), "graph destroy"); b->graph = nullptr; }
  b->graph_sums = nullptr;
  return NT2H_OK;
}

// capture reset + persistent launch once per (config, output buffer)
static int ensure_graph(Batch* b, double* sums) {
  if (b->exec && b->graph_sums == sums) return NT2H_OK;
  int rc = drop_graph(b);
  if (rc) return rc;
  if (!b->graph_stream) NT2H_CHECK(hipStreamCreateWithFlags(&b->graph_stream, hipStreamNonBlocking), "graph stream");
  NT2H_CHECK(hipStreamBeginCapture(b->graph_stream, hipStreamCaptureModeThreadLocal), "begin capture");
  rc = enqueue(b, sums, b->graph_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->graph_stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return hip_fail(e, "end capture");
  b->graph = g;
  NT2H_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
  b->graph_sums = sums;
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

/* Create a batch over `count` tables (host arrays of device pointers and
 * lengths).  The tables must stay allocated while the batch is used. */
int nt2h_batch_create(nt2h_batch** out, int count, const float* const* x, const float* const* y,
                      const size_t* lens, int policy) {
  if (!out || count <= 0 || !x || !y || !lens) return fail(NT2H_EINVAL, "nt2h_batch_create: bad argument");
  if (policy != NT2H_EXACT && policy != NT2H_FAST) return fail(NT2H_EINVAL, "nt2h_batch_create: bad policy");
  Batch* b = new Batch();
  b->count = count;
  b->policy = policy;
  for (int i = 0; i < count; ++i) {
    if ((!x[i] || !y[i]) && lens[i]) { delete b; return fail(NT2H_EINVAL, "nt2h_batch_create: null table"); }
    b->hx.push_back(x[i]);
    b->hy.push_back(y[i]);
    b->hlen.push_back(lens[i]);
  }
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
    b->cus = p.multiProcessorCount;
  int rc = layout(b);
  if (rc) { delete b; return rc; }
  *out = reinterpret_cast<nt2h_batch*>(b);
  return NT2H_OK;
}

int nt2h_batch_set_config(nt2h_batch* hb, int block, int ept, int wg_per_cu) {
  Batch* b = reinterpret_cast<Batch*>(hb);
  if (!b || (block != 128 && block != 256 && block != 512) || ept < 8 || ept % 8 || ept > 256 ||
      wg_per_cu < 1 || wg_per_cu > 32)
    return fail(NT2H_EINVAL, "nt2h_batch_set_config: bad argument");
  b->cfg = BatchConfig{block, ept, wg_per_cu};
  int rc = drop_graph(b);
  if (rc) return rc;
  return layout(b);
}

int nt2h_batch_get_config(nt2h_batch* hb, int* block, int* ept, int* wg_per_cu) {
  Batch* b = reinterpret_cast<Batch*>(hb);
  if (!b) return fail(NT2H_EINVAL, "nt2h_batch_get_config: null batch");
  if (block) *block = b->cfg.block;
  if (ept) *ept = b->cfg.ept;
  if (wg_per_cu) *wg_per_cu = b->cfg.wg_per_cu;
  return NT2H_OK;
}

/* sums: device array of `count` doubles.  Replays the captured graph on
 * `stream` (captured on a private stream the first time, and again when the
 * configuration or the output buffer changes). */
int nt2h_batch_run(nt2h_batch* hb, double* sums, void* stream) {
  Batch* b = reinterpret_cast<Batch*>(hb);
  if (!b || !sums) return fail(NT2H_EINVAL, "nt2h_batch_run: bad argument");
  int rc = ensure_graph(b, sums);
  if (rc) return rc;
  NT2H_CHECK(hipGraphLaunch(b->exec, as_stream(stream)), "graph launch");
  return NT2H_OK;
}

/* The plain schedule: one reduction launch per table. */
int nt2h_batch_run_plain(nt2h_batch* hb, double* sums, void* stream) {
  Batch* b = reinterpret_cast<Batch*>(hb);
  if (!b || !sums) return fail(NT2H_EINVAL, "nt2h_batch_run_plain: bad argument");
  for (int i = 0; i < b->count; ++i) {
    int rc = nt2h_expr4_sum_f32(b->hlen[i], b->hx[i], b->hy[i], sums + i, b->policy, stream);
    if (rc) return rc;
  }
  return NT2H_OK;
}

/* Time each candidate configuration (3 graph replays, HIP events) and keep
 * the fastest; *best_ms (optional, host) receives its time per run. */
int nt2h_batch_autotune(nt2h_batch* hb, double* sums, double* best_ms) {
  Batch* b = reinterpret_cast<Batch*>(hb);
  if (!b || !sums) return fail(NT2H_EINVAL, "nt2h_batch_autotune: bad argument");
  const BatchConfig cands[] = {{256, 16, 4}, {256, 16, 8}, {256, 32, 8}, {256, 64, 4},
                               {512, 16, 4}, {512, 32, 2}, {128, 32, 8}, {256, 8, 8}};
  hipStream_t s;
  NT2H_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "autotune stream");
  hipEvent_t e0, e1;
  NT2H_CHECK(hipEventCreate(&e0), "event");
  NT2H_CHECK(hipEventCreate(&e1), "event");
  float best = 1e30f;
  BatchConfig bestc = b->cfg;
  int rc = NT2H_OK;
  for (const BatchConfig& c : cands) {
    if ((rc = nt2h_batch_set_config(hb, c.block, c.ept, c.wg_per_cu))) break;
    if ((rc = nt2h_batch_run(hb, sums, s))) break;  // warm-up, capture
    NT2H_CHECK(hipEventRecord(e0, s), "event");
    for (int r = 0; r < 3; ++r)
      if ((rc = nt2h_batch_run(hb, sums, s))) break;
    if (rc) break;
    NT2H_CHECK(hipEventRecord(e1, s), "event");
    NT2H_CHECK(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    NT2H_CHECK(hipEventElapsedTime(&ms, e0, e1), "event");
    ms /= 3.0f;
    if (ms < best) { best = ms; bestc = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  rc = nt2h_batch_set_config(hb, bestc.block, bestc.ept, bestc.wg_per_cu);
  if (best_ms) *best_ms = best;
  return rc;
}

int nt2h_batch_destroy(nt2h_batch* hb) {
  Batch* b = reinterpret_cast<Batch*>(hb);
  if (!b) return NT2H_OK;
  (void)hipDeviceSynchronize();
  drop_graph(b);
  if (b->graph_stream) (void)hipStreamDestroy(b->graph_stream);
  (void)hipFree(b->tables);
  (void)hipFree(b->tile_table);
  (void)hipFree(b->tile_partial);
  (void)hipFree(b->slow);
  (void)hipFree(b->state);
  delete b;
  return NT2H_OK;
}

}  // extern "C"
