// Serving engine for Black-Scholes pricing requests with continuous batching.
//
// Requests (f32 or f64 tables of S, X, T resident in device memory, scalars
// r and v) arrive at any time.  The scheduler thread admits them in arrival
// order against a device-memory budget (input tables plus the output tables
// the server allocates), and whenever a launch slot is free it packs every
// admitted request that is not yet running into one batched launch: the
// kernel's blocks are (request, tile) pairs found by a binary search over the
// batch's tile prefix, and f32 and f64 requests share the launch.  Up to
// max_inflight batches run concurrently on separate streams, so requests that
// arrive while others are on the GPU overlap with them.  A completion thread
// waits on each batch's event in launch order and stamps the completion time;
// latency = completion - arrival.  The plain schedule (one launch and one
// wait per request, in arrival order) is the same engine with one slot and
// batches of one.
//
// Every request is evaluated per option with the same device code as the
// standalone kernel (bs_device.hpp) and the same NT2 pack grouping from the
// request's first option, so its results are bit-identical alone or in a mix.
//
// No counterpart in the reference: NT2 evaluates statements synchronously on
// the calling thread.  A host back end (config.backend = NT2H_BACKEND_HOST)
// runs the same scheduler on the CPU with host memory, for testing admission,
// budget, failure and cancellation handling without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bs_device.hpp"
#include "nt2/hip/bs_accurate.hpp"
#include "nt2/hip/bs_accurate64.hpp"
#include "common.hpp"
#include "nt2h.h"

#pragma clang fp contract(off)

using nt2::hip::exact_policy;
using nt2::hip::fast_policy;

namespace nt2h {

using clk = std::chrono::steady_clock;
static inline int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// batched kernel
// ---------------------------------------------------------------------------
struct ReqDesc {
  const void* S;
  const void* X;
  const void* T;
  void* call;
  void* put;
  uint64_t n;
  double r, v;
  uint32_t tile0;  // first tile of this request in the batch
  uint32_t f64;
};

constexpr int kSrvBlock = 256;
constexpr int kSrvTile = kSrvBlock * 4;  // 4 options per thread
constexpr uint32_t kSrvGridMax = 256 * 32;  // workgroups of a batch launch (32 per CU)

typedef float srv_f32x4 __attribute__((ext_vector_type(4)));

// One launch prices every request of one precision in a batch (a batch is
// launched as its f32 group then its f64 group, so that each kernel holds the
// registers of its own evaluation only).  The group's options are cut in tiles
// of kSrvTile (a request's tiles are consecutive, starting at tile0);
// workgroup k walks the contiguous tile range [k tpb, (k+1) tpb), so the
// owning request is found by one binary search and then advances linearly.
// P: the NT2 policy of the requests when ACC is false; ACC: the accurate
// evaluations (FP64-internal for f32 requests, double-double tails for f64
// ones); F64: the group's precision.
template <class P, bool ACC, bool F64>
__global__ __launch_bounds__(kSrvBlock) void serve_bs_k(const ReqDesc* __restrict__ reqs, int nreq,
                                                      uint32_t ntiles, uint32_t tpb) {
  _Pragma("clang fp contract(off)")
  namespace ba = nt2::hip::bsacc;
  namespace bb = nt2::hip::bsacc64;
  constexpr bool kTables32 = ACC && !F64, kTables64 = ACC && F64;
  __shared__ double tab[kTables32 ? (ba::kTailDeg + 1) * ba::kTailN : (kTables64 ? bb::kTabDoubles : 1)];
  __shared__ double ex2[kTables32 ? 32 : 1];
  if (kTables32) ba::load_tables(tab, ex2);
  if (kTables64) bb::load_tables(tab);
  const uint32_t b0 = blockIdx.x * tpb;
  const uint32_t b1 = min(b0 + tpb, ntiles);
  if (b0 >= b1) return;
  int lo = 0, hi = nreq - 1;  // last request with tile0 <= b0
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (reqs[mid].tile0 <= b0) lo = mid; else hi = mid - 1;
  }
  for (uint32_t b = b0; b < b1; ++b) {
    while (lo + 1 < nreq && reqs[lo + 1].tile0 <= b) ++lo;
    const ReqDesc& d = reqs[lo];
    const uint64_t base = (uint64_t)(b - d.tile0) * kSrvTile;
    if constexpr (!F64) {
      const float* S = (const float*)d.S;
      const float* X = (const float*)d.X;
      const float* T = (const float*)d.T;
      float* C = (float*)d.call;
      float* Pt = (float*)d.put;
      const float r = (float)d.r, v = (float)d.v;
      const bool vec = ((reinterpret_cast<uintptr_t>(S) | reinterpret_cast<uintptr_t>(X) |
                         reinterpret_cast<uintptr_t>(T) | reinterpret_cast<uintptr_t>(C) |
                         reinterpret_cast<uintptr_t>(Pt)) & 15) == 0;
      const uint64_t i4 = base + 4 * threadIdx.x;
      if constexpr (ACC) {
        const ba::Params prm = ba::make_params(r, v);
        if (vec && i4 + 4 <= d.n) {
          const srv_f32x4 s4 = __builtin_nontemporal_load(reinterpret_cast<const srv_f32x4*>(S + i4));
          const srv_f32x4 x4 = __builtin_nontemporal_load(reinterpret_cast<const srv_f32x4*>(X + i4));
          const srv_f32x4 t4 = __builtin_nontemporal_load(reinterpret_cast<const srv_f32x4*>(T + i4));
          float c[4], p[4];
          ba::bs_accurate4(s4, x4, t4, prm, tab, ex2, c, p);
          __builtin_nontemporal_store(srv_f32x4{c[0], c[1], c[2], c[3]}, reinterpret_cast<srv_f32x4*>(C + i4));
          __builtin_nontemporal_store(srv_f32x4{p[0], p[1], p[2], p[3]}, reinterpret_cast<srv_f32x4*>(Pt + i4));
        } else if (vec) {  // the request's last, partial float4: this thread's options only
          for (uint64_t i = i4; i < i4 + 4 && i < d.n; ++i)
            ba::bs_accurate(S[i], X[i], T[i], prm, tab, ex2, C[i], Pt[i]);
        } else {
          for (uint64_t i = base + threadIdx.x; i < base + kSrvTile && i < d.n; i += kSrvBlock)
            ba::bs_accurate(S[i], X[i], T[i], prm, tab, ex2, C[i], Pt[i]);
        }
      } else {
        const float drift = (v * v) * 0.5f + r, mr = -r;
        for (uint64_t i = base + threadIdx.x; i < base + kSrvTile && i < d.n; i += kSrvBlock) {
          float c, p;
          bs_one_f32<P, NT2H_BS_TEXTBOOK>(S[i], X[i], T[i], mr, v, drift, c, p);
          C[i] = c;
          Pt[i] = p;
        }
      }
    } else {
      const double* S = (const double*)d.S;
      const double* X = (const double*)d.X;
      const double* T = (const double*)d.T;
      double* C = (double*)d.call;
      double* Pt = (double*)d.put;
      const double r = d.r, v = d.v;
      if constexpr (ACC) {
        const bb::Params prm = bb::make_params(r, v);
        const double* ex2h = tab + bb::kGRows * bb::kGN;
        for (uint64_t i = base + threadIdx.x; i < base + kSrvTile && i < d.n; i += kSrvBlock) {
          double c, p;
          bb::bs_price(S[i], X[i], T[i], prm, tab, ex2h, ex2h + 32, c, p);
          C[i] = c;
          Pt[i] = p;
        }
      } else {
        const double drift = (v * v) * 0.5 + r, mr = -r;
        // NT2 packs of 4 from the request's first option; its last n%4 options
        // take NT2's scalar erfc
        const uint64_t npk = d.n / 4 * 4;
        for (uint64_t i = base + threadIdx.x; i < base + kSrvTile && i < d.n; i += kSrvBlock) {
          double c, p;
          if (i < npk) bs_one_f64<P, false>(S[i], X[i], T[i], mr, v, drift, c, p);
          else bs_one_f64<P, true>(S[i], X[i], T[i], mr, v, drift, c, p);
          C[i] = c;
          Pt[i] = p;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host back end (CPU): textbook formula in the request's precision with the
// C library; used to exercise the scheduler without a GPU
// ---------------------------------------------------------------------------
template <class T>
static void host_bs(const ReqDesc& d) {
  const T* S = (const T*)d.S;
  const T* X = (const T*)d.X;
  const T* Tt = (const T*)d.T;
  T* C = (T*)d.call;
  T* Pp = (T*)d.put;
  const T r = (T)d.r, v = (T)d.v;
  for (uint64_t i = 0; i < d.n; ++i) {
    const T vst = v * std::sqrt(Tt[i]);
    const T d1 = (std::log(S[i] / X[i]) + (v * v * (T)0.5 + r) * Tt[i]) / vst;
    const T d2 = d1 - vst;
    const T e = X[i] * std::exp(-r * Tt[i]);
    const T s2 = (T)0.70710678118654752440;
    C[i] = S[i] * (T)0.5 * std::erfc(-d1 * s2) - e * (T)0.5 * std::erfc(-d2 * s2);
    Pp[i] = e * (T)0.5 * std::erfc(d2 * s2) - S[i] * (T)0.5 * std::erfc(d1 * s2);
  }
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
enum ReqState { RS_PENDING = 0, RS_ADMITTED = 1, RS_RUNNING = 2, RS_DONE = 3, RS_FAILED = 4,
                RS_CANCELLED = 5 };

// One device allocation holds the outputs of every request of a batch that
// did not bring its own; it is returned when the last of them is released.
struct OutBlock {
  void* base = nullptr;
  int refs = 0;
  int cls = 0;  // size class: 2^cls bytes
};

struct Request {
  uint64_t id;
  ReqDesc d;
  size_t bytes;        // budget charge: inputs + outputs
  bool charged;        // `bytes` is counted in the server's used budget
  bool own_out;        // outputs allocated by the server (in blk)
  OutBlock* blk = nullptr;
  int state;
  int status;
  int64_t t_arrival, t_admit, t_launch, t_done;
  int slot;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
  ReqDesc* d_desc = nullptr;   // device descriptors
  ReqDesc* h_desc = nullptr;   // pinned staging
  int cap = 0;
  bool busy = false;
  bool launched = false;       // a batch is recorded on the stream, awaiting completion
  bool fail_next = false;      // test hook: report this slot's launched batch as failed
  int64_t t_freed = 0, t_launched = 0;  // timing breakdown (NT2H_SERVER_TIMING)
  std::vector<uint64_t> reqs;
  std::thread host_worker;
  std::thread done_thr;        // this slot's completion thread
};

struct Server {
  nt2h_server_config cfg;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_slot;
  std::deque<uint64_t> pending;            // arrival order
  std::deque<uint64_t> admitted;           // admitted, not yet launched
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::vector<Slot> slots;
  size_t used = 0, peak_used = 0;
  uint64_t next_id = 1;
  bool stop = false;
  bool sched_exited = false;
  std::thread sched;
  int alloc_count = 0;
  int launch_count = 0;
  // Output blocks are cached on the host side in power-of-two size classes
  // (from 1 MiB): a block whose requests are all released (their batch has
  // completed, so the device no longer touches it) goes back to its class and
  // the next batch of that class takes it without any HIP call.  Only a class
  // with no free block reaches hipMalloc, and blocks go back to the device
  // (hipFree) only while the cache holds more than the budget.  (Stream-ordered
  // pool allocations cost ~14 us per hipFreeAsync and slowed every batch.)
  std::map<int, std::vector<void*>> free_blocks;
  size_t cached_bytes = 0;
  // statistics
  std::vector<double> lat_us;
  uint64_t options_done = 0, launches = 0, batched_reqs = 0;
  int64_t t_first_arrival = -1, t_last_done = -1;
  uint64_t failed = 0, cancelled = 0;
  // per-batch time breakdown, printed at destroy when NT2H_SERVER_TIMING is set:
  // slot freed -> next launch, host launch work, launch -> completion seen,
  // completion seen -> lock taken (ns totals over `timed` batches)
  int64_t tm_idle = 0, tm_launch = 0, tm_gpu = 0, tm_lock = 0;
  uint64_t timed = 0;
};

static size_t elem(const ReqDesc& d) { return d.f64 ? 8 : 4; }

static int class_of(size_t bytes) {
  int c = 20;
  while ((size_t(1) << c) < bytes) ++c;
  return c;
}

static void block_release(Server* s, void* p) {
  if (s->cfg.backend == NT2H_BACKEND_HOST) std::free(p);
  else (void)hipFree(p);
}

// a block of at least `bytes` (class *cls); called with mu held
static int srv_alloc(Server* s, size_t bytes, void** p, int* cls) {
  ++s->alloc_count;
  if (s->cfg.fail_alloc_every > 0 && s->alloc_count % s->cfg.fail_alloc_every == 0)
    return NT2H_ENOMEM;
  const int c = class_of(bytes);
  *cls = c;
  auto it = s->free_blocks.find(c);
  if (it != s->free_blocks.end() && !it->second.empty()) {
    *p = it->second.back();
    it->second.pop_back();
    s->cached_bytes -= size_t(1) << c;
    return NT2H_OK;
  }
  if (s->cfg.backend == NT2H_BACKEND_HOST) {
    *p = std::malloc(size_t(1) << c);
    return *p ? NT2H_OK : NT2H_ENOMEM;
  }
  if (hipMalloc(p, size_t(1) << c) != hipSuccess) {
    (void)hipGetLastError();
    // give the cached blocks back and try once more
    for (auto& kv : s->free_blocks)
      for (void* q : kv.second) block_release(s, q);
    s->free_blocks.clear();
    s->cached_bytes = 0;
    if (hipMalloc(p, size_t(1) << c) != hipSuccess) {
      (void)hipGetLastError();
      return NT2H_ENOMEM;
    }
  }
  return NT2H_OK;
}

// drop the request's reference to its output block; the last reference puts
// the block back in the cache.  Called with mu held, once the device is done
// with the block.
static void srv_free(Server* s, Request* r) {
  if (!r->own_out) return;
  OutBlock* b = r->blk;
  if (--b->refs == 0) {
    s->free_blocks[b->cls].push_back(b->base);
    s->cached_bytes += size_t(1) << b->cls;
    while (s->cached_bytes > s->cfg.budget_bytes) {  // trim the largest classes first
      auto it = s->free_blocks.rbegin();
      while (it != s->free_blocks.rend() && it->second.empty()) ++it;
      if (it == s->free_blocks.rend()) break;
      block_release(s, it->second.back());
      it->second.pop_back();
      s->cached_bytes -= size_t(1) << it->first;
    }
    delete b;
  }
  r->blk = nullptr;
  r->d.call = r->d.put = nullptr;
  r->own_out = false;
}

// a request that will not complete: its outputs go back to the cache and its
// bytes to the budget; called with mu held
static void fail_request(Server* s, Request* r, int status, int64_t t) {
  r->state = RS_FAILED;
  r->status = status;
  r->t_done = t;
  srv_free(s, r);
  if (r->charged) {
    s->used -= r->bytes;
    r->charged = false;
  }
  ++s->failed;
}

// one precision group of a batch: nr descriptors (device) covering `tiles` tiles
template <class P, bool ACC>
static hipError_t launch_group(bool f64, const ReqDesc* dd, int nr, uint32_t tiles, hipStream_t st) {
  const uint32_t grid = std::min<uint32_t>(tiles, kSrvGridMax);
  const uint32_t tpb = (tiles + grid - 1) / grid;
  if (f64) hipLaunchKernelGGL((serve_bs_k<P, ACC, true>), dim3(grid), dim3(kSrvBlock), 0, st, dd, nr, tiles, tpb);
  else hipLaunchKernelGGL((serve_bs_k<P, ACC, false>), dim3(grid), dim3(kSrvBlock), 0, st, dd, nr, tiles, tpb);
  return hipGetLastError();
}

// launch every admitted request (up to max_batch) on a free slot; called with mu held
static void launch_batch(Server* s, int si, std::unique_lock<std::mutex>& lk) {
  Slot& sl = s->slots[si];
  const int64_t t_start = now_ns();
  if (sl.t_freed) s->tm_idle += t_start - sl.t_freed;
  std::vector<Request*> batch;
  while (!s->admitted.empty() && (int)batch.size() < s->cfg.max_batch) {
    Request* r = s->reqs[s->admitted.front()].get();
    s->admitted.pop_front();
    if (r->state != RS_ADMITTED) continue;
    batch.push_back(r);
  }
  if (batch.empty()) return;
  // output tables: one allocation for the batch, 256-byte aligned call and put per request
  auto out_bytes = [](const Request* r) { return (r->d.n * elem(r->d) + 255) / 256 * 256; };
  size_t total = 0;
  int owners = 0;
  for (Request* r : batch) {
    r->slot = si;
    if (!r->d.call) { total += 2 * out_bytes(r); ++owners; }
  }
  void* base = nullptr;
  int cls = 0;
  const int arc = owners ? srv_alloc(s, total, &base, &cls) : NT2H_OK;
  OutBlock* blk = nullptr;
  if (owners && arc == NT2H_OK) {
    blk = new OutBlock();
    blk->base = base;
    blk->refs = owners;
    blk->cls = cls;
  }
  std::vector<Request*> ok;
  size_t off = 0;
  for (Request* r : batch) {
    if (!r->d.call) {
      if (arc) {
        fail_request(s, r, arc, now_ns());
        continue;
      }
      r->d.call = (char*)base + off;
      r->d.put = (char*)base + off + out_bytes(r);
      off += 2 * out_bytes(r);
      r->own_out = true;
      r->blk = blk;
    }
    ok.push_back(r);
  }
  s->cv_done.notify_all();
  s->cv_slot.notify_all();
  if (ok.empty()) return;
  if ((int)ok.size() > sl.cap) {
    int drc = NT2H_OK;
    if (s->cfg.backend == NT2H_BACKEND_HOST) {
      delete[] sl.h_desc;
      sl.h_desc = new ReqDesc[ok.size()];
    } else {
      // the slot is idle (its last batch completed), so its old descriptors can go
      if (sl.h_desc) (void)hipHostFree(sl.h_desc);
      if (sl.d_desc) (void)hipFree(sl.d_desc);
      sl.h_desc = nullptr;
      sl.d_desc = nullptr;
      if (hipHostMalloc((void**)&sl.h_desc, sizeof(ReqDesc) * ok.size(), 0) != hipSuccess ||
          hipMalloc((void**)&sl.d_desc, sizeof(ReqDesc) * ok.size()) != hipSuccess) {
        (void)hipGetLastError();
        if (sl.h_desc) (void)hipHostFree(sl.h_desc);
        sl.h_desc = nullptr;
        sl.d_desc = nullptr;
        drc = NT2H_ENOMEM;
      }
    }
    if (drc) {
      sl.cap = 0;
      const int64_t t = now_ns();
      for (Request* r : ok) fail_request(s, r, drc, t);
      s->cv_done.notify_all();
      s->cv_slot.notify_all();
      return;
    }
    sl.cap = (int)ok.size();
  }
  // the f32 group first, then the f64 group, each in arrival order; tiles are
  // numbered within a group
  std::stable_partition(ok.begin(), ok.end(), [](const Request* r) { return !r->d.f64; });
  uint32_t tiles[2] = {0u, 0u};
  int ngroup[2] = {0, 0};
  const int64_t t = now_ns();
  sl.reqs.clear();
  for (size_t k = 0; k < ok.size(); ++k) {
    Request* r = ok[k];
    const int g = r->d.f64 ? 1 : 0;
    r->d.tile0 = tiles[g];
    tiles[g] += (uint32_t)((r->d.n + kSrvTile - 1) / kSrvTile);
    ++ngroup[g];
    sl.h_desc[k] = r->d;
    r->state = RS_RUNNING;
    r->t_launch = t;
    sl.reqs.push_back(r->id);
  }
  sl.busy = true;
  ++s->launches;
  s->batched_reqs += ok.size();
  ++s->launch_count;
  sl.fail_next = s->cfg.fail_launch_every > 0 && s->launch_count % s->cfg.fail_launch_every == 0;
  const int nreq = (int)ok.size();
  if (s->cfg.backend == NT2H_BACKEND_HOST) {
    std::vector<ReqDesc> descs(sl.h_desc, sl.h_desc + nreq);
    if (sl.host_worker.joinable()) sl.host_worker.join();
    sl.host_worker = std::thread([descs]() {
      for (const ReqDesc& d : descs) d.f64 ? host_bs<double>(d) : host_bs<float>(d);
    });
  } else {
    lk.unlock();
    hipError_t e = hipMemcpyAsync(sl.d_desc, sl.h_desc, sizeof(ReqDesc) * nreq, hipMemcpyHostToDevice, sl.stream);
    for (int g = 0; g < 2 && e == hipSuccess; ++g) {
      if (!ngroup[g]) continue;
      const ReqDesc* dd = sl.d_desc + (g ? ngroup[0] : 0);
      if (s->cfg.policy == NT2H_EXACT) e = launch_group<exact_policy, false>(g == 1, dd, ngroup[g], tiles[g], sl.stream);
      else if (s->cfg.policy == NT2H_FAST) e = launch_group<fast_policy, false>(g == 1, dd, ngroup[g], tiles[g], sl.stream);
      else e = launch_group<exact_policy, true>(g == 1, dd, ngroup[g], tiles[g], sl.stream);
    }
    if (e == hipSuccess) e = hipEventRecord(sl.event, sl.stream);
    lk.lock();
    if (e != hipSuccess) {
      // nothing will signal this batch: wait for whatever did launch, then
      // fail its requests (their output block goes back to the cache)
      (void)hipStreamSynchronize(sl.stream);
      (void)hipGetLastError();
      set_error(std::string("nt2h_server: batch launch failed: ") + hipGetErrorString(e));
      const int64_t t = now_ns();
      for (uint64_t id : sl.reqs) fail_request(s, s->reqs[id].get(), NT2H_EHIP, t);
      sl.reqs.clear();
      sl.busy = false;
      sl.fail_next = false;
      s->cv_done.notify_all();
      s->cv_slot.notify_all();
      return;
    }
  }
  // publish to the slot's completion thread only once the batch's event is recorded
  sl.t_launched = now_ns();
  s->tm_launch += sl.t_launched - t_start;
  ++s->timed;
  sl.launched = true;
  s->cv_done.notify_all();
}

static void sched_loop(Server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // admission in arrival order against the budget
    while (!s->pending.empty()) {
      Request* r = s->reqs[s->pending.front()].get();
      if (r->state == RS_CANCELLED) { s->pending.pop_front(); continue; }
      if (s->used + r->bytes > s->cfg.budget_bytes) break;
      s->pending.pop_front();
      s->used += r->bytes;
      r->charged = true;
      s->peak_used = std::max(s->peak_used, s->used);
      r->state = RS_ADMITTED;
      r->t_admit = now_ns();
      s->admitted.push_back(r->id);
    }
    int free_slot = -1;
    for (int i = 0; i < (int)s->slots.size(); ++i)
      if (!s->slots[i].busy) { free_slot = i; break; }
    if (!s->admitted.empty() && free_slot >= 0) {
      launch_batch(s, free_slot, lk);
      continue;
    }
    if (s->stop && s->pending.empty() && s->admitted.empty()) break;
    s->cv_sched.wait(lk);
  }
  s->sched_exited = true;
  s->cv_done.notify_all();
}

// One completion thread per slot, so that a long batch on one slot never
// holds back the completion (and the reuse) of a shorter one on another.
static void done_loop(Server* s, int si) {
  Slot& sl = s->slots[si];
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    while (!sl.launched && !s->sched_exited) s->cv_done.wait(lk);
    if (!sl.launched && s->sched_exited) break;
    lk.unlock();
    int status = NT2H_OK;
    if (s->cfg.backend == NT2H_BACKEND_HOST) {
      if (sl.host_worker.joinable()) sl.host_worker.join();
    } else {
      // poll the event (a blocking hipEventSynchronize sleeps in the driver and
      // wakes up ~100 us late, which a serving loop cannot afford)
      hipError_t e;
      while ((e = hipEventQuery(sl.event)) == hipErrorNotReady) std::this_thread::yield();
      if (e != hipSuccess) status = NT2H_EHIP;
    }
    const int64_t t = now_ns();
    lk.lock();
    s->tm_gpu += t - sl.t_launched;
    s->tm_lock += now_ns() - t;
    if (sl.fail_next) status = NT2H_EHIP;  // test hook
    sl.fail_next = false;
    sl.launched = false;
    for (uint64_t id : sl.reqs) {
      Request* r = s->reqs[id].get();
      if (status != NT2H_OK) {
        fail_request(s, r, status, t);
        continue;
      }
      r->state = RS_DONE;
      r->status = NT2H_OK;
      r->t_done = t;
      s->lat_us.push_back((r->t_done - r->t_arrival) * 1e-3);
      s->options_done += r->d.n;
      s->t_last_done = std::max(s->t_last_done, t);
    }
    sl.reqs.clear();
    sl.busy = false;
    sl.t_freed = now_ns();
    s->cv_sched.notify_all();
    s->cv_slot.notify_all();
  }
}

static int validate(const nt2h_bs_request* q) {
  if (!q) return NT2H_EINVAL;
  if (q->n == 0 || !q->S || !q->X || !q->T) return NT2H_EINVAL;
  if (q->dtype != NT2H_F32 && q->dtype != NT2H_F64) return NT2H_EINVAL;
  if ((q->call == nullptr) != (q->put == nullptr)) return NT2H_EINVAL;
  if (!(q->v > 0.0) || !std::isfinite(q->r) || !std::isfinite(q->v)) return NT2H_EINVAL;
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_server_create(const nt2h_server_config* cfg, nt2h_server** out) {
  if (!cfg || !out || cfg->budget_bytes == 0 || cfg->max_inflight < 1 || cfg->max_inflight > 16 ||
      cfg->max_batch < 1 || (cfg->policy != NT2H_EXACT && cfg->policy != NT2H_FAST && cfg->policy != NT2H_ACCURATE) ||
      cfg->fail_alloc_every < 0 || cfg->fail_launch_every < 0 ||
      (cfg->backend != NT2H_BACKEND_HIP && cfg->backend != NT2H_BACKEND_HOST))
    return fail(NT2H_EINVAL, "nt2h_server_create: bad configuration");
  Server* s = new Server();
  s->cfg = *cfg;
  s->slots.resize(cfg->max_inflight);
  if (cfg->backend == NT2H_BACKEND_HIP) {
    for (Slot& sl : s->slots) {
      if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&sl.event, hipEventDisableTiming) != hipSuccess) {
        delete s;
        return fail(NT2H_EHIP, "nt2h_server_create: stream/event creation failed");
      }
    }
  }
  s->sched = std::thread(sched_loop, s);
  for (int i = 0; i < (int)s->slots.size(); ++i) s->slots[i].done_thr = std::thread(done_loop, s, i);
  *out = reinterpret_cast<nt2h_server*>(s);
  return NT2H_OK;
}

int nt2h_server_submit(nt2h_server* hs, const nt2h_bs_request* q, uint64_t* id) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s || !id) return fail(NT2H_EINVAL, "nt2h_server_submit: null argument");
  if (validate(q)) return fail(NT2H_EINVAL, "nt2h_server_submit: malformed request");
  const size_t es = q->dtype == NT2H_F64 ? 8 : 4;
  const size_t bytes = 5 * q->n * es;
  if (bytes > s->cfg.budget_bytes) return fail(NT2H_EBUDGET, "nt2h_server_submit: request exceeds the memory budget");
  auto r = std::make_unique<Request>();
  r->d = ReqDesc{q->S, q->X, q->T, q->call, q->put, q->n, q->r, q->v, 0u,
                 (uint32_t)(q->dtype == NT2H_F64)};
  r->bytes = bytes;
  r->charged = false;
  r->own_out = false;
  r->state = RS_PENDING;
  r->status = NT2H_EAGAIN;
  r->t_arrival = q->t_arrival_ns > 0 ? q->t_arrival_ns : now_ns();
  r->t_admit = r->t_launch = r->t_done = 0;
  r->slot = -1;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(NT2H_ESTATE, "nt2h_server_submit: server is shutting down");
  r->id = s->next_id++;
  *id = r->id;
  if (s->t_first_arrival < 0) s->t_first_arrival = r->t_arrival;
  s->pending.push_back(r->id);
  s->reqs[r->id] = std::move(r);
  s->cv_sched.notify_all();
  return NT2H_OK;
}

/* 0 when done; NT2H_EAGAIN while pending or running; the failure code otherwise */
int nt2h_server_poll(nt2h_server* hs, uint64_t id) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(NT2H_EINVAL, "nt2h_server_poll: unknown request");
  const Request* r = it->second.get();
  if (r->state == RS_DONE) return NT2H_OK;
  if (r->state == RS_FAILED) return r->status;
  if (r->state == RS_CANCELLED) return NT2H_ECANCELLED;
  return NT2H_EAGAIN;
}

int nt2h_server_wait(nt2h_server* hs, uint64_t id, double timeout_s) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(NT2H_EINVAL, "nt2h_server_wait: unknown request");
  Request* r = it->second.get();
  const auto deadline = clk::now() + std::chrono::duration<double>(timeout_s);
  while (r->state < RS_DONE) {
    if (s->cv_slot.wait_until(lk, deadline) == std::cv_status::timeout && r->state < RS_DONE)
      return NT2H_EAGAIN;
  }
  if (r->state == RS_DONE) return NT2H_OK;
  if (r->state == RS_CANCELLED) return NT2H_ECANCELLED;
  return r->status;
}

/* outputs of a completed request (device pointers; owned by the server until released) */
int nt2h_server_result(nt2h_server* hs, uint64_t id, void** call, void** put, int64_t* latency_ns) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(NT2H_EINVAL, "nt2h_server_result: unknown request");
  Request* r = it->second.get();
  if (r->state != RS_DONE) return r->state == RS_FAILED ? r->status : NT2H_EAGAIN;
  if (call) *call = r->d.call;
  if (put) *put = r->d.put;
  if (latency_ns) *latency_ns = r->t_done - r->t_arrival;
  return NT2H_OK;
}

/* cancel a request that has not been launched yet */
int nt2h_server_cancel(nt2h_server* hs, uint64_t id) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(NT2H_EINVAL, "nt2h_server_cancel: unknown request");
  Request* r = it->second.get();
  if (r->state == RS_PENDING) {
    r->state = RS_CANCELLED;
    ++s->cancelled;
  } else if (r->state == RS_ADMITTED) {
    r->state = RS_CANCELLED;
    s->used -= r->bytes;
    r->charged = false;
    ++s->cancelled;
  } else {
    return fail(NT2H_ESTATE, "nt2h_server_cancel: request already running or finished");
  }
  s->cv_sched.notify_all();
  s->cv_slot.notify_all();
  return NT2H_OK;
}

/* return the request's memory to the budget (and its outputs to the pool) */
int nt2h_server_release(nt2h_server* hs, uint64_t id) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(NT2H_EINVAL, "nt2h_server_release: unknown request");
  Request* r = it->second.get();
  if (r->state == RS_PENDING || r->state == RS_ADMITTED || r->state == RS_RUNNING)
    return fail(NT2H_ESTATE, "nt2h_server_release: request not finished");
  srv_free(s, r);
  if (r->charged) {
    s->used -= r->bytes;
    r->charged = false;
  }
  s->reqs.erase(it);
  s->cv_sched.notify_all();
  return NT2H_OK;
}

/* block until every submitted request has finished (or failed/cancelled) */
int nt2h_server_drain(nt2h_server* hs, double timeout_s) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  const auto deadline = clk::now() + std::chrono::duration<double>(timeout_s);
  auto busy = [&]() {
    for (auto& kv : s->reqs)
      if (kv.second->state < RS_DONE) return true;
    return false;
  };
  while (busy()) {
    if (s->cv_slot.wait_until(lk, deadline) == std::cv_status::timeout && busy()) return NT2H_EAGAIN;
  }
  return NT2H_OK;
}

int nt2h_server_get_stats(nt2h_server* hs, nt2h_server_stats* st) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s || !st) return NT2H_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  std::vector<double> l = s->lat_us;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) {
    if (l.empty()) return 0.0;
    size_t k = (size_t)std::ceil(q * l.size()) - 1;
    return l[std::min(k, l.size() - 1)];
  };
  st->completed = l.size();
  st->failed = s->failed;
  st->cancelled = s->cancelled;
  st->options = s->options_done;
  st->launches = s->launches;
  st->mean_batch = s->launches ? (double)s->batched_reqs / s->launches : 0.0;
  st->p50_us = pct(0.50);
  st->p99_us = pct(0.99);
  st->max_us = l.empty() ? 0.0 : l.back();
  double sum = 0.0;
  for (double v : l) sum += v;
  st->mean_us = l.empty() ? 0.0 : sum / l.size();
  const double span = (s->t_last_done > s->t_first_arrival && s->t_first_arrival >= 0)
                          ? (s->t_last_done - s->t_first_arrival) * 1e-9 : 0.0;
  st->span_s = span;
  st->options_per_s = span > 0 ? s->options_done / span : 0.0;
  st->used_bytes = s->used;
  st->peak_bytes = s->peak_used;
  return NT2H_OK;
}

/* reset latency/throughput statistics (requests already finished stay) */
int nt2h_server_reset_stats(nt2h_server* hs) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_us.clear();
  s->options_done = s->launches = s->batched_reqs = 0;
  s->t_first_arrival = s->t_last_done = -1;
  s->failed = s->cancelled = 0;
  s->peak_used = s->used;
  return NT2H_OK;
}

/* Replay a request trace: item k is submitted at t0 + arrival_s[k] (busy-wait
 * pacing on the calling thread) and released as soon as it completes.  A
 * closed loop of this call gives the served latency/throughput figures. */
int nt2h_server_run_trace(nt2h_server* hs, int count, const nt2h_bs_request* items,
                          const double* arrival_s, double timeout_s) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s || count < 0 || (count && (!items || !arrival_s))) return fail(NT2H_EINVAL, "nt2h_server_run_trace: bad argument");
  std::vector<uint64_t> ids;
  ids.reserve(count);
  const int64_t t0 = now_ns();
  size_t next_release = 0;
  int64_t last_poll = 0;
  for (int k = 0; k < count; ++k) {
    const int64_t due = t0 + (int64_t)(arrival_s[k] * 1e9);
    for (int64_t now = now_ns(); now < due; now = now_ns()) {
      // release finished requests while waiting (their outputs leave the
      // budget), at most every 20 us so that the pacing loop does not keep
      // the server's lock from the scheduler and completion threads
      if (now - last_poll < 20000 || next_release >= ids.size()) continue;
      last_poll = now;
      while (next_release < ids.size() && nt2h_server_poll(hs, ids[next_release]) != NT2H_EAGAIN) {
        nt2h_server_release(hs, ids[next_release]);
        ++next_release;
      }
    }
    nt2h_bs_request q = items[k];
    q.t_arrival_ns = due;
    uint64_t id = 0;
    int rc = nt2h_server_submit(hs, &q, &id);
    if (rc) return rc;
    ids.push_back(id);
  }
  for (; next_release < ids.size(); ++next_release) {
    int rc = nt2h_server_wait(hs, ids[next_release], timeout_s);
    if (rc == NT2H_EAGAIN) return fail(NT2H_EAGAIN, "nt2h_server_run_trace: timeout");
    nt2h_server_release(hs, ids[next_release]);
  }
  return NT2H_OK;
}

int nt2h_server_destroy(nt2h_server* hs) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return NT2H_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& kv : s->reqs)
      if (kv.second->state == RS_PENDING) kv.second->state = RS_CANCELLED;
    s->cv_sched.notify_all();
    s->cv_done.notify_all();
  }
  s->sched.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->cv_done.notify_all();
  }
  for (Slot& sl : s->slots) {
    if (sl.done_thr.joinable()) sl.done_thr.join();
    if (sl.host_worker.joinable()) sl.host_worker.join();
  }
  for (auto& kv : s->reqs) srv_free(s, kv.second.get());
  for (Slot& sl : s->slots) {
    if (s->cfg.backend == NT2H_BACKEND_HOST) {
      delete[] sl.h_desc;
    } else {
      if (sl.stream) (void)hipStreamSynchronize(sl.stream);
      if (sl.h_desc) (void)hipHostFree(sl.h_desc);
      if (sl.d_desc) (void)hipFree(sl.d_desc);
      if (sl.event) (void)hipEventDestroy(sl.event);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
  }
  for (auto& kv : s->free_blocks)  // every slot stream has completed
    for (void* q : kv.second) block_release(s, q);
  if (std::getenv("NT2H_SERVER_TIMING") && s->timed)
    std::fprintf(stderr, "nt2h_server: %llu batches, per batch (us): slot free -> launch %.1f, launch %.1f, "
                 "launch -> done %.1f, done -> lock %.1f\n", (unsigned long long)s->timed,
                 s->tm_idle * 1e-3 / s->timed, s->tm_launch * 1e-3 / s->timed, s->tm_gpu * 1e-3 / s->timed,
                 s->tm_lock * 1e-3 / s->timed);
  delete s;
  return NT2H_OK;
}

}  // extern "C"
