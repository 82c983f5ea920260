// rf_serve.cpp -- the request server: continuous batching under an HBM budget.
//
// Requests are planned at submission (the plan depends on the request alone,
// so its result does not depend on what it is batched with) and queued in
// arrival order.  One worker thread loops:
//   * retire: a slot whose batch event has completed returns its requests'
//     inputs and the batch scratch to the pool and publishes completions;
//   * admit: every free slot takes queued requests from the head while their
//     inputs and outputs fit in the pool (and the batch limits allow), lays
//     out the batch, allocates its scratch, and enqueues on the slot's stream
//     input generation or upload, the product pipeline, checksums, result
//     copies and an event.
// Requests are never reordered, so a large request at the head waits for
// memory rather than being overtaken indefinitely.  Slots are separate
// streams (default priority), so consecutive batches overlap on the device.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rf_hostpar.h"
#include "rf_internal.h"
#include "rf_util.h"

using namespace rfi;

namespace {

// First-fit allocator over one device block (offsets, 256-byte granules),
// coalescing freed neighbours.
class Pool {
 public:
  void init(size_t cap) {
    free_.clear();
    free_[0] = cap;
    used_ = peak_ = 0;
  }
  bool alloc(size_t n, size_t* off) {
    n = align256(std::max<size_t>(n, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, o = it->first + n;
      free_.erase(it);
      if (rest) free_[o] = rest;
      used_ += n;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = align256(std::max<size_t>(n, 1));
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0, peak_ = 0;
};

constexpr size_t NONE = ~(size_t)0;

// Host copies of RF_IN_HOST inputs into a slot's staging buffer and of
// RF_OUT_HOST results out of it: on the copy pool from 8 MiB on, below that
// on the worker thread (a batch's small copies do not repay the hand-off).
void serve_copy(const std::vector<CopySeg>& cps) {
  size_t bytes = 0;
  for (const CopySeg& c : cps) bytes += c.n;
  if (bytes >= ((size_t)8 << 20)) {
    par_copy(cps.data(), (int)cps.size());
    return;
  }
  for (const CopySeg& c : cps) memcpy(c.dst, c.src, c.n);
}


struct Req {
  uint64_t id = 0, tag = 0, seed = 0;
  int r = 0, d = 0, input = 0, output = 0;
  std::vector<rf_serve_entry> ents;  // A then B
  uint64_t* out_host = nullptr;
  uint64_t out_cap = 0;
  Product P;
  uint64_t rn = 0;  // limbs per C entry (0: the product is zero)
  size_t in_bytes = 0, out_bytes = 0, alone_bytes = 0;
  size_t in_off = NONE, out_off = NONE;
  size_t st_in = NONE, st_out = NONE;  // host inputs / outputs in the slot's pinned staging buffer
  size_t gen_spans = 0, gen_blocks = 0, sum_spans = 0, sum_blocks = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  uint64_t checksum = 0;
  int status = RF_OK;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;
  std::vector<Req*> reqs;
  size_t scratch_off = NONE, scratch_bytes = 0;
  char* h_stage = nullptr;
  size_t stage_cap = 0;
  uint64_t* d_sums = nullptr;
  uint64_t* h_sums = nullptr;
};

size_t blocks_of(uint64_t limbs) { return (size_t)((limbs + RF_UTIL_LIMBS - 1) / RF_UTIL_LIMBS); }

// Host inputs and results pass through the slot's pinned staging buffer
// (one host copy and one DMA transfer per request each way, instead of a
// transfer per entry from pageable memory); a batch stages at most this
// many bytes of them, unless its first request alone needs more.
constexpr size_t kSlotHostBytes = (size_t)512 << 20;

// Device bytes of the span/block tables a set of requests needs (four
// 256-aligned arrays: generation spans and blocks, checksum spans and blocks).
size_t util_bytes(size_t spans, size_t blocks) {
  return align256(spans * sizeof(RfSpan)) + align256(blocks * sizeof(RfBlock)) + 512;
}

}  // namespace

struct rf_server {
  rf_server_config cfg{};
  int device = 0;
  char* pool_base = nullptr;
  Pool pool;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, Req*> kept;
  std::vector<rf_completion> done;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1;
  rf_server_stats stats{};
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

rf_completion completion_of(const Req* q) {
  rf_completion c{};
  c.id = q->id;
  c.tag = q->tag;
  c.status = q->status;
  c.out_limbs = q->rn;
  c.checksum = q->checksum;
  c.t_submit = q->t_submit;
  c.t_admit = q->t_admit;
  c.t_done = q->t_done;
  return c;
}

// Returns a request's input and output blocks (caller holds the lock).
void free_req_mem(rf_server* S, Req* q, bool keep_out) {
  if (q->in_off != NONE) S->pool.release(q->in_off, q->in_bytes), q->in_off = NONE;
  if (!keep_out && q->out_off != NONE) S->pool.release(q->out_off, q->out_bytes), q->out_off = NONE;
}

void retire(rf_server* S, Slot& sl, int status) {
  // host results: from the slot's staging buffer to the callers' arrays
  // (the batch has completed, and the slot is not reused before this returns)
  if (status == RF_OK) {
    std::vector<CopySeg> cps;
    for (Req* q : sl.reqs)
      if (q->status == RF_OK && q->st_out != NONE)
        for (int e = 0; e < q->r * q->d; e++)
          cps.push_back(CopySeg{q->out_host + (size_t)e * q->out_cap, sl.h_stage + q->st_out + (size_t)e * q->rn * 8,
                                q->rn * 8});
    serve_copy(cps);
  }
  const double t = S->now();
  std::lock_guard<std::mutex> lk(S->mu);
  for (size_t i = 0; i < sl.reqs.size(); i++) {
    Req* q = sl.reqs[i];
    q->t_done = t;
    if (status != RF_OK) q->status = status;
    if ((q->output & RF_OUT_CHECKSUM) && q->status == RF_OK) q->checksum = sl.h_sums[i];
    const bool keep = (q->output & RF_OUT_KEEP) && q->status == RF_OK;
    free_req_mem(S, q, keep);
    S->done.push_back(completion_of(q));
    S->stats.completed++;
    if (keep)
      S->kept[q->id] = q;
    else
      delete q;
  }
  if (sl.scratch_off != NONE) S->pool.release(sl.scratch_off, sl.scratch_bytes);
  sl.scratch_off = NONE;
  sl.reqs.clear();
  sl.busy = false;
  S->cv_done.notify_all();
}

int ensure_slot_stage(Slot& sl, size_t bytes) {
  if (bytes <= sl.stage_cap) return RF_OK;
  const size_t old_cap = sl.stage_cap;
  if (sl.h_stage) HIPCHK(hipHostFree(sl.h_stage));
  sl.h_stage = nullptr;
  sl.stage_cap = old_cap;  // the growth below starts from the old size
  // pinning memory is slow (about 0.1 s per GB here): grow geometrically so
  // a slot re-pins only a few times
  const size_t cap = std::max({bytes, 2 * sl.stage_cap, (size_t)1 << 20});
  sl.stage_cap = 0;
  HIPCHK(hipHostMalloc((void**)&sl.h_stage, cap, hipHostMallocDefault));
  sl.stage_cap = cap;
  return RF_OK;
}

size_t host_io_bytes(const Req* q) {
  return (q->input == RF_IN_HOST ? align256(q->in_bytes) : 0) +
         ((q->output & RF_OUT_HOST) && q->rn ? align256(q->out_bytes) : 0);
}

// Takes queued requests for a free slot and enqueues their batch.  Returns
// true if a batch was launched (or failed and was retired).
bool admit(rf_server* S, Slot& sl) {
  std::vector<Req*> take;
  std::vector<Product> prods;
  Layout L;
  size_t ub = 0, soff = NONE, sbytes = 0;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    size_t bytes = 0, hbytes = 0;
    while (!S->queue.empty() && (int)take.size() < S->cfg.max_batch) {
      Req* q = S->queue.front();
      if (!take.empty() && bytes + q->alone_bytes > S->cfg.max_batch_bytes) break;
      const size_t qh = host_io_bytes(q);
      if (!take.empty() && hbytes + qh > kSlotHostBytes) break;
      hbytes += qh;
      size_t io = NONE, oo = NONE;
      if (q->in_bytes && !S->pool.alloc(q->in_bytes, &io)) break;
      if (q->out_bytes && !S->pool.alloc(q->out_bytes, &oo)) {
        if (io != NONE) S->pool.release(io, q->in_bytes);
        break;
      }
      q->in_off = io;
      q->out_off = oo;
      S->queue.pop_front();
      take.push_back(q);
      bytes += q->alone_bytes;
    }
    // scratch for the whole batch; shrink the batch until it fits
    while (!take.empty()) {
      prods.clear();
      size_t spans = 0, blocks = 0;
      for (Req* q : take) {
        prods.push_back(q->P);
        spans += q->gen_spans + q->sum_spans;
        blocks += q->gen_blocks + q->sum_blocks;
      }
      L = layout_of(prods);
      ub = util_bytes(spans, blocks);
      sbytes = L.total() + ub;
      if (S->pool.alloc(sbytes, &soff)) break;
      Req* q = take.back();
      take.pop_back();
      free_req_mem(S, q, false);
      S->queue.push_front(q);
    }
    if (take.empty()) return false;
    S->stats.batches++;
    S->stats.products_in_batches += take.size();
  }
  const double ta = S->now();
  sl.reqs = take;
  sl.scratch_off = soff;
  sl.scratch_bytes = sbytes;
  sl.busy = true;
  char* base = S->pool_base;
  // device addresses of inputs and outputs; synthetic and checksum tables
  std::vector<RfSpan> spans;
  std::vector<RfBlock> blocks;
  std::vector<RfSpan> sspans;
  std::vector<RfBlock> sblocks;
  for (size_t i = 0; i < take.size(); i++) {
    Req* q = take[i];
    q->t_admit = ta;
    Product& P = prods[i];
    uint64_t pos = 0;
    const int rd = q->r * q->d;
    for (int e = 0; e < rd + q->d * q->d; e++) {
      Entry& E = e < rd ? P.A[e] : P.B[e - rd];
      E.dptr = E.nlimbs ? (const uint64_t*)(base + q->in_off) + pos : nullptr;
      if (E.nlimbs && q->input == RF_IN_SYNTH) {
        spans.push_back(RfSpan{(uint64_t*)(base + q->in_off) + pos, q->ents[e].bits,
                               rf_gen_key(q->seed, (uint32_t)e), 0, 0});
        for (uint64_t f = 0; f < E.nlimbs; f += RF_UTIL_LIMBS)
          blocks.push_back(RfBlock{(uint32_t)(spans.size() - 1), 0, f});
      }
      pos += E.nlimbs;
    }
    P.out.clear();
    for (int e = 0; e < rd; e++) {
      uint64_t* o = q->rn ? (uint64_t*)(base + q->out_off) + (size_t)e * q->rn : nullptr;
      P.out.push_back(o);
      if (q->rn && (q->output & RF_OUT_CHECKSUM)) {
        sspans.push_back(RfSpan{o, q->rn, (uint64_t)e << 40, (uint32_t)i, 0});
        for (uint64_t f = 0; f < q->rn; f += RF_UTIL_LIMBS)
          sblocks.push_back(RfBlock{(uint32_t)(sspans.size() - 1), 0, f});
      }
    }
  }
  // the engine's tables stay in place until this batch is enqueued
  // (hw_mpz_clear waits for the device under the exclusive lock)
  std::shared_lock<std::shared_mutex> tl(engine_tables_lock());
  Pipeline pp;
  int rc = engine_ready() ? build_pipeline(prods, pp)
                          : fail(RF_ERR_NOT_INIT, "hw_mpz_clear() was called; requests wait for hw_mpz_setup()");
  // staging image: [task image | gen spans | gen blocks | sum spans | sum blocks]
  const size_t o_gs = align256(pp.image.size());
  const size_t o_gb = o_gs + align256(spans.size() * sizeof(RfSpan));
  const size_t o_ss = o_gb + align256(blocks.size() * sizeof(RfBlock));
  const size_t o_sb = o_ss + align256(sspans.size() * sizeof(RfSpan));
  const size_t o_end = o_sb + align256(sblocks.size() * sizeof(RfBlock));
  // then [host inputs | host results] of each request that has them
  size_t o_host = o_end;
  for (Req* q : take) {
    if (q->input == RF_IN_HOST && q->in_bytes) q->st_in = o_host, o_host += align256(q->in_bytes);
    if ((q->output & RF_OUT_HOST) && q->rn) q->st_out = o_host, o_host += align256(q->out_bytes);
  }
  if (!rc) rc = ensure_slot_stage(sl, o_host);
  if (!rc) {
    char* h = sl.h_stage;
    memcpy(h, pp.image.data(), pp.image.size());
    if (!spans.empty()) memcpy(h + o_gs, spans.data(), spans.size() * sizeof(RfSpan));
    if (!blocks.empty()) memcpy(h + o_gb, blocks.data(), blocks.size() * sizeof(RfBlock));
    if (!sspans.empty()) memcpy(h + o_ss, sspans.data(), sspans.size() * sizeof(RfSpan));
    if (!sblocks.empty()) memcpy(h + o_sb, sblocks.data(), sblocks.size() * sizeof(RfBlock));
    // RF_IN_HOST inputs into the staging buffer, split over the copy pool
    std::vector<CopySeg> cps;
    for (Req* q : take) {
      if (q->st_in == NONE) continue;
      uint64_t pos = 0;
      for (size_t e = 0; e < q->ents.size(); e++) {
        const uint64_t n = q->ents[e].sign ? (q->ents[e].bits + 63) / 64 : 0;
        if (n) cps.push_back(CopySeg{h + q->st_in + pos * 8, q->ents[e].limbs, n * 8});
        pos += n;
      }
    }
    serve_copy(cps);
  }
  // device side: the task image sits in the layout's task area, the util
  // tables after the layout
  char* sbase = base + soff;
  const ScratchPtrs SP = scratch_ptrs(sbase, pp.L);
  char* d_util = sbase + pp.L.total();
  auto enqueue = [&]() -> int {
    if (!pp.image.empty()) HIPCHK(hipMemcpyAsync(SP.tasks, sl.h_stage, pp.image.size(), hipMemcpyHostToDevice, sl.st));
    if (o_end > o_gs)
      HIPCHK(hipMemcpyAsync(d_util, sl.h_stage + o_gs, o_end - o_gs, hipMemcpyHostToDevice, sl.st));
    for (Req* q : take)
      if (q->st_in != NONE)
        HIPCHK(hipMemcpyAsync(base + q->in_off, sl.h_stage + q->st_in, q->in_bytes, hipMemcpyHostToDevice, sl.st));
    HIPCHK(rf_launch_gen((const RfSpan*)(d_util), (const RfBlock*)(d_util + (o_gb - o_gs)), (int)blocks.size(),
                         sl.st));
    int r2 = launch_pipeline(pp, sbase, sl.st, false);
    if (r2) return r2;
    if (!sspans.empty()) {
      HIPCHK(rf_launch_fill(sl.d_sums, take.size(), 0, sl.st));
      HIPCHK(rf_launch_checksum((const RfSpan*)(d_util + (o_ss - o_gs)), (const RfBlock*)(d_util + (o_sb - o_gs)),
                                (int)sblocks.size(), sl.d_sums, sl.st));
      HIPCHK(hipMemcpyAsync(sl.h_sums, sl.d_sums, take.size() * 8, hipMemcpyDeviceToHost, sl.st));
    } else {
      // no result limbs to sum (every product zero): the checksums are 0,
      // not what an earlier batch of this slot left in h_sums
      std::fill(sl.h_sums, sl.h_sums + take.size(), (uint64_t)0);
    }
    for (size_t i = 0; i < take.size(); i++) {  // results to staging; to the callers at retirement
      Req* q = take[i];
      if (q->st_out != NONE)
        HIPCHK(hipMemcpyAsync(sl.h_stage + q->st_out, prods[i].out[0], q->out_bytes, hipMemcpyDeviceToHost, sl.st));
    }
    HIPCHK(hipEventRecord(sl.ev, sl.st));
    return RF_OK;
  };
  if (!rc) rc = enqueue();
  if (rc) {
    (void)hipStreamSynchronize(sl.st);
    retire(S, sl, rc);
  }
  return true;
}

void worker_main(rf_server* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    bool progress = false;
    for (Slot& sl : S->slots) {
      if (!sl.busy) continue;
      const hipError_t e = hipEventQuery(sl.ev);
      if (e == hipErrorNotReady) continue;
      retire(S, sl, e == hipSuccess ? RF_OK : fail(RF_ERR_HIP, "batch failed: %s", hipGetErrorString(e)));
      progress = true;
    }
    for (Slot& sl : S->slots)
      if (!sl.busy) {
        if (admit(S, sl))
          progress = true;
        else
          break;
      }
    if (progress) continue;
    bool inflight = false;
    for (Slot& sl : S->slots) inflight = inflight || sl.busy;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop && S->queue.empty() && !inflight) break;
    if (inflight) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    } else if (S->queue.empty()) {
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
    } else {
      // the head does not fit: memory is held by kept results
      S->cv_work.wait_for(lk, std::chrono::milliseconds(1));
    }
  }
}

void server_free(rf_server* S) {
  for (Slot& sl : S->slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.st) (void)hipStreamDestroy(sl.st);
    if (sl.h_stage) (void)hipHostFree(sl.h_stage);
    if (sl.h_sums) (void)hipHostFree(sl.h_sums);
    if (sl.d_sums) (void)hipFree(sl.d_sums);
  }
  for (auto& kv : S->kept) delete kv.second;
  for (Req* q : S->queue) delete q;
  if (S->pool_base) (void)hipFree(S->pool_base);
  delete S;
}

}  // namespace

extern "C" {

int rf_server_create(const rf_server_config* cfg, rf_server** out) {
  if (!out || !cfg) return fail(RF_ERR_ARG, "rf_server_create: null argument");
  *out = nullptr;
  {
    std::lock_guard<std::mutex> lk(engine_mutex());
    if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  }
  if (cfg->hbm_budget < ((size_t)1 << 20)) return fail(RF_ERR_ARG, "HBM budget below 1 MiB");
  rf_server* S = new rf_server();
  S->cfg = *cfg;
  if (S->cfg.streams <= 0) S->cfg.streams = 3;
  if (S->cfg.max_batch <= 0) S->cfg.max_batch = 256;
  if (S->cfg.max_batch_bytes == 0) S->cfg.max_batch_bytes = S->cfg.hbm_budget / S->cfg.streams;
  S->stats.budget = S->cfg.hbm_budget;
  S->t0 = std::chrono::steady_clock::now();
  hipError_t e = hipGetDevice(&S->device);
  if (e == hipSuccess) e = hipMalloc((void**)&S->pool_base, S->cfg.hbm_budget);
  if (e != hipSuccess) {
    int rc = fail(RF_ERR_NOMEM, "server pool of %llu bytes: %s", (unsigned long long)S->cfg.hbm_budget,
                  hipGetErrorString(e));
    server_free(S);
    return rc;
  }
  S->pool.init(S->cfg.hbm_budget);
  S->slots.resize(S->cfg.streams);
  for (Slot& sl : S->slots) {
    e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void**)&sl.d_sums, (size_t)S->cfg.max_batch * 8);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_sums, (size_t)S->cfg.max_batch * 8, hipHostMallocDefault);
    if (e != hipSuccess) {
      int rc = fail(RF_ERR_HIP, "server slot: %s", hipGetErrorString(e));
      server_free(S);
      return rc;
    }
  }
  S->worker = std::thread(worker_main, S);
  *out = S;
  return RF_OK;
}

int rf_server_submit(rf_server* S, const rf_request* rq, uint64_t* id) {
  if (!S || !rq) return fail(RF_ERR_ARG, "rf_server_submit: null argument");
  const int r = rq->r, d = rq->d;
  if (r <= 0 || d <= 0 || (!rq->A && !rq->left) || !rq->B) return fail(RF_ERR_ARG, "bad request r=%d d=%d", r, d);
  if (rq->input != RF_IN_HOST && rq->input != RF_IN_SYNTH) return fail(RF_ERR_ARG, "bad input mode %d", rq->input);
  if ((rq->output & RF_OUT_HOST) && !rq->out_host) return fail(RF_ERR_ARG, "RF_OUT_HOST without out_host");
  Req* q = new Req();
  q->tag = rq->tag;
  q->seed = rq->seed;
  q->r = r;
  q->d = d;
  q->input = rq->input;
  q->output = rq->output;
  q->out_host = rq->out_host;
  q->out_cap = rq->out_cap;
  q->P.r = r;
  q->P.d = d;
  // A resident left operand replaces A: its entries stand as zeros here (no
  // limbs, nothing generated or copied) and its spectra are read in place.
  q->P.left = rq->left;
  uint64_t in_limbs = 0;
  for (int e = 0; e < r * d + d * d; e++) {
    rf_serve_entry x = e < r * d ? (rq->left ? rf_serve_entry{nullptr, 0, 0, 0} : rq->A[e]) : rq->B[e - r * d];
    if (x.bits == 0) x.sign = 0;
    x.sign = x.sign > 0 ? 1 : (x.sign < 0 ? -1 : 0);
    if (q->input == RF_IN_HOST && x.sign && !x.limbs) {
      delete q;
      return fail(RF_ERR_ARG, "entry %d: no limbs", e);
    }
    q->ents.push_back(x);
    Entry E{};
    E.sign = x.sign;
    E.bits = x.sign ? x.bits : 0;
    E.nlimbs = (E.bits + 63) / 64;
    in_limbs += E.nlimbs;
    if (E.nlimbs && q->input == RF_IN_SYNTH) q->gen_spans++, q->gen_blocks += blocks_of(E.nlimbs);
    (e < r * d ? q->P.A : q->P.B).push_back(E);
  }
  int rc;
  {
    std::shared_lock<std::shared_mutex> tl(engine_tables_lock());
    rc = engine_ready() ? plan_product(q->P) : fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  }
  if (rc) {
    delete q;
    return rc;
  }
  q->rn = q->P.plan.k ? q->P.plan.rn : 0;
  if ((q->output & RF_OUT_HOST) && q->out_cap < q->rn) {
    delete q;
    return fail(RF_ERR_CAPACITY, "out_cap %llu < %llu limbs", (unsigned long long)rq->out_cap,
                (unsigned long long)q->rn);
  }
  if (q->rn && (q->output & RF_OUT_CHECKSUM)) q->sum_spans = r * d, q->sum_blocks = (size_t)r * d * blocks_of(q->rn);
  q->in_bytes = in_limbs * 8;
  q->out_bytes = (size_t)r * d * q->rn * 8;
  std::vector<Product> one{q->P};
  const Layout L = layout_of(one);
  q->alone_bytes = (q->in_bytes ? align256(q->in_bytes) : 0) + (q->out_bytes ? align256(q->out_bytes) : 0) +
                   align256(L.total() + util_bytes(q->gen_spans + q->sum_spans, q->gen_blocks + q->sum_blocks));
  if (q->alone_bytes > S->cfg.hbm_budget) {
    const size_t need = q->alone_bytes;
    delete q;
    return fail(RF_ERR_CAPACITY, "request needs %zu bytes of device memory, budget %llu", need,
                (unsigned long long)S->cfg.hbm_budget);
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    q->id = S->next_id++;
    q->t_submit = S->now();
    S->queue.push_back(q);
    S->stats.submitted++;
    if (id) *id = q->id;
  }
  S->cv_work.notify_one();
  return RF_OK;
}

int rf_server_poll(rf_server* S, rf_completion* out, int max, int wait_ms) {
  if (!S || (max > 0 && !out)) return fail(RF_ERR_ARG, "rf_server_poll: bad arguments");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done.empty() && wait_ms > 0)
    S->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !S->done.empty(); });
  const int n = std::min<int>(max, (int)S->done.size());
  for (int i = 0; i < n; i++) out[i] = S->done[i];
  S->done.erase(S->done.begin(), S->done.begin() + n);
  return n;
}

int rf_server_read(rf_server* S, uint64_t id, uint64_t* dst, uint64_t cap) {
  if (!S || !dst) return fail(RF_ERR_ARG, "rf_server_read: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->kept.find(id);
  if (it == S->kept.end()) return fail(RF_ERR_ARG, "request %llu holds no result", (unsigned long long)id);
  Req* q = it->second;
  const uint64_t n = (uint64_t)q->r * q->d * q->rn;
  if (cap < n) return fail(RF_ERR_CAPACITY, "result needs %llu limbs", (unsigned long long)n);
  if (n) HIPCHK(hipMemcpy(dst, S->pool_base + q->out_off, n * 8, hipMemcpyDeviceToHost));
  return RF_OK;
}

int rf_server_release(rf_server* S, uint64_t id) {
  if (!S) return fail(RF_ERR_ARG, "rf_server_release: null server");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->kept.find(id);
    if (it == S->kept.end()) return fail(RF_ERR_ARG, "request %llu holds no result", (unsigned long long)id);
    Req* q = it->second;
    if (q->out_off != NONE) S->pool.release(q->out_off, q->out_bytes);
    delete q;
    S->kept.erase(it);
  }
  S->cv_work.notify_one();
  return RF_OK;
}

int rf_server_stats_get(rf_server* S, rf_server_stats* st) {
  if (!S || !st) return fail(RF_ERR_ARG, "rf_server_stats_get: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->stats;
  st->peak_bytes = S->pool.peak();
  return RF_OK;
}

double rf_server_clock(rf_server* S) { return S ? S->now() : 0.0; }

int rf_server_destroy(rf_server* S) {
  if (!S) return RF_OK;
  {
    // Kept results are dropped first, so that every queued request (each
    // fits the budget alone) can still be admitted and finished.
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto& kv : S->kept) {
      if (kv.second->out_off != NONE) S->pool.release(kv.second->out_off, kv.second->out_bytes);
      delete kv.second;
    }
    S->kept.clear();
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  server_free(S);
  return RF_OK;
}

int rf_gen_entries_device(uint64_t* const* dst, const uint64_t* bits, const uint64_t* seeds,
                          const uint32_t* entry, int n) {
  if (n < 0 || (n > 0 && (!dst || !bits || !seeds || !entry))) return fail(RF_ERR_ARG, "rf_gen_entries_device: bad arguments");
  std::vector<RfSpan> spans;
  std::vector<RfBlock> blocks;
  for (int e = 0; e < n; e++) {
    if (bits[e] == 0) continue;
    if (!dst[e]) return fail(RF_ERR_ARG, "rf_gen_entries_device: null destination");
    spans.push_back(RfSpan{dst[e], bits[e], rf_gen_key(seeds[e], entry[e]), 0, 0});
    const uint64_t nl = (bits[e] + 63) / 64;
    for (uint64_t f = 0; f < nl; f += RF_UTIL_LIMBS) blocks.push_back(RfBlock{(uint32_t)(spans.size() - 1), 0, f});
  }
  if (blocks.empty()) return RF_OK;
  const size_t b0 = align256(spans.size() * sizeof(RfSpan));
  char* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, b0 + blocks.size() * sizeof(RfBlock)));
  hipError_t e = hipMemcpy(d, spans.data(), spans.size() * sizeof(RfSpan), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + b0, blocks.data(), blocks.size() * sizeof(RfBlock), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = rf_launch_gen((const RfSpan*)d, (const RfBlock*)(d + b0), (int)blocks.size(), nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(d);
  if (e != hipSuccess) return fail(RF_ERR_HIP, "rf_gen_entries_device: %s", hipGetErrorString(e));
  return RF_OK;
}

int rf_gen_entry_device(uint64_t* dst, uint64_t bits, uint64_t seed, uint32_t entry) {
  if (!dst || bits == 0) return fail(RF_ERR_ARG, "rf_gen_entry_device: bad arguments");
  return rf_gen_entries_device(&dst, &bits, &seed, &entry, 1);
}

int rf_checksums_device(const uint64_t* const* src, const uint64_t* n, const uint32_t* entry, const uint32_t* slot,
                        int count, uint64_t* sums, int nslots) {
  if (count < 0 || nslots <= 0 || !sums || (count > 0 && (!src || !n || !entry || !slot)))
    return fail(RF_ERR_ARG, "rf_checksums_device: bad arguments");
  std::vector<RfSpan> spans;
  std::vector<RfBlock> blocks;
  for (int c = 0; c < count; c++) {
    if (slot[c] >= (uint32_t)nslots) return fail(RF_ERR_ARG, "rf_checksums_device: slot out of range");
    if (n[c] == 0) continue;
    spans.push_back(RfSpan{(uint64_t*)src[c], n[c], (uint64_t)entry[c] << 40, slot[c], 0});
    for (uint64_t f = 0; f < n[c]; f += RF_UTIL_LIMBS) blocks.push_back(RfBlock{(uint32_t)(spans.size() - 1), 0, f});
  }
  const size_t b0 = align256(std::max<size_t>(spans.size(), 1) * sizeof(RfSpan));
  const size_t b1 = b0 + align256(std::max<size_t>(blocks.size(), 1) * sizeof(RfBlock));
  char* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, b1 + (size_t)nslots * 8));
  hipError_t e = hipSuccess;
  if (!spans.empty()) e = hipMemcpy(d, spans.data(), spans.size() * sizeof(RfSpan), hipMemcpyHostToDevice);
  if (e == hipSuccess && !blocks.empty())
    e = hipMemcpy(d + b0, blocks.data(), blocks.size() * sizeof(RfBlock), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = rf_launch_fill((uint64_t*)(d + b1), (uint64_t)nslots, 0, nullptr);
  if (e == hipSuccess)
    e = rf_launch_checksum((const RfSpan*)d, (const RfBlock*)(d + b0), (int)blocks.size(), (uint64_t*)(d + b1), nullptr);
  if (e == hipSuccess) e = hipMemcpy(sums, d + b1, (size_t)nslots * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(RF_ERR_HIP, "rf_checksums_device: %s", hipGetErrorString(e));
  return RF_OK;
}

int rf_checksum_device(const uint64_t* limbs, uint64_t n, uint32_t entry, uint64_t* sum) {
  if (!limbs || !sum) return fail(RF_ERR_ARG, "rf_checksum_device: bad arguments");
  const uint32_t slot = 0;
  return rf_checksums_device(&limbs, &n, &entry, &slot, 1, sum, 1);
}

}  // extern "C"
