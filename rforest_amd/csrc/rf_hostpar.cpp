// Copy pool: a few worker threads that split large host copies with the
// caller.  A copy of a megabyte out of (or into) pinned memory runs at one
// core's memory bandwidth; the drop-in makes two per call.
//
// Protocol.  The job (fn, ctx, n) and its task and completion counters live
// in the pool.  A caller publishes a job under the mutex once no worker
// still holds the previous one (busy == 0), then bumps the generation.  A
// worker that sees a new generation takes the mutex, copies the current job
// and counts itself busy before releasing it, so a job is never rewritten
// under a worker; a worker arriving after its job finished finds the task
// counter spent and leaves.  The caller returns only when every task is done
// and no worker is busy.
#include "rf_hostpar.h"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace rfi {
namespace {

constexpr int kWorkers = 3;
constexpr auto kSpin = std::chrono::microseconds(2000);  // spinning after the last job, then sleep
constexpr size_t kPiece = (size_t)64 << 10;              // bytes per copy task

inline void relax() { __builtin_ia32_pause(); }

struct Pool {
  std::mutex mu;          // guards the job fields and the sleep handshake
  std::mutex caller_mu;   // one job at a time
  std::condition_variable cv;
  std::atomic<uint64_t> gen{0};   // job generation
  std::atomic<uint64_t> wake{0};  // wake-ups without a job
  std::atomic<int> sleepers{0};
  std::atomic<int> busy{0};
  std::atomic<int> next{0}, done{0};
  std::atomic<bool> stop{false};
  void (*fn)(void*, int) = nullptr;
  void* ctx = nullptr;
  int n = 0;
  std::vector<std::thread> th;

  Pool() {
    // fewer workers (none: the caller does every task) when the process
    // may not start more threads
    for (int w = 0; w < kWorkers; w++) {
      try {
        th.emplace_back([this] { loop(); });
      } catch (...) {
        break;
      }
    }
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    for (auto& t : th) t.join();
  }

  void work(void (*f)(void*, int), void* c, int cnt) {
    for (int i; (i = next.fetch_add(1)) < cnt;) {
      f(c, i);
      done.fetch_add(1);
    }
  }

  void loop() {
    uint64_t seen = gen.load(), seen_wake = wake.load();
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      if (stop.load()) return;
      if (gen.load() == seen) {
        if (wake.load() != seen_wake) {
          seen_wake = wake.load();
          t0 = std::chrono::steady_clock::now();
        }
        if (std::chrono::steady_clock::now() - t0 < kSpin) {
          for (int k = 0; k < 64; k++) relax();
          continue;
        }
        std::unique_lock<std::mutex> lk(mu);
        sleepers.fetch_add(1);
        cv.wait(lk, [&] { return stop.load() || gen.load() != seen || wake.load() != seen_wake; });
        sleepers.fetch_sub(1);
        seen_wake = wake.load();
        t0 = std::chrono::steady_clock::now();
        continue;
      }
      void (*f)(void*, int);
      void* c;
      int cnt;
      {
        std::lock_guard<std::mutex> lk(mu);
        seen = gen.load();
        f = fn;
        c = ctx;
        cnt = n;
        busy.fetch_add(1);
      }
      work(f, c, cnt);
      busy.fetch_sub(1);
      t0 = std::chrono::steady_clock::now();
    }
  }

  void run(int cnt, void (*f)(void*, int), void* c) {
    std::lock_guard<std::mutex> one(caller_mu);
    {
      std::unique_lock<std::mutex> lk(mu);
      while (busy.load() != 0) {  // a worker still leaving the previous job
        lk.unlock();
        relax();
        lk.lock();
      }
      fn = f;
      ctx = c;
      n = cnt;
      next.store(0);
      done.store(0);
      gen.fetch_add(1);
    }
    if (sleepers.load() > 0) cv.notify_all();
    work(f, c, cnt);
    while (done.load() < cnt) relax();
    while (busy.load() != 0) relax();
  }

  void poke() {
    if (sleepers.load() == 0) return;
    {
      std::lock_guard<std::mutex> lk(mu);
      wake.fetch_add(1);
    }
    cv.notify_all();
  }
};

Pool& pool() {
  static Pool p;
  return p;
}

struct CopyJob {
  char* dst;
  const char* src;
  size_t n;
};

void copy_piece(void* vp, int i) {
  const CopyJob& j = *(const CopyJob*)vp;
  const size_t off = (size_t)i * kPiece;
  const size_t len = j.n - off < kPiece ? j.n - off : kPiece;
  memcpy(j.dst + off, j.src + off, len);
}

// The segments laid end to end: piece i copies bytes [i, i + 1) * kPiece of
// the concatenation, whatever segments (many small ones, or part of a large
// one) that range covers.
struct SegJob {
  const CopySeg* segs;
  const size_t* start;  // byte offset of each segment, and the total at [nseg]
  int nseg;
};

void seg_piece(void* vp, int i) {
  const SegJob& j = *(const SegJob*)vp;
  const size_t lo = (size_t)i * kPiece;
  const size_t hi = lo + kPiece < j.start[j.nseg] ? lo + kPiece : j.start[j.nseg];
  int s = (int)(std::upper_bound(j.start, j.start + j.nseg + 1, lo) - j.start) - 1;
  for (; s < j.nseg && j.start[s] < hi; s++) {
    const size_t a = (lo > j.start[s] ? lo : j.start[s]) - j.start[s];
    const size_t b = (hi < j.start[s + 1] ? hi : j.start[s + 1]) - j.start[s];
    if (b > a) memcpy((char*)j.segs[s].dst + a, (const char*)j.segs[s].src + a, b - a);
  }
}

}  // namespace

void par_copy(const CopySeg* segs, int nseg) {
  if (nseg <= 0) return;
  std::vector<size_t> start(nseg + 1, 0);
  for (int s = 0; s < nseg; s++) start[s + 1] = start[s] + segs[s].n;
  SegJob j{segs, start.data(), nseg};
  par_for((int)((start[nseg] + kPiece - 1) / kPiece), seg_piece, &j);
}

void par_for(int n, void (*fn)(void* ctx, int i), void* ctx) {
  if (n <= 0) return;
  if (n == 1) {
    fn(ctx, 0);
    return;
  }
  pool().run(n, fn, ctx);
}

void par_wake() { pool().poke(); }

void par_memcpy(void* dst, const void* src, size_t n) {
  if (n <= kPiece) {
    memcpy(dst, src, n);
    return;
  }
  CopyJob j{(char*)dst, (const char*)src, n};
  par_for((int)((n + kPiece - 1) / kPiece), copy_piece, &j);
}

}  // namespace rfi
