// rf_internal.h -- engine internals shared by the library's host sources
// (rf_engine.cpp: one-shot calls, batches, operands; rf_serve.cpp: the
// request server).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "../../include/rforest_hip.h"
#include "rf_kernels.h"
#include "rf_plan.h"

// A resident left operand: r*d spectra, [entry][prime][N] like a batch's
// spectra, computed once under a plan fixed at creation.
struct rf_operand {
  int r = 0, d = 0;
  RfPlan plan{};
  uint64_t la_bits = 0;     // largest entry, bits (0: the operand is zero)
  std::vector<int> signs;   // per entry
  double* spec = nullptr;   // device, r*d*k*N doubles
  // set when spec lies in an allocation shared by the operands made together
  // (freed with the last of them); otherwise spec is the operand's own
  std::shared_ptr<void> slab;
};

namespace rfi {

int fail(int code, const char* fmt, ...);
#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      return fail(RF_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                          \
  } while (0)

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- batch description shared by all entry points ----
struct Entry {
  uint64_t limb_off;     // host entries: offset in the staged limb buffer
  const uint64_t* dptr;  // device pointer to the magnitude limbs
  uint64_t nlimbs;
  uint64_t bits;  // upper bound on the magnitude's bit length
  int sign;
};
struct Product {
  int r, d;
  const rf_operand* left = nullptr;  // resident A: its spectra replace A's forward transforms
  std::vector<Entry> A, B;
  std::vector<uint64_t*> out;  // device pointers, plan.rn limbs each
  RfPlan plan;
  uint64_t spec_off, coef_off;  // in doubles, within their regions
};

// Scratch layout of one pipeline run (after the caller's `base` bytes).
struct Layout {
  uint64_t spec_n = 0, coef_n = 0;  // doubles
  uint32_t nblk = 0;                // CRT blocks in total
  size_t nfwd = 0, ninv = 0, nrow = 0, ncol = 0, ncrt = 0;
  size_t spec_b = 0, coef_b = 0, tasks_b = 0, pre_b = 0, fun_b = 0;
  size_t total() const { return spec_b + coef_b + tasks_b + pre_b + fun_b; }
};

enum Kind { K_FWD, K_FWD_COL, K_FWD_ROW, K_INV, K_INV_COL, K_CRT, K_APPLY };
int stage_of(int kind);
struct LaunchRec {
  int kind;    // Kind
  int param;   // lgN for transforms, k for CRT
  size_t off;  // byte offset of its task table in the image
  int n;       // tasks
};
struct Pipeline {
  Layout L;
  std::vector<char> image;
  std::vector<LaunchRec> launches;
};

struct ScratchPtrs {
  double* spec;
  double* coef;
  char* tasks;
  uint8_t* pre;
  uint8_t* fun;
};

// The library's global lock, its readiness and its device tables.
std::mutex& engine_mutex();
// Held shared while planning and enqueueing against the tables without the
// engine mutex (server workers); hw_mpz_clear takes it exclusively.
std::shared_mutex& engine_tables_lock();
bool engine_ready();
// Device tables for a launch (a copy: long-table pointers may appear later).
RfDevTables engine_tables();
// Makes the twiddle tables transforms above 2^RF_LGN_TW points need (no-op
// below); callable without the engine mutex.
int engine_long_tables(int lgN);
const RfPrimes& engine_primes();
// A resident operand from entries already in device memory (task tables at
// arena byte `base`), the pipeline of planned products run on the engine's
// stream with its scratch at arena byte `base` (the arena must already hold
// base + L.total() bytes; returns with the stream busy), and an operand's
// release.
// fixed: the plan to use (a cyclic one, say) instead of the planner's.
int make_operand(std::vector<Entry>& A, int r, int d, uint64_t b_max_bits, size_t base, rf_operand** out,
                 const RfPlan* fixed = nullptr);
// Arena bytes an operand of nent entries needs for its transforms' task
// tables at most (reserve them with its staged inputs: the arena moves when
// it grows).
size_t operand_task_bound(int nent);
int pipeline_run(std::vector<Product>& prods, const Layout& L, size_t base);
void free_operand(rf_operand* op);
// Resident 1x1 operands for many device entries at once: ents[i] under the
// plan fixed[i] when given, otherwise planned for right operands of
// b_max_bits[i] bits; the spectra share one device allocation and are
// transformed by one set of launches (task tables at arena byte `base`).
// On failure nothing is left allocated and out is empty.
int make_operands_batch(const std::vector<Entry>& ents, const std::vector<uint64_t>& b_max_bits,
                        const std::vector<const RfPlan*>& fixed, size_t base, std::vector<rf_operand*>& out);
// Arena bytes make_operands_batch needs for the task tables of nent entries
// of at most k primes (an upper bound).
size_t operands_batch_task_bound(size_t nent);
// A device allocation shared by several owners (hipFree with the last).
std::shared_ptr<void> device_slab(size_t bytes, void** ptr);
// Host-side set-up only (primes; no device): lets planners run without a GPU.
int engine_host_init();
// The engine's stream, and its device arena and pinned staging buffer grown
// to at least `bytes` (both reused by every call made under the lock; a
// grown buffer replaces the old one only after the stream is idle).
hipStream_t engine_stream();
int engine_arena(size_t bytes, char** base);
int engine_stage(size_t bytes, char** host);

// Plans one product.  make_tables: create the twiddle tables of order 2^24
// when the plan needs them (false: a plan that needs missing tables fails
// with RF_ERR_CAPACITY instead, for callers that may not allocate).
int plan_product(Product& P, bool make_tables = true);
// Task counts per launch group, keyed as build_pipeline keys its launches
// (forward: fwd_key; inverse and inverse columns: lgN; CRT: primes).
struct TaskGroups {
  std::map<int, size_t> fwd, inv, col, crt;
  size_t row = 0;
};
Layout layout_of(std::vector<Product>& prods, TaskGroups* groups = nullptr);
int build_pipeline(std::vector<Product>& prods, Pipeline& pp);
ScratchPtrs scratch_ptrs(char* base, const Layout& L);
// Enqueues every kernel of a built pipeline whose task image already sits in
// the scratch region at `base`.  No host synchronisation, no allocation.
int launch_pipeline(const Pipeline& pp, char* base, hipStream_t st, bool prof);
// The same with the task image at `tasks` (device) instead of inside the
// scratch region, so that several built pipelines can share one scratch
// region and be enqueued back to back.
int launch_pipeline_tasks(const Pipeline& pp, char* base, const char* tasks, hipStream_t st, bool prof);
uint64_t bitlen_limbs(const uint64_t* x, uint64_t n);
// CRT blocks of one C entry, and its CRT tasks (coefficient residues at
// coef_off doubles, k primes x N, into out): pre bytes and aggregate slots
// are numbered on from npre and nblk_total.
uint32_t crt_blocks(const RfPlan& pl);
void crt_tasks(const RfPlan& pl, uint64_t coef_off, uint64_t* out, uint32_t& npre, uint32_t& nblk_total,
               std::vector<RfCrtTask>& v);
// Two-row split of a plan's transforms (column + row passes).
bool plan_split(const RfPlan& pl);
void tc_to_signmag(std::vector<uint64_t>& x, bool& neg);

}  // namespace rfi
