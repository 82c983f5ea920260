// rf_engine.cpp -- host orchestration and the C ABI of librforest_hip.so.
//
// One process-wide context holds the prime set, the constant and twiddle
// tables in device memory, a stream, a grow-only device arena and a pinned
// staging buffer.  A call (one product or a batch) plans every product,
// lays out its buffers in the arena, writes task tables, and enqueues
//   k_fwd (per transform length) -> k_inv (per length) -> k_crt (per prime
//   count) -> k_carry -> k_apply
// on the stream.  Host mpz entry points add one H2D copy of the limbs and
// one D2H copy of the results around that.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <map>
#include <atomic>
#include <mutex>
#include <unordered_map>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/rforest_hip.h"
#include "../../include/rforest_hip_stages.h"
#include "rf_hostpar.h"
#include "rf_internal.h"
#include "rf_kernels.h"
#include "rf_noalloc.h"
#include "rf_plan.h"
#include "rf_util.h"

extern "C" {
int hw_disable_fft = 0;
int mpzfft_threads = 1;
}


namespace rfi {

struct Ctx {
  bool ready = false;       // GPU state initialised
  bool host_ready = false;  // prime set initialised (no GPU needed)
  RfPrimes primes;
  RfConsts* d_consts = nullptr;
  double* d_twf = nullptr;
  double* d_twi = nullptr;
  RfDevTables tables{};
  hipStream_t stream = nullptr;
  // Marks the end of the last asynchronous batch on a caller's stream; the
  // library's stream waits on it, so every later user of the arena and the
  // staging buffer (all ordered after the library's stream) waits too.
  hipEvent_t foreign_ev = nullptr;
  char* d_arena = nullptr;
  size_t arena_cap = 0;
  char* h_stage = nullptr;
  char* d_stage = nullptr;  // the staging buffer's device address
  size_t stage_cap = 0;
  // Device blocks of a product tree (its inputs, each level's outputs, its
  // root), kept from one tree to the next (up to 512 MiB in all) so that a
  // tree of the same shape allocates nothing: (pointer, bytes), handed out in
  // order from tree_next.
  std::vector<std::pair<uint64_t*, size_t>> tree_blk;
  size_t tree_next = 0;
  // profiling
  bool prof = false;
  struct Ev {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Ev> pending;
  double prof_ms[RF_NSTAGES] = {0};
  uint64_t prof_n[RF_NSTAGES] = {0};
};

Ctx g;
std::mutex g_mu;
// Incremented by every set-up: a prepared batch's captured graph holds the
// table addresses of the set-up it was captured under, and is captured again
// when the generation has changed.
uint64_t g_gen = 0;
// Held shared by request-server workers while they plan and enqueue against
// the engine's tables without the engine mutex; hw_mpz_clear holds it
// exclusively while it waits for the device and frees the tables.
std::shared_mutex g_tables_rw;
thread_local std::string g_err;

// Twiddle tables beyond the set-up ones (2^RF_LGN_TW points, 38 MB): order
// 2^RF_LGN_MID for transforms up to that length (RF_TW_PRIMES x (2^20 +
// 2^21) pairs, 1.2 GB) and order 2^RF_LGN_MAX above it (x (2^23 + 2^24), 9.7
// GB), each made on the first plan that needs it and kept until
// hw_mpz_clear.  Server workers launch without the engine mutex, so the
// pointers are atomics and g_tw_mu alone serialises their creation.  All
// tables are powers of one root of order 2^RF_S per prime, so they agree on
// every index they share and a transform may mix row passes over one with
// column passes over another.
std::mutex g_tw_mu;
std::atomic<double*> g_twf_mid{nullptr}, g_twi_mid{nullptr};
std::atomic<double*> g_twf_long{nullptr}, g_twi_long{nullptr};

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}



int host_init_locked() {
  if (g.host_ready) return RF_OK;
  rf_primes_init(&g.primes);
  if (g.primes.k < RF_MAXP) return fail(RF_ERR_ARG, "prime generation found only %d primes", g.primes.k);
  g.host_ready = true;
  return RF_OK;
}

// Tuned planner rules: $RF_TUNE_FILE if set ("none": no rules), otherwise
// tuned_plans.txt beside the library when it exists.
void load_tuned_plans() {
  std::string path;
  const char* env = getenv("RF_TUNE_FILE");
  if (env) {
    if (strcmp(env, "none") == 0) return;
    path = env;
  } else {
    Dl_info info;
    if (!dladdr((void*)&hw_mpz_setup, &info) || !info.dli_fname) return;
    path = info.dli_fname;
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuned_plans.txt";
  }
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  fclose(f);
  (void)rf_tune_load(path.c_str());
}

int setup_locked() {
  if (g.ready) return RF_OK;
  std::unique_lock<std::shared_mutex> wl(g_tables_rw);  // server workers see all of it or none
  int rc0 = host_init_locked();
  if (rc0) return rc0;
  RfConsts* hc = (RfConsts*)calloc(1, sizeof(RfConsts));
  if (!hc) return fail(RF_ERR_NOMEM, "host allocation failed");
  rf_consts_init(&g.primes, hc);
  double roots[2 * RF_TW_PRIMES] = {0};
  rf_twiddle_roots(&g.primes, RF_LGN_TW, roots, roots + RF_TW_PRIMES);
  HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&g.foreign_ev, hipEventDisableTiming));
  HIPCHK(hipMalloc((void**)&g.d_consts, sizeof(RfConsts)));
  HIPCHK(hipMemcpy(g.d_consts, hc, sizeof(RfConsts), hipMemcpyHostToDevice));
  free(hc);
  // Twiddle tables for transforms up to 2^RF_LGN_TW points (RF_TW_PRIMES x
  // (2^15 + 2^16) pairs, 38 MB) are generated on the device from each
  // prime's root of order 2^RF_LGN_TW; longer ones are made on first use.
  const size_t twf_n = (size_t)RF_TW_PRIMES * RF_FWD_TW * 2, twi_n = (size_t)RF_TW_PRIMES * RF_INV_TW * 2;
  HIPCHK(hipMalloc((void**)&g.d_twf, twf_n * sizeof(double)));
  HIPCHK(hipMalloc((void**)&g.d_twi, twi_n * sizeof(double)));
  double* d_roots = nullptr;
  HIPCHK(hipMalloc((void**)&d_roots, sizeof roots));
  HIPCHK(hipMemcpy(d_roots, roots, sizeof roots, hipMemcpyHostToDevice));
  HIPCHK(rf_launch_twiddles(g.d_consts, d_roots, d_roots + RF_TW_PRIMES, RF_TW_PRIMES, RF_LGN_TW, g.d_twf, g.d_twi,
                            g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  HIPCHK(hipFree(d_roots));
  g.tables.consts = g.d_consts;
  g.tables.tw_fwd = g.d_twf;
  g.tables.tw_inv = g.d_twi;
  HIPCHK(rf_noalloc_init());  // this device's address of the no-allocation scratch
  g.ready = true;
  g_gen++;
  load_tuned_plans();
  return RF_OK;
}

// Frees the tables and buffers.  Work the library enqueued on any stream
// (asynchronous batches on callers' streams, prepared batches, server
// slots) may still read them, so the whole device is waited for first;
// server workers are held off by g_tables_rw meanwhile and find the engine
// not ready afterwards.
void clear_locked() {
  if (!g.ready) return;
  std::unique_lock<std::shared_mutex> wl(g_tables_rw);
  (void)hipDeviceSynchronize();
  for (auto& e : g.pending) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  g.pending.clear();
  hipFree(g.d_consts);
  hipFree(g.d_twf);
  hipFree(g.d_twi);
  {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    hipFree(g_twf_mid.exchange(nullptr));
    hipFree(g_twi_mid.exchange(nullptr));
    hipFree(g_twf_long.exchange(nullptr));
    hipFree(g_twi_long.exchange(nullptr));
  }
  if (g.d_arena) hipFree(g.d_arena);
  for (auto& b : g.tree_blk) hipFree(b.first);
  if (g.h_stage) hipHostFree(g.h_stage);
  if (g.foreign_ev) hipEventDestroy(g.foreign_ev);
  hipStreamDestroy(g.stream);
  g = Ctx();
}

// The tables a launch passes to its kernels (middle and long pointers null
// until made).
RfDevTables tables() {
  RfDevTables t = g.tables;
  t.tw_fwd_mid = g_twf_mid.load(std::memory_order_acquire);
  t.tw_inv_mid = g_twi_mid.load(std::memory_order_acquire);
  t.tw_fwd_long = g_twf_long.load(std::memory_order_acquire);
  t.tw_inv_long = g_twi_long.load(std::memory_order_acquire);
  return t;
}

// Whether the tables a transform of 2^lgN points reads exist.
bool tables_exist(int lgN) {
  if (lgN > RF_LGN_MID) return g_twi_long.load(std::memory_order_acquire) != nullptr;
  if (lgN > RF_LGN_TW) return g_twi_mid.load(std::memory_order_acquire) != nullptr;
  return true;
}

int ensure_long_tables(int lgN) {
  if (tables_exist(lgN)) return RF_OK;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  if (tables_exist(lgN)) return RF_OK;
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  const bool mid = lgN <= RF_LGN_MID;
  const int lgmax = mid ? RF_LGN_MID : RF_LGN_MAX;
  std::atomic<double*>& pf = mid ? g_twf_mid : g_twf_long;
  std::atomic<double*>& pi = mid ? g_twi_mid : g_twi_long;
  double roots[2 * RF_TW_PRIMES] = {0};
  rf_twiddle_roots(&g.primes, lgmax, roots, roots + RF_TW_PRIMES);
  const size_t twf_n = (size_t)RF_TW_PRIMES * ((size_t)1 << (lgmax - 1)) * 2,
               twi_n = (size_t)RF_TW_PRIMES * ((size_t)1 << lgmax) * 2;
  double *f = nullptr, *i = nullptr, *d_roots = nullptr;
  hipStream_t st = nullptr;
  hipError_t e = hipMalloc((void**)&f, twf_n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&i, twi_n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&d_roots, sizeof roots);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemcpyAsync(d_roots, roots, sizeof roots, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = rf_launch_twiddles(g.d_consts, d_roots, d_roots + RF_TW_PRIMES, RF_TW_PRIMES, lgmax, f, i, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (st) hipStreamDestroy(st);
  hipFree(d_roots);
  if (e != hipSuccess) {
    hipFree(f);
    hipFree(i);
    return fail(e == hipErrorOutOfMemory ? RF_ERR_NOMEM : RF_ERR_HIP,
                "twiddle tables for transforms of 2^%d points: %s", lgN, hipGetErrorString(e));
  }
  pf.store(f, std::memory_order_release);
  pi.store(i, std::memory_order_release);
  return RF_OK;
}

int ensure_arena(size_t bytes) {
  if (bytes <= g.arena_cap) return RF_OK;
  HIPCHK(hipStreamSynchronize(g.stream));
  if (g.d_arena) HIPCHK(hipFree(g.d_arena));
  g.d_arena = nullptr;
  size_t cap = std::max(bytes, g.arena_cap * 3 / 2);
  hipError_t e = hipMalloc((void**)&g.d_arena, cap);
  if (e != hipSuccess) {
    g.arena_cap = 0;
    return fail(RF_ERR_NOMEM, "device arena of %zu bytes: %s", cap, hipGetErrorString(e));
  }
  g.arena_cap = cap;
  return RF_OK;
}

int ensure_stage(size_t bytes) {
  if (bytes <= g.stage_cap) return RF_OK;
  HIPCHK(hipStreamSynchronize(g.stream));
  const size_t cap = std::max(bytes, g.stage_cap * 3 / 2);
  if (g.h_stage) HIPCHK(hipHostFree(g.h_stage));
  g.h_stage = g.d_stage = nullptr;
  g.stage_cap = 0;
  hipError_t e = hipHostMalloc((void**)&g.h_stage, cap, hipHostMallocDefault);
  if (e != hipSuccess) {
    g.h_stage = nullptr;
    return fail(RF_ERR_NOMEM, "pinned staging of %zu bytes: %s", cap, hipGetErrorString(e));
  }
  e = hipHostGetDevicePointer((void**)&g.d_stage, g.h_stage, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(g.h_stage);
    g.h_stage = g.d_stage = nullptr;
    return fail(RF_ERR_HIP, "staging device address: %s", hipGetErrorString(e));
  }
  g.stage_cap = cap;
  return RF_OK;
}

// ---- profiling around launches ----
struct ProfScope {
  int stage;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(int s, hipStream_t stream) : stage(s), st(stream) {
    if (!g.prof || stage < 0) return;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, st);
  }
  ~ProfScope() {
    if (!g.prof || stage < 0) return;
    hipEventRecord(b, st);
    g.pending.push_back({stage, a, b});
  }
};

void prof_collect() {
  for (auto& e : g.pending) {
    hipEventSynchronize(e.b);
    float ms = 0;
    hipEventElapsedTime(&ms, e.a, e.b);
    g.prof_ms[e.stage] += ms;
    g.prof_n[e.stage]++;
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  g.pending.clear();
}


// Tuning rules (rf_tune_*), read by every planning call; their own lock, since
// the request server plans outside the library lock.
std::mutex g_tune_mu;
std::vector<RfTuneRule> g_tune;

bool tuned_plan(uint64_t LA, uint64_t LB, int r, int d, RfPlan* out) {
  if (r != d || LA == 0 || LB == 0) return false;
  std::vector<RfTuneRule> rules;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& x : g_tune)
      if (x.r == r) rules.push_back(x);
  }
  if (rules.empty()) return false;
  bool found = false;
  double best = 0;
  for (int t = 1; t <= RF_TMAX; t++) {
    const uint64_t b = 64 * (uint64_t)t;
    const uint64_t nA = (LA + b - 1) / b, nB = (LB + b - 1) / b;
    int lgN = RF_LGN_MIN;
    while (lgN < RF_LGN_MAX && ((uint64_t)1 << lgN) < nA + nB - 1) lgN++;
    RfPlan pl;
    if (rf_plan_fixed(&g.primes, LA, LB, r, d, t, lgN, &pl) != 0) continue;
    for (const auto& x : rules) {
      if (x.lgN != lgN || x.t != t || (x.split && lgN != RF_LGN_LDS)) continue;
      if (!found || x.cost < best) {
        best = x.cost;
        found = true;
        pl.split = x.split ? 1 : 0;
        *out = pl;
      }
    }
  }
  return found;
}

int tables_for(int lgN, bool make) {
  if (make) return ensure_long_tables(lgN);
  if (!tables_exist(lgN))
    return fail(RF_ERR_CAPACITY, "a plan of 2^%d points needs twiddle tables that are not made (no allocation allowed)",
                lgN);
  return RF_OK;
}

// Plans of products without a resident operand depend only on (LA, LB, r,
// d) and the tuning rules; repeated shapes (a drop-in called in a loop, a
// tree level of equal sizes) reuse them.  Cleared whenever the rules change.
struct PlanKey {
  uint64_t la, lb;
  int r, d;
  bool operator==(const PlanKey& o) const { return la == o.la && lb == o.lb && r == o.r && d == o.d; }
};
struct PlanKeyHash {
  size_t operator()(const PlanKey& k) const {
    return std::hash<uint64_t>()(k.la * 0x9E3779B97F4A7C15ull ^ k.lb) ^ ((size_t)k.r << 40) ^ ((size_t)k.d << 48);
  }
};
std::mutex g_plan_mu;
std::unordered_map<PlanKey, RfPlan, PlanKeyHash> g_plan_cache;

void plan_cache_clear() {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_plan_cache.clear();
}

int plan_product(Product& P, bool make_tables) {
  uint64_t LA = 0, LB = 0;
  for (auto& e : P.A) LA = std::max(LA, e.sign ? e.bits : 0);
  for (auto& e : P.B) LB = std::max(LB, e.sign ? e.bits : 0);
  if (P.left) {
    const rf_operand& op = *P.left;
    if (op.r != P.r || op.d != P.d) return fail(RF_ERR_ARG, "resident operand is %dx%d, product %dx%d", op.r, op.d, P.r, P.d);
    if (op.la_bits == 0 || LB == 0) {  // a zero side: the product is zero
      rf_plan_fixed(&g.primes, 0, 0, P.r, P.d, 1, RF_LGN_MIN, &P.plan);
      return RF_OK;
    }
    const RfPlan& f = op.plan;
    if (f.cyclic) {  // the operand's wrap-around plan, if the right side's chunks fit it
      const uint64_t nB = (LB + 64 * (uint64_t)f.t - 1) / (64 * (uint64_t)f.t);
      if (nB > (uint64_t)f.nB)
        return fail(RF_ERR_ARG, "right operand of %llu bits does not fit the resident operand's plan",
                    (unsigned long long)LB);
      P.plan = f;
      return RF_OK;
    }
    if (rf_plan_fixed_k(&g.primes, op.la_bits, LB, P.r, P.d, f.t, f.lgN, f.k, &P.plan) != 0 || P.plan.k != f.k)
      return fail(RF_ERR_ARG, "right operand of %llu bits does not fit the resident operand's plan",
                  (unsigned long long)LB);
    return RF_OK;
  }
  const PlanKey key{LA, LB, P.r, P.d};
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plan_cache.find(key);
    if (it != g_plan_cache.end()) {
      P.plan = it->second;
      return tables_for(P.plan.lgN, make_tables);
    }
  }
  if (!(LA && LB && tuned_plan(LA, LB, P.r, P.d, &P.plan)) &&
      rf_plan_choose(&g.primes, LA, LB, P.r, P.d, &P.plan) != 0)
    return fail(RF_ERR_UNSUPPORTED,
                "no transform plan for entries of %llu x %llu bits (r=%d, d=%d)",
                (unsigned long long)LA, (unsigned long long)LB, P.r, P.d);
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (g_plan_cache.size() >= 4096) g_plan_cache.clear();
    g_plan_cache[key] = P.plan;
  }
  return tables_for(P.plan.lgN, make_tables);
}

uint64_t out_limbs_for(uint64_t la_bits, uint64_t lb_bits, int d) {
  if (la_bits == 0 || lb_bits == 0) return 0;
  int cl = 0;
  while ((1 << cl) < d) cl++;
  return (la_bits + lb_bits + cl + 1 + 63) / 64;
}


// Forward launch groups: one per transform length, and per chunk width
// where the kernels are specialised on it.
bool plan_split(const RfPlan& pl) { return pl.lgN > RF_LGN_LDS || pl.split; }
int fwd_key(const RfPlan& pl) {
  const bool sp = plan_split(pl);
  return (sp ? 1 << 12 : 0) | (pl.lgN << 4) | ((pl.lgN >= RF_FWD_TSPEC || sp) ? pl.t : 0);
}

uint32_t crt_blocks(const RfPlan& pl) {
  const uint64_t ndig = (pl.rn + pl.t - 1) / pl.t;
  const uint64_t D = RF_CRT_THREADS - rf_crt_halo(pl.k, pl.t);
  return (uint32_t)((ndig + D - 1) / D);
}

void crt_tasks(const RfPlan& pl, uint64_t coef_off, uint64_t* out, uint32_t& npre, uint32_t& nblk_total,
               std::vector<RfCrtTask>& v) {
  const uint32_t nblk = crt_blocks(pl);
  const int halo = rf_crt_halo(pl.k, pl.t);
  for (uint32_t b = 0; b < nblk; b++) {
    RfCrtTask t{};
    t.coef_off = coef_off;
    t.out = out;
    t.pre_off = npre;
    t.fun_idx = nblk_total + b;
    t.N = (uint32_t)1 << pl.lgN;
    t.nout = (uint32_t)pl.nout;
    t.rn = (uint32_t)pl.rn;
    t.digit0 = b * (uint32_t)(RF_CRT_THREADS - halo);
    t.t = (uint8_t)pl.t;
    t.k = (uint8_t)pl.k;
    t.s = (uint8_t)pl.s;
    t.halo = (uint8_t)halo;
    v.push_back(t);
    npre += RF_CRT_THREADS;
  }
  nblk_total += nblk;
}

Layout layout_of(std::vector<Product>& prods, TaskGroups* groups) {
  Layout L;
  // task counts per launch group, as build_pipeline groups them
  std::map<int, size_t> fwd_n, inv_n, col_n, crt_n;
  size_t row_n = 0;
  for (auto& P : prods) {
    const RfPlan& pl = P.plan;
    P.spec_off = L.spec_n;
    P.coef_off = L.coef_n;
    if (pl.k == 0) continue;
    const uint64_t N = (uint64_t)1 << pl.lgN;
    const int rd = P.r * P.d, dd = P.d * P.d;
    const int npoly = P.left ? dd : rd + dd;  // polys transformed here
    L.spec_n += (uint64_t)npoly * pl.k * N;
    L.coef_n += (uint64_t)rd * pl.k * N;
    fwd_n[fwd_key(pl)] += (size_t)npoly * pl.k;
    if (plan_split(pl)) {
      const size_t rows = (size_t)1 << (pl.lgN - RF_LGN_ROW);
      row_n += (size_t)npoly * pl.k * rows;
      inv_n[RF_LGN_ROW] += (size_t)rd * pl.k * rows;
      col_n[pl.lgN] += (size_t)rd * pl.k;
    } else {
      inv_n[pl.lgN] += (size_t)rd * pl.k;
    }
    crt_n[pl.k] += (size_t)rd * crt_blocks(pl);
    L.nblk += (uint32_t)(rd * crt_blocks(pl));
  }
  L.spec_b = align256(L.spec_n * 8);
  L.coef_b = align256(L.coef_n * 8);
  size_t tb = 0;
  for (auto& kv : fwd_n) tb += align256(kv.second * sizeof(RfFwdTask)), L.nfwd += kv.second;
  for (auto& kv : inv_n) tb += align256(kv.second * sizeof(RfInvTask)), L.ninv += kv.second;
  for (auto& kv : col_n) tb += align256(kv.second * sizeof(RfRowTask)), L.ncol += kv.second;
  tb += align256(row_n * sizeof(RfRowTask));
  L.nrow = row_n;
  for (auto& kv : crt_n) tb += align256(kv.second * sizeof(RfCrtTask)), L.ncrt += kv.second;
  L.tasks_b = tb;
  L.pre_b = align256((size_t)L.nblk * RF_CRT_THREADS);
  L.fun_b = align256(L.nblk);
  if (groups) {
    groups->fwd.swap(fwd_n);
    groups->inv.swap(inv_n);
    groups->col.swap(col_n);
    groups->crt.swap(crt_n);
    groups->row = row_n;
  }
  return L;
}

// A built pipeline: task tables in one host byte image (uploaded to the
// scratch region's task area) and the launch sequence over them.  Task
// tables hold absolute device pointers for limbs and outputs and offsets
// into the scratch region for spectra and coefficients, so one image serves
// every launch of the same batch.

int stage_of(int kind) {
  switch (kind) {
    case K_FWD: case K_FWD_COL: case K_FWD_ROW: return RF_STAGE_FWD;
    case K_INV: case K_INV_COL: return RF_STAGE_INV;
    case K_CRT: return RF_STAGE_CRT;
    default: return RF_STAGE_APPLY;
  }
}
template <typename T>
size_t append_tasks(std::vector<char>& img, const std::vector<T>& v) {
  size_t off = align256(img.size());
  img.resize(off + v.size() * sizeof(T));
  if (!v.empty()) memcpy(img.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

// Every task is written once, straight into the image: the groups' sizes
// (from the layout's counts) fix each group's place, in launch order (every
// forward pass, column passes before the row pass, then every inverse pass,
// row passes before the column passes, CRT, carry application), and a
// cursor per group takes the products' tasks in product order.
int build_pipeline(std::vector<Product>& prods, Pipeline& pp) {
  TaskGroups tg;
  pp.L = layout_of(prods, &tg);
  pp.launches.clear();
  size_t off = 0;
  std::map<int, size_t> fwd_at, inv_at, col_at;  // next task's byte offset, per group
  for (auto& kv : tg.fwd) {
    pp.launches.push_back({(kv.first >> 12) ? K_FWD_COL : K_FWD, kv.first & 4095, off, (int)kv.second});
    fwd_at[kv.first] = off;
    off = align256(off + kv.second * sizeof(RfFwdTask));
  }
  size_t row_at = off;
  if (tg.row) {
    pp.launches.push_back({K_FWD_ROW, RF_LGN_ROW, off, (int)tg.row});
    off = align256(off + tg.row * sizeof(RfRowTask));
  }
  for (auto& kv : tg.inv) {
    pp.launches.push_back({K_INV, kv.first, off, (int)kv.second});
    inv_at[kv.first] = off;
    off = align256(off + kv.second * sizeof(RfInvTask));
  }
  for (auto& kv : tg.col) {
    pp.launches.push_back({K_INV_COL, kv.first, off, (int)kv.second});
    col_at[kv.first] = off;
    off = align256(off + kv.second * sizeof(RfRowTask));
  }
  // the CRT tasks of every prime count in one array: one CRT launch per
  // prime count over its range, then one carry application over all of them
  // (k_apply does not depend on the prime count; a level of mixed sizes has
  // several counts, and each launch of a few microseconds would wait on the
  // one before)
  std::map<int, size_t> crt_at;
  const size_t crt0 = off;
  size_t ncrt = 0;
  for (auto& kv : tg.crt) {
    pp.launches.push_back({K_CRT, kv.first, off, (int)kv.second});
    crt_at[kv.first] = off;
    off += kv.second * sizeof(RfCrtTask);
    ncrt += kv.second;
  }
  if (ncrt) pp.launches.push_back({K_APPLY, 0, crt0, (int)ncrt});
  if (off > pp.L.tasks_b) return fail(RF_ERR_ARG, "internal: task tables exceed their layout");
  pp.image.resize(off);
  char* img = pp.image.data();
  uint32_t nblk_total = 0, npre = 0;
  std::vector<RfCrtTask> tv;
  for (auto& P : prods) {
    const RfPlan& pl = P.plan;
    if (pl.k == 0) continue;
    const uint64_t N = (uint64_t)1 << pl.lgN;
    const int r = P.r, d = P.d, k = pl.k;
    // poly e < r*d is A's entry e, the rest are B's; with a resident A the
    // batch's spectra hold B only and A's offsets count from its own base
    const int rd = r * d;
    auto spec_of = [&](int poly, int q) {
      const int pos = P.left ? poly - rd : poly;
      return P.spec_off + ((uint64_t)pos * k + q) * N;
    };
    auto aspec_of = [&](int poly, int q) {
      return P.left ? ((uint64_t)poly * k + q) * N : spec_of(poly, q);
    };
    const double* abase = P.left ? P.left->spec : nullptr;
    size_t& fa = fwd_at[fwd_key(pl)];
    for (int e = P.left ? rd : 0; e < r * d + d * d; e++) {
      const Entry& E = e < r * d ? P.A[e] : P.B[e - r * d];
      const bool isB = e >= r * d;
      for (int q = 0; q < k; q++, fa += sizeof(RfFwdTask)) {
        RfFwdTask t{};
        t.limbs = E.dptr;
        t.spec_off = spec_of(e, q);
        t.nlimbs = (uint32_t)E.nlimbs;
        t.nchunk = E.sign ? (uint32_t)std::min<uint64_t>((E.nlimbs + pl.t - 1) / pl.t, N) : 0;
        t.sign = E.sign;
        t.q = (uint16_t)q;
        t.t = (uint8_t)pl.t;
        t.scale = isB ? (uint8_t)pl.lgN : 0;
        memcpy(img + fa, &t, sizeof t);
      }
    }
    // Longer transforms: 2^(lgN-13) rows of 2^13 residues each.
    const bool split = plan_split(pl);
    const uint32_t rows = split ? 1u << (pl.lgN - RF_LGN_ROW) : 1u;
    const uint64_t rowlen = split ? (uint64_t)1 << RF_LGN_ROW : N;
    if (split)
      for (int e = P.left ? rd : 0; e < r * d + d * d; e++)
        for (int q = 0; q < k; q++)
          for (uint32_t h = 0; h < rows; h++, row_at += sizeof(RfRowTask)) {
            const RfRowTask t{spec_of(e, q) + h * rowlen, h, (uint16_t)q, (uint16_t)pl.lgN};
            memcpy(img + row_at, &t, sizeof t);
          }
    size_t& ia = inv_at[split ? RF_LGN_ROW : pl.lgN];
    size_t* ca = split ? &col_at[pl.lgN] : nullptr;
    for (int q = 0; q < k; q++)
      for (int i = 0; i < r; i++)
        for (int j = 0; j < d; j++) {
          const uint64_t out_off = P.coef_off + ((uint64_t)(i * d + j) * k + q) * N;
          for (uint32_t h = 0; h < rows; h++, ia += sizeof(RfInvTask)) {
            RfInvTask t{};
            t.a_off = aspec_of(i * d, q) + h * rowlen;
            t.a_stride = (uint32_t)(k * N);
            t.b_off = spec_of(r * d + j, q) + h * rowlen;
            t.b_stride = (uint32_t)(d * k * N);
            t.out_off = out_off + h * rowlen;
            t.q = (uint16_t)q;
            t.d = (uint16_t)d;
            t.a_base = abase;  // a resident operand's spectra, in the same launch as the rest
            memcpy(img + ia, &t, sizeof t);
          }
          if (ca) {
            const RfRowTask t{out_off, 0, (uint16_t)q, (uint16_t)pl.lgN};
            memcpy(img + *ca, &t, sizeof t);
            *ca += sizeof t;
          }
        }
    size_t& ta = crt_at[k];
    for (int e = 0; e < r * d; e++) {
      tv.clear();
      crt_tasks(pl, P.coef_off + (uint64_t)e * k * N, P.out[e], npre, nblk_total, tv);
      memcpy(img + ta, tv.data(), tv.size() * sizeof(RfCrtTask));
      ta += tv.size() * sizeof(RfCrtTask);
    }
  }
  return RF_OK;
}


ScratchPtrs scratch_ptrs(char* base, const Layout& L) {
  ScratchPtrs s;
  s.spec = (double*)base;
  s.coef = (double*)(base + L.spec_b);
  s.tasks = base + L.spec_b + L.coef_b;
  s.pre = (uint8_t*)(s.tasks + L.tasks_b);
  s.fun = s.pre + L.pre_b;
  return s;
}

int launch_pipeline(const Pipeline& pp, char* base, hipStream_t st, bool prof) {
  return launch_pipeline_tasks(pp, base, scratch_ptrs(base, pp.L).tasks, st, prof);
}

int launch_pipeline_tasks(const Pipeline& pp, char* base, const char* tasks, hipStream_t st, bool prof) {
  const ScratchPtrs S = scratch_ptrs(base, pp.L);
  for (const LaunchRec& l : pp.launches) {
    ProfScope ps(prof ? stage_of(l.kind) : -1, st);
    const char* t = tasks + l.off;
    switch (l.kind) {
      case K_FWD:  // param: fwd_key
        HIPCHK(rf_launch_fwd(tables(), l.param >> 4, l.param & 15, (const RfFwdTask*)t, l.n, S.spec, st));
        break;
      case K_FWD_COL:
        HIPCHK(rf_launch_fwd_col(tables(), l.param >> 4, l.param & 15, (const RfFwdTask*)t, l.n, S.spec, st));
        break;
      case K_FWD_ROW:
        HIPCHK(rf_launch_fwd_row(tables(), (const RfRowTask*)t, l.n, S.spec, st));
        break;
      case K_INV:
        HIPCHK(rf_launch_inv(tables(), l.param, (const RfInvTask*)t, l.n, S.spec, S.spec,
                             S.coef, st));
        break;
      case K_INV_COL:
        HIPCHK(rf_launch_inv_col(tables(), l.param, (const RfRowTask*)t, l.n, S.coef, st));
        break;
      case K_CRT:
        HIPCHK(rf_launch_crt(tables(), l.param, (const RfCrtTask*)t, l.n, S.coef, S.pre, S.fun, st));
        break;
      case K_APPLY:
        HIPCHK(rf_launch_apply((const RfCrtTask*)t, l.n, S.pre, S.fun, st));
        break;
    }
  }
  return RF_OK;
}

// Build, upload the task image through the pinned staging buffer, launch.
// The arena must hold base + L.total() bytes (scratch starts at base).
int run_pipeline(std::vector<Product>& prods, const Layout& L, size_t base, hipStream_t st) {
  if (g.arena_cap < base + L.total()) return fail(RF_ERR_NOMEM, "arena too small for pipeline");
  Pipeline pp;
  int rc = build_pipeline(prods, pp);
  if (rc) return rc;
  char* sbase = g.d_arena + base;
  const ScratchPtrs S = scratch_ptrs(sbase, pp.L);
  rc = ensure_stage(pp.image.size());
  if (rc) return rc;
  // Earlier copies out of the staging buffer, and kernels over the arena,
  // must finish before either is reused: they may be on the library's
  // stream or on st.
  HIPCHK(hipStreamSynchronize(st));
  if (st != g.stream) HIPCHK(hipStreamSynchronize(g.stream));
  if (!pp.image.empty()) {
    memcpy(g.h_stage, pp.image.data(), pp.image.size());
    HIPCHK(hipMemcpyAsync(S.tasks, g.h_stage, pp.image.size(), hipMemcpyHostToDevice, st));
  }
  return launch_pipeline(pp, sbase, st, g.prof);
}

uint64_t bitlen_limbs(const uint64_t* x, uint64_t n) {
  while (n && x[n - 1] == 0) n--;
  return n ? 64 * n - (uint64_t)__builtin_clzll(x[n - 1]) : 0;
}

// Host-array product: the engine behind every host entry point.  Entry e of
// A is ap[e] (an[e] limbs, sign as[e]); likewise B.  Per output entry e the
// sink receives its rn limbs of two's complement, read from the staging
// buffer (valid until the engine's next staging use; rn = 0: a zero entry).
using TcSink = int (*)(void* ctx, int e, const uint64_t* tc, size_t rn);

// One product through the staging buffer, laid out [task image | inputs]
// like the start of the device arena, so that one copy uploads both; the
// results come back without a copy of their own: the CRT and carry kernels
// store them straight into the staging buffer (pinned host memory, through
// its device address), over the task image and the inputs, which the copy
// has consumed by then.  The call waits once and hands the results to
// `sink` from the staging buffer.
int host_product(int r, int d, const uint64_t* const* ap, const uint64_t* an, const int* as,
                 const uint64_t* const* bp, const uint64_t* bn, const int* bs, TcSink sink, void* ctx,
                 const rf_operand* left = nullptr) {
  if (r <= 0 || d <= 0) return fail(RF_ERR_ARG, "bad dimensions r=%d d=%d", r, d);
  par_wake();  // copy workers asleep after an idle spell start spinning now
  Product P;
  P.r = r;
  P.d = d;
  P.left = left;  // when set, A's arrays are not read
  uint64_t total_limbs = 0;
  for (int e = left ? r * d : 0; e < r * d + d * d; e++) {
    const bool isA = e < r * d;
    const int ei = isA ? e : e - r * d;
    const uint64_t* p = isA ? ap[ei] : bp[ei];
    uint64_t n = isA ? an[ei] : bn[ei];
    int s = isA ? as[ei] : bs[ei];
    Entry E{};
    E.bits = s ? bitlen_limbs(p, n) : 0;
    E.nlimbs = (E.bits + 63) / 64;
    E.sign = E.bits ? (s < 0 ? -1 : 1) : 0;
    E.limb_off = total_limbs;
    total_limbs += (E.nlimbs + 1) & ~(uint64_t)1;  // entries start 16-byte aligned
    (isA ? P.A : P.B).push_back(E);
  }
  int rc = plan_product(P);
  if (rc) return rc;
  const RfPlan pl = P.plan;
  if (pl.k == 0) {  // a zero operand: all products are zero
    for (int e = 0; e < r * d && !rc; e++) rc = sink(ctx, e, nullptr, 0);
    return rc;
  }
  std::vector<Product> prods{P};
  const Layout L = layout_of(prods);
  const size_t in_b = align256(total_limbs * 8);
  const size_t out_n = (size_t)r * d * pl.rn;
  const size_t in_off = align256(L.tasks_b);
  // device arena: [task image | inputs | pipeline scratch]
  rc = ensure_arena(in_off + in_b + L.total());
  if (rc) return rc;
  rc = ensure_stage(std::max(in_off + in_b, out_n * 8));
  if (rc) return rc;
  // an earlier call's kernels may still write results into the staging buffer
  HIPCHK(hipStreamSynchronize(g.stream));
  uint64_t* hs = (uint64_t*)(g.h_stage + in_off);
  Product& Q = prods[0];
  std::vector<CopySeg> segs;
  for (int e = left ? r * d : 0; e < r * d + d * d; e++) {
    const bool isA = e < r * d;
    const int ei = isA ? e : e - r * d;
    const Entry& E = isA ? Q.A[ei] : Q.B[ei];
    if (E.nlimbs) segs.push_back({hs + E.limb_off, isA ? ap[ei] : bp[ei], E.nlimbs * 8});
  }
  par_copy(segs.data(), (int)segs.size());
  char* d_tasks = g.d_arena;
  uint64_t* d_in = (uint64_t*)(g.d_arena + in_off);
  for (auto* v : {&Q.A, &Q.B})
    for (auto& E : *v) E.dptr = d_in + E.limb_off;
  uint64_t* d_res = (uint64_t*)g.d_stage;
  for (int e = 0; e < r * d; e++) Q.out.push_back(d_res + (size_t)e * pl.rn);
  Pipeline pp;
  rc = build_pipeline(prods, pp);
  if (rc) return rc;
  if (pp.image.size() > in_off) return fail(RF_ERR_ARG, "internal: task image exceeds its staging room");
  memcpy(g.h_stage, pp.image.data(), pp.image.size());
  HIPCHK(hipMemcpyAsync(d_tasks, g.h_stage, in_off + total_limbs * 8, hipMemcpyHostToDevice, g.stream));
  rc = launch_pipeline_tasks(pp, g.d_arena + in_off + in_b, d_tasks, g.stream, g.prof);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g.stream));
  if (g.prof) prof_collect();
  for (int e = 0; e < r * d && !rc; e++) rc = sink(ctx, e, (const uint64_t*)g.h_stage + (size_t)e * pl.rn, pl.rn);
  return rc;
}

// Sign and magnitude of rn-limb two's complement results, several at once
// and in pieces over the copy pool.  A negative tc's magnitude is 2^(64 rn) - tc: zero below its lowest nonzero
// limb z, -tc[z] at z and ~tc[i] above, so every piece is independent, and
// its length is known before any limb is written (limb i > z of the
// magnitude is zero exactly when tc[i] is all ones).
struct TcOut {
  const uint64_t* tc;
  size_t rn;
  uint64_t* dst = nullptr;  // n limbs of room (set after tc_prepare)
  bool neg = false;
  size_t z = 0;  // lowest nonzero limb of a negative tc
  size_t n = 0;  // the magnitude's length without high zero limbs
};
constexpr size_t RF_TC_PIECE = 8192;  // limbs per task

void tc_prepare(TcOut& o) {
  o.neg = o.rn && (o.tc[o.rn - 1] >> 63);
  o.z = 0;
  size_t n = o.rn;
  if (o.neg) {
    while (o.tc[o.z] == 0) o.z++;  // stops: the top limb of a negative tc is nonzero
    while (n > o.z + 1 && o.tc[n - 1] == ~(uint64_t)0) n--;
  } else {
    while (n && o.tc[n - 1] == 0) n--;
  }
  o.n = n;
}

// Writes every prepared result's n magnitude limbs to its dst.
void tc_convert_many(const std::vector<TcOut>& v) {
  std::vector<size_t> first(v.size() + 1, 0);
  for (size_t k = 0; k < v.size(); k++) first[k + 1] = first[k] + (v[k].n + RF_TC_PIECE - 1) / RF_TC_PIECE;
  struct Job {
    const std::vector<TcOut>* v;
    const size_t* first;
  } job{&v, first.data()};
  par_for((int)first[v.size()], [](void* vp, int i) {
    const Job& j = *(const Job*)vp;
    size_t k = 0;
    while (j.first[k + 1] <= (size_t)i) k++;
    const TcOut& o = (*j.v)[k];
    const size_t lo = ((size_t)i - j.first[k]) * RF_TC_PIECE;
    const size_t hi = std::min(o.n, lo + RF_TC_PIECE);
    if (!o.neg) {
      memcpy(o.dst + lo, o.tc + lo, (hi - lo) * 8);
      return;
    }
    for (size_t l = lo; l < hi; l++) o.dst[l] = l < o.z ? 0 : (l == o.z ? 0 - o.tc[l] : ~o.tc[l]);
  }, &job);
}

void tc_to_signmag(std::vector<uint64_t>& x, bool& neg) {
  neg = !x.empty() && (x.back() >> 63);
  if (neg) {
    uint64_t br = 1;
    for (auto& v : x) {
      uint64_t y = ~v + br;
      br = (br && y == 0);
      v = y;
    }
  }
  while (!x.empty() && x.back() == 0) x.pop_back();
}

int mpz_product(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d, const rf_operand* left = nullptr) {
  std::vector<const uint64_t*> ap(r * d), bp(d * d);
  std::vector<uint64_t> an(r * d), bn(d * d);
  std::vector<int> as(r * d), bs(d * d);
  for (int e = 0; e < (left ? 0 : r * d); e++) {
    ap[e] = (const uint64_t*)A[e]->_mp_d;
    an[e] = mpz_size(A[e]);
    as[e] = mpz_sgn(A[e]);
  }
  for (int e = 0; e < d * d; e++) {
    bp[e] = (const uint64_t*)B[e]->_mp_d;
    bn[e] = mpz_size(B[e]);
    bs[e] = mpz_sgn(B[e]);
  }
  // results go straight from the staging buffer into C (no aliasing of C
  // with A or B, src/include/hwmpz.h:128): the sink makes room in each
  // entry, then all entries convert at once over the copy pool while the
  // staging buffer still holds them (the engine's lock is held throughout)
  struct Out {
    mpz_t* C;
    std::vector<TcOut> outs;
    std::vector<int> ents;
  } out{C, {}, {}};
  auto sink = [](void* ctx, int e, const uint64_t* tc, size_t rn) -> int {
    Out& o = *(Out*)ctx;
    if (rn == 0) {
      mpz_set_ui(o.C[e], 0);
      return RF_OK;
    }
    _mpz_realloc(o.C[e], (mp_size_t)rn);  // src/lib/hwmpz.c:114
    TcOut t;
    t.tc = tc;
    t.rn = rn;
    tc_prepare(t);
    t.dst = (uint64_t*)o.C[e]->_mp_d;
    o.outs.push_back(t);
    o.ents.push_back(e);
    return RF_OK;
  };
  int rc = host_product(r, d, ap.data(), an.data(), as.data(), bp.data(), bn.data(), bs.data(), sink, &out, left);
  if (rc) return rc;
  tc_convert_many(out.outs);
  for (size_t k = 0; k < out.outs.size(); k++) {
    const TcOut& t = out.outs[k];
    out.C[out.ents[k]]->_mp_size = t.n == 0 ? 0 : (t.neg ? -(int)t.n : (int)t.n);  // src/lib/hwmpz.c:116-122
  }
  return RF_OK;
}

// ---- products beyond one device plan ----
//
// The reference's CPU mpz_rmatrix_mult_fft (src/lib/hwmpz.c:65-130) takes
// entries of any size; the device plans stop at 2^32-bit entries (2^24
// points of 512-bit chunks), and a large plan may not fit in device memory.
// Then every entry x is cut into pieces of P bits carrying its sign,
// x = sum_u x_u 2^(u P) with x_u = sgn(x) (floor(|x| / 2^(u P)) mod 2^P), and
// C = sum_s 2^(s P) D_s with D_s = sum_{u+v=s} A_u B_v, where A_u B_v is an
// ordinary (device) product of the piece matrices.  C is formed by Horner's
// rule from the top diagonal down.
constexpr uint64_t RF_PIECE_MAX = (uint64_t)1 << 31;  // pieces 2^31 x 2^31 bits: 2^23-point plans
constexpr uint64_t RF_PIECE_MIN = (uint64_t)1 << 16;
uint64_t g_piece_override = 0;  // rf_stage_dropin_piece_bits: cut above this size (0: only when needed)

uint64_t max_bits(mpz_t* X, int n) {
  uint64_t m = 0;
  for (int e = 0; e < n; e++)
    if (mpz_sgn(X[e])) m = std::max<uint64_t>(m, mpz_sizeinbase(X[e], 2));
  return m;
}

struct MpzVec {
  std::vector<__mpz_struct> v;
  explicit MpzVec(size_t n) : v(n) {
    for (auto& z : v) mpz_init(&z);
  }
  ~MpzVec() {
    for (auto& z : v) mpz_clear(&z);
  }
  mpz_t* data() { return reinterpret_cast<mpz_t*>(v.data()); }
  __mpz_struct* operator[](size_t i) { return &v[i]; }
};

// piece u (P bits, the entry's sign) of x into out
void mpz_piece(__mpz_struct* out, const __mpz_struct* x, uint64_t u, uint64_t P, __mpz_struct* tmp) {
  mpz_tdiv_q_2exp(tmp, x, u * P);
  mpz_tdiv_r_2exp(out, tmp, P);
}

int mpz_product_pieces(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d, uint64_t P) {
  const uint64_t LA = max_bits(A, r * d), LB = max_bits(B, d * d);
  const uint64_t na = (LA + P - 1) / P, nb = (LB + P - 1) / P;
  if (na == 0 || nb == 0) {
    for (int e = 0; e < r * d; e++) mpz_set_ui(C[e], 0);
    return RF_OK;
  }
  const size_t rd = (size_t)r * d, dd = (size_t)d * d;
  MpzVec Bp(nb * dd), Au(rd), T(rd), tmp(1);
  for (uint64_t v = 0; v < nb; v++)
    for (size_t e = 0; e < dd; e++) mpz_piece(Bp[v * dd + e], B[e], v, P, tmp[0]);
  MpzVec D((na + nb - 1) * rd);  // diagonal sums D_s, entry after entry
  for (uint64_t u = 0; u < na; u++) {
    for (size_t e = 0; e < rd; e++) mpz_piece(Au[e], A[e], u, P, tmp[0]);
    for (uint64_t v = 0; v < nb; v++) {
      int rc = mpz_product(T.data(), Au.data(), r, Bp.data() + v * dd, d);
      if (rc) return rc;
      for (size_t e = 0; e < rd; e++) mpz_add(D[(u + v) * rd + e], D[(u + v) * rd + e], T[e]);
    }
  }
  for (size_t e = 0; e < rd; e++) {
    mpz_set(C[e], D[(na + nb - 2) * rd + e]);
    for (uint64_t s = na + nb - 1; s-- > 0;) {
      if (s == na + nb - 2) continue;
      mpz_mul_2exp(C[e], C[e], P);
      mpz_add(C[e], C[e], D[s * rd + e]);
    }
  }
  return RF_OK;
}

// The drop-in: one device product when a plan exists and fits, otherwise the
// product in pieces (halved further while the device runs out of memory).
int mpz_product_any(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d) {
  const uint64_t L = std::max(max_bits(A, r * d), max_bits(B, d * d));
  uint64_t P = g_piece_override;
  if (!P || L <= P) {
    const int rc = mpz_product(C, A, r, B, d);
    if (rc != RF_ERR_UNSUPPORTED && rc != RF_ERR_NOMEM) return rc;
    P = std::min(RF_PIECE_MAX, ((L + 1) / 2 + 63) & ~(uint64_t)63);
  }
  for (;;) {
    const int rc = mpz_product_pieces(C, A, r, B, d, P);
    if (rc != RF_ERR_NOMEM || P <= RF_PIECE_MIN) return rc;
    P = std::max(RF_PIECE_MIN, (P / 2 + 63) & ~(uint64_t)63);
  }
}

int engine_host_init() { return host_init_locked(); }
hipStream_t engine_stream() { return g.stream; }
int engine_arena(size_t bytes, char** base) {
  int rc = ensure_arena(bytes);
  if (rc) return rc;
  *base = (char*)g.d_arena;
  return RF_OK;
}
int engine_stage(size_t bytes, char** host) {
  int rc = ensure_stage(bytes);
  if (rc) return rc;
  *host = (char*)g.h_stage;
  return RF_OK;
}

std::mutex& engine_mutex() { return g_mu; }
std::shared_mutex& engine_tables_lock() { return g_tables_rw; }
bool engine_ready() { return g.ready; }
RfDevTables engine_tables() { return tables(); }
int engine_long_tables(int lgN) { return ensure_long_tables(lgN); }
const RfPrimes& engine_primes() { return g.primes; }

}  // namespace rfi

using namespace rfi;

extern "C" {

void hw_mpz_setup(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (hw_disable_fft) return;
  int rc = setup_locked();
  if (rc) {
    fprintf(stderr, "hw_mpz_setup: %s\n", g_err.c_str());
    abort();
  }
}

void hw_mpz_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_locked();
}

int rf_mpz_rmatrix_mult(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (r <= 0 || d <= 0) return fail(RF_ERR_ARG, "bad dimensions r=%d d=%d", r, d);
  return mpz_product_any(C, A, r, B, d);
}

int rf_stage_tc_to_signmag(const uint64_t* tc, uint64_t rn, uint64_t* dst, uint64_t cap, long* nsigned) {
  if ((rn && (!tc || !dst)) || !nsigned) return fail(RF_ERR_ARG, "rf_stage_tc_to_signmag: null argument");
  std::vector<TcOut> v(1);
  v[0].tc = tc;
  v[0].rn = rn;
  tc_prepare(v[0]);
  if (v[0].n > cap) return fail(RF_ERR_CAPACITY, "magnitude needs %zu limbs", v[0].n);
  v[0].dst = dst;
  tc_convert_many(v);
  *nsigned = v[0].neg ? -(long)v[0].n : (long)v[0].n;
  return RF_OK;
}

int rf_stage_copy_segments(void* const* dst, const void* const* src, const uint64_t* n, int nseg) {
  if (nseg < 0 || (nseg && (!dst || !src || !n))) return fail(RF_ERR_ARG, "rf_stage_copy_segments: bad arguments");
  std::vector<CopySeg> segs(nseg);
  for (int s = 0; s < nseg; s++) segs[s] = CopySeg{dst[s], src[s], (size_t)n[s]};
  par_copy(segs.data(), nseg);
  return RF_OK;
}

int rf_stage_dropin_piece_bits(uint64_t bits) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (bits && bits < 64) return fail(RF_ERR_ARG, "piece size below 64 bits");
  g_piece_override = bits ? (bits + 63) & ~(uint64_t)63 : 0;
  return RF_OK;
}

void mpz_rmatrix_mult_fft(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d, mpz_t w) {
  (void)w;
  int rc = rf_mpz_rmatrix_mult(C, A, r, B, d);
  if (rc) {
    fprintf(stderr, "mpz_rmatrix_mult_fft: %s\n", g_err.c_str());
    abort();
  }
}

void cu_mpz_rmatrix_mult_fft(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d, mpz_t w) {
  mpz_rmatrix_mult_fft(C, A, r, B, d, w);
}

int mpz_rmatrix_gpu_fft_suitable(mpz_t* A, int r, mpz_t* B, int d) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (host_init_locked()) return 0;
  uint64_t LA = 0, LB = 0;
  for (int e = 0; e < r * d; e++) LA = std::max<uint64_t>(LA, mpz_sgn(A[e]) ? mpz_sizeinbase(A[e], 2) : 0);
  for (int e = 0; e < d * d; e++) LB = std::max<uint64_t>(LB, mpz_sgn(B[e]) ? mpz_sizeinbase(B[e], 2) : 0);
  RfPlan pl;
  return rf_plan_choose(&g.primes, LA, LB, r, d, &pl) == 0;
}

const char* rf_last_error(void) { return g_err.c_str(); }

int rf_rmatrix_mult_limbs(uint64_t* c, const size_t* coff, const size_t* ccap, long* csize,
                          const uint64_t* a, const size_t* aoff, const size_t* an, const int* as,
                          const uint64_t* b, const size_t* boff, const size_t* bn, const int* bs,
                          int r, int d) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (r <= 0 || d <= 0) return fail(RF_ERR_ARG, "bad dimensions r=%d d=%d", r, d);
  std::vector<const uint64_t*> ap(r * d), bp(d * d);
  std::vector<uint64_t> anv(r * d), bnv(d * d);
  for (int e = 0; e < r * d; e++) ap[e] = a + aoff[e], anv[e] = an[e];
  for (int e = 0; e < d * d; e++) bp[e] = b + boff[e], bnv[e] = bn[e];
  // each result's length is known before it is written: the capacity check
  // comes first, then all results convert at once over the copy pool
  // straight into the caller's array (the staging buffer still holds them)
  struct Ctx {
    uint64_t* c;
    const size_t *coff, *ccap;
    long* csize;
    std::vector<TcOut> outs;
    std::vector<int> ents;
  } cx{c, coff, ccap, csize, {}, {}};
  auto sink = [](void* vp, int e, const uint64_t* tc, size_t rn) -> int {
    Ctx& x = *(Ctx*)vp;
    x.csize[e] = 0;
    if (rn == 0) return RF_OK;
    TcOut t;
    t.tc = tc;
    t.rn = rn;
    tc_prepare(t);
    if (t.n > x.ccap[e]) return fail(RF_ERR_CAPACITY, "output %d needs %zu limbs", e, t.n);
    t.dst = x.c + x.coff[e];
    x.outs.push_back(t);
    x.ents.push_back(e);
    return RF_OK;
  };
  int rc = host_product(r, d, ap.data(), anv.data(), as, bp.data(), bnv.data(), bs, sink, &cx);
  if (rc) return rc;
  tc_convert_many(cx.outs);
  for (size_t k = 0; k < cx.outs.size(); k++) {
    const TcOut& t = cx.outs[k];
    cx.csize[cx.ents[k]] = t.n == 0 ? 0 : (t.neg ? -(long)t.n : (long)t.n);
  }
  return RF_OK;
}

uint64_t rf_out_limbs(uint64_t la_limbs, uint64_t lb_limbs, int d) {
  return out_limbs_for(64 * la_limbs, 64 * lb_limbs, d);
}

}  // extern "C"

namespace rfi {
int products_from_dev(const rf_dev_product* prods, int nprod, std::vector<Product>& ps, bool make_tables = true) {
  if (nprod < 0 || (nprod > 0 && !prods)) return fail(RF_ERR_ARG, "bad product list");
  ps.assign(nprod, Product());
  for (int i = 0; i < nprod; i++) {
    const rf_dev_product& dp = prods[i];
    Product& P = ps[i];
    P.r = dp.r;
    P.d = dp.d;
    P.left = dp.left;
    if (P.r <= 0 || P.d <= 0 || (!dp.A && !dp.left) || !dp.B || !dp.C)
      return fail(RF_ERR_ARG, "product %d: bad descriptor", i);
    for (int e = dp.left ? P.r * P.d : 0; e < P.r * P.d + P.d * P.d; e++) {
      const rf_dev_entry& x = e < P.r * P.d ? dp.A[e] : dp.B[e - P.r * P.d];
      Entry E{};
      E.dptr = x.limbs;
      E.nlimbs = x.nlimbs;
      E.sign = x.nlimbs ? (x.sign > 0 ? 1 : (x.sign < 0 ? -1 : 0)) : 0;
      E.bits = E.sign ? 64 * x.nlimbs : 0;
      (e < P.r * P.d ? P.A : P.B).push_back(E);
    }
    int rc = plan_product(P, make_tables);
    if (rc) return rc;
    const uint64_t rn = P.plan.k ? P.plan.rn : 0;  // a zero product writes no limbs
    for (int e = 0; e < P.r * P.d; e++) {
      if (dp.C[e].cap < rn)
        return fail(RF_ERR_CAPACITY, "product %d output %d needs %llu limbs", i, e, (unsigned long long)rn);
      P.out.push_back(dp.C[e].limbs);
    }
  }
  return RF_OK;
}
}  // namespace

// A prepared batch runs as `lanes` independent pipelines over disjoint
// product sets (balanced by planned cost), each on its own stream and
// scratch: the lanes' kernels overlap, so one lane's memory-bound phases
// fill another's compute-bound ones and launch gaps are hidden.
#define RF_MAX_LANES 4
struct rf_batch {
  int lanes = 1;
  std::vector<std::vector<Product>> prods;  // per lane
  std::vector<Pipeline> pp;                 // per lane
  std::vector<size_t> off;                  // per lane scratch offset
  char* scratch = nullptr;
  hipStream_t side[RF_MAX_LANES] = {};
  hipEvent_t fork = nullptr, join[RF_MAX_LANES] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  uint64_t graph_gen = 0;  // engine generation the graph was captured under
};

namespace {
// Enqueues every lane: lane 0 on st, lane i on side[i], forked from and
// joined back into st with events (also under stream capture).
int batch_enqueue(rf_batch* b, hipStream_t st, bool prof) {
  if (b->lanes > 1) HIPCHK(hipEventRecord(b->fork, st));
  for (int l = 0; l < b->lanes; l++) {
    hipStream_t ls = l == 0 ? st : b->side[l];
    if (l > 0) HIPCHK(hipStreamWaitEvent(ls, b->fork, 0));
    int rc = launch_pipeline(b->pp[l], b->scratch + b->off[l], ls, prof);
    if (rc) return rc;
    if (l > 0) {
      HIPCHK(hipEventRecord(b->join[l], ls));
      HIPCHK(hipStreamWaitEvent(st, b->join[l], 0));
    }
  }
  return RF_OK;
}

void batch_free(rf_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->scratch) (void)hipFree(b->scratch);
  for (int l = 1; l < RF_MAX_LANES; l++) {
    if (b->side[l]) (void)hipStreamDestroy(b->side[l]);
    if (b->join[l]) (void)hipEventDestroy(b->join[l]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
}
}  // namespace

extern "C" {

int rf_batch_mult_device(const rf_dev_product* prods, int nprod, void* stream, int async) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  hipStream_t st = stream ? (hipStream_t)stream : g.stream;
  std::vector<Product> ps;
  int rc = products_from_dev(prods, nprod, ps);
  if (rc) return rc;
  const Layout L = layout_of(ps);
  rc = ensure_arena(L.total());
  if (rc) return rc;
  rc = run_pipeline(ps, L, 0, st);
  if (rc) return rc;
  if (!async) {
    HIPCHK(hipStreamSynchronize(st));
    if (g.prof) prof_collect();
  } else if (st != g.stream) {
    // the pipeline still reads the arena and the task image: order every
    // later arena or staging user after it
    HIPCHK(hipEventRecord(g.foreign_ev, st));
    HIPCHK(hipStreamWaitEvent(g.stream, g.foreign_ev, 0));
  }
  return RF_OK;
}

int rf_batch_mult_device_noalloc(const rf_dev_product* prods, int nprod, void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  char* scratch = nullptr;
  HIPCHK(rf_noalloc_scratch(&scratch));
  hipStream_t st = stream ? (hipStream_t)stream : g.stream;
  std::vector<Product> ps;
  int rc = products_from_dev(prods, nprod, ps, /*make_tables=*/false);
  if (rc) return rc;
  Pipeline pp;
  rc = build_pipeline(ps, pp);
  if (rc) return rc;
  if (pp.L.total() > RF_NOALLOC_SCRATCH_BYTES)
    return fail(RF_ERR_CAPACITY, "products need %zu bytes of scratch; the fixed scratch holds %zu",
                pp.L.total(), (size_t)RF_NOALLOC_SCRATCH_BYTES);
  // the task image goes from host memory straight into the scratch's task
  // area; the call returns after completion, so the image outlives the copy
  if (!pp.image.empty())
    HIPCHK(hipMemcpyAsync(scratch_ptrs(scratch, pp.L).tasks, pp.image.data(), pp.image.size(),
                          hipMemcpyHostToDevice, st));
  rc = launch_pipeline(pp, scratch, st, g.prof);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(st));
  if (g.prof) prof_collect();
  return RF_OK;
}

int rf_batch_create(const rf_dev_product* prods, int nprod, rf_batch** out) {
  return rf_batch_create_lanes(prods, nprod, 1, out);
}

int rf_batch_create_lanes(const rf_dev_product* prods, int nprod, int lanes, rf_batch** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!out) return fail(RF_ERR_ARG, "rf_batch_create: null output");
  *out = nullptr;
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (lanes < 1 || lanes > RF_MAX_LANES) return fail(RF_ERR_ARG, "lanes must be 1..%d", RF_MAX_LANES);
  std::vector<Product> all;
  int rc = products_from_dev(prods, nprod, all);
  if (rc) return rc;
  rf_batch* b = new rf_batch();
  b->lanes = std::max(1, std::min<int>(lanes, (int)all.size()));
  // greedy balance by planned cost, largest first, order kept within a lane
  std::vector<size_t> order(all.size());
  for (size_t i = 0; i < all.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return all[x].plan.cost * all[x].r * all[x].d > all[y].plan.cost * all[y].r * all[y].d; });
  std::vector<std::vector<size_t>> pick(b->lanes);
  std::vector<double> load(b->lanes, 0.0);
  for (size_t i : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    pick[l].push_back(i);
    load[l] += all[i].plan.cost * all[i].r * all[i].d;
  }
  b->prods.resize(b->lanes);
  b->pp.resize(b->lanes);
  b->off.resize(b->lanes);
  size_t total = 0;
  for (int l = 0; l < b->lanes && !rc; l++) {
    std::sort(pick[l].begin(), pick[l].end());
    for (size_t i : pick[l]) b->prods[l].push_back(all[i]);
    rc = build_pipeline(b->prods[l], b->pp[l]);
    b->off[l] = total;
    total += align256(b->pp[l].L.total());
  }
  if (!rc) {
    hipError_t e = hipMalloc((void**)&b->scratch, std::max<size_t>(total, 256));
    if (e != hipSuccess) rc = fail(RF_ERR_NOMEM, "batch scratch of %zu bytes: %s", total, hipGetErrorString(e));
  }
  for (int l = 0; l < b->lanes && !rc; l++) {
    if (!b->pp[l].image.empty()) {
      const ScratchPtrs S = scratch_ptrs(b->scratch + b->off[l], b->pp[l].L);
      hipError_t e = hipMemcpy(S.tasks, b->pp[l].image.data(), b->pp[l].image.size(), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = fail(RF_ERR_HIP, "task upload: %s", hipGetErrorString(e));
    }
    if (!rc && l > 0) {
      hipError_t e = hipStreamCreateWithFlags(&b->side[l], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming);
      if (e != hipSuccess) rc = fail(RF_ERR_HIP, "batch lane: %s", hipGetErrorString(e));
    }
  }
  if (!rc && b->lanes > 1) {
    hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e != hipSuccess) rc = fail(RF_ERR_HIP, "batch fork event: %s", hipGetErrorString(e));
  }
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return RF_OK;
}

int rf_batch_launch(rf_batch* b, void* stream, int use_graph) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!b) return fail(RF_ERR_ARG, "rf_batch_launch: null batch");
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  hipStream_t st = stream ? (hipStream_t)stream : g.stream;
  if (!use_graph || g.prof) return batch_enqueue(b, st, g.prof);
  if (b->exec && (b->graph_stream != st || b->graph_gen != g_gen)) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  if (!b->exec) {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, st, false);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(st, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(RF_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(gr);
      return fail(RF_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    b->graph = gr;
    b->graph_stream = st;
    b->graph_gen = g_gen;
  }
  HIPCHK(hipGraphLaunch(b->exec, st));
  return RF_OK;
}

void rf_batch_destroy(rf_batch* b) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!b) return;
  if (g.ready) (void)hipDeviceSynchronize();
  batch_free(b);
}

}  // extern "C"

namespace {
// Forward transforms of a left operand's entries (device limbs in A[e].dptr)
// into op->spec; task tables go to the arena at byte offset `base`.
int operand_transform(rf_operand* op, const std::vector<Entry>& A, size_t base) {
  const RfPlan& pl = op->plan;
  const int k = pl.k, lgN = pl.lgN;
  const uint64_t N = (uint64_t)1 << lgN;
  const bool split = plan_split(pl);
  const uint32_t rows = split ? 1u << (lgN - RF_LGN_ROW) : 0u;
  std::vector<RfFwdTask> ft;
  std::vector<RfRowTask> rt;
  for (size_t e = 0; e < A.size(); e++)
    for (int q = 0; q < k; q++) {
      RfFwdTask t{};
      t.limbs = A[e].dptr;
      t.spec_off = ((uint64_t)e * k + q) * N;
      t.nlimbs = (uint32_t)A[e].nlimbs;
      t.nchunk = A[e].sign ? (uint32_t)std::min<uint64_t>((A[e].nlimbs + pl.t - 1) / pl.t, N) : 0;
      t.sign = A[e].sign;
      t.q = (uint16_t)q;
      t.t = (uint8_t)pl.t;
      t.scale = 0;
      ft.push_back(t);
      for (uint32_t h = 0; h < rows; h++) rt.push_back(RfRowTask{t.spec_off + ((uint64_t)h << RF_LGN_ROW), h, (uint16_t)q, (uint16_t)lgN});
    }
  const size_t fb = align256(ft.size() * sizeof(RfFwdTask)), rb = align256(rt.size() * sizeof(RfRowTask));
  int rc = ensure_arena(base + fb + rb);
  if (rc) return rc;
  rc = ensure_stage(fb + rb);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g.stream));
  memcpy(g.h_stage, ft.data(), ft.size() * sizeof(RfFwdTask));
  if (!rt.empty()) memcpy(g.h_stage + fb, rt.data(), rt.size() * sizeof(RfRowTask));
  RfFwdTask* d_ft = (RfFwdTask*)(g.d_arena + base);
  RfRowTask* d_rt = (RfRowTask*)(g.d_arena + base + fb);
  HIPCHK(hipMemcpyAsync(d_ft, g.h_stage, fb + rb, hipMemcpyHostToDevice, g.stream));
  if (split) {
    HIPCHK(rf_launch_fwd_col(tables(), lgN, pl.t, d_ft, (int)ft.size(), op->spec, g.stream));
    HIPCHK(rf_launch_fwd_row(tables(), d_rt, (int)rt.size(), op->spec, g.stream));
  } else {
    HIPCHK(rf_launch_fwd(tables(), lgN, pl.t, d_ft, (int)ft.size(), op->spec, g.stream));
  }
  HIPCHK(hipStreamSynchronize(g.stream));
  return RF_OK;
}

// Plans and allocates an operand for entries A (bits/sign filled), then
// transforms them.
int operand_make(std::vector<Entry>& A, int r, int d, uint64_t b_max_bits, size_t base, rf_operand** out,
                 const RfPlan* fixed = nullptr) {
  rf_operand* op = new rf_operand();
  op->r = r;
  op->d = d;
  for (auto& e : A) {
    op->la_bits = std::max(op->la_bits, e.sign ? e.bits : 0);
    op->signs.push_back(e.sign);
  }
  int rc = RF_OK;
  if (op->la_bits && b_max_bits) {
    if (fixed)
      op->plan = *fixed;
    else if (!tuned_plan(op->la_bits, b_max_bits, r, d, &op->plan) &&
             rf_plan_choose(&g.primes, op->la_bits, b_max_bits, r, d, &op->plan) != 0)
      rc = fail(RF_ERR_UNSUPPORTED, "no transform plan for a %dx%d operand of %llu bits and right entries of %llu bits",
                r, d, (unsigned long long)op->la_bits, (unsigned long long)b_max_bits);
    if (!rc) rc = ensure_long_tables(op->plan.lgN);
    if (!rc) {
      const size_t bytes = (size_t)r * d * op->plan.k * ((size_t)8 << op->plan.lgN);
      hipError_t e = hipMalloc((void**)&op->spec, bytes);
      if (e != hipSuccess) rc = fail(RF_ERR_NOMEM, "operand spectra of %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    if (!rc) rc = operand_transform(op, A, base);
  } else {
    op->la_bits = 0;  // a zero operand (or zero right operands): every product is zero
  }
  if (rc) {
    if (op->spec) (void)hipFree(op->spec);
    delete op;
    return rc;
  }
  *out = op;
  return RF_OK;
}
}  // namespace

extern "C" {

int rf_operand_create_device(const rf_dev_entry* A, int r, int d, uint64_t b_max_bits, rf_operand** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!out) return fail(RF_ERR_ARG, "rf_operand_create_device: null output");
  *out = nullptr;
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (r <= 0 || d <= 0 || !A) return fail(RF_ERR_ARG, "bad operand r=%d d=%d", r, d);
  std::vector<Entry> E(r * d);
  for (int e = 0; e < r * d; e++) {
    E[e].dptr = A[e].limbs;
    E[e].nlimbs = A[e].nlimbs;
    E[e].sign = A[e].nlimbs ? (A[e].sign > 0 ? 1 : (A[e].sign < 0 ? -1 : 0)) : 0;
    E[e].bits = E[e].sign ? 64 * A[e].nlimbs : 0;
  }
  return operand_make(E, r, d, b_max_bits, 0, out);
}

int rf_operand_create(mpz_t* A, int r, int d, uint64_t b_max_bits, rf_operand** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!out) return fail(RF_ERR_ARG, "rf_operand_create: null output");
  *out = nullptr;
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (r <= 0 || d <= 0 || !A) return fail(RF_ERR_ARG, "bad operand r=%d d=%d", r, d);
  std::vector<Entry> E(r * d);
  uint64_t total = 0;
  for (int e = 0; e < r * d; e++) {
    E[e].nlimbs = mpz_size(A[e]);
    E[e].sign = mpz_sgn(A[e]);
    E[e].bits = E[e].sign ? mpz_sizeinbase(A[e], 2) : 0;
    E[e].limb_off = total;
    total += (E[e].nlimbs + 1) & ~(uint64_t)1;  // entries start 16-byte aligned
  }
  const size_t in_b = align256(total * 8 + 8);
  // the arena must not grow (and move) once the limbs are in it
  int rc = ensure_arena(in_b + operand_task_bound(r * d));
  if (rc) return rc;
  rc = ensure_stage(in_b);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g.stream));
  uint64_t* hs = (uint64_t*)g.h_stage;
  for (int e = 0; e < r * d; e++)
    if (E[e].nlimbs) memcpy(hs + E[e].limb_off, A[e]->_mp_d, E[e].nlimbs * 8);
  uint64_t* d_in = (uint64_t*)g.d_arena;
  if (total) HIPCHK(hipMemcpyAsync(d_in, hs, total * 8, hipMemcpyHostToDevice, g.stream));
  for (auto& e : E) e.dptr = d_in + e.limb_off;
  // the transform reuses the staging buffer only after synchronising the stream
  return operand_make(E, r, d, b_max_bits, in_b, out);
}

int rf_operand_mult(const rf_operand* left, mpz_t* C, mpz_t* B) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (!left) return fail(RF_ERR_ARG, "rf_operand_mult: null operand");
  return mpz_product(C, nullptr, left->r, B, left->d, left);
}

int rf_operand_plan(const rf_operand* op, int* t, int* lgN, int* k) {
  if (!op) return fail(RF_ERR_ARG, "rf_operand_plan: null operand");
  *t = op->plan.t;
  *lgN = op->plan.lgN;
  *k = op->plan.k;
  return RF_OK;
}

void rf_operand_destroy(rf_operand* op) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!op) return;
  if (g.ready) (void)hipDeviceSynchronize();
  if (op->slab)
    op->slab.reset();
  else if (op->spec)
    (void)hipFree(op->spec);
  delete op;
}

}  // extern "C"

namespace {
// A d x d matrix of sign/magnitude entries in device memory.
using Mat = std::vector<Entry>;

// Blocks kept after a tree, at most this many bytes in all: a larger tree's
// allocation cost is small beside its transforms, and its blocks go back.
constexpr size_t kTreeKeepBytes = (size_t)512 << 20;

// After a tree: its blocks stay for the next one when they total at most
// kTreeKeepBytes, else all are freed.
void tree_blocks_trim() {
  size_t total = 0;
  for (auto& b : g.tree_blk) total += b.second;
  if (total <= kTreeKeepBytes) return;
  (void)hipStreamSynchronize(g.stream);
  for (auto& b : g.tree_blk) (void)hipFree(b.first);
  g.tree_blk.clear();
}

// The next device block of the current tree, of at least `bytes`: the block
// an earlier tree took at the same point when it is large enough, else a new
// one in its place (nothing of the current tree lives there yet).
int tree_block(size_t bytes, uint64_t** out) {
  *out = nullptr;
  if (bytes == 0) return RF_OK;
  if (g.tree_next < g.tree_blk.size() && g.tree_blk[g.tree_next].second >= bytes) {
    *out = g.tree_blk[g.tree_next++].first;
    return RF_OK;
  }
  const size_t cap = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
  if (g.tree_next < g.tree_blk.size()) {
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipFree(g.tree_blk[g.tree_next].first));
    g.tree_blk[g.tree_next] = {nullptr, 0};
  } else {
    g.tree_blk.push_back({nullptr, 0});
  }
  uint64_t* p = nullptr;
  hipError_t e = hipMalloc((void**)&p, cap);
  if (e != hipSuccess) return fail(RF_ERR_NOMEM, "product tree block of %zu bytes: %s", cap, hipGetErrorString(e));
  g.tree_blk[g.tree_next++] = {p, cap};
  *out = p;
  return RF_OK;
}

// A d x d matrix of sign/magnitude entries in device memory.
using Mat = std::vector<Entry>;

// One level of a product tree, planned before the level below has run: pairs
// (M_2i, M_2i+1) multiplied in one pipeline, outputs converted in place to
// sign and magnitude (the next level's input format); an odd last matrix is
// carried up.  A level's inputs are planned from their limb bounds (the
// conversion leaves every output zero-padded to its bound) with a provisional
// sign; the true signs reach the forward tasks on the device through sign
// patches from the conversion's metadata, written before the level runs.
struct TreeLevel {
  std::vector<Product> ps;
  Pipeline pp;
  std::vector<RfConvTask> conv;
  std::vector<RfSignPatch> patch;
  uint32_t meta0 = 0;   // the level's first metadata slot (one per output entry)
  uint64_t max_rn = 0;  // longest output, limbs
};

// src[m][e]: the metadata slot holding the sign of entry e of matrix m, or -1
// (the sign is the host's).  Plans and builds the level over `cur` into L and
// makes `cur`, `src` the level's outputs (plus a carried matrix).
int tree_plan_level(std::vector<Mat>& cur, std::vector<std::vector<int32_t>>& src, int d, uint32_t& nmeta,
                    TreeLevel& L) {
  const int np = (int)cur.size() / 2, dd = d * d;
  L.ps.resize(np);
  size_t total = 0;
  std::vector<size_t> base(np);
  for (int i = 0; i < np; i++) {
    Product& P = L.ps[i];
    P.r = P.d = d;
    P.A = cur[2 * i];
    P.B = cur[2 * i + 1];
    int rc = plan_product(P);
    if (rc) return rc;
    const uint64_t rn = P.plan.k ? P.plan.rn : 0;
    base[i] = total;
    total += (size_t)dd * rn;
    L.max_rn = std::max(L.max_rn, rn);
  }
  uint64_t* buf = nullptr;
  int rc = tree_block(total * 8, &buf);
  if (rc) return rc;
  L.meta0 = nmeta;
  nmeta += (uint32_t)(np * dd);
  for (int i = 0; i < np; i++)
    for (int e = 0; e < dd; e++) {
      const uint64_t rn = L.ps[i].plan.k ? L.ps[i].plan.rn : 0;
      L.ps[i].out.push_back(rn ? buf + base[i] + (size_t)e * rn : nullptr);
      L.conv.push_back(RfConvTask{L.ps[i].out.back(), rn});
    }
  rc = build_pipeline(L.ps, L.pp);
  if (rc) return rc;
  // forward tasks in build_pipeline's order: per group, product by product,
  // A's entries then B's, prime by prime
  std::map<int, size_t> goff, gcnt;
  for (const LaunchRec& l : L.pp.launches) {
    if (l.kind == K_FWD) goff[l.param] = l.off;
    if (l.kind == K_FWD_COL) goff[(1 << 12) | l.param] = l.off;
  }
  L.patch.reserve((size_t)np * 2 * dd * (L.ps.empty() ? 0 : L.ps[0].plan.k));
  for (int i = 0; i < np; i++) {
    const RfPlan& pl = L.ps[i].plan;
    if (!pl.k) continue;
    const int key = fwd_key(pl);
    size_t& c = gcnt[key];
    const size_t g0 = goff[key] + offsetof(RfFwdTask, sign);
    for (int e = 0; e < 2 * dd; e++) {
      const int32_t sv = e < dd ? src[2 * i][e] : src[2 * i + 1][e - dd];
      if (sv < 0) {
        c += pl.k;
        continue;
      }
      for (int q = 0; q < pl.k; q++, c++) L.patch.push_back(RfSignPatch{g0 + c * sizeof(RfFwdTask), (uint32_t)sv, 0});
    }
  }
  // an output entry, a sum of d products, has at most bits(A) + bits(B) +
  // ceil(log2 d) bits: the bound the next level plans with (its buffer of rn
  // limbs holds the magnitude zero-padded)
  int cl = 0;
  while ((1 << cl) < d) cl++;
  std::vector<Mat> next;
  std::vector<std::vector<int32_t>> nsrc;
  for (int i = 0; i < np; i++) {
    Mat M(dd);
    std::vector<int32_t> sv(dd, -1);
    const uint64_t rn = L.ps[i].plan.k ? L.ps[i].plan.rn : 0;
    uint64_t la = 0, lb = 0;
    for (auto& x : L.ps[i].A) la = std::max(la, x.sign ? x.bits : 0);
    for (auto& x : L.ps[i].B) lb = std::max(lb, x.sign ? x.bits : 0);
    const uint64_t bits = rn ? std::min<uint64_t>(la + lb + cl, 64 * rn) : 0;
    for (int e = 0; e < dd; e++) {
      M[e].dptr = L.ps[i].out[e];
      M[e].nlimbs = (bits + 63) / 64;
      M[e].sign = rn ? 1 : 0;  // provisional: patched on the device
      M[e].bits = bits;
      if (rn) sv[e] = (int32_t)(L.meta0 + i * dd + e);
    }
    next.push_back(M);
    nsrc.push_back(sv);
  }
  if (cur.size() % 2) {
    next.push_back(cur.back());
    nsrc.push_back(src.back());
  }
  cur.swap(next);
  src.swap(nsrc);
  return RF_OK;
}

// The whole tree, level after level with no wait between them: the host
// plans, builds and stages level l + 1 while the GPU runs level l.  Device
// arena: [pipeline scratch, reused by the levels in turn | per level: task
// image, conversion tasks, sign patches]; the table slices (and their
// staging slices) follow one another until a level does not fit, which waits
// for the levels before it and starts again at the front.  stage_busy: bytes
// at the front of the staging buffer that an earlier copy may still read
// (the tree's leaves).  The root's entries (sign/magnitude) are copied to
// out + e * out_cap in device memory, or with out null to the staging buffer
// (which then holds at least d * d * out_cap limbs; its address is taken
// after the levels, which may have grown and moved it), their signs and
// lengths to sign[e], nlimbs[e].
// The largest scratch and table totals of the trees so far: a tree makes
// room for them before its first level, so that one of the same shape runs
// every level without a wait.
size_t g_tree_scratch = 0, g_tree_tables = 0;

int tree_product(std::vector<Mat> cur, int d, uint64_t* out, uint64_t out_cap, int* sign, uint64_t* nlimbs,
                 size_t stage_busy = 0) {
  const int dd = d * d;
  std::vector<std::vector<int32_t>> src(cur.size(), std::vector<int32_t>(dd, -1));
  // metadata slots: one per output entry of every level
  size_t nslots = 0;
  for (size_t m = cur.size(); m > 1; m = (m + 1) / 2) nslots += (m / 2) * (size_t)dd;
  RfEntryMeta* d_meta = nullptr;
  int rc = tree_block(nslots * sizeof(RfEntryMeta), (uint64_t**)&d_meta);
  if (rc) return rc;
  uint32_t nmeta = 0, root_meta = 0;
  const size_t busy = align256(stage_busy);
  if (cur.size() > 1) {
    rc = ensure_arena(g_tree_scratch + g_tree_tables);
    if (!rc) rc = ensure_stage(busy + g_tree_tables);
    if (rc) return rc;
  }
  size_t S = g_tree_scratch, tab = 0, stg = busy;  // scratch bytes; table and staging cursors
  size_t tables = 0;                                // this tree's table bytes
  bool levels = false;
  while (cur.size() > 1) {
    TreeLevel L;
    rc = tree_plan_level(cur, src, d, nmeta, L);
    if (rc) return rc;
    const size_t img_b = align256(L.pp.image.size());
    const size_t conv_b = align256(L.conv.size() * sizeof(RfConvTask));
    const size_t slice = img_b + conv_b + align256(L.patch.size() * sizeof(RfSignPatch));
    if (L.pp.L.total() > S || S + tab + slice > g.arena_cap || stg + slice > g.stage_cap) {
      // the levels before are done: their tables, the scratch and the staging are free
      HIPCHK(hipStreamSynchronize(g.stream));
      S = std::max(S, L.pp.L.total());
      tab = 0;
      stg = 0;
      rc = ensure_arena(S + slice);
      if (!rc) rc = ensure_stage(slice);
      if (rc) return rc;
    }
    char* dt = g.d_arena + S + tab;
    char* ht = g.h_stage + stg;
    par_memcpy(ht, L.pp.image.data(), L.pp.image.size());
    memcpy(ht + img_b, L.conv.data(), L.conv.size() * sizeof(RfConvTask));
    if (!L.patch.empty()) memcpy(ht + img_b + conv_b, L.patch.data(), L.patch.size() * sizeof(RfSignPatch));
    HIPCHK(hipMemcpyAsync(dt, ht, slice, hipMemcpyHostToDevice, g.stream));
    HIPCHK(rf_launch_sign_patch(dt, (const RfSignPatch*)(dt + img_b + conv_b), (int)L.patch.size(), d_meta,
                                g.stream));
    rc = launch_pipeline_tasks(L.pp, g.d_arena, dt, g.stream, g.prof);
    if (rc) return rc;
    HIPCHK(rf_launch_signmag((const RfConvTask*)(dt + img_b), (int)L.conv.size(), d_meta + L.meta0, L.max_rn,
                             g.stream));
    tab += slice;
    stg += slice;
    tables += slice;
    root_meta = L.meta0;
    levels = true;
  }
  g_tree_scratch = std::max(g_tree_scratch, S);
  g_tree_tables = std::max(g_tree_tables, tables);
  if (levels) {
    // the root is the last level's output; the copy follows every upload on the stream
    HIPCHK(hipMemcpyAsync(g.h_stage, d_meta + root_meta, (size_t)dd * sizeof(RfEntryMeta), hipMemcpyDeviceToHost,
                          g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (g.prof) prof_collect();
    const RfEntryMeta* m = (const RfEntryMeta*)g.h_stage;
    for (int e = 0; e < dd; e++) {
      cur[0][e].sign = m[e].sign;
      cur[0][e].nlimbs = m[e].nlimbs;
    }
  }
  const hipMemcpyKind kind = out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (!out) out = (uint64_t*)g.h_stage;
  for (int e = 0; e < dd && !rc; e++) {
    const Entry& E = cur[0][e];
    if (E.sign && E.nlimbs > out_cap) rc = fail(RF_ERR_CAPACITY, "root entry %d needs %llu limbs", e,
                                                (unsigned long long)E.nlimbs);
    sign[e] = E.sign;
    nlimbs[e] = E.sign ? E.nlimbs : 0;
    if (!rc && E.sign && E.nlimbs) {
      hipError_t x = hipMemcpyAsync(out + (size_t)e * out_cap, E.dptr, E.nlimbs * 8, kind, g.stream);
      if (x != hipSuccess) rc = fail(RF_ERR_HIP, "root copy: %s", hipGetErrorString(x));
    }
  }
  if (!rc) {
    hipError_t x = hipStreamSynchronize(g.stream);
    if (x != hipSuccess) rc = fail(RF_ERR_HIP, "tree: %s", hipGetErrorString(x));
  }
  return rc;
}
}  // namespace

extern "C" {

int rf_tree_product_device(const rf_dev_entry* mats, int n, int d, uint64_t* out, uint64_t out_cap, int* sign,
                           uint64_t* nlimbs) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (n <= 0 || d <= 0 || !mats || !out || !sign || !nlimbs) return fail(RF_ERR_ARG, "rf_tree_product_device: bad arguments");
  std::vector<Mat> cur(n, Mat(d * d));
  for (int m = 0; m < n; m++)
    for (int e = 0; e < d * d; e++) {
      const rf_dev_entry& x = mats[(size_t)m * d * d + e];
      Entry& E = cur[m][e];
      E.dptr = x.limbs;
      E.nlimbs = x.nlimbs;
      E.sign = x.nlimbs ? (x.sign > 0 ? 1 : (x.sign < 0 ? -1 : 0)) : 0;
      E.bits = E.sign ? 64 * x.nlimbs : 0;
    }
  g.tree_next = 0;
  const int rc = tree_product(cur, d, out, out_cap, sign, nlimbs);
  tree_blocks_trim();
  return rc;
}

int rf_mpz_tree_product(mpz_t* C, mpz_t* mats, int n, int d) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g.ready) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (n <= 0 || d <= 0 || !mats || !C) return fail(RF_ERR_ARG, "rf_mpz_tree_product: bad arguments");
  const size_t ne = (size_t)n * d * d;
  std::vector<uint64_t> off(ne);
  uint64_t total = 0, bound = 0;
  std::vector<uint64_t> mx(n, 0);
  for (size_t i = 0; i < ne; i++) {
    off[i] = total;
    const uint64_t nl = mpz_size(mats[i]);
    total += (nl + 1) & ~(uint64_t)1;
    mx[i / (d * d)] = std::max<uint64_t>(mx[i / (d * d)], nl);
  }
  // a node's entries have at most bits(left) + bits(right) + ceil(log2 d)
  // bits, so the root's have at most 64 sum(leaf limbs) + (n - 1) ceil(log2 d)
  int cl = 0;
  while ((1 << cl) < d) cl++;
  for (int m = 0; m < n; m++) bound += mx[m];
  bound += ((uint64_t)(n - 1) * cl + 63) / 64 + 1;
  // the tree's blocks: its leaves, then a block per level; the root's
  // entries come back straight into the staging buffer
  g.tree_next = 0;
  uint64_t* d_in = nullptr;
  int rc = tree_block(total * 8, &d_in);
  if (!rc) rc = ensure_stage(std::max<uint64_t>(total * 8, (size_t)d * d * bound * 8));
  if (rc) return rc;
  std::vector<Mat> cur(n, Mat(d * d));
  HIPCHK(hipStreamSynchronize(g.stream));
  std::vector<CopySeg> segs;
  for (size_t i = 0; i < ne; i++) {
    const uint64_t nl = mpz_size(mats[i]);
    if (nl) segs.push_back({(uint64_t*)g.h_stage + off[i], mats[i]->_mp_d, nl * 8});
    Entry& E = cur[i / (d * d)][i % (d * d)];
    E.dptr = d_in + off[i];
    E.nlimbs = nl;
    E.sign = mpz_sgn(mats[i]);
    E.bits = E.sign ? mpz_sizeinbase(mats[i], 2) : 0;
  }
  par_copy(segs.data(), (int)segs.size());
  if (total) HIPCHK(hipMemcpyAsync(d_in, g.h_stage, total * 8, hipMemcpyHostToDevice, g.stream));
  std::vector<int> sg(d * d);
  std::vector<uint64_t> nls(d * d);
  rc = tree_product(cur, d, nullptr, bound, sg.data(), nls.data(), total * 8);
  tree_blocks_trim();
  if (rc) return rc;
  segs.clear();
  for (int e = 0; e < d * d; e++) {
    if (!sg[e]) {
      mpz_set_ui(C[e], 0);
      continue;
    }
    _mpz_realloc(C[e], (mp_size_t)nls[e]);
    segs.push_back({C[e]->_mp_d, (uint64_t*)g.h_stage + (size_t)e * bound, nls[e] * 8});
    C[e]->_mp_size = sg[e] < 0 ? -(int)nls[e] : (int)nls[e];
  }
  par_copy(segs.data(), (int)segs.size());
  return RF_OK;
}

int rf_plan_query(uint64_t la_bits, uint64_t lb_bits, int r, int d, int* t, int* lgN, int* k) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = host_init_locked();
  if (rc) return rc;
  if (r <= 0 || d <= 0) return fail(RF_ERR_ARG, "bad dimensions r=%d d=%d", r, d);
  RfPlan pl;
  if (!(la_bits && lb_bits && tuned_plan(la_bits, lb_bits, r, d, &pl)) &&
      rf_plan_choose(&g.primes, la_bits, lb_bits, r, d, &pl) != 0)
    return fail(RF_ERR_UNSUPPORTED, "no plan");
  *t = pl.t;
  *lgN = pl.lgN;
  *k = pl.k;
  return RF_OK;
}

int rf_tune_set(int r, int lgN, int t, int split, double cost_us) {
  if (r <= 0 || lgN < RF_LGN_MIN || lgN > RF_LGN_MAX || t < 1 || t > RF_TMAX || !(cost_us >= 0))
    return fail(RF_ERR_ARG, "rf_tune_set: bad measurement");
  plan_cache_clear();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& x : g_tune)
    if (x.r == r && x.lgN == lgN && x.t == t && x.split == (split != 0)) {
      x.cost = cost_us;
      return RF_OK;
    }
  g_tune.push_back(RfTuneRule{r, lgN, t, split != 0, cost_us});
  return RF_OK;
}

void rf_tune_clear(void) {
  plan_cache_clear();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

int rf_tune_load(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return fail(RF_ERR_ARG, "cannot open %s", path);
  char line[256];
  int n = 0, rc = RF_OK;
  while (fgets(line, sizeof line, f)) {
    int r, lgN, t, sp;
    double cost;
    if (line[0] == '#') continue;
    if (sscanf(line, "%d %d %d %d %lf", &r, &lgN, &t, &sp, &cost) != 5) continue;
    rc = rf_tune_set(r, lgN, t, sp, cost);
    if (rc) break;
    n++;
  }
  fclose(f);
  return rc ? rc : n;
}

int rf_primes(uint64_t* out, int max) {
  RfPrimes pr;
  rf_primes_init(&pr);
  for (int i = 0; i < max && i < pr.k; i++) out[i] = pr.p[i];
  return pr.k;
}

void rf_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g.prof = on != 0;
}

int rf_profile_read(double* ms, uint64_t* launches, int nstages) {
  std::lock_guard<std::mutex> lk(g_mu);
  prof_collect();
  for (int i = 0; i < nstages && i < RF_NSTAGES; i++) {
    ms[i] = g.prof_ms[i];
    launches[i] = g.prof_n[i];
  }
  return RF_OK;
}

void rf_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  prof_collect();
  for (int i = 0; i < RF_NSTAGES; i++) g.prof_ms[i] = 0, g.prof_n[i] = 0;
}

// ---- device and pinned memory helpers ----

void* cu_malloc_pinned(size_t n) {
  void* p = nullptr;
  if (hipHostMalloc(&p, n ? n : 1, hipHostMallocDefault) != hipSuccess) {
    fail(RF_ERR_NOMEM, "hipHostMalloc(%zu) failed", n);
    return nullptr;
  }
  return p;
}

void cu_free_pinned(void* p) {
  if (p) (void)hipHostFree(p);
}

void* cu_device_malloc(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, n ? n : 1) != hipSuccess) {
    fail(RF_ERR_NOMEM, "hipMalloc(%zu) failed", n);
    return nullptr;
  }
  return p;
}

void cu_device_free(void* p) {
  if (p) (void)hipFree(p);
}

void cu_memcpy_d2h(void* dst, const void* src, size_t n) {
  hipError_t e = hipMemcpy(dst, src, n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    fprintf(stderr, "cu_memcpy_d2h: %s\n", hipGetErrorString(e));
    abort();
  }
}

int rf_memcpy_h2d(void* dst, const void* src, size_t n) {
  HIPCHK(hipMemcpy(dst, src, n, hipMemcpyHostToDevice));
  return RF_OK;
}

int rf_memcpy_d2d(void* dst, const void* src, size_t n) {
  HIPCHK(hipMemcpy(dst, src, n, hipMemcpyDeviceToDevice));
  return RF_OK;
}

int rf_device_sync(void) {
  HIPCHK(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(g_mu);
  if (g.prof) prof_collect();
  return RF_OK;
}

// ---- stage entry points ----

uint64_t rf_root(int q, int lgN) {
  RfPrimes pr;
  rf_primes_init(&pr);
  if (q < 0 || q >= pr.k || lgN < 0 || lgN > RF_S) return 0;
  uint64_t p = pr.p[q];
  uint64_t w = pr.g[q];
  for (int i = lgN; i < RF_S; i++) w = (uint64_t)(((unsigned __int128)w * w) % p);
  return w;
}

static void canon_out(const double* x, size_t n, uint64_t p, uint64_t* out) {
  for (size_t i = 0; i < n; i++) {
    long long v = (long long)x[i];
    long long m = v % (long long)p;
    if (m < 0) m += (long long)p;
    out[i] = (uint64_t)m;
  }
}

int rf_stage_fwd(const uint64_t* limbs, const size_t* off, const size_t* nlimbs, const int* signs,
                 int n, int t, uint32_t nchunk, int lgN, int q, int scale, uint64_t* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = setup_locked();
  if (rc) return rc;
  if (n <= 0 || t < 1 || t > RF_TMAX || lgN < RF_LGN_MIN || lgN > RF_LGN_MAX || q < 0 || q >= RF_TW_PRIMES ||
      (scale != 0 && scale != lgN) || nchunk > (1u << lgN))
    return fail(RF_ERR_ARG, "rf_stage_fwd: bad arguments");
  rc = ensure_long_tables(lgN);
  if (rc) return rc;
  const size_t N = (size_t)1 << lgN;
  size_t total = 0;
  for (int e = 0; e < n; e++) total = std::max(total, off[e] + nlimbs[e]);
  const bool split = lgN > RF_LGN_LDS;
  const uint32_t rows = split ? 1u << (lgN - RF_LGN_ROW) : 0u;
  const size_t in_b = align256(total * 8 + 8), spec_b = align256((size_t)n * N * 8),
               task_b = align256(n * sizeof(RfFwdTask)) + align256((size_t)n * rows * sizeof(RfRowTask));
  rc = ensure_arena(in_b + spec_b + task_b);
  if (rc) return rc;
  rc = ensure_stage(std::max(in_b + task_b, spec_b));
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g.stream));
  uint64_t* d_in = (uint64_t*)g.d_arena;
  double* d_spec = (double*)(g.d_arena + in_b);
  RfFwdTask* d_task = (RfFwdTask*)(g.d_arena + in_b + spec_b);
  memcpy(g.h_stage, limbs, total * 8);
  RfFwdTask* ht = (RfFwdTask*)(g.h_stage + in_b);
  for (int e = 0; e < n; e++) {
    RfFwdTask tk{};
    tk.limbs = d_in + off[e];
    tk.spec_off = (uint64_t)e * N;
    tk.nlimbs = (uint32_t)nlimbs[e];
    tk.nchunk = nchunk;
    tk.sign = signs[e] < 0 ? -1 : (signs[e] > 0 ? 1 : 0);
    tk.q = (uint16_t)q;
    tk.t = (uint8_t)t;
    tk.scale = (uint8_t)scale;
    ht[e] = tk;
  }
  RfRowTask* hr = (RfRowTask*)((char*)ht + align256(n * sizeof(RfFwdTask)));
  RfRowTask* d_row = (RfRowTask*)((char*)d_task + align256(n * sizeof(RfFwdTask)));
  for (int e = 0; e < n; e++)
    for (uint32_t h = 0; h < rows; h++)
      hr[(size_t)e * rows + h] = RfRowTask{(uint64_t)e * N + ((uint64_t)h << RF_LGN_ROW), h, (uint16_t)q, (uint16_t)lgN};
  HIPCHK(hipMemcpyAsync(d_in, g.h_stage, total * 8, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(d_task, ht, task_b, hipMemcpyHostToDevice, g.stream));
  if (split) {
    HIPCHK(rf_launch_fwd_col(tables(), lgN, t, d_task, n, d_spec, g.stream));
    HIPCHK(rf_launch_fwd_row(tables(), d_row, n * (int)rows, d_spec, g.stream));
  } else {
    HIPCHK(rf_launch_fwd(tables(), lgN, t, d_task, n, d_spec, g.stream));
  }
  HIPCHK(hipStreamSynchronize(g.stream));
  HIPCHK(hipMemcpy(g.h_stage, d_spec, (size_t)n * N * 8, hipMemcpyDeviceToHost));
  canon_out((const double*)g.h_stage, (size_t)n * N, g.primes.p[q], out);
  return RF_OK;
}

int rf_stage_inv(const uint64_t* specA, const uint64_t* specB, int r, int d, int lgN, int q,
                 uint64_t* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = setup_locked();
  if (rc) return rc;
  if (r <= 0 || d <= 0 || lgN < RF_LGN_MIN || lgN > RF_LGN_MAX || q < 0 || q >= RF_TW_PRIMES)
    return fail(RF_ERR_ARG, "rf_stage_inv: bad arguments");
  rc = ensure_long_tables(lgN);
  if (rc) return rc;
  const size_t N = (size_t)1 << lgN;
  const size_t np = (size_t)(r * d + d * d);
  const bool split = lgN > RF_LGN_LDS;
  const uint32_t rows = split ? 1u << (lgN - RF_LGN_ROW) : 1u;
  const uint64_t rowlen = (uint64_t)N / rows;
  const size_t ninv = (size_t)r * d * rows;
  const size_t spec_b = align256(np * N * 8), coef_b = align256((size_t)r * d * N * 8),
               task_b = align256(ninv * sizeof(RfInvTask)) + align256((size_t)r * d * sizeof(RfRowTask));
  rc = ensure_arena(spec_b + coef_b + task_b);
  if (rc) return rc;
  rc = ensure_stage(std::max(spec_b + task_b, coef_b));
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g.stream));
  double* hs = (double*)g.h_stage;
  const uint64_t p = g.primes.p[q];
  // centred representatives, as the forward stage stores them
  for (size_t i = 0; i < (size_t)r * d * N; i++) {
    uint64_t v = specA[i] % p;
    hs[i] = v > p / 2 ? (double)v - (double)p : (double)v;
  }
  for (size_t i = 0; i < (size_t)d * d * N; i++) {
    uint64_t v = specB[i] % p;
    hs[(size_t)r * d * N + i] = v > p / 2 ? (double)v - (double)p : (double)v;
  }
  double* d_spec = (double*)g.d_arena;
  double* d_coef = (double*)(g.d_arena + spec_b);
  RfInvTask* d_task = (RfInvTask*)(g.d_arena + spec_b + coef_b);
  RfInvTask* ht = (RfInvTask*)(g.h_stage + spec_b);
  RfRowTask* hc = (RfRowTask*)((char*)ht + align256(ninv * sizeof(RfInvTask)));
  RfRowTask* d_col = (RfRowTask*)((char*)d_task + align256(ninv * sizeof(RfInvTask)));
  for (int i = 0; i < r; i++)
    for (int j = 0; j < d; j++) {
      for (uint32_t h = 0; h < rows; h++) {
        RfInvTask tk{};
        tk.a_off = (uint64_t)(i * d) * N + h * rowlen;
        tk.a_stride = (uint32_t)N;
        tk.b_off = (uint64_t)(r * d + j) * N + h * rowlen;
        tk.b_stride = (uint32_t)(d * N);
        tk.out_off = (uint64_t)(i * d + j) * N + h * rowlen;
        tk.q = (uint16_t)q;
        tk.d = (uint16_t)d;
        ht[(size_t)(i * d + j) * rows + h] = tk;
      }
      hc[i * d + j] = RfRowTask{(uint64_t)(i * d + j) * N, 0, (uint16_t)q, (uint16_t)lgN};
    }
  HIPCHK(hipMemcpyAsync(d_spec, hs, np * N * 8, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(d_task, ht, task_b, hipMemcpyHostToDevice, g.stream));
  if (split) {
    HIPCHK(rf_launch_inv(tables(), RF_LGN_ROW, d_task, (int)ninv, d_spec, d_spec, d_coef, g.stream));
    HIPCHK(rf_launch_inv_col(tables(), lgN, d_col, r * d, d_coef, g.stream));
  } else {
    HIPCHK(rf_launch_inv(tables(), lgN, d_task, r * d, d_spec, d_spec, d_coef, g.stream));
  }
  HIPCHK(hipStreamSynchronize(g.stream));
  HIPCHK(hipMemcpy(g.h_stage, d_coef, (size_t)r * d * N * 8, hipMemcpyDeviceToHost));
  canon_out((const double*)g.h_stage, (size_t)r * d * N, p, out);
  return RF_OK;
}

int rf_stage_crt(const uint64_t* res, uint32_t nout, int k, int t, uint64_t rn, uint64_t* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = setup_locked();
  if (rc) return rc;
  if (k < 1 || k > 24 || t < 1 || t > RF_TMAX || nout < 1 || rn < 1)
    return fail(RF_ERR_ARG, "rf_stage_crt: bad arguments");
  const size_t coef_b = align256((size_t)k * nout * 8), out_b = align256(rn * 8);
  const uint64_t ndig = (rn + t - 1) / t;
  const int halo = rf_crt_halo(k, t);
  const uint64_t D = RF_CRT_THREADS - halo;
  const uint32_t nblk = (uint32_t)((ndig + D - 1) / D);
  const size_t task_b = align256(nblk * sizeof(RfCrtTask));
  const size_t pre_b = align256((size_t)nblk * RF_CRT_THREADS), fun_b = align256(nblk);
  rc = ensure_arena(coef_b + out_b + task_b + pre_b + fun_b);
  if (rc) return rc;
  rc = ensure_stage(std::max(coef_b + task_b, out_b));
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g.stream));
  double* hs = (double*)g.h_stage;
  for (int j = 0; j < k; j++) {
    const uint64_t p = g.primes.p[j];
    for (uint32_t i = 0; i < nout; i++) {
      uint64_t v = res[(size_t)j * nout + i] % p;
      hs[(size_t)j * nout + i] = v > p / 2 ? (double)v - (double)p : (double)v;
    }
  }
  char* A = g.d_arena;
  double* d_coef = (double*)A;
  uint64_t* d_out = (uint64_t*)(A + coef_b);
  char* d_task = A + coef_b + out_b;
  uint8_t* d_pre = (uint8_t*)(d_task + task_b);
  uint8_t* d_fun = d_pre + pre_b;
  char* ht = g.h_stage + coef_b;
  for (uint32_t b = 0; b < nblk; b++) {
    RfCrtTask tk{};
    tk.coef_off = 0;
    tk.out = d_out;
    tk.pre_off = (uint64_t)b * RF_CRT_THREADS;
    tk.fun_idx = b;
    tk.N = nout;
    tk.nout = nout;
    tk.rn = (uint32_t)rn;
    tk.digit0 = b * (uint32_t)D;
    tk.t = (uint8_t)t;
    tk.k = (uint8_t)k;
    tk.s = (uint8_t)((48 * k + 1 + 63) / 64);
    tk.halo = (uint8_t)halo;
    ((RfCrtTask*)ht)[b] = tk;
  }
  HIPCHK(hipMemcpyAsync(d_coef, hs, (size_t)k * nout * 8, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(d_task, ht, task_b, hipMemcpyHostToDevice, g.stream));
  const RfCrtTask* tasks = (const RfCrtTask*)d_task;
  HIPCHK(rf_launch_crt(tables(), k, tasks, (int)nblk, d_coef, d_pre, d_fun, g.stream));
  HIPCHK(rf_launch_apply(tasks, (int)nblk, d_pre, d_fun, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  HIPCHK(hipMemcpy(out, d_out, rn * 8, hipMemcpyDeviceToHost));
  return RF_OK;
}

}  // extern "C"

namespace rfi {
// Entry points for the library's other host sources (caller holds the lock).
size_t operand_task_bound(int nent) {
  const size_t rows = (size_t)1 << (RF_LGN_MAX - RF_LGN_ROW);
  return align256((size_t)nent * RF_KMAX * sizeof(RfFwdTask)) + align256((size_t)nent * RF_KMAX * rows * sizeof(RfRowTask));
}
int make_operand(std::vector<Entry>& A, int r, int d, uint64_t b_max_bits, size_t base, rf_operand** out,
                 const RfPlan* fixed) {
  return operand_make(A, r, d, b_max_bits, base, out, fixed);
}
int pipeline_run(std::vector<Product>& prods, const Layout& L, size_t base) {
  return run_pipeline(prods, L, base, g.stream);
}
void free_operand(rf_operand* op) {
  if (!op) return;
  if (g.ready) (void)hipStreamSynchronize(g.stream);
  if (op->slab)
    op->slab.reset();
  else if (op->spec)
    (void)hipFree(op->spec);
  delete op;
}

std::shared_ptr<void> device_slab(size_t bytes, void** ptr) {
  *ptr = nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) return nullptr;
  *ptr = p;
  return std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
}

size_t operands_batch_task_bound(size_t nent) {
  const size_t rows = (size_t)1 << (RF_LGN_MAX - RF_LGN_ROW);
  // per fwd_key group alignment: at most 2 * RF_TMAX * (RF_LGN_MAX + 1) groups
  return align256(nent * RF_KMAX * sizeof(RfFwdTask)) + align256(nent * RF_KMAX * rows * sizeof(RfRowTask)) +
         (size_t)512 * RF_TMAX * (RF_LGN_MAX + 2);
}

int make_operands_batch(const std::vector<Entry>& ents, const std::vector<uint64_t>& b_max_bits,
                        const std::vector<const RfPlan*>& fixed, size_t base, std::vector<rf_operand*>& out) {
  out.clear();
  const size_t cnt = ents.size();
  std::vector<rf_operand*> ops(cnt, nullptr);
  std::vector<size_t> off(cnt, 0);
  std::map<std::pair<uint64_t, uint64_t>, RfPlan> planned;
  size_t total = 0;
  int rc = RF_OK;
  for (size_t i = 0; i < cnt && !rc; i++) {
    rf_operand* op = ops[i] = new rf_operand();
    op->r = op->d = 1;
    op->la_bits = ents[i].sign ? ents[i].bits : 0;
    op->signs.push_back(ents[i].sign);
    if (!op->la_bits || !b_max_bits[i]) {
      op->la_bits = 0;  // a zero operand
      continue;
    }
    if (fixed[i]) {
      op->plan = *fixed[i];
    } else {
      // operands made together mostly share their sizes: plan each size once
      const auto key = std::make_pair(op->la_bits, b_max_bits[i]);
      auto it = planned.find(key);
      if (it != planned.end()) {
        op->plan = it->second;
      } else if (!tuned_plan(op->la_bits, b_max_bits[i], 1, 1, &op->plan) &&
                 rf_plan_choose(&g.primes, op->la_bits, b_max_bits[i], 1, 1, &op->plan) != 0) {
        rc = fail(RF_ERR_UNSUPPORTED, "no transform plan for an operand of %llu bits and right entries of %llu bits",
                  (unsigned long long)op->la_bits, (unsigned long long)b_max_bits[i]);
      } else {
        planned[key] = op->plan;
      }
    }
    if (!rc) rc = ensure_long_tables(op->plan.lgN);
    off[i] = total;
    total += align256((size_t)op->plan.k * ((size_t)8 << op->plan.lgN));
  }
  std::shared_ptr<void> slab;
  char* sp = nullptr;
  if (!rc && total) {
    slab = device_slab(total, (void**)&sp);
    if (!slab) rc = fail(RF_ERR_NOMEM, "operand spectra of %zu bytes", total);
  }
  // forward tasks grouped like the product pipeline's (one launch per group)
  std::map<int, std::vector<RfFwdTask>> fwd;
  std::vector<RfRowTask> rows;
  for (size_t i = 0; i < cnt && !rc; i++) {
    rf_operand* op = ops[i];
    if (!op->la_bits) continue;
    op->slab = slab;
    op->spec = (double*)(sp + off[i]);
    const RfPlan& pl = op->plan;
    const uint64_t N = (uint64_t)1 << pl.lgN;
    const uint64_t spec0 = (uint64_t)(off[i] / 8);  // doubles from the slab's start
    for (int q = 0; q < pl.k; q++) {
      RfFwdTask t{};
      t.limbs = ents[i].dptr;
      t.spec_off = spec0 + (uint64_t)q * N;
      t.nlimbs = (uint32_t)ents[i].nlimbs;
      t.nchunk = (uint32_t)std::min<uint64_t>((ents[i].nlimbs + pl.t - 1) / pl.t, N);
      t.sign = ents[i].sign;
      t.q = (uint16_t)q;
      t.t = (uint8_t)pl.t;
      fwd[fwd_key(pl)].push_back(t);
      if (plan_split(pl))
        for (uint32_t h = 0; h < (1u << (pl.lgN - RF_LGN_ROW)); h++)
          rows.push_back(RfRowTask{t.spec_off + ((uint64_t)h << RF_LGN_ROW), h, (uint16_t)q, (uint16_t)pl.lgN});
    }
  }
  if (!rc && !fwd.empty()) {
    std::vector<char> img;
    std::vector<std::pair<int, size_t>> groups;  // (fwd_key, image offset)
    for (auto& kv : fwd) groups.push_back({kv.first, append_tasks(img, kv.second)});
    const size_t roff = append_tasks(img, rows);
    if (base + img.size() > g.arena_cap || img.size() > operands_batch_task_bound(cnt)) {
      rc = fail(RF_ERR_ARG, "internal: operand task tables exceed their room");
    } else {
      rc = ensure_stage(img.size());
      if (!rc) {
        hipError_t e = hipStreamSynchronize(g.stream);
        if (e == hipSuccess) {
          memcpy(g.h_stage, img.data(), img.size());
          e = hipMemcpyAsync(g.d_arena + base, g.h_stage, img.size(), hipMemcpyHostToDevice, g.stream);
        }
        const RfDevTables tb = tables();
        for (size_t gi = 0; gi < groups.size() && e == hipSuccess; gi++) {
          const int key = groups[gi].first;
          const RfFwdTask* tt = (const RfFwdTask*)(g.d_arena + base + groups[gi].second);
          const int nt = (int)fwd[key].size();
          const int lgN = (key >> 4) & 255, t = key & 15;
          e = (key >> 12) ? rf_launch_fwd_col(tb, lgN, t, tt, nt, (double*)sp, g.stream)
                          : rf_launch_fwd(tb, lgN, t, tt, nt, (double*)sp, g.stream);
        }
        if (e == hipSuccess && !rows.empty())
          e = rf_launch_fwd_row(tables(), (const RfRowTask*)(g.d_arena + base + roff), (int)rows.size(), (double*)sp,
                                g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) rc = fail(RF_ERR_HIP, "operand transforms: %s", hipGetErrorString(e));
      }
    }
  }
  if (rc) {
    for (rf_operand* op : ops) {
      if (!op) continue;
      op->slab.reset();
      delete op;
    }
    return rc;
  }
  out = ops;
  return RF_OK;
}
}  // namespace rfi
