// rf_mod.cpp -- remainders X mod D on the device (the reference's remainder
// step, mpzfft_mod_init / mpzfft_mod_mod, src/lib/mpzfft.c:59-100 and
// src/include/mpnfft_mod.h:13-75).
//
// With n = limbs of D, m = most limbs of |X| and B = 2^64, creation computes
// I = floor(B^(m+1) / D) once and keeps I and D as resident transforms.  For
// X = X_hi B^(n-1) + X_lo:
//   Q = floor(X_hi I / B^(m-n+2))  satisfies  X/D - 3 < Q <= X/D,
// since X_hi I / B^(m-n+2) = (X - X_lo)/D - X_hi e / B^(m-n+2) with
// e = B^(m+1)/D - I in [0, 1), X_lo / D < 1 and X_hi < B^(m-n+1).  So
// R = X - Q D lies in [0, 3D) and fits n + 1 limbs, two corrections at most
// bring it into [0, D), and a negative X takes D - R.  Per remainder: one
// product with the resident I (its high limbs are Q, read in place), one with
// the resident D, and the fix-up kernel (rf_util.hip) -- a batch of
// remainders runs each phase as one pipeline.
#include <gmp.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <system_error>
#include <thread>
#include <vector>

#include "rf_internal.h"
#include "rf_util.h"

struct rf_mod {
  uint64_t m = 0, n = 0;  // most limbs of |X|, limbs of D
  uint64_t wrap = 0;      // W when Q D is formed modulo B^W - 1 (div has a cyclic plan), else 0
  rf_operand* inv = nullptr;
  rf_operand* div = nullptr;
  uint64_t* d_dev = nullptr;  // D, n limbs
};

namespace {
using namespace rfi;

// A batch of remainders X_i mod D_i, planned from the sizes and signs of the
// X_i alone, so that the arena can be sized before any input is placed in it.
struct Item {
  int p1 = -1, p2 = -1;  // indices into the two product lists
  uint64_t t_off = 0, qd_off = 0, rt_off = 0;
};
struct ModBatch {
  std::vector<Item> it;
  std::vector<Product> p1, p2;
  Layout L1, L2;
  size_t t_b = 0, qd_b = 0, rt_b = 0, task_b = 0;
  size_t bytes() const { return t_b + qd_b + rt_b + task_b + std::max(L1.total(), L2.total()); }
};

int mod_batch_plan(const rf_mod* const* mods, const rf_dev_entry* xs, int cnt, ModBatch& mb) {
  mb.it.assign(cnt, Item());
  uint64_t t_limbs = 0, qd_limbs = 0, rt_limbs = 0;
  for (int i = 0; i < cnt; i++) {
    const rf_mod* md = mods[i];
    if (!md) return fail(RF_ERR_ARG, "remainder %d: null modulus", i);
    const uint64_t xn = xs[i].nlimbs;
    if (xn > md->m) return fail(RF_ERR_ARG, "remainder %d: %llu limbs, the modulus allows %llu", i,
                                (unsigned long long)xn, (unsigned long long)md->m);
    Item& I = mb.it[i];
    I.rt_off = rt_limbs;
    rt_limbs += md->wrap ? 3 * md->wrap : 2 * (md->n + 1);
    if (xs[i].sign == 0 || xn < md->n) continue;  // X_hi = 0: Q = 0
    Product P;
    P.r = P.d = 1;
    P.left = md->inv;
    Entry E{};
    E.nlimbs = xn - (md->n - 1);
    E.bits = 64 * E.nlimbs;
    E.sign = 1;
    P.B.push_back(E);
    int rc = plan_product(P);
    if (rc) return rc;
    const uint64_t s = md->m - md->n + 2;
    if (P.plan.k == 0 || P.plan.rn <= s) continue;
    I.p1 = (int)mb.p1.size();
    I.t_off = t_limbs;
    t_limbs += P.plan.rn;
    Product Q;
    Q.r = Q.d = 1;
    Q.left = md->div;
    Entry F{};
    F.nlimbs = std::min<uint64_t>(P.plan.rn - s, md->m - md->n + 1);  // Q < B^(m-n+1)
    F.bits = 64 * F.nlimbs;
    F.sign = 1;
    Q.B.push_back(F);
    rc = plan_product(Q);
    if (rc) return rc;
    // only (X - QD) mod B^(n+1) is formed: the recomposition stops after the
    // low n + 1 limbs of QD (carries only move upward, so they are exact),
    // which halves that product's CRT blocks.  A cyclic Q D keeps all its
    // limbs: they are folded modulo B^W - 1.
    if (!md->wrap) Q.plan.rn = std::min<uint64_t>(Q.plan.rn, md->n + 1);
    mb.p1.push_back(std::move(P));
    if (Q.plan.k == 0) continue;
    I.p2 = (int)mb.p2.size();
    I.qd_off = qd_limbs;
    qd_limbs += Q.plan.rn;
    mb.p2.push_back(std::move(Q));
  }
  mb.L1 = layout_of(mb.p1);
  mb.L2 = layout_of(mb.p2);
  mb.t_b = align256(t_limbs * 8 + 8);
  mb.qd_b = align256(qd_limbs * 8 + 8);
  mb.rt_b = align256(rt_limbs * 8 + 8);
  mb.task_b = align256((size_t)cnt * sizeof(RfModTask));
  return RF_OK;
}

// Runs a planned batch with its scratch at arena byte `base` (the arena
// already holds base + mb.bytes()); outputs outs[i], n_i limbs each.
// Returns with the stream idle.
int mod_batch_exec(ModBatch& mb, const rf_mod* const* mods, const rf_dev_entry* xs, int cnt, uint64_t* const* outs,
                   size_t base) {
  char* d_base = nullptr;
  int rc = engine_arena(base + mb.bytes(), &d_base);
  if (rc) return rc;
  uint64_t* T = (uint64_t*)(d_base + base);
  uint64_t* QD = (uint64_t*)(d_base + base + mb.t_b);
  uint64_t* RT = (uint64_t*)(d_base + base + mb.t_b + mb.qd_b);
  RfModTask* d_tasks = (RfModTask*)(d_base + base + mb.t_b + mb.qd_b + mb.rt_b);
  const size_t pipe = base + mb.t_b + mb.qd_b + mb.rt_b + mb.task_b;
  for (int i = 0; i < cnt; i++) {
    const rf_mod* md = mods[i];
    const Item& I = mb.it[i];
    if (I.p1 < 0) continue;
    Product& P = mb.p1[I.p1];
    P.B[0].dptr = xs[i].limbs + (md->n - 1);
    P.out.assign(1, T + I.t_off);
    if (I.p2 >= 0) {
      Product& Q = mb.p2[I.p2];
      Q.B[0].dptr = T + I.t_off + (md->m - md->n + 2);
      Q.out.assign(1, QD + I.qd_off);
    }
  }
  if (!mb.p1.empty()) rc = pipeline_run(mb.p1, mb.L1, pipe);
  if (!rc && !mb.p2.empty()) rc = pipeline_run(mb.p2, mb.L2, pipe);  // waits for the first before reusing staging
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  // the pipelines may have grown (replaced) the staging buffer: fetch it now
  char* h_base = nullptr;
  rc = engine_stage(mb.task_b, &h_base);
  if (rc) return rc;
  RfModTask* ht = (RfModTask*)h_base;
  for (int i = 0; i < cnt; i++) {
    const rf_mod* md = mods[i];
    const Item& I = mb.it[i];
    RfModTask k{};
    k.x = xs[i].limbs;
    k.xn = xs[i].nlimbs;
    k.sign = xs[i].nlimbs ? (xs[i].sign > 0 ? 1 : (xs[i].sign < 0 ? -1 : 0)) : 0;
    k.qd = I.p2 >= 0 ? QD + I.qd_off : nullptr;
    k.qdn = I.p2 >= 0 ? mb.p2[I.p2].plan.rn : 0;
    k.d = md->d_dev;
    k.n = md->n;
    k.r = RT + I.rt_off;
    k.t = k.r + (md->wrap ? md->wrap : md->n + 1);
    k.xf = md->wrap ? k.t + md->wrap : nullptr;
    k.wrap = I.p2 >= 0 ? md->wrap : 0;  // Q = 0: X - 0 needs no folding
    k.out = outs[i];
    ht[i] = k;
  }
  HIPCHK(hipMemcpyAsync(d_tasks, ht, (size_t)cnt * sizeof(RfModTask), hipMemcpyHostToDevice, st));
  HIPCHK(rf_launch_mod_fix(d_tasks, cnt, st));
  HIPCHK(hipStreamSynchronize(st));
  return RF_OK;
}

}  // namespace

namespace {
using namespace rfi;

// The host part of preparing a modulus: n, m and I = floor(B^(m+1) / D)
// (GMP division; no engine state, so callers may run it in parallel).
void mod_reciprocal(uint64_t max_bits, const mpz_t d, uint64_t* n, uint64_t* m, mpz_t I) {
  *n = mpz_size(d);
  *m = std::max<uint64_t>((max_bits + 63) / 64, *n);
  mpz_t Bm;
  mpz_init(Bm);
  mpz_setbit(Bm, 64 * (*m + 1));
  mpz_tdiv_q(I, Bm, d);  // at most m - n + 2 limbs
  mpz_clear(Bm);
}

// The device part (engine mutex held): D and I uploaded and made resident
// transforms.
int mod_create_locked(const mpz_t d, uint64_t n, uint64_t m, const mpz_t I, rf_mod** out) {
  rf_mod* md = new rf_mod();
  md->n = n;
  md->m = m;
  const uint64_t in = mpz_size(I);
  int rc = RF_OK;
  const size_t d_b = align256(md->n * 8), i_b = align256(in * 8);
  char* d_base = nullptr;
  char* h_base = nullptr;
  rc = engine_arena(i_b + operand_task_bound(1), &d_base);  // the transforms' tables follow I
  if (!rc) rc = engine_stage(d_b + i_b, &h_base);
  if (!rc) {
    hipError_t e = hipMalloc((void**)&md->d_dev, md->n * 8);
    if (e != hipSuccess) rc = fail(RF_ERR_NOMEM, "modulus of %llu limbs: %s", (unsigned long long)md->n,
                                   hipGetErrorString(e));
  }
  hipStream_t st = engine_stream();
  if (!rc) {
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
      memcpy(h_base, d->_mp_d, md->n * 8);
      memcpy(h_base + d_b, I->_mp_d, in * 8);
      e = hipMemcpyAsync(md->d_dev, h_base, md->n * 8, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(d_base, h_base + d_b, in * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = fail(RF_ERR_HIP, "uploading the modulus: %s", hipGetErrorString(e));
  }
  const uint64_t q_bits = 64 * (md->m - md->n + 1);  // X_hi and Q limbs
  if (!rc) {
    std::vector<Entry> E(1);
    E[0].dptr = (const uint64_t*)d_base;
    E[0].nlimbs = in;
    E[0].sign = 1;
    E[0].bits = 64 * in;
    rc = make_operand(E, 1, 1, q_bits, i_b, &md->inv);
  }
  if (!rc) {
    std::vector<Entry> E(1);
    E[0].dptr = md->d_dev;
    E[0].nlimbs = md->n;
    E[0].sign = 1;
    E[0].bits = mpz_sizeinbase(d, 2);
    // Q D is needed only modulo B^W - 1 for W >= n + 1 (R < 3D < B^W - 1) and
    // W >= m - n + 1 (Q fits): a cyclic product of about half the length,
    // when the cost model prefers it
    const uint64_t qn = md->m - md->n + 1;
    RfPlan cyc{}, acyc{};
    const bool use_cyc =
        rf_plan_cyclic(&engine_primes(), std::max<uint64_t>(md->n + 1, qn), md->n, qn, &cyc) == 0 &&
        (rf_plan_choose(&engine_primes(), E[0].bits, q_bits, 1, 1, &acyc) != 0 || cyc.cost < acyc.cost);
    if (use_cyc) md->wrap = (uint64_t)cyc.t << cyc.lgN;
    rc = make_operand(E, 1, 1, q_bits, i_b, &md->div, use_cyc ? &cyc : nullptr);
  }
  if (rc) {
    free_operand(md->inv);
    free_operand(md->div);
    if (md->d_dev) (void)hipFree(md->d_dev);
    delete md;
    return rc;
  }
  *out = md;
  return RF_OK;
}

}  // namespace

extern "C" {

int rf_mod_create(uint64_t max_bits, const mpz_t d, rf_mod** out) {
  if (!out) return fail(RF_ERR_ARG, "rf_mod_create: null output");
  *out = nullptr;
  if (!d || mpz_sgn(d) <= 0) return fail(RF_ERR_ARG, "rf_mod_create: the modulus must be positive");
  if (max_bits == 0) return fail(RF_ERR_ARG, "rf_mod_create: max_bits must be positive");
  {
    std::lock_guard<std::mutex> lk(engine_mutex());
    if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  }
  uint64_t n = 0, m = 0;
  mpz_t I;
  mpz_init(I);
  mod_reciprocal(max_bits, d, &n, &m, I);  // outside the engine mutex
  int rc;
  {
    std::lock_guard<std::mutex> lk(engine_mutex());
    rc = engine_ready() ? mod_create_locked(d, n, m, I, out)
                        : fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  }
  mpz_clear(I);
  return rc;
}

int rf_mod_create_batch(const uint64_t* max_bits, const mpz_t* ds, int cnt, rf_mod** outs) {
  if (cnt < 0 || (cnt > 0 && (!max_bits || !ds || !outs))) return fail(RF_ERR_ARG, "rf_mod_create_batch: bad arguments");
  for (int i = 0; i < cnt; i++) outs[i] = nullptr;
  for (int i = 0; i < cnt; i++) {
    if (mpz_sgn(ds[i]) <= 0) return fail(RF_ERR_ARG, "rf_mod_create_batch: modulus %d is not positive", i);
    if (max_bits[i] == 0) return fail(RF_ERR_ARG, "rf_mod_create_batch: max_bits[%d] must be positive", i);
  }
  {
    std::lock_guard<std::mutex> lk(engine_mutex());
    if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  }
  // the GMP divisions dominate a modulus's preparation: run them on up to
  // 16 host threads, then make the resident transforms in order
  std::vector<uint64_t> n(cnt), m(cnt);
  std::vector<__mpz_struct> I(cnt);
  for (int i = 0; i < cnt; i++) mpz_init(&I[i]);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int i = next++; i < cnt; i = next++) mod_reciprocal(max_bits[i], ds[i], &n[i], &m[i], &I[i]);
  };
  const int nth = std::max(1, std::min({cnt, 16, (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nth; t++) pool.emplace_back(work);
  } catch (const std::system_error&) {
    // fewer helper threads: the calling thread takes what is left
  }
  work();
  for (auto& t : pool) t.join();
  int rc = RF_OK;
  {
    std::lock_guard<std::mutex> lk(engine_mutex());
    // hw_mpz_clear may have run while the divisions were unlocked
    if (!engine_ready()) rc = fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
    for (int i = 0; i < cnt && !rc; i++) rc = mod_create_locked(ds[i], n[i], m[i], &I[i], &outs[i]);
  }
  for (int i = 0; i < cnt; i++) mpz_clear(&I[i]);
  if (rc)
    for (int i = 0; i < cnt; i++)
      if (outs[i]) {
        rf_mod_destroy(outs[i]);
        outs[i] = nullptr;
      }
  return rc;
}

uint64_t rf_mod_limbs(const rf_mod* mod) { return mod ? mod->n : 0; }
uint64_t rf_mod_wrap_limbs(const rf_mod* mod) { return mod ? mod->wrap : 0; }

int rf_mod_mod_device(const rf_mod* const* mods, const rf_dev_entry* xs, int cnt, uint64_t* const* outs) {
  std::lock_guard<std::mutex> lk(engine_mutex());
  if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (cnt < 0 || (cnt > 0 && (!mods || !xs || !outs))) return fail(RF_ERR_ARG, "rf_mod_mod_device: bad arguments");
  if (cnt == 0) return RF_OK;
  for (int i = 0; i < cnt; i++)
    if (!outs[i]) return fail(RF_ERR_ARG, "remainder %d: null output", i);
  ModBatch mb;
  int rc = mod_batch_plan(mods, xs, cnt, mb);
  if (rc) return rc;
  return mod_batch_exec(mb, mods, xs, cnt, outs, 0);
}

int rf_mod_mod(const rf_mod* mod, mpz_t rop, const mpz_t op) {
  std::lock_guard<std::mutex> lk(engine_mutex());
  if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  if (!mod || !rop || !op) return fail(RF_ERR_ARG, "rf_mod_mod: null argument");
  const uint64_t xn = mpz_size(op), n = mod->n;
  if (xn > mod->m) return fail(RF_ERR_ARG, "rf_mod_mod: %llu limbs, the modulus allows %llu",
                               (unsigned long long)xn, (unsigned long long)mod->m);
  const size_t x_b = align256(xn * 8 + 8), o_b = align256(n * 8);
  rf_dev_entry x{};
  x.nlimbs = xn;
  x.sign = mpz_sgn(op);
  const rf_mod* mods[1] = {mod};
  ModBatch mb;
  int rc = mod_batch_plan(mods, &x, 1, mb);
  if (rc) return rc;
  char* d_base = nullptr;
  char* h_base = nullptr;
  rc = engine_arena(x_b + o_b + mb.bytes(), &d_base);  // sized before the input goes in
  if (!rc) rc = engine_stage(std::max(std::max(x_b, o_b), mb.task_b), &h_base);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  if (xn) {
    memcpy(h_base, op->_mp_d, xn * 8);
    HIPCHK(hipMemcpyAsync(d_base, h_base, xn * 8, hipMemcpyHostToDevice, st));
  }
  x.limbs = (const uint64_t*)d_base;
  uint64_t* outs[1] = {(uint64_t*)(d_base + x_b)};
  rc = mod_batch_exec(mb, mods, &x, 1, outs, x_b + o_b);
  if (rc) return rc;
  rc = engine_stage(o_b, &h_base);  // the staging buffer may have been replaced
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h_base, d_base + x_b, n * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  uint64_t nn = n;
  const uint64_t* r = (const uint64_t*)h_base;
  while (nn && r[nn - 1] == 0) nn--;  // src/lib/mpzfft.c:95-99
  if (nn == 0) {
    mpz_set_ui(rop, 0);
    return RF_OK;
  }
  _mpz_realloc(rop, (mp_size_t)nn);
  memcpy(rop->_mp_d, r, nn * 8);
  rop->_mp_size = (int)nn;
  return RF_OK;
}

void rf_mod_destroy(rf_mod* mod) {
  if (!mod) return;
  std::lock_guard<std::mutex> lk(engine_mutex());
  free_operand(mod->inv);
  free_operand(mod->div);
  if (mod->d_dev) (void)hipFree(mod->d_dev);
  delete mod;
}

}  // extern "C"
