// rf_mod.cpp -- remainders X mod D on the device (the reference's remainder
// step, mpzfft_mod_init / mpzfft_mod_mod, src/lib/mpzfft.c:59-100 and
// src/include/mpnfft_mod.h:13-75).
//
// With n = limbs of D, m = most limbs of |X| and B = 2^64, creation computes
// I with floor(B^(m+1) / D) - 2 <= I <= floor(B^(m+1) / D) once and keeps I
// and D as resident transforms.  For X = X_hi B^(n-1) + X_lo:
//   Q = floor(X_hi I / B^(m-n+2))  satisfies  X/D - 3 < Q <= X/D,
// since X_hi I / B^(m-n+2) = (X - X_lo)/D - X_hi e / B^(m-n+2) with
// e = B^(m+1)/D - I in [0, 3), X_lo / D < 1 and X_hi < B^(m-n+1), so the
// middle term is below 3 / B.  So
// R = X - Q D lies in [0, 3D) and fits n + 1 limbs, two corrections at most
// bring it into [0, D), and a negative X takes D - R.  Per remainder: one
// product with the resident I (its high limbs are Q, read in place), one with
// the resident D, and the fix-up kernel (rf_util.hip) -- a batch of
// remainders runs each phase as one pipeline.
#include <gmp.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <tuple>
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
  std::shared_ptr<void> d_slab;  // the allocation d_dev lies in (shared by a batch's moduli)
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

// ---- preparing moduli: the reciprocal I on the device ----
//
// For D of n limbs and dividends of at most m limbs, I approximates
// B^(m+1) / D from below: with D' = D 2^s (top bit set) padded with zero
// limbs to n' = max(n, m + 2 - n) limbs, D'' = D' B^(n'-n), the
// approximate-reciprocal recursion (Brent & Zimmermann, Modern Computer
// Arithmetic, Algorithm 3.5; Lemma 3.6: A X < B^(2n') <= A (X + 2)) gives X
// on n' + 1 limbs, and I = floor(X 2^s / B^o) with o = n' + n - m - 1 >= 1
// satisfies floor(B^(m+1)/D) - 2 <= I <= floor(B^(m+1)/D): X 2^s / B^o is
// within 2 2^s / B^o <= 1 below B^(m+1) / D.  That keeps the quotient
// estimate's bound X/D - 3 < Q <= X/D (the error term X_hi (e + 2) /
// B^(m-n+2) stays below 1) -- the reference's remainder step also works
// from an approximate inverse (zz_invertappr, src/lib/mpnfft_mod.c:13-90).
// The recursion runs bottom-up, all moduli of a batch level by level: the
// base (at most kRecipBase limbs) by GMP on the host, every level as two
// batched device products (A X_h and Tm X_h) and two multi-limb kernels.
constexpr uint64_t kRecipBase = 64;

struct ModPrep {
  uint64_t n = 0, m = 0, np = 0, il = 0;
  uint32_t s = 0, o = 0;
  std::vector<uint64_t> c;           // size chain: c[0] = n', c[j+1] = c[j] - floor((c[j] - 1) / 2)
  std::vector<uint64_t> dtop;        // the top c.back() limbs of D'' (D'' itself is formed on the device)
  std::vector<uint64_t> xbase;       // X of the base level (c.back() + 1 limbs)
  // arena offsets in limbs
  uint64_t dn_off = 0, t_off = 0, tm_off = 0, u_off = 0, i_off = 0, tn = 0, un = 0;
  std::vector<uint64_t> x_off;       // X_j, j = 0 .. c.size() - 1 (c[j] + 1 limbs)
};

// Host part (no engine state; callers may run it in parallel).
void mod_prep(uint64_t max_bits, const mpz_t d, ModPrep& P) {
  P.n = mpz_size(d);
  P.m = std::max<uint64_t>((max_bits + 63) / 64, P.n);
  const uint64_t top = d->_mp_d[P.n - 1];
  P.s = (uint32_t)__builtin_clzll(top);
  P.np = std::max<uint64_t>(P.n, P.m + 2 - P.n);
  P.o = (uint32_t)(P.np + P.n - P.m - 1);
  P.il = P.m - P.n + 2;
  P.c.assign(1, P.np);
  while (P.c.back() > kRecipBase) P.c.push_back(P.c.back() - (P.c.back() - 1) / 2);
  // base: X = floor((B^(2 c) - 1) / A) for the top c limbs A of
  // D'' = (D << s) B^(n'-n)
  const uint64_t cb = P.c.back();
  P.dtop.assign(cb, 0);
  const uint64_t* dl = (const uint64_t*)d->_mp_d;
  for (uint64_t q = 0; q < cb; q++) {
    const uint64_t j = P.np - cb + q;  // limb of D''
    if (j < P.np - P.n) continue;
    const uint64_t i = j - (P.np - P.n);
    P.dtop[q] = P.s ? (dl[i] << P.s) | (i ? dl[i - 1] >> (64 - P.s) : 0) : dl[i];
  }
  mpz_t A, num, X;
  mpz_init(A);
  mpz_init(num);
  mpz_init(X);
  mpz_import(A, cb, -1, 8, 0, 0, P.dtop.data());
  mpz_setbit(num, 128 * cb);
  mpz_sub_ui(num, num, 1);
  mpz_tdiv_q(X, num, A);
  P.xbase.assign(cb + 1, 0);
  size_t cnt = 0;
  mpz_export(P.xbase.data(), &cnt, -1, 8, 0, 0, X);
  mpz_clear(A);
  mpz_clear(num);
  mpz_clear(X);
}

// Device part (engine mutex held): the reciprocals of all moduli, then D and
// I as resident transforms.
// recips (tests): the reciprocals I are copied out instead, and no modulus is made.
int mod_create_batch_locked(const mpz_t* ds, std::vector<ModPrep>& prep, rf_mod** outs,
                            std::vector<std::vector<uint64_t>>* recips) {
  const int cnt = (int)prep.size();
  size_t levels = 0;
  for (auto& P : prep) levels = std::max(levels, P.c.size() - 1);
  // plan every level's two products first: their sizes fix the scratch
  std::vector<std::vector<Product>> p1(levels), p2(levels);
  std::vector<std::vector<int>> who(levels);  // modulus of each product, per level
  for (int i = 0; i < cnt; i++) {
    ModPrep& P = prep[i];
    for (size_t j = 0; j + 1 < P.c.size(); j++) {
      const uint64_t h = P.c[j + 1];
      Product a;
      a.r = a.d = 1;
      a.A.push_back(Entry{0, nullptr, P.c[j], 64 * P.c[j], 1});
      a.B.push_back(Entry{0, nullptr, h + 1, 64 * (h + 1), 1});
      int rc = plan_product(a);
      if (rc) return rc;
      Product b;
      b.r = b.d = 1;
      b.A.push_back(Entry{0, nullptr, h + 1, 64 * (h + 1), 1});
      b.B.push_back(Entry{0, nullptr, h + 1, 64 * (h + 1), 1});
      rc = plan_product(b);
      if (rc) return rc;
      P.tn = std::max<uint64_t>(P.tn, a.plan.rn);
      P.un = std::max<uint64_t>(P.un, b.plan.rn);
      p1[j].push_back(std::move(a));
      p2[j].push_back(std::move(b));
      who[j].push_back(i);
    }
  }
  // arena layout (limbs): the base X's (uploaded), D'' (formed on the device
  // from D), X levels, T, Tm, U, I per modulus; then the reciprocal and
  // normalisation task tables, every level's pipeline task images, and the
  // pipelines' shared scratch
  uint64_t lim = 0;
  auto take = [&](uint64_t n) {
    const uint64_t o = lim;
    lim += (n + 31) & ~(uint64_t)31;  // 256-byte steps
    return o;
  };
  for (auto& P : prep) {
    P.x_off.assign(P.c.size(), 0);
    P.x_off.back() = take(P.c.back() + 1);
  }
  const uint64_t up_limbs = lim;  // uploaded region
  for (auto& P : prep) {
    P.dn_off = take(P.np);
    for (size_t j = 0; j + 1 < P.c.size(); j++) P.x_off[j] = take(P.c[j] + 1);
    P.t_off = take(std::max<uint64_t>(P.tn, 1));
    P.tm_off = take(P.c.size() > 1 ? P.c[1] + 1 : 1);
    P.u_off = take(std::max<uint64_t>(P.un, 1));
    P.i_off = take(P.il);
  }
  size_t nrt = 0;
  for (size_t j = 0; j < levels; j++) nrt += who[j].size();
  const size_t rt_b = align256(nrt * sizeof(RfRecipTask)) + align256(cnt * sizeof(RfRecipFinal)) +
                      align256(cnt * sizeof(RfNormTask));
  // each level's two pipelines get task images of their own, so that all
  // levels are enqueued at once; their scratch (spectra, coefficients,
  // carry codes) is shared, one level after the other on the stream
  std::vector<size_t> img1(levels), img2(levels);
  size_t img_b = 0;
  size_t pipe_b = operands_batch_task_bound(2 * (size_t)cnt);
  for (size_t j = 0; j < levels; j++) {
    const Layout l1 = layout_of(p1[j]), l2 = layout_of(p2[j]);
    img1[j] = img_b;
    img_b += align256(l1.tasks_b);
    img2[j] = img_b;
    img_b += align256(l2.tasks_b);
    pipe_b = std::max(pipe_b, std::max(l1.total(), l2.total()));
  }
  const size_t limbs_b = align256(lim * 8);
  char* A = nullptr;
  int rc = engine_arena(limbs_b + rt_b + img_b + pipe_b, &A);
  if (rc) return rc;
  uint64_t* L = (uint64_t*)A;
  RfRecipTask* d_rt = (RfRecipTask*)(A + limbs_b);
  RfRecipFinal* d_fin = (RfRecipFinal*)(A + limbs_b + align256(nrt * sizeof(RfRecipTask)));
  RfNormTask* d_norm = (RfNormTask*)(A + limbs_b + align256(nrt * sizeof(RfRecipTask)) +
                                     align256(cnt * sizeof(RfRecipFinal)));
  char* d_img = A + limbs_b + rt_b;
  const size_t pipe = limbs_b + rt_b + img_b;
  // D limbs (kept for the fix-up) in one allocation shared by the moduli
  uint64_t d_total = 0;
  std::vector<uint64_t> d_off(cnt);
  for (int i = 0; i < cnt; i++) d_off[i] = d_total, d_total += (prep[i].n + 31) & ~(uint64_t)31;
  uint64_t* dslab_p = nullptr;
  std::shared_ptr<void> dslab = device_slab(d_total * 8, (void**)&dslab_p);
  if (!dslab) return fail(RF_ERR_NOMEM, "moduli of %llu limbs", (unsigned long long)d_total);
  // the levels' products, their pipelines and reciprocal tasks
  std::vector<RfRecipTask> rts;
  std::vector<size_t> rt_first(levels + 1, 0);
  std::vector<Pipeline> pp1(levels), pp2(levels);
  for (size_t j = levels; j-- > 0;) {  // bottom-up order, the order of the launches
    rt_first[j] = rts.size();
    for (size_t k = 0; k < who[j].size(); k++) {
      const ModPrep& P = prep[who[j][k]];
      const uint64_t h = P.c[j + 1];
      Product& a = p1[j][k];
      a.A[0].dptr = L + P.dn_off + (P.np - P.c[j]);
      a.B[0].dptr = L + P.x_off[j + 1];
      a.out.assign(1, L + P.t_off);
      Product& b = p2[j][k];
      b.A[0].dptr = L + P.tm_off;
      b.B[0].dptr = L + P.x_off[j + 1];
      b.out.assign(1, L + P.u_off);
      RfRecipTask t{};
      t.a = L + P.dn_off + (P.np - P.c[j]);
      t.xh = L + P.x_off[j + 1];
      t.t = L + P.t_off;
      t.tm = L + P.tm_off;
      t.u = L + P.u_off;
      t.x = L + P.x_off[j];
      t.n = P.c[j];
      t.h = h;
      t.l = P.c[j] - h;
      t.tn = a.plan.rn;
      t.un = b.plan.rn;
      rts.push_back(t);
    }
    rc = build_pipeline(p1[j], pp1[j]);
    if (!rc) rc = build_pipeline(p2[j], pp2[j]);
    if (rc) return rc;
  }
  std::vector<RfRecipFinal> fins(cnt);
  std::vector<RfNormTask> norms(cnt);
  uint64_t max_il = 0, max_np = 0;
  for (int i = 0; i < cnt; i++) {
    const ModPrep& P = prep[i];
    fins[i] = RfRecipFinal{L + P.x_off[0], P.c[0] + 1, L + P.i_off, P.il, P.s, P.o};
    norms[i] = RfNormTask{dslab_p + d_off[i], P.n, L + P.dn_off, P.np, P.s, 0};
    max_il = std::max(max_il, P.il);
    max_np = std::max(max_np, P.np);
  }
  // host image: [uploaded limbs | recip, final and normalisation tasks |
  // level task images | D limbs], one staging copy and four transfers
  const size_t img_off = limbs_b + rt_b, dl_off = img_off + img_b;
  char* H = nullptr;
  rc = engine_stage(dl_off + align256(d_total * 8), &H);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  for (auto& P : prep) memcpy(H + P.x_off.back() * 8, P.xbase.data(), P.xbase.size() * 8);
  if (!rts.empty()) memcpy(H + limbs_b, rts.data(), rts.size() * sizeof(RfRecipTask));
  memcpy(H + limbs_b + align256(nrt * sizeof(RfRecipTask)), fins.data(), fins.size() * sizeof(RfRecipFinal));
  memcpy(H + limbs_b + align256(nrt * sizeof(RfRecipTask)) + align256(cnt * sizeof(RfRecipFinal)), norms.data(),
         norms.size() * sizeof(RfNormTask));
  for (size_t j = 0; j < levels; j++) {
    if (!pp1[j].image.empty()) memcpy(H + img_off + img1[j], pp1[j].image.data(), pp1[j].image.size());
    if (!pp2[j].image.empty()) memcpy(H + img_off + img2[j], pp2[j].image.data(), pp2[j].image.size());
  }
  for (int i = 0; i < cnt; i++) memcpy(H + dl_off + d_off[i] * 8, ds[i]->_mp_d, prep[i].n * 8);
  HIPCHK(hipMemcpyAsync(A, H, up_limbs * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(A + limbs_b, H + limbs_b, rt_b + img_b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dslab_p, H + dl_off, d_total * 8, hipMemcpyHostToDevice, st));
  HIPCHK(rf_launch_mod_norm(d_norm, cnt, max_np, st));
  // the levels, bottom-up, enqueued without waiting
  for (size_t j = levels; j-- > 0;) {
    rc = launch_pipeline_tasks(pp1[j], A + pipe, d_img + img1[j], st, false);
    if (rc) return rc;
    HIPCHK(rf_launch_recip_adjust(d_rt + rt_first[j], (int)who[j].size(), st));
    rc = launch_pipeline_tasks(pp2[j], A + pipe, d_img + img2[j], st, false);
    if (rc) return rc;
    HIPCHK(rf_launch_recip_combine(d_rt + rt_first[j], (int)who[j].size(), st));
  }
  HIPCHK(rf_launch_recip_final(d_fin, cnt, max_il, st));
  if (recips) {
    HIPCHK(hipStreamSynchronize(st));
    recips->assign(cnt, {});
    for (int i = 0; i < cnt; i++) {
      (*recips)[i].resize(prep[i].il);
      HIPCHK(hipMemcpy((*recips)[i].data(), L + prep[i].i_off, prep[i].il * 8, hipMemcpyDeviceToHost));
    }
    return RF_OK;
  }
  // D and I as resident transforms, all in one set of launches each way
  std::vector<Entry> ents;
  std::vector<uint64_t> bmax;
  std::vector<RfPlan> cyc(cnt);
  std::vector<const RfPlan*> fixed;
  std::vector<uint64_t> wrap(cnt, 0);
  // the cyclic-or-whole decision depends on (n, m, bits of D) only: moduli
  // of a level mostly share it
  std::map<std::tuple<uint64_t, uint64_t, uint64_t>, std::pair<bool, RfPlan>> decided;
  for (int i = 0; i < cnt; i++) {
    const ModPrep& P = prep[i];
    const uint64_t q_bits = 64 * (P.m - P.n + 1);  // X_hi and Q limbs
    ents.push_back(Entry{0, L + P.i_off, P.il, 64 * P.il, 1});
    bmax.push_back(q_bits);
    fixed.push_back(nullptr);
    // Q D is needed only modulo B^W - 1 for W >= n + 1 (R < 3D < B^W - 1) and
    // W >= m - n + 1 (Q fits): a cyclic product of about half the length,
    // when the cost model prefers it
    const uint64_t qn = P.m - P.n + 1, dbits = mpz_sizeinbase(ds[i], 2);
    const auto key = std::make_tuple(P.n, P.m, dbits);
    auto it = decided.find(key);
    if (it == decided.end()) {
      RfPlan c{}, acyc{};
      const bool u = rf_plan_cyclic(&engine_primes(), std::max<uint64_t>(P.n + 1, qn), P.n, qn, &c) == 0 &&
                     (rf_plan_choose(&engine_primes(), dbits, q_bits, 1, 1, &acyc) != 0 || c.cost < acyc.cost);
      it = decided.emplace(key, std::make_pair(u, c)).first;
    }
    const bool use_cyc = it->second.first;
    cyc[i] = it->second.second;
    if (use_cyc) wrap[i] = (uint64_t)cyc[i].t << cyc[i].lgN;
    ents.push_back(Entry{0, dslab_p + d_off[i], P.n, dbits, 1});
    bmax.push_back(q_bits);
    fixed.push_back(use_cyc ? &cyc[i] : nullptr);
  }
  std::vector<rf_operand*> ops;
  rc = make_operands_batch(ents, bmax, fixed, pipe, ops);  // waits for the reciprocals first (same stream)
  if (rc) return rc;
  for (int i = 0; i < cnt; i++) {
    rf_mod* md = new rf_mod();
    md->n = prep[i].n;
    md->m = prep[i].m;
    md->wrap = wrap[i];
    md->inv = ops[2 * i];
    md->div = ops[2 * i + 1];
    md->d_dev = dslab_p + d_off[i];
    md->d_slab = dslab;
    outs[i] = md;
  }
  return RF_OK;
}

int mod_create_batch(const uint64_t* max_bits, const mpz_t* ds, int cnt, rf_mod** outs,
                     std::vector<std::vector<uint64_t>>* recips) {
  std::vector<ModPrep> prep(cnt);
  // host preparation (GMP base cases) on up to 16 threads for large levels
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int i = next++; i < cnt; i = next++) mod_prep(max_bits[i], ds[i], prep[i]);
  };
  const int nth = cnt < 32 ? 1 : std::max(1, std::min({cnt / 16, 16, (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nth; t++) pool.emplace_back(work);
  } catch (const std::system_error&) {
    // fewer helper threads: the calling thread takes what is left
  }
  work();
  for (auto& t : pool) t.join();
  std::lock_guard<std::mutex> lk(engine_mutex());
  // hw_mpz_clear may have run meanwhile
  if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called");
  return mod_create_batch_locked(ds, prep, outs, recips);
}

}  // namespace

extern "C" {

int rf_mod_create(uint64_t max_bits, const mpz_t d, rf_mod** out) {
  if (!out) return fail(RF_ERR_ARG, "rf_mod_create: null output");
  *out = nullptr;
  if (!d || mpz_sgn(d) <= 0) return fail(RF_ERR_ARG, "rf_mod_create: the modulus must be positive");
  if (max_bits == 0) return fail(RF_ERR_ARG, "rf_mod_create: max_bits must be positive");
  return mod_create_batch(&max_bits, (const mpz_t*)d, 1, out, nullptr);
}

int rf_mod_create_batch(const uint64_t* max_bits, const mpz_t* ds, int cnt, rf_mod** outs) {
  if (cnt < 0 || (cnt > 0 && (!max_bits || !ds || !outs))) return fail(RF_ERR_ARG, "rf_mod_create_batch: bad arguments");
  for (int i = 0; i < cnt; i++) outs[i] = nullptr;
  for (int i = 0; i < cnt; i++) {
    if (mpz_sgn(ds[i]) <= 0) return fail(RF_ERR_ARG, "rf_mod_create_batch: modulus %d is not positive", i);
    if (max_bits[i] == 0) return fail(RF_ERR_ARG, "rf_mod_create_batch: max_bits[%d] must be positive", i);
  }
  if (cnt == 0) return RF_OK;
  return mod_create_batch(max_bits, ds, cnt, outs, nullptr);
}

int rf_stage_mod_reciprocal(const uint64_t* d, uint64_t dn, uint64_t max_bits, uint64_t* out, uint64_t cap,
                            uint64_t* nlimbs) {
  if (!d || dn == 0 || d[dn - 1] == 0 || max_bits == 0 || !out || !nlimbs)
    return fail(RF_ERR_ARG, "rf_stage_mod_reciprocal: bad arguments");
  mpz_t D;
  mpz_init(D);
  mpz_import(D, dn, -1, 8, 0, 0, d);
  std::vector<std::vector<uint64_t>> rec;
  const int rc = mod_create_batch(&max_bits, (const mpz_t*)D, 1, nullptr, &rec);
  mpz_clear(D);
  if (rc) return rc;
  if (rec[0].size() > cap) return fail(RF_ERR_CAPACITY, "the reciprocal has %zu limbs", rec[0].size());
  memcpy(out, rec[0].data(), rec[0].size() * 8);
  *nlimbs = rec[0].size();
  return RF_OK;
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
  mod->d_slab.reset();  // D's limbs, freed with the batch's last modulus
  delete mod;
}

}  // extern "C"
