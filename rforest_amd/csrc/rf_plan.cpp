// rf_plan.cpp -- host-side prime generation, planning and table construction.
#define RF_HOST_ONLY
#include "rf_plan.h"

#include <vector>

#include <string.h>

#include "rf_field.h"

namespace {

bool is_prime_u64(uint64_t n) {
  if (n < 2) return false;
  static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t q : small) {
    if (n % q == 0) return n == q;
  }
  uint64_t d = n - 1;
  int r = 0;
  while ((d & 1) == 0) d >>= 1, r++;
  // These bases make Miller-Rabin deterministic for all n < 2^64.
  for (uint64_t a : small) {
    uint64_t x = rf_h_powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int i = 1; i < r && comp; i++) {
      x = rf_h_mulmod(x, x, n);
      if (x == n - 1) comp = false;
    }
    if (comp) return false;
  }
  return true;
}

// Root of order 2^RF_S: smallest u >= 2 with w = u^((p-1)/2^S) and
// w^(2^(S-1)) != 1 (the rule of src/lib/fft62.c:40-50 with S for 44).
uint64_t root_2S(uint64_t p) {
  for (uint64_t u = 2;; u++) {
    uint64_t w = rf_h_powmod(u, (p - 1) >> RF_S, p);
    if (rf_h_powmod(w, (uint64_t)1 << (RF_S - 1), p) != 1) return w;
  }
}

// x[0..n) *= m, returns carry out.
uint64_t mw_mul_1(uint64_t* x, int n, uint64_t m) {
  rf_u128 c = 0;
  for (int i = 0; i < n; i++) {
    c += (rf_u128)x[i] * m;
    x[i] = (uint64_t)c;
    c >>= 64;
  }
  return (uint64_t)c;
}

int ceil_log2(uint64_t x) {
  int l = 0;
  while (l < 64 && ((uint64_t)1 << l) < x) l++;
  return l;
}

}  // namespace

void rf_primes_init(RfPrimes* pr) {
  memset(pr, 0, sizeof *pr);
  // p = c 2^S + 1 < RF_PMAX, scanning c downwards from the top.
  for (uint64_t c = (RF_PMAX - 1) >> RF_S; pr->k < RF_MAXP; c--) {
    uint64_t p = (c << RF_S) + 1;
    if (p >= RF_PMAX) continue;
    if (p <= ((uint64_t)1 << 47)) break;
    if (is_prime_u64(p)) {
      pr->p[pr->k] = p;
      pr->g[pr->k] = root_2S(p);
      pr->k++;
    }
  }
}

int rf_floor_log2_P(const RfPrimes* pr, int k) {
  uint64_t P[RF_MAXP + 2] = {1};
  for (int j = 0; j < k; j++) mw_mul_1(P, RF_MAXP + 2, pr->p[j]);
  for (int l = RF_MAXP + 1; l >= 0; l--)
    if (P[l]) return 64 * l + 63 - __builtin_clzll(P[l]);
  return 0;
}

bool rf_bound_holds(const RfPrimes* pr, int k, int b, uint64_t f) {
  enum { W = 2 * RF_TMAX + 4 > RF_MAXP + 2 ? 2 * RF_TMAX + 4 : RF_MAXP + 2 };
  uint64_t P[W] = {1}, m[W] = {0}, sq[W] = {0};
  for (int j = 0; j < k; j++) mw_mul_1(P, W, pr->p[j]);
  int nl = (b + 63) / 64;
  if (2 * nl + 2 > W) return false;
  for (int l = 0; l < nl; l++) m[l] = (l + 1) * 64 <= b ? ~(uint64_t)0 : (((uint64_t)1 << (b % 64)) - 1);
  for (int i = 0; i < nl; i++) {
    rf_u128 c = 0;
    for (int j = 0; j < nl; j++) {
      c += (rf_u128)m[i] * m[j] + sq[i + j];
      sq[i + j] = (uint64_t)c;
      c >>= 64;
    }
    sq[i + nl] = (uint64_t)c;
  }
  if (mw_mul_1(sq, W, f)) return false;
  for (int l = W - 1; l >= 0; l--)
    if (sq[l] != P[l]) return sq[l] < P[l];
  return false;
}

int rf_plan_fixed(const RfPrimes* pr, size_t LA, size_t LB, int r, int d, int t, int lgN,
                  RfPlan* out) {
  return rf_plan_fixed_k(pr, LA, LB, r, d, t, lgN, 0, out);
}

int rf_plan_fixed_k(const RfPrimes* pr, size_t LA, size_t LB, int r, int d, int t, int lgN, int kmin,
                    RfPlan* out) {
  const size_t b = 64 * (size_t)t;
  RfPlan pl;
  memset(&pl, 0, sizeof pl);
  pl.t = t;
  pl.lgN = lgN;
  pl.split = 0;
  size_t nA = (LA + b - 1) / b, nB = (LB + b - 1) / b;
  pl.nA = (int)nA;
  pl.nB = (int)nB;
  if (nA == 0 || nB == 0) {  // product is zero
    pl.k = 0;
    pl.nout = 0;
    pl.rn = 0;
    *out = pl;
    return 0;
  }
  size_t nout = nA + nB - 1;
  if (lgN > RF_LGN_MAX || lgN < RF_LGN_MIN || nout > ((size_t)1 << lgN)) return -1;
  uint64_t f = 2 * (uint64_t)d * (nA < nB ? nA : nB);
  int k = 0;
  for (int kk = 1; kk <= pr->k; kk++)
    if (rf_bound_holds(pr, kk, (int)b, f)) {
      k = kk;
      break;
    }
  if (k == 0 || kmin > pr->k) return -1;
  if (k < kmin) k = kmin;  // more primes than the bound needs are always valid
  if (k > RF_KMAX) return -1;
  if (((size_t)d * k << lgN) > 0xffffffffu) return -1;  // spectrum strides are 32-bit
  pl.k = k;
  pl.nout = (int)nout;
  // Coefficients lie in (-P/2, P/2) with P < 2^(48 k): 48 k + 1 bits of two's
  // complement suffice, and the CRT kernel is specialised on k alone.
  pl.s = (48 * k + 1 + 63) / 64;
  if (pl.s > RF_MAXS) return -1;
  pl.M = (pl.s + t - 1) / t;
  size_t bits = LA + LB + (size_t)ceil_log2((uint64_t)d) + 1;
  pl.rn = (bits + 63) / 64;
  // Cost estimate in VALU instructions (full-rate, per wave lane): transforms
  // at ~9 per butterfly, pack at ~4 per 32-bit word, pointwise ~8 per term,
  // Garner ~5 k^2 and base conversion ~6 k s per coefficient.  HBM bytes are
  // weighed at 39 instructions per 8 bytes (39.3 T lane-op/s vs 8 TB/s).
  double N = (double)((size_t)1 << lgN);
  double polys = (double)(r * d + d * d), outs = (double)(r * d);
  double words = 2.0 * t;
  double ntt = N / 2 * lgN * 9.0;
  double ops = polys * k * (ntt + (double)nA * (4 * words + 20)) +
               outs * k * (ntt + N * 8 * d) + outs * (double)nout * (5.0 * k * k + 6.0 * k * pl.s);
  double bytes = 8 * N * k * (polys + outs) * 2;
  // Transforms longer than one workgroup's LDS take a column pass and a row
  // pass: one more read and write of every residue, and a few more ops.
  if (lgN > RF_LGN_LDS) bytes *= 2, ops *= 1.1;
  if (lgN > RF_LGN_ROW + 4) bytes *= 1.5;  // a second column pass
  if (lgN > RF_LGN_ROW + 8) bytes *= 1.33;  // a third
  double c = ops > bytes * 39 / 8 ? ops : bytes * 39 / 8;
  if (lgN == RF_LGN_LDS) c *= 1.25;  // one workgroup per CU at 128 KiB of LDS
  pl.cost = c;
  *out = pl;
  return 0;
}

int rf_plan_choose(const RfPrimes* pr, size_t LA, size_t LB, int r, int d, RfPlan* out) {
  bool found = false;
  RfPlan best;
  for (int t = 1; t <= RF_TMAX; t++) {
    size_t b = 64 * (size_t)t;
    size_t nA = (LA + b - 1) / b, nB = (LB + b - 1) / b;
    if (nA == 0 || nB == 0) return rf_plan_fixed(pr, LA, LB, r, d, t, RF_LGN_MIN, out);
    int lgN = ceil_log2(nA + nB - 1);
    if (lgN < RF_LGN_MIN) lgN = RF_LGN_MIN;
    RfPlan pl;
    if (rf_plan_fixed(pr, LA, LB, r, d, t, lgN, &pl) != 0) continue;
    if (!found || pl.cost < best.cost) best = pl, found = true;
  }
  if (!found) return -1;
  *out = best;
  return 0;
}

int rf_plan_general(const RfPrimes* pr, size_t bits, unsigned terms, RfPlan* out) {
  if (bits == 0 || terms == 0) return -1;
  bool found = false;
  RfPlan best;
  memset(&best, 0, sizeof best);
  for (int t = 1; t <= RF_TMAX; t++) {
    const size_t b = 64 * (size_t)t;
    // |x| |y| < 2^bits gives bits(x) + bits(y) <= bits + 1, so the chunk
    // counts satisfy nx + ny - 1 <= ceil((bits + 1) / b) + 1 = points, and
    // min(nx, ny) <= ceil(points / 2)
    const size_t points = (bits + 1 + b - 1) / b + 1;
    int lgN = ceil_log2(points);
    if (lgN < RF_LGN_MIN) lgN = RF_LGN_MIN;
    if (lgN > RF_LGN_MAX) continue;
    const uint64_t f = 2 * (uint64_t)terms * ((points + 1) / 2);
    int k = 0;
    for (int kk = 1; kk <= pr->k; kk++)
      if (rf_bound_holds(pr, kk, (int)b, f)) {
        k = kk;
        break;
      }
    if (k == 0 || k > RF_KMAX) continue;
    RfPlan pl;
    memset(&pl, 0, sizeof pl);
    pl.t = t;
    pl.lgN = lgN;
    pl.k = k;
    pl.s = (48 * k + 1 + 63) / 64;
    if (pl.s > RF_MAXS) continue;
    pl.M = (pl.s + t - 1) / t;
    pl.nA = pl.nB = (int)points;
    pl.nout = (int)points;
    pl.rn = (bits + 63) / 64 + 1;
    // residues moved per transform, the dominant cost of every operation
    const int passes = lgN <= RF_LGN_LDS ? 1 : 2 + (lgN - RF_LGN_ROW - 1) / 4;  // row pass + column passes
    pl.cost = (double)k * (double)((size_t)1 << lgN) * passes;
    if (!found || pl.cost < best.cost) best = pl, found = true;
  }
  if (!found) return -1;
  *out = best;
  return 0;
}

int rf_plan_cyclic(const RfPrimes* pr, uint64_t min_limbs, uint64_t a_limbs, uint64_t b_limbs, RfPlan* out) {
  if (min_limbs == 0 || a_limbs == 0 || b_limbs == 0) return -1;
  bool found = false;
  RfPlan best;
  memset(&best, 0, sizeof best);
  for (int t = 1; t <= RF_TMAX; t++) {
    int lgN = ceil_log2((min_limbs + t - 1) / t);
    if (lgN < RF_LGN_MIN) lgN = RF_LGN_MIN;
    if (lgN + 1 > RF_LGN_MAX) continue;
    const uint64_t N = (uint64_t)1 << lgN;
    if ((a_limbs + t - 1) / t > N || (b_limbs + t - 1) / t > N) continue;
    // the acyclic plan of twice the length has the same chunks, so the same
    // coefficient bound (prime count) and CRT shape
    RfPlan pl;
    if (rf_plan_fixed_k(pr, 64 * a_limbs, 64 * b_limbs, 1, 1, t, lgN + 1, 0, &pl) != 0) continue;
    pl.lgN = lgN;
    pl.split = 0;
    pl.cyclic = 1;
    pl.nout = (int)N;
    pl.rn = (size_t)t * N + t + 1;
    pl.cost *= 0.5;
    if (!found || pl.cost < best.cost) best = pl, found = true;
  }
  if (!found) return -1;
  *out = best;
  return 0;
}

void rf_consts_init(const RfPrimes* pr, RfConsts* c) {
  memset(c, 0, sizeof *c);
  for (int j = 0; j < pr->k; j++) {
    uint64_t p = pr->p[j];
    c->pu[j] = p;
    c->p[j] = (double)p;
    c->pinv[j] = 1.0 / (double)p;
    uint64_t prod = 1;
    for (int l = 0; l < j; l++) {
      uint64_t v = pr->p[l] % p;
      c->pm[j][l][0] = (double)v;
      c->pm[j][l][1] = (double)v / (double)p;
      prod = rf_h_mulmod(prod, v, p);
    }
    uint64_t inv = rf_h_invmod(prod, p);
    c->ci[j][0] = (double)inv;
    c->ci[j][1] = (double)inv / (double)p;
    for (int v = 0; v <= RF_LGN_MAX; v++) {
      uint64_t S = v == 0 ? 1 : rf_h_invmod(((uint64_t)1 << v) % p, p);
      for (int w = 0; w < RF_MAXWORDS; w++) {
        uint64_t K = rf_h_mulmod(rf_h_powmod(2, (uint64_t)32 * w, p), S, p);
        c->pk[j][v][w][0] = (double)(K & 0xffff);
        c->pk[j][v][w][1] = (double)(K & 0xffff0000u);          // k1 2^16, exact
        c->pk[j][v][w][2] = (double)(K & 0xffff00000000ull);    // k2 2^32, exact
      }
    }
  }
  for (int k = 0; k <= pr->k; k++) {
    uint64_t P[RF_MAXS + 1] = {1};
    for (int j = 0; j < k; j++) mw_mul_1(P, RF_MAXS + 1, pr->p[j]);
    for (int l = 0; l < RF_MAXS; l++) c->P[k][l] = P[l];
  }
}

void rf_twiddle_roots(const RfPrimes* pr, int lgmax, double* om, double* omi) {
  for (int q = 0; q < pr->k && q < RF_TW_PRIMES; q++) {
    const uint64_t p = pr->p[q];
    const uint64_t w = rf_h_powmod(pr->g[q], (uint64_t)1 << (RF_S - lgmax), p);  // order 2^lgmax
    om[q] = (double)w;
    omi[q] = (double)rf_h_invmod(w, p);
  }
}
