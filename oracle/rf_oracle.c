/*
 * rf_oracle.c -- CPU restatement of the reference's exact matrix product and
 * its stages (see rf_oracle.h).  TEST INFRASTRUCTURE ONLY: loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as a checker; the
 * product library never links it.
 *
 * Straightforward, unoptimised code: 128-bit integer arithmetic, O(N log N)
 * radix-2 transforms, schoolbook multiword sums.  Each function cites the
 * reference code it states.
 */
#include "rf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

uint64_t orc_mulmod(uint64_t x, uint64_t y, uint64_t p) { return (uint64_t)(((u128)x * y) % p); }

uint64_t orc_powmod(uint64_t x, uint64_t e, uint64_t p)
{
  uint64_t acc = 1 % p;
  x %= p;
  while (e) {
    if (e & 1) acc = orc_mulmod(acc, x, p);
    x = orc_mulmod(x, x, p);
    e >>= 1;
  }
  return acc;
}

uint64_t orc_invmod(uint64_t x, uint64_t p) { return orc_powmod(x, p - 2, p); }

/* src/lib/fft62.c:26-35 */
unsigned orc_log2(size_t n)
{
  unsigned l = 0;
  if (n == 0) return 0;
  while (((size_t)1 << l) < n) l++;
  return l;
}

/* src/lib/fft62.c:140-168: array decomposition lgN = lgK + lgM with
   lgK = lgN / 2 above the threshold 2^16; admissible sizes are multiples of
   2^lgM whose row count is itself admissible for 2^lgK. */
size_t orc_next_size(size_t n, unsigned lgN)
{
  if (n == 0) n = 1;
  if (lgN <= 16) return n;
  unsigned lgK = lgN / 2, lgM = lgN - lgK;
  size_t rows = ((n - 1) >> lgM) + 1;
  return orc_next_size(rows, lgK) << lgM;
}

/* src/lib/mpnfft.c:18-78 (gpu: src/lib/mpnfft.c:664-708).  The reference
   works in IEEE doubles with frexp; this keeps that arithmetic so the chosen
   (r, lgN, points) are the reference's. */
void orc_ref_params(size_t bits, size_t terms, unsigned k, const uint64_t* p, int gpu,
                    unsigned* r_out, unsigned* lgN_out, size_t* points_out)
{
  double bound = 0.4999;
  for (unsigned i = 0; i < k; i++) bound *= (double)p[i];
  double bound2 = bound / (double)terms;
  unsigned r = 62 * k / 2;
  for (;;) {
    size_t points = (bits - 1) / r + 2;
    double thing = bound2 / (double)((points + 1) / 2);
    int e;
    frexp(thing, &e);
    unsigned rr = ((unsigned)e - 1) / 2;
    if (rr == r) {
      unsigned lgN = orc_log2(points);
      *r_out = r;
      *lgN_out = lgN;
      *points_out = gpu ? ((size_t)1 << lgN) : orc_next_size(points, lgN);
      return;
    }
    r = rr;
  }
}

/* src/lib/fft62.c:40-50 and 70-82, with FFT62_MAX_LGN replaced by s. */
uint64_t orc_root(uint64_t p, unsigned s, unsigned lgN)
{
  uint64_t g = 0;
  for (uint64_t u = 2;; u++) {
    uint64_t w = orc_powmod(u, (p - 1) >> s, p);
    if (orc_powmod(w, (uint64_t)1 << (s - 1), p) != 1) {
      g = w;
      break;
    }
  }
  for (unsigned i = lgN; i < s; i++) g = orc_mulmod(g, g, p);
  return g;
}

/* Bit field [pos, pos + len) of {up, un}, len <= 64; zero beyond un. */
static uint64_t get_bits(const uint64_t* up, size_t un, size_t pos, unsigned len)
{
  size_t li = pos / 64;
  unsigned sh = pos % 64;
  uint64_t lo = li < un ? up[li] : 0;
  uint64_t hi = li + 1 < un ? up[li + 1] : 0;
  uint64_t v = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
  return len == 64 ? v : (v & (((uint64_t)1 << len) - 1));
}

/* src/lib/split.c:10-81 */
void orc_split(uint64_t* rp, size_t rn, const uint64_t* up, size_t un, unsigned b)
{
  unsigned t = (b + 63) / 64;
  for (size_t i = 0; i < rn; i++)
    for (unsigned l = 0; l < t; l++) {
      unsigned len = (l + 1) * 64 <= b ? 64 : b - l * 64;
      rp[i * t + l] = get_bits(up, un, (size_t)i * b + (size_t)l * 64, len);
    }
}

/* src/lib/reduce.c:10-28 and src/lib/split_reduce.c:27-38 (lgS = 0):
   sign * chunk mod p, canonical. */
void orc_reduce(uint64_t* out, const uint64_t* chunks, size_t rn, unsigned t, int sign, uint64_t p)
{
  uint64_t B = (uint64_t)(((u128)1 << 64) % p);
  for (size_t i = 0; i < rn; i++) {
    uint64_t acc = 0;
    for (unsigned l = t; l-- > 0;)
      acc = (uint64_t)(((u128)acc * B + chunks[i * t + l]) % p);
    out[i] = (sign < 0 && acc) ? p - acc : acc;
  }
}

static size_t bitrev(size_t x, unsigned bits)
{
  size_t r = 0;
  for (unsigned i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

/* Transform of src/include/fft62.h:35-54 (Z = 1): radix-2 decimation in
   frequency on natural-order input gives the bit-reversed output order the
   reference defines. */
void orc_ntt_fwd(uint64_t* y, size_t yn, const uint64_t* x, size_t xn, unsigned lgN, uint64_t p,
                 uint64_t w)
{
  size_t N = (size_t)1 << lgN;
  uint64_t* a = calloc(N, sizeof(uint64_t));
  for (size_t i = 0; i < xn; i++) a[i] = x[i] % p;
  for (size_t half = N / 2; half >= 1; half /= 2) {
    /* w_len = root of order 2*half */
    uint64_t wl = orc_powmod(w, N / (2 * half), p);
    for (size_t s = 0; s < N; s += 2 * half) {
      uint64_t z = 1;
      for (size_t j = 0; j < half; j++) {
        uint64_t u = a[s + j], v = a[s + j + half];
        a[s + j] = (u + v) % p;
        a[s + j + half] = orc_mulmod((u + p - v) % p, z, p);
        z = orc_mulmod(z, wl, p);
      }
    }
    if (half == 1) break;
  }
  memcpy(y, a, yn * sizeof(uint64_t));
  free(a);
}

/* Inverse of orc_ntt_fwd (src/lib/fft62.c:922-1043 and the 2^-lgN scaling of
   src/lib/mpnfft.c:418-428): undo the bit reversal, transform with w^-1. */
void orc_ntt_inv(uint64_t* out, const uint64_t* y, unsigned lgN, uint64_t p, uint64_t w, int scale)
{
  size_t N = (size_t)1 << lgN;
  uint64_t wi = orc_invmod(w, p);
  uint64_t* a = malloc(N * sizeof(uint64_t));
  for (size_t m = 0; m < N; m++) a[bitrev(m, lgN)] = y[m] % p;
  /* natural-order spectrum -> DIF with w^-1 gives bit-reversed N*a */
  uint64_t* t = malloc(N * sizeof(uint64_t));
  orc_ntt_fwd(t, N, a, N, lgN, p, wi);
  uint64_t ninv = scale ? orc_invmod((uint64_t)(N % p), p) : 1;
  for (size_t m = 0; m < N; m++) out[bitrev(m, lgN)] = orc_mulmod(t[m], ninv, p);
  free(a);
  free(t);
}

/* src/lib/mpnfft.c:559-651 */
void orc_pointwise(uint64_t* C, const uint64_t* A, const uint64_t* B, unsigned r, unsigned d,
                   size_t n, uint64_t p)
{
  for (unsigned i = 0; i < r; i++)
    for (unsigned j = 0; j < d; j++)
      for (size_t h = 0; h < n; h++) {
        u128 acc = 0;
        for (unsigned l = 0; l < d; l++) {
          acc += (u128)(A[(i * d + l) * n + h] % p) * (B[(l * d + j) * n + h] % p);
          acc %= p;
        }
        C[(i * d + j) * n + h] = (uint64_t)acc;
      }
}

/* multiword helpers on little-endian u64 limbs */
static void mw_mul_add_1(uint64_t* x, unsigned n, uint64_t m, uint64_t a)
{
  u128 c = a;
  for (unsigned i = 0; i < n; i++) {
    c += (u128)x[i] * m;
    x[i] = (uint64_t)c;
    c >>= 64;
  }
}

static int mw_cmp(const uint64_t* x, const uint64_t* y, unsigned n)
{
  for (unsigned i = n; i-- > 0;)
    if (x[i] != y[i]) return x[i] > y[i] ? 1 : -1;
  return 0;
}

static void mw_sub(uint64_t* x, const uint64_t* y, unsigned n)
{
  uint64_t br = 0;
  for (unsigned i = 0; i < n; i++) {
    uint64_t xi = x[i], yi = y[i];
    uint64_t d = xi - yi - br;
    br = (xi < yi) || (xi == yi && br);
    x[i] = d;
  }
}

/* src/lib/crt.c:11-94: Garner's mixed-radix CRT, then the symmetric lift into
   (-P/2, P/2) written in two's complement. */
void orc_crt(uint64_t* rp, unsigned s, const uint64_t* res, size_t un, unsigned k,
             const uint64_t* p)
{
  uint64_t P[ORC_MAXP + 2], Ph[ORC_MAXP + 2], X[ORC_MAXP + 2];
  uint64_t cinv[ORC_MAXP];
  memset(P, 0, sizeof P);
  P[0] = 1;
  for (unsigned j = 0; j < k; j++) mw_mul_add_1(P, s, p[j], 0);
  for (unsigned l = 0; l < s; l++) Ph[l] = (P[l] >> 1) | (l + 1 < s ? P[l + 1] << 63 : 0);
  for (unsigned j = 1; j < k; j++) {
    uint64_t m = 1;
    for (unsigned l = 0; l < j; l++) m = orc_mulmod(m, p[l] % p[j], p[j]);
    cinv[j] = orc_invmod(m, p[j]);
  }
  uint64_t v[ORC_MAXP];
  for (size_t i = 0; i < un; i++) {
    for (unsigned j = 0; j < k; j++) {
      uint64_t x = res[(size_t)j * un + i] % p[j];
      uint64_t acc = 0; /* v_0 + p_0 (v_1 + ... ) mod p_j */
      for (unsigned l = j; l-- > 0;) acc = (orc_mulmod(acc, p[l] % p[j], p[j]) + v[l] % p[j]) % p[j];
      v[j] = orc_mulmod((x + p[j] - acc) % p[j], j ? cinv[j] : 1, p[j]);
    }
    memset(X, 0, sizeof X);
    for (unsigned j = k; j-- > 0;) mw_mul_add_1(X, s, j + 1 < k ? p[j] : 0, 0), mw_mul_add_1(X, s, 1, v[j]);
    if (mw_cmp(X, Ph, s) > 0) mw_sub(X, P, s);
    memcpy(rp + i * s, X, s * sizeof(uint64_t));
  }
}

/* src/lib/recompose.c:11-107 (r0 = 0). */
void orc_recompose(uint64_t* rp, size_t rn, const uint64_t* up, size_t un, unsigned s, unsigned b)
{
  /* Accumulate positive and negative parts separately, then subtract. */
  uint64_t* pos = calloc(rn + 1, sizeof(uint64_t));
  uint64_t* neg = calloc(rn + 1, sizeof(uint64_t));
  uint64_t* tmp = malloc((s + 2) * sizeof(uint64_t));
  for (size_t i = 0; i < un; i++) {
    const uint64_t* d = up + i * s;
    int negative = (d[s - 1] >> 63) & 1;
    /* magnitude of d as s limbs */
    uint64_t br = 1;
    for (unsigned l = 0; l < s; l++) {
      uint64_t x = negative ? ~d[l] : d[l];
      if (negative) {
        uint64_t y = x + br;
        br = (y < x);
        x = y;
      }
      tmp[l] = x;
    }
    tmp[s] = 0;
    tmp[s + 1] = 0;
    size_t sh = i * (size_t)b, li = sh / 64;
    unsigned bs = sh % 64;
    if (li >= rn) continue;
    uint64_t* acc = negative ? neg : pos;
    uint64_t carry = 0;
    for (size_t l = 0; li + l < rn + 1; l++) {
      uint64_t lo = l < s + 2 ? tmp[l] : 0;
      uint64_t prev = (l >= 1 && l - 1 < s + 2) ? tmp[l - 1] : 0;
      uint64_t part = bs ? (lo << bs) | (prev >> (64 - bs)) : lo;
      if (l > s + 2 && carry == 0) break;
      u128 t = (u128)acc[li + l] + part + carry;
      acc[li + l] = (uint64_t)t;
      carry = (uint64_t)(t >> 64);
    }
  }
  uint64_t br = 0;
  for (size_t l = 0; l < rn; l++) {
    uint64_t x = pos[l], y = neg[l];
    rp[l] = x - y - br;
    br = (x < y) || (x == y && br);
  }
  free(pos);
  free(neg);
  free(tmp);
}

size_t orc_ref_out_limbs(size_t points, unsigned b, unsigned k)
{
  return ((points - 1) * b + 62 * k + 2) / 64 + 1;
}

/* src/include/fft62.h:24-27 */
static const uint64_t REF_PRIMES[4] = {0x3fffc00000000001ULL, 0x3ffdf00000000001ULL,
                                       0x3ffd900000000001ULL, 0x3ffd500000000001ULL};

void orc_ref_plan(orc_plan* pl, size_t max_limbs, unsigned d, size_t* points)
{
  unsigned r, lgN;
  size_t bits = 64 * (2 * max_limbs + 1); /* src/lib/hwmpz.c:54-63 */
  orc_ref_params(bits, d, 4, REF_PRIMES, 0, &r, &lgN, points);
  memset(pl, 0, sizeof *pl);
  pl->b = r;
  pl->lgN = lgN;
  pl->k = 4;
  for (unsigned j = 0; j < 4; j++) {
    pl->p[j] = REF_PRIMES[j];
    pl->w[j] = orc_root(REF_PRIMES[j], 44, lgN);
  }
}

/* log2 of P rounded down, exactly: bit length of prod p_j, minus one. */
static unsigned floor_log2_P(const orc_plan* pl)
{
  uint64_t P[ORC_MAXP + 2];
  memset(P, 0, sizeof P);
  P[0] = 1;
  unsigned s = pl->k + 1;
  for (unsigned j = 0; j < pl->k; j++) mw_mul_add_1(P, s, pl->p[j], 0);
  for (unsigned l = s; l-- > 0;)
    if (P[l]) return 64 * l + 63 - (unsigned)__builtin_clzll(P[l]);
  return 0;
}

/* Exact test of f * (2^b - 1)^2 < P, in multiword integers. */
static int bound_holds(const orc_plan* pl, uint64_t f)
{
  enum { W = 2 * ORC_MAXP + 4 };
  uint64_t P[W], m[W], sq[W];
  memset(P, 0, sizeof P);
  P[0] = 1;
  for (unsigned j = 0; j < pl->k; j++) mw_mul_add_1(P, W, pl->p[j], 0);
  unsigned b = pl->b, nl = (b + 63) / 64;
  if (2 * nl + 2 > W) return 0;
  memset(m, 0, sizeof m);
  for (unsigned l = 0; l < nl; l++) m[l] = (l + 1) * 64 <= b ? ~(uint64_t)0 : (((uint64_t)1 << (b % 64)) - 1);
  memset(sq, 0, sizeof sq);
  for (unsigned i = 0; i < nl; i++) {
    u128 c = 0;
    for (unsigned j = 0; j < nl; j++) {
      c += (u128)m[i] * m[j] + sq[i + j];
      sq[i + j] = (uint64_t)c;
      c >>= 64;
    }
    sq[i + nl] = (uint64_t)c;
  }
  mw_mul_add_1(sq, W, f, 0);
  return mw_cmp(sq, P, W) < 0;
}

static size_t bitlen(const uint64_t* x, size_t n)
{
  while (n && x[n - 1] == 0) n--;
  return n ? 64 * n - (size_t)__builtin_clzll(x[n - 1]) : 0;
}

/* src/lib/hwmpz.c:65-130 with plan pl, full-length transforms. */
int orc_rmatrix_mult(uint64_t* c, const size_t* coff, const size_t* ccap, long* csize,
                     const uint64_t* a, const size_t* aoff, const size_t* an, const int* as,
                     const uint64_t* bb, const size_t* boff, const size_t* bn, const int* bs,
                     unsigned r, unsigned d, const orc_plan* pl)
{
  const unsigned b = pl->b, k = pl->k, lgN = pl->lgN;
  const size_t N = (size_t)1 << lgN;
  const unsigned t = (b + 63) / 64;
  size_t na = 0, nb = 0, la = 0, lb = 0;
  for (unsigned e = 0; e < r * d; e++) {
    size_t L = as[e] ? bitlen(a + aoff[e], an[e]) : 0;
    if (L > la) la = L;
  }
  for (unsigned e = 0; e < d * d; e++) {
    size_t L = bs[e] ? bitlen(bb + boff[e], bn[e]) : 0;
    if (L > lb) lb = L;
  }
  na = (la + b - 1) / b;
  nb = (lb + b - 1) / b;
  /* bound: na + nb - 1 <= N and 2 * d * min(na, nb) * (2^b - 1)^2 < P */
  if (na && nb) {
    size_t m = na < nb ? na : nb;
    if (na + nb - 1 > N) return -1;
    if (!bound_holds(pl, 2 * (uint64_t)d * m)) return -1;
  }
  size_t n_out = (na && nb) ? na + nb - 1 : 0;
  unsigned s = (floor_log2_P(pl) + 2 + 63) / 64;

  uint64_t* specA = calloc((size_t)r * d * N, sizeof(uint64_t));
  uint64_t* specB = calloc((size_t)d * d * N, sizeof(uint64_t));
  uint64_t* specC = calloc((size_t)r * d * N, sizeof(uint64_t));
  uint64_t* res = calloc((size_t)k * (n_out ? n_out : 1), sizeof(uint64_t));
  uint64_t* crt = calloc((size_t)s * (n_out ? n_out : 1), sizeof(uint64_t));
  uint64_t* chunks = calloc((size_t)t * (na > nb ? na : nb) + 1, sizeof(uint64_t));
  uint64_t* tmp = calloc(N, sizeof(uint64_t));
  uint64_t* coefC = calloc((size_t)r * d * k * (n_out ? n_out : 1), sizeof(uint64_t));

  for (unsigned q = 0; q < k; q++) {
    uint64_t p = pl->p[q], w = pl->w[q];
    for (unsigned e = 0; e < r * d; e++) {
      memset(tmp, 0, N * sizeof(uint64_t));
      if (as[e] && na) {
        orc_split(chunks, na, a + aoff[e], an[e], b);
        orc_reduce(tmp, chunks, na, t, as[e], p);
      }
      orc_ntt_fwd(specA + (size_t)e * N, N, tmp, N, lgN, p, w);
    }
    for (unsigned e = 0; e < d * d; e++) {
      memset(tmp, 0, N * sizeof(uint64_t));
      if (bs[e] && nb) {
        orc_split(chunks, nb, bb + boff[e], bn[e], b);
        orc_reduce(tmp, chunks, nb, t, bs[e], p);
      }
      orc_ntt_fwd(specB + (size_t)e * N, N, tmp, N, lgN, p, w);
    }
    orc_pointwise(specC, specA, specB, r, d, N, p);
    for (unsigned e = 0; e < r * d; e++) {
      orc_ntt_inv(tmp, specC + (size_t)e * N, lgN, p, w, 1);
      if (n_out) memcpy(coefC + ((size_t)e * k + q) * n_out, tmp, n_out * sizeof(uint64_t));
    }
  }

  int rc = 0;
  for (unsigned e = 0; e < r * d; e++) {
    uint64_t* out = c + coff[e];
    if (n_out == 0) {
      csize[e] = 0;
      continue;
    }
    /* exact value has at most la + lb + log2(d) bits; one more for the sign */
    size_t rn = (la + lb + orc_log2(d) + 1) / 64 + 1;
    uint64_t* tw = calloc(rn, sizeof(uint64_t));
    orc_crt(crt, s, coefC + (size_t)e * k * n_out, n_out, k, pl->p);
    orc_recompose(tw, rn, crt, n_out, s, b);
    int neg = (tw[rn - 1] >> 63) & 1;
    if (neg) { /* sign epilogue, src/lib/hwmpz.c:116-122 */
      uint64_t br = 1;
      for (size_t l = 0; l < rn; l++) {
        uint64_t x = ~tw[l], y = x + br;
        br = (y < x);
        tw[l] = y;
      }
    }
    size_t n = rn;
    while (n && tw[n - 1] == 0) n--;
    if (n > ccap[e]) {
      rc = -2;
    } else {
      memcpy(out, tw, n * sizeof(uint64_t));
      csize[e] = neg ? -(long)n : (long)n;
    }
    free(tw);
  }
  free(specA);
  free(specB);
  free(specC);
  free(res);
  free(crt);
  free(chunks);
  free(tmp);
  free(coefC);
  return rc;
}
