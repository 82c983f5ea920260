/*
 * Test-infrastructure driver for the reference CPU path, compiled together
 * with the reference's own C sources (read in place, never copied) into
 * oracle/_ref/libzzref.so by oracle/Makefile.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as a checker / baseline.  Nothing in the product
 * links it.
 *
 * The reference's driver mpz_rmatrix_mult_fft (src/lib/hwmpz.c:65-130) cannot
 * be linked as-is because hwmpz.c's setup also initialises the CUDA path
 * (src/lib/hwmpz.c:24-40).  zzref_rmatrix_mult_fft below performs the same
 * sequence of reference calls over the reference's own zz_moduli_t with 4
 * primes (src/lib/hwmpz.c:13), so its output and its timing are the
 * reference CPU path's.  The zzref_* stage hooks expose single reference
 * functions so the repository's own C restatement (oracle/rf_oracle.c) can be
 * pinned against them stage by stage.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <gmp.h>

#include "mpnfft.h"
#include "mpzfft_moduli.h"
#include "split_reduce.h"
#include "zzcrt.h"
#include "recompose.h"
#include "crt_recompose.h"
#include "fft62.h"
#include "mod62.h"

#define EXPORT __attribute__((visibility("default")))
#define ZZREF_PRIMES 4 /* src/lib/hwmpz.c:13 */

static zz_moduli_t g_moduli;
static int g_ready;

EXPORT int zzref_setup(void)
{
  if (!g_ready) {
    zz_moduli_init(&g_moduli, ZZ_MAX_PRIMES); /* src/lib/hwmpz.c:28 */
    g_ready = 1;
  }
  return 0;
}

EXPORT void zzref_clear(void)
{
  if (g_ready) {
    zz_moduli_clear(&g_moduli);
    g_ready = 0;
  }
}

EXPORT uint64_t zzref_prime(unsigned i) { zzref_setup(); return g_moduli.p[i]; }

/* Root of order 2^lgN the reference uses for prime i (fft62_mod_t.rtab). */
EXPORT uint64_t zzref_root(unsigned i, unsigned lgN)
{
  zzref_setup();
  return g_moduli.fft62_mod[i].rtab[lgN];
}

/* zz_mpnfft_params_init (src/lib/mpnfft.c:18-78); gpu != 0 selects the
   cu_ variant with points = N (src/lib/mpnfft.c:664-708). */
EXPORT void zzref_params(size_t bits, size_t terms, unsigned num_primes, int gpu,
                         unsigned* r, unsigned* lgN, size_t* points)
{
  zz_mpnfft_params_t P;
  zzref_setup();
  if (gpu)
    cu_zz_mpnfft_params_init(&P, bits, terms, num_primes, &g_moduli);
  else
    zz_mpnfft_params_init(&P, bits, terms, num_primes, &g_moduli);
  *r = P.r;
  *lgN = P.lgN;
  *points = P.points;
}

EXPORT size_t zzref_next_size(size_t n, unsigned lgN) { return fft62_next_size(n, lgN); }
EXPORT unsigned zzref_log2(size_t n) { return fft62_log2(n); }

/* zz_split_reduce (src/lib/split_reduce.c:15-66): out is num_primes x rn,
   prime-major; values in [0, 2p). */
EXPORT void zzref_split_reduce(uint64_t* out, size_t rn, int sign, int lgS, unsigned num_primes,
                               const uint64_t* up, size_t un, unsigned b)
{
  uint64_t* rp[ZZ_MAX_PRIMES];
  zzref_setup();
  for (unsigned j = 0; j < num_primes; j++) rp[j] = out + j * rn;
  zz_split_reduce(rp, rn, sign, lgS, &g_moduli, num_primes, (const mp_limb_t*)up, un, b, 1);
}

/* fft62_fft (src/include/fft62.h:155-160) for prime i. xp is modified only
   if yp == xp. */
EXPORT void zzref_fft(uint64_t* yp, size_t yn, uint64_t* xp, size_t xn, unsigned lgN, unsigned i)
{
  zzref_setup();
  fft62_fft(yp, yn, xp, xn, lgN, &g_moduli.fft62_mod[i], 1);
}

/* fft62_ifft (src/include/fft62.h:221-226) for prime i: outputs N*a. */
EXPORT void zzref_ifft(uint64_t* yp, size_t yn, uint64_t* xp, unsigned lgN, unsigned i)
{
  zzref_setup();
  fft62_ifft(yp, yn, xp, lgN, &g_moduli.fft62_mod[i], 1);
}

/* zz_crt (src/lib/crt.c:11-94): residues prime-major (num_primes x un) in
   [0, 4p); output un x num_primes limbs. */
EXPORT void zzref_crt(uint64_t* rp, const uint64_t* res, size_t un, unsigned num_primes)
{
  uint64_t* up[ZZ_MAX_PRIMES];
  zzref_setup();
  for (unsigned j = 0; j < num_primes; j++) up[j] = (uint64_t*)res + j * un;
  zz_crt((mp_limb_t*)rp, up, un, &g_moduli, num_primes);
}

/* zz_recompose (src/lib/recompose.c:11-107). */
EXPORT void zzref_recompose(uint64_t* rp, size_t rn, uint64_t* up, size_t un, unsigned s, unsigned r,
                            unsigned r0)
{
  zz_recompose((mp_limb_t*)rp, rn, (mp_limb_t*)up, un, s, r, r0);
}

/* zz_crt_recompose (src/lib/crt_recompose.c:15-156). */
EXPORT void zzref_crt_recompose(uint64_t* rp, size_t rn, unsigned r, const uint64_t* res, size_t un,
                                unsigned num_primes)
{
  uint64_t* up[ZZ_MAX_PRIMES];
  zzref_setup();
  for (unsigned j = 0; j < num_primes; j++) up[j] = (uint64_t*)res + j * un;
  zz_crt_recompose((mp_limb_t*)rp, rn, r, up, un, &g_moduli, num_primes, 1);
}

static size_t max_size(mpz_t* X, int n)
{
  size_t m = 0;
  for (int i = 0; i < n; i++)
    if (mpz_size(X[i]) > m) m = mpz_size(X[i]);
  return m;
}

/* The reference CPU product C (r x d) = A (r x d) * B (d x d), following
   src/lib/hwmpz.c:54-130 call for call. */
EXPORT void zzref_rmatrix_mult_fft(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d)
{
  zz_mpnfft_params_t params;
  zzref_setup();
  size_t ma = max_size(A, r * d), mb = max_size(B, d * d);
  size_t bits = 64 * (2 * (ma > mb ? ma : mb) + 1); /* src/lib/hwmpz.c:54-63 */
  zz_mpnfft_params_init(&params, bits, d, ZZREF_PRIMES, &g_moduli);

  zz_mpnfft_poly_t* AT = malloc(sizeof(zz_mpnfft_poly_t) * r * d);
  zz_mpnfft_poly_t* BT = malloc(sizeof(zz_mpnfft_poly_t) * d * d);
  for (int i = 0; i < r * d; i++) zz_mpnfft_poly_init(AT[i], &params);
  for (int i = 0; i < d * d; i++) zz_mpnfft_poly_init(BT[i], &params);

  for (int i = 0; i < r * d; i++)
    zz_mpnfft_mpn_to_poly(AT[i], A[i]->_mp_d, mpz_size(A[i]), mpz_sgn(A[i]), 0, 0, 1);
  for (int i = 0; i < d * d; i++)
    zz_mpnfft_mpn_to_poly(BT[i], B[i]->_mp_d, mpz_size(B[i]), mpz_sgn(B[i]), 0, 0, 1);

  for (int i = 0; i < r * d; i++) zz_mpnfft_poly_fft(AT[i], AT[i], 1);
  for (int i = 0; i < d * d; i++) zz_mpnfft_poly_fft(BT[i], BT[i], 1);

  zz_mpnfft_poly_matrix_mul(AT, AT, BT, (unsigned)r, (unsigned)d, (unsigned)d, 1);
  for (int i = 0; i < d * d; i++) zz_mpnfft_poly_clear(BT[i]);
  free(BT);

  for (int i = 0; i < r * d; i++) zz_mpnfft_poly_ifft(AT[i], AT[i], 1, 1);

  for (int i = 0; i < r * d; i++) {
    if (AT[i]->size == 0) {
      mpz_set_ui(C[i], 0);
    } else {
      size_t n = ((AT[i]->size - 1) * params.r + 62 * params.num_primes + 2) / 64 + 1;
      mpz_realloc(C[i], n);
      zz_mpnfft_poly_to_mpn(C[i]->_mp_d, n, AT[i], 1);
      int neg = 0;
      if (C[i]->_mp_d[n - 1] >> 63) {
        mpn_neg(C[i]->_mp_d, C[i]->_mp_d, n);
        neg = 1;
      }
      while (n > 0 && C[i]->_mp_d[n - 1] == 0) n--;
      C[i]->_mp_size = neg ? -(int)n : (int)n;
    }
    zz_mpnfft_poly_clear(AT[i]);
  }
  free(AT);
}

/* Same product on plain limb arrays, for callers without mpz_t (ctypes).
   Entry e of A is limbs a + aoff[e], length an[e], sign as[e] (-1, 0, 1).
   Output entry e goes to c + coff[e] with capacity ccap[e]; its signed size
   (like _mp_size) is written to csize[e].  Returns 0, or -1 if an output
   capacity is too small. */
EXPORT int zzref_rmatrix_mult_limbs(uint64_t* c, const size_t* coff, const size_t* ccap, long* csize,
                                    const uint64_t* a, const size_t* aoff, const size_t* an,
                                    const int* as, const uint64_t* bb, const size_t* boff,
                                    const size_t* bn, const int* bs, int r, int d)
{
  mpz_t* A = malloc(sizeof(mpz_t) * r * d);
  mpz_t* B = malloc(sizeof(mpz_t) * d * d);
  mpz_t* C = malloc(sizeof(mpz_t) * r * d);
  int rc = 0;
  for (int i = 0; i < r * d; i++) {
    mpz_init(A[i]);
    mpz_init(C[i]);
    mpz_import(A[i], an[i], -1, 8, 0, 0, a + aoff[i]);
    if (as[i] < 0) mpz_neg(A[i], A[i]);
  }
  for (int i = 0; i < d * d; i++) {
    mpz_init(B[i]);
    mpz_import(B[i], bn[i], -1, 8, 0, 0, bb + boff[i]);
    if (bs[i] < 0) mpz_neg(B[i], B[i]);
  }
  zzref_rmatrix_mult_fft(C, A, r, B, d);
  for (int i = 0; i < r * d; i++) {
    size_t n = mpz_size(C[i]);
    if (n > ccap[i]) {
      rc = -1;
    } else {
      memcpy(c + coff[i], C[i]->_mp_d, n * 8);
      csize[i] = mpz_sgn(C[i]) < 0 ? -(long)n : (long)n;
    }
    mpz_clear(C[i]);
    mpz_clear(A[i]);
  }
  for (int i = 0; i < d * d; i++) mpz_clear(B[i]);
  free(A);
  free(B);
  free(C);
  return rc;
}
