/* test_dropin.c -- a C caller of librforest_hip.so through the drop-in
 * interface alone, the way rforest calls the reference's library: setup, a
 * sweep of seeded signed products through mpz_rmatrix_mult_fft and
 * cu_mpz_rmatrix_mult_fft, each compared with schoolbook GMP
 * (mpz_mul / mpz_addmul), then clear.  The reference's own end-to-end
 * harness (tests/test_gpu_matmul.c:18-89) checks its GPU product against its
 * naive product the same way, on unsigned 50 kbit 2x2 inputs with a time
 * seed; this one is seeded, signed, and covers several shapes and sizes.
 *
 * Build (C11, against include/rforest_hip.h only):
 *   gcc -std=gnu11 -O2 -Iinclude -I<gmp include> tests/c/test_dropin.c \
 *       -Lrforest_amd -lrforest_hip -l:libgmp.so.10 -o tests/c/test_dropin
 * With -DRF_WITH_REFERENCE_HEADER and the reference's src/include on the
 * include path, the reference's hwmpz.h is included beside rforest_hip.h:
 * both declare the drop-in functions, and C accepts the pair only if the
 * declarations agree.
 *
 * Usage: test_dropin [seed]   (exit status 0: every product exact)
 */
#include <stdio.h>
#include <stdlib.h>

#include <gmp.h>

#include "rforest_hip.h"
#ifdef RF_WITH_REFERENCE_HEADER
#include "hwmpz.h"
#endif

static void schoolbook(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d) {
  mpz_t w;
  mpz_init(w);
  for (int i = 0; i < r; i++)
    for (int j = 0; j < d; j++) {
      mpz_mul(w, A[i * d], B[j]);
      for (int k = 1; k < d; k++) mpz_addmul(w, A[i * d + k], B[k * d + j]);
      mpz_set(C[i * d + j], w);
    }
  mpz_clear(w);
}

/* a signed entry of up to `bits` bits (top bit set), or zero now and then */
static void entry(mpz_t x, gmp_randstate_t rs, unsigned long bits) {
  if (gmp_urandomm_ui(rs, 16) == 0) {
    mpz_set_ui(x, 0);
    return;
  }
  mpz_urandomb(x, rs, bits);
  mpz_setbit(x, bits - 1);
  if (gmp_urandomb_ui(rs, 1)) mpz_neg(x, x);
}

static mpz_t* vec(int n) {
  mpz_t* v = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)n);
  for (int i = 0; i < n; i++) mpz_init(v[i]);
  return v;
}

static void vec_free(mpz_t* v, int n) {
  for (int i = 0; i < n; i++) mpz_clear(v[i]);
  free(v);
}

int main(int argc, char** argv) {
  const unsigned long seed = argc > 1 ? strtoul(argv[1], NULL, 10) : 2024;
  static const struct {
    int r, d;
    unsigned long bits;
  } shapes[] = {{2, 2, 50000}, {2, 2, 4096}, {1, 3, 100000}, {3, 3, 20000}, {4, 4, 1048576},
                {2, 2, 1048576}, {3, 2, 777}, {1, 1, 64}, {4, 4, 3000000}};
  gmp_randstate_t rs;
  gmp_randinit_default(rs);
  gmp_randseed_ui(rs, seed);
  hw_mpz_setup();
  int failures = 0;
  for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; s++) {
    const int r = shapes[s].r, d = shapes[s].d;
    mpz_t *A = vec(r * d), *B = vec(d * d), *C = vec(r * d), *G = vec(r * d), *Want = vec(r * d);
    mpz_t w;
    mpz_init(w);
    for (int e = 0; e < r * d; e++) entry(A[e], rs, shapes[s].bits);
    for (int e = 0; e < d * d; e++) entry(B[e], rs, shapes[s].bits);
    schoolbook(Want, A, r, B, d);
    mpz_rmatrix_mult_fft(C, A, r, B, d, w);
    cu_mpz_rmatrix_mult_fft(G, A, r, B, d, w);
    int bad = 0;
    for (int e = 0; e < r * d; e++) bad += mpz_cmp(C[e], Want[e]) != 0 || mpz_cmp(G[e], Want[e]) != 0;
    const int suitable = mpz_rmatrix_gpu_fft_suitable(A, r, B, d);
    printf("%s %dx%d . %dx%d, %lu-bit signed entries (gpu suitable %d)\n", bad ? "FAIL" : "ok", r, d, d, d,
           shapes[s].bits, suitable);
    fflush(stdout);
    failures += bad + (suitable != 1);
    mpz_clear(w);
    vec_free(A, r * d);
    vec_free(B, d * d);
    vec_free(C, r * d);
    vec_free(G, r * d);
    vec_free(Want, r * d);
  }
  hw_mpz_clear();
  gmp_randclear(rs);
  printf("%s: %d failures\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
