/*
 * rf_oracle.h -- CPU restatement of the reference's exact matrix product
 * mpz_rmatrix_mult_fft (src/lib/hwmpz.c:54-130) and of each of its stages.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load liboracle.so, and only as a checker.
 * The product library never links or calls it.
 *
 * Every function is generic in the prime set (any odd primes p < 2^62 with
 * 2^s | p-1), so the same code states the reference's parameterisation
 * (four 62-bit primes, src/include/fft62.h:24-31) and the GPU library's.
 * Values are canonical ([0, p) residues, two's complement limbs); the
 * reference's lazy representatives ([0, 2p), [0, 4p)) reduce to these.
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXP 64

/* ceil(log2(n)), 0 for n == 0 (src/lib/fft62.c:26-35). */
unsigned orc_log2(size_t n);

/* Smallest admissible truncated-transform size >= n for length 2^lgN
   (src/lib/fft62.c:149-168, array threshold 16 from src/include/fft62.h:61). */
size_t orc_next_size(size_t n, unsigned lgN);

/* The reference's parameter choice (src/lib/mpnfft.c:18-78; gpu != 0 gives
   the points = N variant of src/lib/mpnfft.c:664-708). */
void orc_ref_params(size_t bits, size_t terms, unsigned k, const uint64_t* p, int gpu,
                    unsigned* r, unsigned* lgN, size_t* points);

/* x^e mod p, x * y mod p, 1/x mod p (p prime). */
uint64_t orc_powmod(uint64_t x, uint64_t e, uint64_t p);
uint64_t orc_mulmod(uint64_t x, uint64_t y, uint64_t p);
uint64_t orc_invmod(uint64_t x, uint64_t p);

/* Root of order 2^lgN: g = smallest u >= 2 whose power u^((p-1)/2^s) has
   order exactly 2^s, then squared s - lgN times (src/lib/fft62.c:40-50,
   70-82 with s = FFT62_MAX_LGN = 44). */
uint64_t orc_root(uint64_t p, unsigned s, unsigned lgN);

/* Chunk i = bits [i*b, (i+1)*b) of {up, un}, t = ceil(b/64) limbs each,
   rn chunks (src/lib/split.c:10-81). */
void orc_split(uint64_t* rp, size_t rn, const uint64_t* up, size_t un, unsigned b);

/* out[i] = sign * chunk_i mod p in [0, p) (src/lib/reduce.c:10-28,
   src/lib/split_reduce.c:27-38 with lgS = 0). */
void orc_reduce(uint64_t* out, const uint64_t* chunks, size_t rn, unsigned t, int sign, uint64_t p);

/* y_j = sum_{i < xn} x_i w^(i * rev_lgN(j)) mod p for j < yn; w has order
   2^lgN (the transform of src/include/fft62.h:35-54 with Z = 1). */
void orc_ntt_fwd(uint64_t* y, size_t yn, const uint64_t* x, size_t xn, unsigned lgN, uint64_t p,
                 uint64_t w);

/* Inverse of orc_ntt_fwd on a full bit-reversed spectrum y[0..N): writes
   N * a_i (scale == 0, as fft62_ifft, src/include/fft62.h:185-226) or a_i
   (scale != 0, as zz_mpnfft_poly_ifft with scale, src/lib/mpnfft.c:418-428). */
void orc_ntt_inv(uint64_t* a, const uint64_t* y, unsigned lgN, uint64_t p, uint64_t w, int scale);

/* Per point h < n: C[i][j] = sum_k A[i][k] * B[k][j] mod p, polys stored
   entry-major with stride n (src/lib/mpnfft.c:559-651). */
void orc_pointwise(uint64_t* C, const uint64_t* A, const uint64_t* B, unsigned r, unsigned d,
                   size_t n, uint64_t p);

/* Garner CRT of k residue arrays (prime-major, res[j*un + i], any value,
   reduced mod p_j first) to s-limb two's complement in (-P/2, P/2),
   rp[i*s + l] (src/lib/crt.c:11-94, with s = k there). */
void orc_crt(uint64_t* rp, unsigned s, const uint64_t* res, size_t un, unsigned k,
             const uint64_t* p);

/* rp = sum_i d_i 2^(i*b) mod 2^(64*rn), d_i = s-limb two's complement at
   up + i*s (src/lib/recompose.c:11-107 with r0 = 0). */
void orc_recompose(uint64_t* rp, size_t rn, const uint64_t* up, size_t un, unsigned s, unsigned b);

/* One plan: chunk width b, transform length 2^lgN, k primes with roots of
   order 2^lgN (w[j]). */
typedef struct {
  unsigned b, lgN, k;
  uint64_t p[ORC_MAXP];
  uint64_t w[ORC_MAXP];
} orc_plan;

/* Fills the reference's plan (4 primes of src/include/fft62.h:24-27,
   parameters of src/lib/mpnfft.c:18-78 with bits from
   src/lib/hwmpz.c:54-63); returns the reference's output limb count rule
   input "points" in *points. */
void orc_ref_plan(orc_plan* pl, size_t max_limbs, unsigned d, size_t* points);

/* Exact product C (r x d) = A (r x d) * B (d x d) through the six stages
   with plan pl (src/lib/hwmpz.c:65-130).  Entry e of A: limbs a + aoff[e],
   an[e] limbs (magnitude), sign as[e].  Output entry e: magnitude limbs at
   c + coff[e] (capacity ccap[e]), signed size in csize[e] (as _mp_size).
   Returns 0; -1 if the plan's bound does not hold for these inputs; -2 if an
   output capacity is too small. */
int orc_rmatrix_mult(uint64_t* c, const size_t* coff, const size_t* ccap, long* csize,
                     const uint64_t* a, const size_t* aoff, const size_t* an, const int* as,
                     const uint64_t* bb, const size_t* boff, const size_t* bn, const int* bs,
                     unsigned r, unsigned d, const orc_plan* pl);

/* Reference's own output limb count for one entry of points coefficients
   (src/lib/hwmpz.c:113): ((points-1)*b + 62*k + 2)/64 + 1. */
size_t orc_ref_out_limbs(size_t points, unsigned b, unsigned k);

#ifdef __cplusplus
}
#endif
#endif
