/*
 * rforest_hip_stages.h -- stage entry points of librforest_hip.so.
 *
 * Each function runs exactly one stage of the product pipeline through the
 * same kernels mpz_rmatrix_mult_fft uses, on host buffers, so each stage can
 * be checked on its own.  No reference counterpart as separate symbols; the
 * stages correspond to the reference's zz_split_reduce + fft62_fft
 * (src/lib/split_reduce.c:15-66, src/lib/fft62.c:490-615), pointwise product
 * + fft62_ifft (src/lib/mpnfft.c:559-651, src/lib/fft62.c:922-1043) and
 * zz_crt + zz_recompose (src/lib/crt.c:11-94, src/lib/recompose.c:11-107).
 *
 * All residues crossing this interface are canonical, in [0, p).
 */
#ifndef RFOREST_HIP_STAGES_H
#define RFOREST_HIP_STAGES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pack + forward transform of n entries at prime index q: entry e has
   magnitude limbs + off[e], nlimbs[e] limbs, sign signs[e]; chunks of 64*t
   bits, the first nchunk of them (<= 2^lgN) are packed, scaled by 2^-scale
   (scale = 0 or lgN), prime index q < 24 (the primes with twiddle tables),
   1 <= lgN <= 24 (above 2^21 points the first call makes the long twiddle
   tables).  Writes n x 2^lgN spectrum values; position m holds
   sum_i a_i omega^(i rev(m)), omega = rf_root(q, lgN). */
int rf_stage_fwd(const uint64_t* limbs, const size_t* off, const size_t* nlimbs, const int* signs,
                 int n, int t, uint32_t nchunk, int lgN, int q, int scale, uint64_t* out);

/* Pointwise C = A . B (A: r x d spectra, B: d x d spectra, each 2^lgN values,
   row-major) then the inverse transform of each C entry without the 1/N
   factor: writes r*d x 2^lgN values N * c_i. */
int rf_stage_inv(const uint64_t* specA, const uint64_t* specB, int r, int d, int lgN, int q,
                 uint64_t* out);

/* CRT of k residue arrays (prime-major, k x nout, residues of the first k
   primes) into signed coefficients, then their sum at bit offsets 64*t*i,
   written as rn limbs of two's complement. */
int rf_stage_crt(const uint64_t* res, uint32_t nout, int k, int t, uint64_t rn, uint64_t* out);

/* Test hook of the drop-in (mpz_rmatrix_mult_fft, rf_mpz_rmatrix_mult):
   products whose entries exceed `bits` bits are computed from pieces of
   `bits` bits (the path otherwise taken only above the device plans' 2^32-bit
   entries or when device memory runs out); 0 restores the default.  Lets the
   piecewise recombination be checked at small sizes. */
int rf_stage_dropin_piece_bits(uint64_t bits);

/* The reciprocal a remainder modulus keeps (rf_mod_create): for D (dn
   limbs, top limb nonzero) and dividends of at most max_bits bits (m =
   max(ceil(max_bits / 64), dn) limbs), I with
   floor(2^(64 (m+1)) / D) - 2 <= I <= floor(2^(64 (m+1)) / D), computed by the
   device's Newton recursion; written to out (cap limbs), its length
   (m - dn + 2 limbs, high zeros included) to *nlimbs. */
int rf_stage_mod_reciprocal(const uint64_t* d, uint64_t dn, uint64_t max_bits, uint64_t* out, uint64_t cap,
                            uint64_t* nlimbs);

/* The host epilogue of every host entry point: the rn-limb two's complement
   tc as sign and magnitude, the magnitude's limbs (without high zeros) to dst
   (cap limbs of room), its length times the sign to *nsigned; converted in
   pieces over the library's copy threads.  RF_ERR_CAPACITY when the
   magnitude needs more than cap limbs (nothing written then).  Host only:
   no GPU work, usable before hw_mpz_setup. */
int rf_stage_tc_to_signmag(const uint64_t* tc, uint64_t rn, uint64_t* dst, uint64_t cap, long* nsigned);

/* The library's segmented host copy (operands into the staging buffer,
   product-tree leaves and roots): dst[s][0, n[s]) = src[s][0, n[s]) for
   every segment, in 64 KiB pieces of the segments' concatenation over the
   copy threads.  Host only: no GPU work, usable before hw_mpz_setup. */
int rf_stage_copy_segments(void* const* dst, const void* const* src, const uint64_t* n, int nseg);

/* Root of order 2^lgN used for prime index q. */
uint64_t rf_root(int q, int lgN);

#ifdef __cplusplus
}
#endif
#endif
