/*
 * rforest_hip.h -- C ABI of the MI355X (gfx950) exact big-integer matrix
 * product library librforest_hip.so.
 *
 * Drop-in section: the same prototypes and semantics as the reference's
 * src/include/hwmpz.h, so an rforest build links this library instead of the
 * reference's and calls nothing else.  Plain-array and device-resident entry
 * points follow for callers without GMP types (ctypes, JNI, cgo).
 *
 * Results are exact integers, identical to GMP schoolbook mpz_mul/mpz_addmul.
 */
#ifndef RFOREST_HIP_H
#define RFOREST_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <gmp.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Drop-in (replaces src/include/hwmpz.h:31, :74-75, :127-139)
 * ------------------------------------------------------------------------- */

/* src/include/hwmpz.h:31 -- when nonzero, hw_mpz_setup does nothing. */
extern int hw_disable_fft;
/* src/lib/hwmpz.c:16 -- host threads for CPU work; accepted, unused. */
extern int mpzfft_threads;

/* src/include/hwmpz.h:74 (src/lib/hwmpz.c:24-40, :51): one-time setup of the
   prime tables and GPU state; idempotent. */
void hw_mpz_setup(void);
/* src/include/hwmpz.h:75 (src/lib/hwmpz.c:42-52): releases GPU state. */
void hw_mpz_clear(void);

/* src/include/hwmpz.h:129 (src/lib/hwmpz.c:65-130): C (r x d) := A (r x d) *
   B (d x d), all row-major arrays of initialised mpz_t; C entries are
   reallocated as needed.  No aliasing.  w is unused scratch.  Any signs and
   zeros.  On failure prints the reason and aborts (the reference exits). */
void mpz_rmatrix_mult_fft(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d, mpz_t w);

/* src/include/hwmpz.h:134 (src/lib/hwmpz.c:132-247): the reference's GPU
   twin; here the same GPU path as mpz_rmatrix_mult_fft. */
void cu_mpz_rmatrix_mult_fft(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d, mpz_t w);

/* src/include/hwmpz.h:139 (src/lib/hwmpz.c:249-255): 1 if the GPU path can
   take this product, 0 otherwise. */
int mpz_rmatrix_gpu_fft_suitable(mpz_t* A, int r, mpz_t* B, int d);

/* ---------------------------------------------------------------------------
 * Error-returning variants (no reference counterpart: the reference asserts,
 * exits or throws, src/lib/hwmpz.c:69,146; src/lib/cufftwrapper.cu:19)
 * ------------------------------------------------------------------------- */
#define RF_OK 0
#define RF_ERR_NOT_INIT (-1)
#define RF_ERR_UNSUPPORTED (-2) /* no proven transform plan for these sizes */
#define RF_ERR_HIP (-3)         /* a HIP runtime call failed */
#define RF_ERR_ARG (-4)
#define RF_ERR_NOMEM (-5)
#define RF_ERR_CAPACITY (-6)    /* an output buffer is too small */

/* Same as mpz_rmatrix_mult_fft; returns RF_OK or an RF_ERR_* code. */
int rf_mpz_rmatrix_mult(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d);

/* Description of the last error on this process. */
const char* rf_last_error(void);

/* Same product on plain host limb arrays.  Entry e of A: magnitude limbs
   a + aoff[e], an[e] limbs, sign as[e] in {-1, 0, 1}; likewise B.  Output e:
   magnitude written to c + coff[e] (capacity ccap[e] limbs), signed size
   (as _mp_size) to csize[e]. */
int rf_rmatrix_mult_limbs(uint64_t* c, const size_t* coff, const size_t* ccap, long* csize,
                          const uint64_t* a, const size_t* aoff, const size_t* an, const int* as,
                          const uint64_t* b, const size_t* boff, const size_t* bn, const int* bs,
                          int r, int d);

/* ---------------------------------------------------------------------------
 * Device-resident batches (no reference counterpart; a product-tree level
 * runs as one call, src/lib/hwmpz.c:132-247 handles one product per call)
 * ------------------------------------------------------------------------- */

/* One matrix entry in device memory: magnitude limbs (little-endian, no
   requirement that the top limb is nonzero) and sign. */
typedef struct rf_dev_entry {
  const uint64_t* limbs; /* device pointer */
  uint64_t nlimbs;
  int32_t sign;          /* -1, 0, +1 */
  int32_t reserved;
} rf_dev_entry;

/* One output entry in device memory.  The library writes the value as
   two's complement in exactly rf_out_limbs(...) limbs. */
typedef struct rf_dev_out {
  uint64_t* limbs;       /* device pointer */
  uint64_t cap;          /* capacity in limbs */
} rf_dev_out;

typedef struct rf_operand rf_operand;

typedef struct rf_dev_product {
  int32_t r, d;
  const rf_dev_entry* A; /* host array of r*d descriptors (ignored when left is set) */
  const rf_dev_entry* B; /* host array of d*d descriptors */
  const rf_dev_out* C;   /* host array of r*d descriptors */
  const rf_operand* left; /* NULL, or a resident left operand used as A (same r, d) */
} rf_dev_product;

/* Output limbs (two's complement, sign included) written for every C entry of
   a product whose A entries have at most la_limbs limbs and B entries at most
   lb_limbs: ceil((64 (la + lb) + ceil(log2 d) + 1) / 64). */
uint64_t rf_out_limbs(uint64_t la_limbs, uint64_t lb_limbs, int d);

/* Runs all products in one pipeline on the library's stream (stream == NULL)
   or on the given hipStream_t; returns after enqueueing when async != 0,
   otherwise after completion. */
int rf_batch_mult_device(const rf_dev_product* prods, int nprod, void* stream, int async);

/* A prepared batch: the products, their plans, task tables and scratch
   memory fixed once, for callers that run the same level shape repeatedly
   (serving loops, benchmarks, graph replay).  Inputs are read and outputs
   written at the device addresses given at creation, on every launch. */
typedef struct rf_batch rf_batch;
int rf_batch_create(const rf_dev_product* prods, int nprod, rf_batch** out);
/* Enqueues the whole pipeline on stream (NULL: the library's stream) and
   returns; with use_graph != 0 the launch sequence is captured once into a
   hipGraph and replayed.  Completion: rf_device_sync or the stream. */
int rf_batch_launch(rf_batch* b, void* stream, int use_graph);
void rf_batch_destroy(rf_batch* b);

/* ---------------------------------------------------------------------------
 * Resident operands: a left matrix transformed once and kept in the NTT
 * domain in device memory, reused by every product it enters.  The
 * reference's counterpart is an array of mpzfft_t filled by mpzfft_fft and
 * combined by mpzfft_matrix_mul (src/include/mpzfft.h:52-118,
 * src/lib/mpzfft.c:10-23); here the plan (chunk width, transform length,
 * primes) is fixed at creation for right operands of up to b_max_bits bits
 * per entry, and a product whose right operand does not fit that plan is
 * refused with RF_ERR_ARG.
 * ------------------------------------------------------------------------- */
/* From r*d device entries (the limbs are only read during the call). */
int rf_operand_create_device(const rf_dev_entry* A, int r, int d, uint64_t b_max_bits, rf_operand** out);
/* From r*d host mpz_t entries. */
int rf_operand_create(mpz_t* A, int r, int d, uint64_t b_max_bits, rf_operand** out);
/* C (r x d) = left (r x d) * B (d x d) on mpz_t, like mpz_rmatrix_mult_fft. */
int rf_operand_mult(const rf_operand* left, mpz_t* C, mpz_t* B);
/* The fixed plan: chunk width 64*t bits, transform 2^lgN, k primes. */
int rf_operand_plan(const rf_operand* op, int* t, int* lgN, int* k);
void rf_operand_destroy(rf_operand* op);

/* ---------------------------------------------------------------------------
 * Device and pinned memory (src/include/hwmpz.h:38-46 declares the same
 * helpers for the reference's CUDA path; callers of rf_batch_mult_device use
 * them to place entries in HBM without another runtime in the process)
 * ------------------------------------------------------------------------- */
void* cu_malloc_pinned(size_t n);
void cu_free_pinned(void* p);
void* cu_device_malloc(size_t n);
void cu_device_free(void* p);
void cu_memcpy_d2h(void* dst, const void* src, size_t n);
/* Host to device copy; returns RF_OK or RF_ERR_HIP. */
int rf_memcpy_h2d(void* dst, const void* src, size_t n);
/* Device to device copy (a plain copy kernel of known size: n bytes read,
   n bytes written); returns RF_OK or RF_ERR_HIP. */
int rf_memcpy_d2d(void* dst, const void* src, size_t n);
/* Waits for all work the library enqueued (hipDeviceSynchronize). */
int rf_device_sync(void);

/* ---------------------------------------------------------------------------
 * Introspection (no reference counterpart)
 * ------------------------------------------------------------------------- */

/* The plan the library would use: chunk width 64*t bits, transform length
   2^lgN, k primes.  Returns RF_OK or RF_ERR_UNSUPPORTED. */
int rf_plan_query(uint64_t la_bits, uint64_t lb_bits, int r, int d, int* t, int* lgN, int* k);

/* Copies up to max primes of the library's prime set (all 1 mod 2^24,
   between 2^47 and 2^51 / 9); returns how many exist. */
int rf_primes(uint64_t* out, int max);

/* Per-kernel timing: when enabled, every launch of the named stage is
   bracketed by HIP events on the stream it runs on. */
#define RF_STAGE_FWD 0
#define RF_STAGE_INV 1
#define RF_STAGE_CRT 2
#define RF_STAGE_CARRY 3
#define RF_STAGE_APPLY 4
#define RF_NSTAGES 5
void rf_profile_enable(int on);
/* Total milliseconds and launch count per stage since the last reset;
   waits for outstanding events. */
int rf_profile_read(double* ms, uint64_t* launches, int nstages);
void rf_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif
