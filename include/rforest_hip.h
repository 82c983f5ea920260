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
   prime tables and GPU state (38 MB of twiddle tables for transforms up to
   2^16 points; longer transforms add 1.2 GB, then 9.7 GB, on the first plan
   that needs them); idempotent. */
void hw_mpz_setup(void);
/* src/include/hwmpz.h:75 (src/lib/hwmpz.c:42-52): releases GPU state. */
void hw_mpz_clear(void);

/* src/include/hwmpz.h:129 (src/lib/hwmpz.c:65-130): C (r x d) := A (r x d) *
   B (d x d), all row-major arrays of initialised mpz_t; C entries are
   reallocated as needed.  No aliasing.  w is unused scratch.  Any signs and
   zeros, and any sizes, like the reference's CPU path: entries beyond one
   device plan (2^32 bits) or beyond device memory are cut into pieces whose
   products run on the device and are recombined on the host.  On failure
   (not set up, a HIP error, host or device memory exhausted even for
   2^16-bit pieces) prints the reason and aborts (the reference asserts or
   exits, src/lib/hwmpz.c:69,146). */
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

/* Same as mpz_rmatrix_mult_fft (any sizes); returns RF_OK or an RF_ERR_* code. */
int rf_mpz_rmatrix_mult(mpz_t* C, mpz_t* A, int r, mpz_t* B, int d);

/* Description of the last error on this process. */
const char* rf_last_error(void);

/* Same product on plain host limb arrays.  Entry e of A: magnitude limbs
   a + aoff[e], an[e] limbs, sign as[e] in {-1, 0, 1}; likewise B.  Output e:
   magnitude written to c + coff[e] (capacity ccap[e] limbs), signed size
   (as _mp_size) to csize[e].  One device plan per call: RF_ERR_UNSUPPORTED
   for entries beyond 2^32 bits (the mpz_t entry points cut those into
   pieces). */
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

/* No counterpart in the reference: for callers that cannot allocate device
   memory (serving loops under a fixed memory budget, harnesses that forbid
   hipMalloc).  The same products as rf_batch_mult_device, with
   the pipeline's scratch (spectra, coefficient residues, task tables, carry
   codes) in a fixed device array of the library (64 MiB per GPU, held from
   module load until the process ends) instead of the engine's arena: the
   call makes no device or pinned allocation of any kind.  Products whose
   scratch exceeds the array, or whose plans would need twiddle tables not
   made yet (those of transforms above 2^16 points are made by the first
   allocating call that plans one), are refused with RF_ERR_CAPACITY.  The array keeps the last call's contents and every call
   in the process on that GPU shares it; calls are serialised by the
   library's lock and each returns after its work completes on `stream`
   (NULL: the library's stream), so they never overlap -- two calls running
   at once on one GPU would race on it. */
int rf_batch_mult_device_noalloc(const rf_dev_product* prods, int nprod, void* stream);

/* A prepared batch: the products, their plans, task tables and scratch
   memory fixed once, for callers that run the same level shape repeatedly
   (serving loops, benchmarks, graph replay).  Inputs are read and outputs
   written at the device addresses given at creation, on every launch. */
typedef struct rf_batch rf_batch;
int rf_batch_create(const rf_dev_product* prods, int nprod, rf_batch** out);
/* The same with the products split into `lanes` (1..4) independent
   pipelines, balanced by planned cost, each on its own stream so that their
   kernels may overlap (rf_batch_create uses one lane). */
int rf_batch_create_lanes(const rf_dev_product* prods, int nprod, int lanes, rf_batch** out);
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
 * Integers held in the transform domain (general transform API).  Mirrors
 * the reference's mpzfft interface (src/include/mpzfft.h:13-120,
 * src/lib/mpzfft.c:10-55): a params object fixes, for an output bound `bits`
 * and `terms` products, the chunk width, transform length and primes
 * (mpzfft_params_init, src/lib/mpnfft.c:18-78: any sum of at most `terms`
 * products x*y with |x1*y1| + ... < 2^bits comes back exactly); an rf_fft
 * value (mpzfft_t) is one integer's spectra, k primes x 2^lgN residues, in
 * device memory.  Values combine by set/add/sub/neg, pointwise products and
 * matrix products without leaving the device, and rf_fft_inverse turns one
 * back into an integer.  Unlike the reference, rf_fft_inverse leaves its
 * operand intact.  Every call is synchronous; results of arithmetic whose
 * inputs exceed the params bound are unspecified, as in the reference.
 * Aliasing of outputs and inputs is allowed everywhere.
 * ------------------------------------------------------------------------- */
typedef struct rf_fft_params rf_fft_params;
typedef struct rf_fft rf_fft;
/* mpzfft_params_init (src/include/mpzfft.h:17-22); RF_ERR_UNSUPPORTED when no
   plan within 2^24 points and 24 primes holds the bound. */
int rf_fft_params_create(uint64_t bits, unsigned terms, rf_fft_params** out);
/* The plan: chunk width 64*t bits, transform 2^lgN, k primes. */
int rf_fft_params_plan(const rf_fft_params* params, int* t, int* lgN, int* k);
void rf_fft_params_destroy(rf_fft_params* params); /* mpzfft_params_clear */
/* mpzfft_init (src/include/mpzfft.h:35-38): a value of zero, device memory
   allocated; it keeps a copy of the plan, so params may be destroyed first. */
int rf_fft_create(const rf_fft_params* params, rf_fft** out);
void rf_fft_destroy(rf_fft* f); /* mpzfft_clear */
/* mpzfft_fft (src/lib/mpzfft.c:10-23): rop := transform of op.  RF_ERR_ARG if
   op has more bits than the params bound. */
int rf_fft_forward(rf_fft* rop, const mpz_t op);
/* The same from device limbs (read during the call only). */
int rf_fft_forward_device(rf_fft* rop, const rf_dev_entry* op);
/* mpzfft_ifft (src/lib/mpzfft.c:26-55): rop := the integer op holds. */
int rf_fft_inverse(mpz_t rop, const rf_fft* op);
/* The same into device memory: magnitude limbs at out (cap limbs), sign and
   limb count to *sign, *nlimbs (host); RF_ERR_CAPACITY if cap is too small
   (rf_fft_out_limbs limbs always suffice). */
int rf_fft_inverse_device(const rf_fft* op, uint64_t* out, uint64_t cap, int* sign, uint64_t* nlimbs);
uint64_t rf_fft_out_limbs(const rf_fft* f);
/* mpzfft_set/add/sub/neg/mul (src/include/mpzfft.h:60-107); all operands
   must come from the same plan (RF_ERR_ARG otherwise). */
int rf_fft_set(rf_fft* rop, const rf_fft* op);
int rf_fft_add(rf_fft* rop, const rf_fft* op1, const rf_fft* op2);
int rf_fft_sub(rf_fft* rop, const rf_fft* op1, const rf_fft* op2);
int rf_fft_neg(rf_fft* rop, const rf_fft* op);
int rf_fft_mul(rf_fft* rop, const rf_fft* op1, const rf_fft* op2);
/* mpzfft_matrix_mul (src/include/mpzfft.h:109-118): rop (dim1 x dim3) :=
   op1 (dim1 x dim2) * op2 (dim2 x dim3), row-major arrays of values. */
int rf_fft_matrix_mul(rf_fft* const* rop, rf_fft* const* op1, rf_fft* const* op2, unsigned dim1, unsigned dim2,
                      unsigned dim3);

/* ---------------------------------------------------------------------------
 * Remainders X mod D (the reference's remainder step, mpzfft_mod_init /
 * mpzfft_mod_mod, src/lib/mpzfft.c:59-100, src/include/mpnfft_mod.h:53-75):
 * a modulus object prepared once for dividends of up to max_bits bits keeps
 * D and I = floor(2^(64(m+1)) / D) (m = limbs of the largest dividend) as
 * resident transforms; each remainder is then two products with them (the
 * quotient estimate from the high limbs of X I, then Q D) and a device
 * fix-up to [0, D).  Unlike the reference, I is exact (GMP division at
 * creation) rather than approximate; the results are the same.
 * ------------------------------------------------------------------------- */
typedef struct rf_mod rf_mod;
/* mpzfft_mod_init: D > 0 (RF_ERR_ARG otherwise), dividends |X| < 2^max_bits. */
int rf_mod_create(uint64_t max_bits, const mpz_t d, rf_mod** out);
/* mpzfft_mod_init for a whole level of moduli (a remainder tree's nodes):
   outs[i] for ds[i] with dividends |X| < 2^max_bits[i].  The reciprocal
   divisions, which dominate a modulus's preparation, run on up to 16 host
   threads.  On an error no modulus is left created and every outs[i] is
   NULL. */
int rf_mod_create_batch(const uint64_t* max_bits, const mpz_t* ds, int cnt, rf_mod** outs);
/* Limbs of D, which every remainder of this modulus is written in. */
uint64_t rf_mod_limbs(const rf_mod* mod);
/* No counterpart in the reference (a query of this implementation's plan):
   W when the product Q D is formed modulo 2^(64 W) - 1 by a cyclic
   convolution (then folded with X; W >= limbs of D + 1), 0 when it is formed
   whole. */
uint64_t rf_mod_wrap_limbs(const rf_mod* mod);
/* mpzfft_mod_mod: rop := op mod D in [0, D) (any sign of op). */
int rf_mod_mod(const rf_mod* mod, mpz_t rop, const mpz_t op);
/* A batch (a remainder-tree level) in one call: outs[i] (device,
   rf_mod_limbs(mods[i]) limbs) := xs[i] mod D_i for device entries xs[i]. */
int rf_mod_mod_device(const rf_mod* const* mods, const rf_dev_entry* xs, int cnt, uint64_t* const* outs);
void rf_mod_destroy(rf_mod* mod); /* mpzfft_mod_clear */

/* ---------------------------------------------------------------------------
 * Product trees kept in device memory (no reference counterpart: rforest's
 * trees call mpz_rmatrix_mult_fft level by level through mpz_t).  M_0 M_1 ...
 * M_{n-1} of n square d x d matrices, level by level: each level's pairs run
 * as one batched pipeline and its results are converted in place from two's
 * complement to sign and magnitude, which is the next level's input format,
 * so only the entries' signs and lengths visit the host between levels.
 * ------------------------------------------------------------------------- */
/* mats: n*d*d device entries (matrix after matrix, row-major).  The root's
   magnitudes go to out + e*out_cap (device), its signs and limb counts to
   sign[e], nlimbs[e] (host); RF_ERR_CAPACITY if an entry exceeds out_cap. */
int rf_tree_product_device(const rf_dev_entry* mats, int n, int d, uint64_t* out, uint64_t out_cap, int* sign,
                           uint64_t* nlimbs);
/* C (d x d) = mats[0..d*d) * mats[d*d..2*d*d) * ... (n matrices of mpz_t). */
int rf_mpz_tree_product(mpz_t* C, mpz_t* mats, int n, int d);

/* ---------------------------------------------------------------------------
 * Request server (no reference counterpart; the reference serves one
 * synchronous call at a time; its transform-domain reuse is mpzfft_fft,
 * src/lib/mpzfft.c:10-23).  Continuous batching: a worker thread admits
 * queued requests in arrival order whenever a stream slot is free and the
 * request's device memory (inputs, outputs and its share of the batch's
 * spectra, coefficients and task tables) fits in the HBM budget, runs each
 * admitted group as one pipeline on its slot's stream, and retires requests
 * as their batch's event completes; batches on different slots overlap on
 * the device.  Each result is the exact product whatever else shares its
 * batch.
 * ------------------------------------------------------------------------- */
#define RF_IN_HOST 0        /* entry limbs are host pointers, copied at admission */
#define RF_IN_SYNTH 1       /* entries generated at admission from the request's seed
                               (the synthetic stream of rf_util.h; load generation) */
#define RF_OUT_KEEP 1       /* keep the result in device memory until rf_server_release */
#define RF_OUT_HOST 2       /* copy the result to out_host (r*d*out_cap limbs, two's complement) */
#define RF_OUT_CHECKSUM 4   /* compute the result's 64-bit checksum (rf_util.h) */

typedef struct rf_serve_entry {
  const uint64_t* limbs; /* RF_IN_HOST: ceil(bits / 64) magnitude limbs; RF_IN_SYNTH: unused */
  uint64_t bits;         /* magnitude bit length (RF_IN_SYNTH: exact) */
  int32_t sign;          /* -1, 0, +1 */
  int32_t pad;
} rf_serve_entry;

typedef struct rf_request {
  int32_t r, d;
  int32_t input;               /* RF_IN_* */
  int32_t output;              /* RF_OUT_* flags */
  const rf_serve_entry* A;     /* r*d, copied at submission (ignored when left is set) */
  const rf_serve_entry* B;     /* d*d */
  uint64_t seed;               /* RF_IN_SYNTH */
  uint64_t* out_host;          /* RF_OUT_HOST */
  uint64_t out_cap;            /* RF_OUT_HOST: limbs per C entry */
  uint64_t tag;                /* returned in the completion */
  /* NULL, or a resident left operand (rf_operand_create*, same r and d) used
     as A: its NTT-domain spectra are read in place by every request naming
     it, so only B is transferred and transformed.  The caller keeps it alive
     until every request naming it has completed; B must fit the operand's
     plan (RF_ERR_ARG at submission otherwise). */
  const rf_operand* left;
} rf_request;

typedef struct rf_completion {
  uint64_t id, tag;
  int32_t status;              /* RF_OK or an RF_ERR_* code */
  int32_t pad;
  uint64_t out_limbs;          /* limbs per C entry of the result */
  uint64_t checksum;           /* RF_OUT_CHECKSUM */
  double t_submit, t_admit, t_done;  /* seconds on the server clock */
} rf_completion;

typedef struct rf_server_config {
  uint64_t hbm_budget;         /* bytes of device memory the server may use */
  int32_t streams;             /* batches in flight at once (default 3) */
  int32_t max_batch;           /* products per batch (default 256) */
  uint64_t max_batch_bytes;    /* device bytes per batch (default budget / streams) */
} rf_server_config;

typedef struct rf_server_stats {
  uint64_t submitted, completed, batches, products_in_batches;
  uint64_t peak_bytes, budget;
} rf_server_stats;

typedef struct rf_server rf_server;
int rf_server_create(const rf_server_config* cfg, rf_server** out);
/* Queues one request (thread-safe, returns at once).  RF_ERR_CAPACITY if the
   request alone cannot fit the budget, RF_ERR_UNSUPPORTED if it has no plan. */
int rf_server_submit(rf_server* s, const rf_request* req, uint64_t* id);
/* Moves up to max completions into out; waits up to wait_ms for the first.
   Returns the number moved, or a negative RF_ERR_* code. */
int rf_server_poll(rf_server* s, rf_completion* out, int max, int wait_ms);
/* Copies a completed RF_OUT_KEEP result (r*d entries of out_limbs limbs,
   two's complement) to host memory; cap is the room in limbs. */
int rf_server_read(rf_server* s, uint64_t id, uint64_t* dst, uint64_t cap);
/* Frees a completed RF_OUT_KEEP result. */
int rf_server_release(rf_server* s, uint64_t id);
int rf_server_stats_get(rf_server* s, rf_server_stats* st);
/* Seconds since the server was created (the clock of the completions). */
double rf_server_clock(rf_server* s);
/* Finishes every queued and running request, then frees the server. */
int rf_server_destroy(rf_server* s);

/* The synthetic stream and checksum of rf_util.h on device memory (for
   checking the server's results against other paths). */
int rf_gen_entry_device(uint64_t* dst, uint64_t bits, uint64_t seed, uint32_t entry);
int rf_checksum_device(const uint64_t* limbs, uint64_t n, uint32_t entry, uint64_t* sum);
/* The same for many entries in one launch: entry e of the synthetic stream
   of request seeds[i] with bits[i] bits into dst[i] (bits 0: nothing);
   checksum shares of count C entries (n[c] limbs at src[c], C-entry index
   entry[c]) summed into sums[slot[c]], nslots sums. */
int rf_gen_entries_device(uint64_t* const* dst, const uint64_t* bits, const uint64_t* seeds,
                          const uint32_t* entry, int n);
int rf_checksums_device(const uint64_t* const* src, const uint64_t* n, const uint32_t* entry, const uint32_t* slot,
                        int count, uint64_t* sums, int nslots);

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

/* Planner tuning (measured by the autotuner, rforest_amd/tune.py): a square
   r x r product planned with chunk width 64*t and 2^lgN points (split != 0:
   2^14 points as two rows of 2^13) took cost_us microseconds per product.
   For square products the planner then picks the measured valid plan of
   least cost, falling back to its cost model when none is measured.
   rf_tune_load reads "r lgN t split cost_us" lines (# comments) and returns
   the number read; hw_mpz_setup loads tuned_plans.txt from the library's
   directory when it exists. */
int rf_tune_set(int r, int lgN, int t, int split, double cost_us);
void rf_tune_clear(void);
int rf_tune_load(const char* path);

/* Copies up to max primes of the library's prime set (all 1 mod 2^24,
   between 2^47 and 2^51 / 9); returns how many exist. */
int rf_primes(uint64_t* out, int max);

/* Per-kernel timing: when enabled, every launch of the named stage is
   bracketed by HIP events on the stream it runs on. */
#define RF_STAGE_FWD 0
#define RF_STAGE_INV 1
#define RF_STAGE_CRT 2
#define RF_STAGE_CARRY 3  /* kept for the slot layout: block carries are now part of RF_STAGE_APPLY */
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
