// rf_kernels.h -- task descriptors and launchers for the product pipeline.
//
// Device data layout (all in one arena per call or batch):
//   limbs   u64   input magnitudes, entry after entry
//   spec    f64   [poly][prime][N] spectra; polys are a product's A entries
//                 (row-major r x d) then its B entries (d x d)
//   coef    f64   [C entry][prime][N] coefficient residues after the inverse
//   out     u64   [C entry][rn] two's complement result limbs
// Residues are integers in doubles, |x| <= p/2 + 8 whenever stored.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_plan.h"

// Global-memory views of device pointers.  Pointers that arrive inside task
// descriptors or table structs are generic to the compiler, and loads through
// them become flat_* instructions, whose completion the LDS wait counter also
// tracks: every wait for an LDS exchange would then wait for outstanding
// twiddle or limb loads too.  Loads through these types are global_*.
// (The host pass of a .hip source parses device code too; it has no such
// address space, and the qualifier applies to the device pass only.)
#if defined(__HIP_DEVICE_COMPILE__)
#define RF_GLOBAL __attribute__((address_space(1)))
#else
#define RF_GLOBAL
#endif
typedef const double2 RF_GLOBAL* rf_gtw;       // twiddle tables (w, w / p)
typedef const uint64_t RF_GLOBAL* rf_glimbs;   // entry magnitudes
typedef const ulonglong2 RF_GLOBAL* rf_glimbs2;
typedef const double RF_GLOBAL* rf_gres;       // stored residues
template <typename T>
__device__ __forceinline__ T rf_ld(const T RF_GLOBAL* p) {
  return *p;
}
// Stores of results written once and read by a later launch (spectra,
// coefficient residues), marked non-temporal: they stream past the L2
// instead of evicting the twiddle tables and limbs the running kernel
// re-reads from it.
typedef double rf_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void rf_st_nt(double* p, double x) {
  __builtin_nontemporal_store(x, (double RF_GLOBAL*)p);
}
__device__ __forceinline__ void rf_st_nt2(double* p, double x, double y) {
  const rf_d2v v = {x, y};
  __builtin_nontemporal_store(v, (rf_d2v RF_GLOBAL*)p);
}

// One forward task: pack chunks of one entry modulo prime q, transform.
struct RfFwdTask {
  const uint64_t* limbs;  // entry magnitude (device pointer)
  uint64_t spec_off;      // spectrum at spec[spec_off ... + N)
  uint32_t nlimbs;    // magnitude length in limbs
  uint32_t nchunk;    // chunks to pack (<= N; the rest is zero)
  int32_t sign;       // -1, 0, +1
  uint16_t q;         // prime index
  uint8_t t;          // limbs per chunk
  uint8_t scale;      // 0: none; v: multiply by 2^-v (B operands, v = lgN)
};

// One inverse task: C[i][j] at prime q = sum_l A[i][l] B[l][j], inverse
// transform, coefficient residues out.
struct RfInvTask {
  uint64_t a_off;     // offset of A[i][0] at prime q from a_base (or from the launch's A base)
  uint64_t b_off;     // spec offset of B[0][j] at prime q
  uint64_t out_off;   // coef offset of C[i][j] at prime q
  uint32_t a_stride;  // spec distance between A[i][l] and A[i][l+1]
  uint32_t b_stride;  // spec distance between B[l][j] and B[l+1][j]
  uint16_t q;
  uint16_t d;         // terms; 0: inverse of the one spectrum at a_off (b unused)
  const double* a_base;  // A spectra (a resident operand's), or nullptr: the launch's A base
};

// Transforms longer than 2^RF_LGN_LDS points, N = 2^(13 + n1), view the
// array as 2^n1 rows of 2^13 (i = row 2^13 + col).  Forward: a column pass
// (pack + the first stages, one column per thread: all n1 for n1 <= 4,
// otherwise the top n1 mod 4 (or 4), then middle passes of 4 row bits each)
// then a row pass (the last 13 stages of each row in LDS).  Inverse: a row pass (pointwise product + the
// first 13 stages per row, an RfInvTask per row) then column passes over the
// row bits (4 at a time).  All passes work in place on the same array.
struct RfRowTask {
  uint64_t off;       // first residue of the row (forward rows) or of the array (inverse columns)
  uint32_t row;       // row index (forward rows; its bits precede the row's own in the twiddle index)
  uint16_t q;
  uint16_t lg;        // log2 of the whole transform's length (forward rows: picks the twiddle table)
};
#define RF_COL_THREADS 256

// One CRT/recompose block of one C entry: RF_CRT_THREADS coefficients, one
// per thread, covering the digits [digit0, digit0 + RF_CRT_THREADS - halo);
// the first `halo` coefficients only feed those digits' overlaps, so
// consecutive blocks recompute halo coefficients instead of any thread
// computing two.
#define RF_CRT_THREADS 256
struct RfCrtTask {
  uint64_t coef_off;  // coef offset of prime 0 of this entry; prime j at + j * N
  uint64_t* out;      // output limbs of this entry (device pointer)
  uint64_t pre_off;   // per-digit prefix carry function codes (bytes)
  uint32_t fun_idx;   // index of this block's aggregate in the block arrays
  uint32_t N;         // transform length (prime stride in coef)
  uint32_t nout;      // number of coefficients
  uint32_t rn;        // output limbs
  uint32_t digit0;    // first digit of the block
  uint8_t t;          // limbs per digit
  uint8_t k;          // primes
  uint8_t s;          // CRT limbs per coefficient
  uint8_t halo;       // coefficients before digit0 the block recomputes
};

// Halo of a CRT block for k primes and t limbs per digit: ceil(S/t) + 2
// coefficients with S = ceil((48 k + 1) / 64) limbs per coefficient.
static inline int rf_crt_halo(int k, int t) {
  const int S = (48 * k + 1 + 63) / 64;
  return (S + t - 1) / t + 2;
}

// A descriptor read by a whole wave: loaded as 32-bit words and broadcast
// from the first lane, so its fields live in scalar registers and branches
// on them stay uniform.  Only for descriptors all lanes of the wave share.
template <typename T>
__device__ __forceinline__ T rf_wave_uniform_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size must be a multiple of 4");
  union {
    T t;
    uint32_t w[sizeof(T) / 4];
  } u;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; i++) u.w[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return u.t;
}

// Device-side pointers to the constant tables.
struct RfDevTables {
  const RfConsts* consts;
  const double* tw_fwd;  // RF_TW_PRIMES * RF_FWD_TW * 2
  const double* tw_inv;  // RF_TW_PRIMES * RF_INV_TW * 2
  const double* tw_fwd_mid;   // RF_TW_PRIMES * RF_FWD_TW_MID * 2 (null until a transform above 2^RF_LGN_TW)
  const double* tw_inv_mid;   // RF_TW_PRIMES * RF_INV_TW_MID * 2
  const double* tw_fwd_long;  // RF_TW_PRIMES * RF_FWD_TW_LONG * 2 (null until a transform above 2^RF_LGN_MID)
  const double* tw_inv_long;  // RF_TW_PRIMES * RF_INV_TW_LONG * 2
};

// Every task of one forward launch has the same chunk width t (the kernels
// for lengths from 2^RF_FWD_TSPEC on are specialised on it).
#define RF_FWD_TSPEC 11
// Fills the twiddle tables (maximum length 2^lgmax) of the first nprimes
// primes (RF_TW_PRIMES at most) on the device: om[q], omi[q] = omega of order
// 2^lgmax and its inverse.
hipError_t rf_launch_twiddles(const RfConsts* consts, const double* om, const double* omi, int nprimes, int lgmax,
                              double* fwd, double* inv, hipStream_t st);
hipError_t rf_launch_fwd(const RfDevTables& tb, int lgN, int t, const RfFwdTask* tasks, int ntasks,
                         double* spec, hipStream_t st);
// a_off of the tasks counts from aspec (the batch's spectra or a resident
// operand's), b_off from bspec.
hipError_t rf_launch_inv(const RfDevTables& tb, int lgN, const RfInvTask* tasks, int ntasks,
                         const double* aspec, const double* bspec, double* coef, hipStream_t st);
// Column and row passes of the longer transforms (lgN in RF_LGN_LDS+1 .. RF_LGN_MAX).
hipError_t rf_launch_fwd_col(const RfDevTables& tb, int lgN, int t, const RfFwdTask* tasks, int ntasks,
                             double* spec, hipStream_t st);
hipError_t rf_launch_fwd_row(const RfDevTables& tb, const RfRowTask* tasks, int ntasks, double* spec,
                             hipStream_t st);
hipError_t rf_launch_inv_col(const RfDevTables& tb, int lgN, const RfRowTask* tasks, int ntasks,
                             double* coef, hipStream_t st);
// Blocks with the same k must be launched together (the kernel is specialised on k).
// Element-wise arithmetic on stored spectra of k primes x 2^lgN residues
// (one value of the general transform API, [prime][N]).  Every stored
// spectrum holds 2^-lgN times the transform of its integer, so a product
// multiplies by 2^lgN once more.
enum RfEwOp { RF_EW_SET, RF_EW_ADD, RF_EW_SUB, RF_EW_NEG, RF_EW_MUL };
hipError_t rf_launch_ew(const RfDevTables& tb, int op, int k, int lgN, const double* a, const double* b,
                        double* out, hipStream_t st);
// out[i][j] = sum_l a[i][l] * b[l][j] per residue, for n1 x n2 times n2 x n3
// matrices of spectra given as device arrays of pointers (row-major).
hipError_t rf_launch_ew_matmul(const RfDevTables& tb, int k, int lgN, const double* const* a,
                               const double* const* b, double* const* out, int n1, int n2, int n3,
                               hipStream_t st);

hipError_t rf_launch_crt(const RfDevTables& tb, int k, const RfCrtTask* tasks, int ntasks,
                         const double* coef, uint8_t* pre, uint8_t* fun, hipStream_t st);
hipError_t rf_launch_apply(const RfCrtTask* tasks, int ntasks, const uint8_t* pre, const uint8_t* fun,
                           hipStream_t st);
