// rf_plan.h -- prime set, transform plans and device constant tables.
//
// A plan fixes, for one r x d times d x d product, the chunk width b = 64 t
// bits, the transform length N = 2^lgN and the number k of primes.  The
// planner proves for each candidate that the exact convolution coefficients
// fit: with nA, nB chunks per entry, nA + nB - 1 <= N and
// 2 * d * min(nA, nB) * (2^b - 1)^2 < P = p_0 ... p_{k-1}, checked in exact
// multiword integer arithmetic (the reference checks a floating-point
// estimate instead, src/lib/mpnfft.c:44-62).
#pragma once
#include <stddef.h>
#include <stdint.h>

#define RF_MAXP 32        // primes generated at setup
// Upper limit of the primes: 2^51 / 9, so that p 2^-52 < 1/18.  The
// transforms reduce every eighth stage.  From |x| <= p/2 + 8, a stage whose
// twiddle is 1 and is not multiplied (block 0 of a transform's first pass, at
// most four stages, all at the start of a window) at most doubles the bound B;
// any other stage adds at most p (1/2 + B 2^-52).  Within a window values stay
// below 16 (p/2 + 8) (1 + 1/18)^4 + 2.2 p < 12.2 p + 160 < 2^51.5, inside the
// |a| < 2^53 that rf_mulmod and rf_reduce need.
#define RF_PMAX ((((uint64_t)1) << 51) / 9)
#define RF_S 24           // every prime is 1 mod 2^RF_S
#define RF_LGN_MIN 1
#define RF_LGN_LDS 14     // largest one-workgroup transform (2^14 doubles = 128 KiB of LDS)
#define RF_LGN_ROW 13     // longer transforms: columns of 2^(lgN-13) in registers, rows of 2^13 in LDS
#define RF_LGN_MAX 24     // largest transform: 2^11 rows of 2^13 (column passes of <= 16 residues)
#define RF_LGN_TW 16      // twiddle tables made at set-up; longer transforms use tables of 2^RF_LGN_MID
#define RF_LGN_MID 21     // or 2^RF_LGN_MAX, each generated on the first plan that needs it
#define RF_KMAX 24        // most primes a plan may use (the CRT kernel's instantiations)
#define RF_TW_PRIMES RF_KMAX  // primes with twiddle tables
#define RF_TMAX 8         // chunk width up to 8 limbs = 512 bits
#define RF_MAXWORDS (2 * RF_TMAX)  // 32-bit words per chunk
#define RF_MAXS 24        // CRT limbs per coefficient, >= ceil((48 RF_MAXP + 2) / 64)

struct RfPrimes {
  int k;
  uint64_t p[RF_MAXP];
  uint64_t g[RF_MAXP];  // root of order 2^RF_S: smallest u >= 2 with u^((p-1)/2^S) of full order
};

struct RfPlan {
  int t;        // limbs per chunk; b = 64 t
  int lgN;      // transform length 2^lgN
  int k;        // primes used (the first k of RfPrimes)
  int s;        // CRT output limbs per coefficient
  int nA, nB;   // chunks per A / B entry (max over the product's entries)
  int nout;     // nA + nB - 1 (0 if either side is zero)
  int M;        // coefficients overlapping one output digit: ceil(s / t)
  size_t rn;    // output limbs per C entry (two's complement, sign included)
  double cost;  // planner's estimate, arbitrary units
  int split;    // 2^14 points as two rows of 2^13 (column + row passes) instead of one workgroup
  int cyclic;   // 1: cyclic convolution of 2^lgN chunks, the product modulo 2^(64 t 2^lgN) - 1
};

// A tuning measurement: a square r x r product whose plan has chunk width
// 64 t and 2^lgN points (with the two-row split of 2^14 points if split)
// took `cost` microseconds per product in a batch of such products.  The
// planner picks, among the valid plans of a product that have measurements,
// the cheapest.
struct RfTuneRule {
  int r, lgN, t, split;
  double cost;
};

// Generates the RF_MAXP largest primes 2^47 < p < RF_PMAX with p = 1 mod 2^RF_S.
void rf_primes_init(RfPrimes* pr);

// Exact test of f * (2^b - 1)^2 < p_0 ... p_{k-1}.
bool rf_bound_holds(const RfPrimes* pr, int k, int b, uint64_t f);

// floor(log2(p_0 ... p_{k-1})).
int rf_floor_log2_P(const RfPrimes* pr, int k);

// Chooses the cheapest proven plan for an r x d times d x d product whose A
// entries have at most LA bits and B entries at most LB bits.  Returns 0, or
// -1 if no plan with a transform of at most 2^RF_LGN_MAX points exists.
int rf_plan_choose(const RfPrimes* pr, size_t LA, size_t LB, int r, int d, RfPlan* out);

// Fills a plan for a fixed (t, lgN); returns 0 or -1 if it cannot hold.
int rf_plan_fixed(const RfPrimes* pr, size_t LA, size_t LB, int r, int d, int t, int lgN,
                  RfPlan* out);
// The same with at least kmin primes (a resident operand's spectra fix k).
int rf_plan_fixed_k(const RfPrimes* pr, size_t LA, size_t LB, int r, int d, int t, int lgN, int kmin,
                    RfPlan* out);

// Plan of the general transform API (values in the transform domain, the
// reference's mpzfft_params_init, src/lib/mpnfft.c:18-78): any sum of at
// most `terms` products x y with |x1 y1| + ... < 2^bits is recovered
// exactly.  nA = nB = nout = the point count, rn = ceil(bits / 64) + 1.
int rf_plan_general(const RfPrimes* pr, size_t bits, unsigned terms, RfPlan* out);

// Plan for A B modulo B^W - 1 (B = 2^64) with W = t 2^lgN >= min_limbs, for
// non-negative operands of at most a_limbs and b_limbs limbs (each <= W): a
// cyclic convolution of the operands' 2^lgN chunks, no zero padding.  Every
// coefficient sums at most min(nA, nB) products, the acyclic bound; the
// recomposed sum has rn = W + t + 1 limbs, to be folded modulo B^W - 1.
// Returns 0, or -1 if no chunk width fits.
int rf_plan_cyclic(const RfPrimes* pr, uint64_t min_limbs, uint64_t a_limbs, uint64_t b_limbs, RfPlan* out);

// ---- constant tables shared by all kernels (one copy in device memory) ----
struct RfConsts {
  double p[RF_MAXP];
  double pinv[RF_MAXP];
  // Garner: pm[j][l] = (p_l mod p_j, fl(that / p_j)) for l < j;
  // ci[j] = ((p_0...p_{j-1})^-1 mod p_j, fl(. / p_j)).
  double pm[RF_MAXP][RF_MAXP][2];
  double ci[RF_MAXP][2];
  uint64_t pu[RF_MAXP];                  // the primes as integers
  uint64_t P[RF_MAXP + 1][RF_MAXS];      // P_k = p_0 ... p_{k-1}
  // pack: pk[q][v][w] = the three 16-bit pieces k0, k1, k2 of
  // K = 2^(32 w) * S_v mod p_q = k0 + k1 2^16 + k2 2^32, held as k0, k1 2^16
  // and k2 2^32, where S_0 = 1 and S_v = 2^-v mod p_q for v = lgN (B operands
  // absorb 1/N).
  double pk[RF_MAXP][RF_LGN_MAX + 1][RF_MAXWORDS][3];
};

// Twiddle tables, per prime, as (w, fl(w/p)) pairs, for a maximum length
// 2^L (L = RF_LGN_TW for the tables made at set-up, RF_LGN_MID and
// RF_LGN_MAX for the ones made when a plan first needs them):
//   fwd[q][m] = omega^(rev_{L-1}(m)), m < 2^(L-1), omega of order 2^L; the
//     prefix of length N/2 serves length N, so both tables agree where both
//     exist.
//   inv[q][len + j] = omega_{2 len}^(-j), len = 2^e < 2^L, j < len.
#define RF_FWD_TW (1 << (RF_LGN_TW - 1))
#define RF_INV_TW (1 << RF_LGN_TW)
#define RF_FWD_TW_MID (1 << (RF_LGN_MID - 1))
#define RF_INV_TW_MID (1 << RF_LGN_MID)
#define RF_FWD_TW_LONG (1 << (RF_LGN_MAX - 1))
#define RF_INV_TW_LONG (1 << RF_LGN_MAX)

void rf_consts_init(const RfPrimes* pr, RfConsts* c);
// omega of order 2^lgmax and its inverse, per prime, as exact doubles (the
// device generates the tables from them).
void rf_twiddle_roots(const RfPrimes* pr, int lgmax, double* om, double* omi);
