// rf_util.h -- synthetic limb generation and result checksums (definitions
// shared by the device kernels and the host code that restates them).
//
// Synthetic entry e of a request with seed s, b bits (b >= 1):
//   key   = mix64(s ^ ((e + 1) * 0xD1B54A32D192ED03))
//   limb i = mix64(key + (i + 1) * 0x9E3779B97F4A7C15), i < n = ceil(b / 64);
//   the top limb keeps its low b - 64 (n - 1) bits and has the highest of
//   them set, so the entry has exactly b bits.
// mix64 is the splitmix64 finaliser (Steele, Lea, Flood 2014).
//
// Checksum of a result: the sum mod 2^64, over C entries e and limbs i of its
// rn-limb two's complement form, of mix64(limb + ((e << 40) + i + 1) * 0x9E3779B97F4A7C15).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RF_UTIL_THREADS 256
#define RF_UTIL_LIMBS 2048  // limbs per block

__host__ __device__ inline uint64_t rf_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rf_gen_key(uint64_t seed, uint32_t entry) {
  return rf_mix64(seed ^ ((uint64_t)(entry + 1) * 0xD1B54A32D192ED03ull));
}
__host__ __device__ inline uint64_t rf_gen_limb(uint64_t key, uint64_t i) {
  return rf_mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull);
}
__host__ __device__ inline uint64_t rf_sum_term(uint64_t limb, uint64_t entry_key, uint64_t i) {
  return rf_mix64(limb + (entry_key + i + 1) * 0x9E3779B97F4A7C15ull);
}

// A span: k_gen writes `bits` bits of entry key `key` at ptr; k_checksum
// sums `bits` limbs (a limb count there) at ptr with entry_key = key
// (= e << 40) into sums[slot].
struct RfSpan {
  uint64_t* ptr;
  uint64_t bits;
  uint64_t key;
  uint32_t slot;
  uint32_t pad;
};
struct RfBlock {
  uint32_t span;
  uint32_t pad;
  uint64_t first;  // first limb of the block
};

// Two's complement -> sign and magnitude, in place, one entry per task; the
// entry's sign and magnitude length (leading zero limbs dropped) go to meta.
// Two launches: a scan per entry (sign, length, lowest nonzero limb), then
// the negation of the negative entries in tiles of RF_FLIP_LIMBS limbs over
// as many workgroups as the longest entry (max_n limbs) needs.
struct RfConvTask {
  uint64_t* limbs;
  uint64_t n;  // limbs
};
struct RfEntryMeta {
  int32_t sign;
  uint32_t nlimbs;
  uint32_t l0;  // lowest nonzero limb of a negative entry
  uint32_t pad;
};
#define RF_FLIP_LIMBS 2048
// A sign written on the device into an uploaded task image: the int32 at
// byte `off` takes meta[src].sign (a product tree's forward tasks take the
// signs of the level below, known only once its conversion has run).
struct RfSignPatch {
  uint64_t off;
  uint32_t src;
  uint32_t pad;
};
hipError_t rf_launch_sign_patch(char* image, const RfSignPatch* patches, int n, const RfEntryMeta* meta,
                                hipStream_t st);
hipError_t rf_launch_signmag(const RfConvTask* tasks, int ntasks, RfEntryMeta* meta, uint64_t max_n, hipStream_t st);

// Remainder fix-up, one task per remainder X mod D (D of n limbs, top limb
// nonzero): given QD = Q D for a quotient estimate with X - Q D in [0, 3D),
// R = (X - QD) mod B^(n+1), then R -= D while R >= D (at most twice), then
// R = D - R if X is negative and R != 0; out = R (n limbs, in [0, D)).
// r and t are scratch of n + 1 limbs each.
struct RfModTask {
  const uint64_t* x;   // |X|
  uint64_t xn;
  const uint64_t* qd;  // Q D (non-negative), qdn limbs (0: Q = 0); with wrap, a sum congruent to Q D mod B^wrap - 1
  uint64_t qdn;
  const uint64_t* d;   // D
  uint64_t n;
  uint64_t* r;         // scratch: n + 1 limbs (wrap limbs with wrap)
  uint64_t* t;         // scratch: as r
  uint64_t* xf;        // scratch with wrap: wrap limbs (X folded)
  uint64_t wrap;       // 0, or W: Q D was formed modulo B^W - 1 (W >= n + 1, xn and qdn <= 2 W)
  uint64_t* out;
  int32_t sign;        // sign of X: -1, 0, +1
  int32_t pad;
};
hipError_t rf_launch_mod_fix(const RfModTask* tasks, int ntasks, hipStream_t st);

// One level of the approximate-reciprocal recursion (Brent & Zimmermann,
// Modern Computer Arithmetic, Algorithm 3.5) for a normalised A of n = l + h
// limbs, given X_h ~ B^(2h) / A_h for its top h limbs (h + 1 limbs) and
// T = A X_h (the first product, tn limbs):
//   k_recip_adjust: while T >= B^(n+h): X_h -= 1, T -= A; then
//                   Tm = floor((B^(n+h) - T) / B^l)  (h + 1 limbs)
//   k_recip_combine (after the second product U = Tm X_h, un limbs):
//                   X = X_h B^l + floor(U / B^(2h-l))  (n + 1 limbs)
struct RfRecipTask {
  const uint64_t* a;  // A, n limbs (normalised: top bit set)
  uint64_t* xh;       // X_h, h + 1 limbs (adjusted in place)
  uint64_t* t;        // T = A X_h, tn limbs (overwritten)
  uint64_t* tm;       // Tm, h + 1 limbs
  const uint64_t* u;  // U = Tm X_h, un limbs
  uint64_t* x;        // X, n + 1 limbs
  uint64_t n, h, l, tn, un;
};
hipError_t rf_launch_recip_adjust(const RfRecipTask* tasks, int ntasks, hipStream_t st);
hipError_t rf_launch_recip_combine(const RfRecipTask* tasks, int ntasks, hipStream_t st);
// I = floor(X 2^s / B^o) into out (il limbs): X of xn limbs, 0 <= s < 64, o >= 1.
struct RfRecipFinal {
  const uint64_t* x;
  uint64_t xn;
  uint64_t* out;
  uint64_t il;
  uint32_t s, o;
};
hipError_t rf_launch_recip_final(const RfRecipFinal* tasks, int ntasks, uint64_t max_il, hipStream_t st);
// D'' = (D 2^s) B^(np - n) into out (np limbs): D of n <= np limbs, 0 <= s < 64
// with (D 2^s) < B^n.
struct RfNormTask {
  const uint64_t* d;
  uint64_t n;
  uint64_t* out;
  uint64_t np;
  uint32_t s, pad;
};
hipError_t rf_launch_mod_norm(const RfNormTask* tasks, int ntasks, uint64_t max_np, hipStream_t st);

hipError_t rf_launch_gen(const RfSpan* spans, const RfBlock* blocks, int nblocks, hipStream_t st);
hipError_t rf_launch_checksum(const RfSpan* spans, const RfBlock* blocks, int nblocks, uint64_t* sums,
                              hipStream_t st);
hipError_t rf_launch_fill(uint64_t* dst, uint64_t n, uint64_t v, hipStream_t st);
