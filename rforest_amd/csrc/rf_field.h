// rf_field.h -- exact modular arithmetic on fp64 lanes for primes 2^47 < p < 2^51 / 9.
//
// gfx950 issues v_fma_f64, v_mul_f64 and v_rndne_f64 at full VALU rate, so a
// residue of ~48 bits costs about six full-rate instructions per modular
// product (Shoup's precomputed-quotient method carried out with an exact
// FMA product split).  Residues are integers held exactly in doubles, in a
// signed lazy range; every bound below keeps all intermediate integers under
// 2^53 in magnitude, where doubles are exact.
//
// Invariants (proved in the comments of each function):
//   * rf_mulmod(a, w, wq): |a| < 2^53, 0 <= w < p, wq = fl(w / p)
//       -> result == a*w (mod p) exactly, |result| <= p * (1/2 + |a| 2^-52).
//     For |a| <= 2^51 this is |result| <= p.
//   * rf_mulmod_var(a, b): |a|, |b| <= 2^49 -> result == a*b (mod p) exactly,
//       |result| <= p/2 + |a b| 1.51 2^-52.
//   * rf_reduce(x): |x| < 2^53 -> result == x (mod p), |result| <= p/2 + 8.
#pragma once
#include <stdint.h>

// RF_FIELD_HOST compiles the same functions for the host (plain static
// inline, for exhaustive checks against exact integer arithmetic);
// RF_HOST_ONLY leaves them out.
#if defined(RF_FIELD_HOST)
#define RF_DEV static inline
#elif !defined(RF_HOST_ONLY)
#include <hip/hip_runtime.h>
#define RF_DEV __device__ __forceinline__
#endif
#if defined(RF_FIELD_HOST) || !defined(RF_HOST_ONLY)

// x mod p into |x| <= p/2 + 8.
// q = rint(fl(x * pinv)) with pinv = fl(1/p): |fl(x*pinv) - x/p| <= |x/p| 2^-52
// <= 2^-46 for |x| < 2^53, p > 2^47, so |x/p - q| <= 1/2 + 2^-46 and
// |x - q p| <= p/2 + p 2^-46 < p/2 + 8.  The FMA forms x - q*p exactly and
// rounds once; the exact value is an integer below 2^48, hence representable.
RF_DEV double rf_reduce(double x, double p, double pinv) {
  double q = __builtin_rint(x * pinv);
  return __builtin_fma(-q, p, x);
}

// x mod p for an integer x held exactly, |x| < 2^90 (sums of constants
// times words whose low bits are zero can exceed 2^53 and stay exact).
// |fl(x*pinv) - x/p| <= |x/p| 2^-52 < 2^-9 since |x/p| < 2^43, so
// |x - q p| <= p (1/2 + 2^-9) < 2^48: an integer, which the FMA forms exactly.
RF_DEV double rf_reduce_wide(double x, double p, double pinv) {
  double q = __builtin_rint(x * pinv);
  return __builtin_fma(-q, p, x);
}

// a * w mod p.  h + l == a*w exactly (FMA error-free product); with
// q = rint(fl(a * wq)), |a w / p - q| <= 1/2 + |a| 2^-52, so r = a w - q p
// satisfies |r| <= p (1/2 + |a| 2^-52).  fma(-q, p, h) = r - l exactly because
// |l| <= ulp(h)/2 <= 2^48 (|h| < 2^101) and |r - l| < 2^51 is an integer; the
// final r - l + l is an integer below 2^51, so it is exact too.
RF_DEV double rf_mulmod(double a, double w, double wq, double p) {
  double h = a * w;
  double l = __builtin_fma(a, w, -h);
  double q = __builtin_rint(a * wq);
  double r = __builtin_fma(-q, p, h);
  return r + l;
}

// a * b mod p for two data operands.  bq = fl(b * pinv) = (b / p)(1 + e) with
// |e| <= 2^-52 + 2^-106, and fl(a * bq) rounds once more, so
// |fl(a bq) - a b / p| <= |a b / p| 1.51 2^-52 and q = rint(fl(a bq)) leaves
// |r| = |a b - q p| <= p/2 + |a b| 1.51 2^-52.  For |a|, |b| <= 2^49 that is
// below 1.26 p < 2^49, so the exactness argument of rf_mulmod applies
// (|l| <= 2^45, |r - l| < 2^53 an integer); for reduced operands
// |a|, |b| <= p/2 + 8 it is below 0.53 p.
RF_DEV double rf_mulmod_var(double a, double b, double p, double pinv) {
  double bq = b * pinv;
  double h = a * b;
  double l = __builtin_fma(a, b, -h);
  double q = __builtin_rint(a * bq);
  double r = __builtin_fma(-q, p, h);
  return r + l;
}

// Canonical representative in [0, p) of an integer |x| < p, without compares:
// y = fl(x pinv + 2^-50) has floor -1 for x < 0 and 0 for x >= 0.  With
// pinv = (1/p)(1 + d), |d| <= 2^-53, and 2^47 < p < 2^48: for 0 <= x <= p - 1,
// 0 < x pinv + 2^-50 <= 1 - 1/p + 2^-53 + 2^-50 < 1 - 2^-49, which rounds
// below 1; for -(p - 1) <= x <= -1, -1 < x pinv + 2^-50 <= -2^-48 + 2^-50 < 0.
// fma(-q, p, x) is then x + p or x, an integer below 2^48: exact.
RF_DEV double rf_canon1(double x, double p, double pinv) {
  const double q = __builtin_floor(__builtin_fma(x, pinv, 0x1p-50));
  return __builtin_fma(-q, p, x);
}

// Canonical representative in [0, p) of |x| <= 2p.
RF_DEV double rf_canon(double x, double p) {
  x = (x < 0.0) ? x + p : x;
  x = (x < 0.0) ? x + p : x;
  x = (x >= p) ? x - p : x;
  return x;
}
#endif

// ---- host-side exact integer helpers (table generation) ----
typedef unsigned __int128 rf_u128;

static inline uint64_t rf_h_mulmod(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((rf_u128)a * b) % p);
}
static inline uint64_t rf_h_powmod(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  a %= p;
  while (e) {
    if (e & 1) r = rf_h_mulmod(r, a, p);
    a = rf_h_mulmod(a, a, p);
    e >>= 1;
  }
  return r;
}
static inline uint64_t rf_h_invmod(uint64_t a, uint64_t p) { return rf_h_powmod(a, p - 2, p); }
