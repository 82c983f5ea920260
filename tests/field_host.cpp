// field_host.cpp -- the fp64 modular primitives of rforest_amd/csrc/rf_field.h
// compiled for the host (RF_FIELD_HOST), with checkers that compare them
// against exact 128-bit integer arithmetic at and around their stated bounds.
// Built by tests/test_field.py with g++ -O2 -ffp-contract=off (the library
// builds its kernels with the same flag: every FMA is explicit).
#define RF_FIELD_HOST 1
#include <math.h>
#include <stdint.h>

#include "../rforest_amd/csrc/rf_field.h"

typedef __int128 i128;

namespace {

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform integer in [-m, m] (m < 2^62)
int64_t sym(uint64_t& s, int64_t m) { return (int64_t)(splitmix(s) % (uint64_t)(2 * m + 1)) - m; }

bool is_int(double x) { return x == floor(x); }

i128 mod_p(i128 x, i128 p) {
  i128 r = x % p;
  return r < 0 ? r + p : r;
}

// An integer-valued double (|x| < 2^127) as an exact i128.
i128 to_i128(double x) {
  if (x == 0) return 0;
  int e;
  const double f = frexp(x, &e);          // x = f 2^e, 1/2 <= |f| < 1
  const int64_t m = (int64_t)ldexp(f, 53);  // exact, |m| < 2^53
  const int sh = e - 53;
  return sh >= 0 ? (i128)m * ((i128)1 << sh) : (i128)(m / ((int64_t)1 << -sh));  // exact: x is an integer
}

// Planted mistakes (which = 1), so that a test can show each checker
// rejects a plausible slip: the product without its low half, the quotient
// rounded down, the reduction rounded down, the canonicalisation offset
// 2^-45 instead of 2^-50.
double mut_mulmod(double a, double w, double wq, double p) {
  double h = a * w;
  double q = __builtin_rint(a * wq);
  return __builtin_fma(-q, p, h);
}
double mut_mulmod_var(double a, double b, double p, double pinv) {
  double h = a * b;
  double l = __builtin_fma(a, b, -h);
  double q = __builtin_floor(a * (b * pinv));
  return __builtin_fma(-q, p, h) + l;
}
double mut_reduce(double x, double p, double pinv) { return __builtin_fma(-__builtin_floor(x * pinv), p, x); }
double mut_canon1(double x, double p, double pinv) {
  const double q = __builtin_floor(__builtin_fma(x, pinv, 0x1p-45));
  return __builtin_fma(-q, p, x);
}

double f_mulmod(int which, double a, double w, double wq, double p) {
  return which ? mut_mulmod(a, w, wq, p) : rf_mulmod(a, w, wq, p);
}
double f_mulmod_var(int which, double a, double b, double p, double pinv) {
  return which ? mut_mulmod_var(a, b, p, pinv) : rf_mulmod_var(a, b, p, pinv);
}
double f_reduce(int which, double x, double p, double pinv) {
  return which ? mut_reduce(x, p, pinv) : rf_reduce(x, p, pinv);
}
double f_canon1(int which, double x, double p, double pinv) {
  return which ? mut_canon1(x, p, pinv) : rf_canon1(x, p, pinv);
}

}  // namespace

extern "C" {

double fh_mulmod(double a, double w, double wq, double p) { return rf_mulmod(a, w, wq, p); }
double fh_mulmod_var(double a, double b, double p, double pinv) { return rf_mulmod_var(a, b, p, pinv); }
double fh_reduce(double x, double p, double pinv) { return rf_reduce(x, p, pinv); }
double fh_reduce_wide(double x, double p, double pinv) { return rf_reduce_wide(x, p, pinv); }
double fh_canon1(double x, double p, double pinv) { return rf_canon1(x, p, pinv); }

// Checkers: `which` 0 runs the library's function, 1 its planted mistake.
//
// rf_mulmod(a, w, fl(w/p), p) on n inputs: |a| <= amax (amax = 2^53 - 1 at
// the stated bound), w in [0, p), a fifth of the inputs at the edges
// (|a| = amax, amax - 1; w in {0, 1, p - 1, p - 2, (p - 1) / 2}).  A result
// fails if it is not an integer, not congruent to a w, or above the bound
// p (1/2 + |a| 2^-52).  *worst: the largest |result| / bound seen.
long fh_check_mulmod(int which, uint64_t p_, int64_t amax, long n, uint64_t seed, double* worst) {
  const double p = (double)p_;
  long bad = 0;
  double wr = 0;
  for (long i = 0; i < n; i++) {
    int64_t a = sym(seed, amax);
    uint64_t w = splitmix(seed) % p_;
    const uint64_t e = splitmix(seed) % 10;
    if (e == 0) a = (splitmix(seed) & 1) ? amax : -amax;
    if (e == 1) a = (splitmix(seed) & 1) ? amax - 1 : -(amax - 1);
    if (e == 2) {
      const uint64_t ws[5] = {0, 1, p_ - 1, p_ - 2, (p_ - 1) / 2};
      w = ws[splitmix(seed) % 5];
    }
    const double ad = (double)a, wd = (double)w;
    const double r = f_mulmod(which, ad, wd, wd / p, p);
    const double bound = p * (0.5 + fabs(ad) * 0x1p-52);
    // bound in exact integers: 2^53 |r| <= 2^52 p + 2 |a| p
    const i128 A = a < 0 ? -(i128)a : (i128)a;
    if (!is_int(r) || fabs(r) >= 0x1p62 ||
        ((i128)1 << 53) * (i128)(int64_t)fabs(r) > ((i128)1 << 52) * (i128)p_ + 2 * A * (i128)p_) {
      bad++;
      continue;
    }
    if (mod_p((i128)(int64_t)r - (i128)a * (i128)w, (i128)p_) != 0) bad++;
    if (fabs(r) / bound > wr) wr = fabs(r) / bound;
  }
  *worst = wr;
  return bad;
}

// rf_mulmod_var(a, b, p, fl(1/p)) on |a|, |b| <= amax (2^49 at the stated
// bound), edges included; bound p/2 + |a b| 1.51 2^-52.
long fh_check_mulmod_var(int which, uint64_t p_, int64_t amax, long n, uint64_t seed, double* worst) {
  const double p = (double)p_, pinv = 1.0 / p;
  long bad = 0;
  double wr = 0;
  for (long i = 0; i < n; i++) {
    int64_t a = sym(seed, amax), b = sym(seed, amax);
    const uint64_t e = splitmix(seed) % 8;
    if (e == 0) a = (splitmix(seed) & 1) ? amax : -amax;
    if (e == 1) b = (splitmix(seed) & 1) ? amax : -amax;
    if (e == 2) a = amax, b = -amax;
    const double r = f_mulmod_var(which, (double)a, (double)b, p, pinv);
    const double bound = p * 0.5 + fabs((double)a) * fabs((double)b) * 1.51 * 0x1p-52;
    // bound in exact integers: 100 2^52 |r| <= 100 2^51 p + 151 |a b|
    const i128 AB = (i128)a * (i128)b < 0 ? -((i128)a * (i128)b) : (i128)a * (i128)b;
    if (!is_int(r) || fabs(r) >= 0x1p55 ||
        100 * ((i128)1 << 52) * (i128)(int64_t)fabs(r) > 100 * ((i128)1 << 51) * (i128)p_ + 151 * AB) {
      bad++;
      continue;
    }
    if (mod_p((i128)(int64_t)r - (i128)a * (i128)b, (i128)p_) != 0) bad++;
    if (fabs(r) / bound > wr) wr = fabs(r) / bound;
  }
  *worst = wr;
  return bad;
}

// rf_reduce on |x| <= xmax (2^53 - 1 at the stated bound): congruent, and
// |result| <= p/2 + 8.
long fh_check_reduce(int which, uint64_t p_, int64_t xmax, long n, uint64_t seed, double* worst) {
  const double p = (double)p_, pinv = 1.0 / p;
  long bad = 0;
  double wr = 0;
  for (long i = 0; i < n; i++) {
    int64_t x = sym(seed, xmax);
    const uint64_t e = splitmix(seed) % 8;
    if (e == 0) x = (splitmix(seed) & 1) ? xmax : -xmax;
    if (e == 1) {  // near odd multiples of p/2, where the rounding of x/p is closest to a tie
      const int64_t m = (int64_t)(splitmix(seed) % (uint64_t)(xmax / (int64_t)p_ + 1));
      x = (int64_t)((2 * m + 1) * (p_ / 2)) + sym(seed, 16);
      if (x > xmax) x = xmax;
      if (splitmix(seed) & 1) x = -x;
    }
    const double r = f_reduce(which, (double)x, p, pinv);
    const double bound = p / 2 + 8;
    if (!is_int(r) || fabs(r) > bound) {
      bad++;
      continue;
    }
    if (mod_p((i128)(int64_t)r - (i128)x, (i128)p_) != 0) bad++;
    if (fabs(r) / bound > wr) wr = fabs(r) / bound;
  }
  *worst = wr;
  return bad;
}

// rf_reduce_wide on exact integers x = m 2^s with |x| < 2^emax (2^90 at the
// stated bound; m a 53-bit integer): congruent, |result| <= p (1/2 + 2^-9).
long fh_check_reduce_wide(uint64_t p_, int emax, long n, uint64_t seed, double* worst) {
  const double p = (double)p_, pinv = 1.0 / p;
  long bad = 0;
  double wr = 0;
  for (long i = 0; i < n; i++) {
    const int s = (int)(splitmix(seed) % (uint64_t)(emax - 53 + 1));
    int64_t m = (int64_t)(splitmix(seed) >> 11);  // below 2^53
    if (splitmix(seed) % 8 == 0) m = ((int64_t)1 << 53) - 1;
    if (splitmix(seed) & 1) m = -m;
    const double x = ldexp((double)m, s);
    const double r = rf_reduce_wide(x, p, pinv);
    const double bound = p * (0.5 + 0x1p-9);
    if (!is_int(r) || fabs(r) > bound) {
      bad++;
      continue;
    }
    const i128 X = to_i128(x);
    if (mod_p((i128)(int64_t)r - X, (i128)p_) != 0) bad++;
    if (fabs(r) / bound > wr) wr = fabs(r) / bound;
  }
  *worst = wr;
  return bad;
}

// rf_canon1 on |x| <= xmax (p - 1 at the stated bound): result in [0, p)
// and congruent to x.
long fh_check_canon1(int which, uint64_t p_, int64_t xmax, long n, uint64_t seed) {
  const double p = (double)p_, pinv = 1.0 / p;
  long bad = 0;
  for (long i = 0; i < n; i++) {
    int64_t x = sym(seed, xmax);
    const uint64_t e = splitmix(seed) % 8;
    if (e == 0) x = (splitmix(seed) & 1) ? xmax : -xmax;
    if (e == 1) x = sym(seed, 2);
    const double r = f_canon1(which, (double)x, p, pinv);
    if (!is_int(r) || r < 0 || r >= p || mod_p((i128)(int64_t)r - (i128)x, (i128)p_) != 0) bad++;
  }
  return bad;
}

}  // extern "C"
