// rf_crt.hip -- Garner CRT and recomposition of the product pipeline (gfx950).
//
//   k_crt    Garner CRT of each coefficient + digit sums of the overlapping
//            coefficients + block scan of carry transfer functions
//   k_apply  per block: the carry into it from the earlier blocks' aggregates,
//            then the resolved carries added into the output limbs
//
// The carry-propagating sum of src/lib/recompose.c:11-107 and the serial
// fix-up of src/lib/crt_recompose.c:103-153 become a segmented scan: every
// output digit's carry behaviour is a function {-1, 0, 1} -> {-1, 0, 1} of its
// incoming carry, and such functions compose associatively.
#include "rf_field.h"
#include "rf_kernels.h"

namespace {

// ---- CRT, digit sums and carry transfer functions ----
//
// A carry transfer function maps an incoming carry c in {-1, 0, 1} to the
// outgoing one, also in {-1, 0, 1}; code = sum_c (f(c) + 1) << (2 (c + 1)).
__device__ __forceinline__ int fn_apply(uint32_t code, int c) {
  return (int)((code >> (2 * (c + 1))) & 3u) - 1;
}
__device__ __forceinline__ uint32_t fn_compose(uint32_t g, uint32_t f) {  // g o f
  // field c of g o f is field (field c of f) of g: both hold carry + 1
  uint32_t h = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) h |= __builtin_amdgcn_ubfe(g, 2 * __builtin_amdgcn_ubfe(f, 2 * c, 2), 2) << (2 * c);
  return h;
}
constexpr uint32_t FN_ID = (0u << 0) | (1u << 2) | (2u << 4);
// Whether f maps every incoming carry to the same outgoing one.
__device__ __forceinline__ bool fn_constant(uint32_t f) { return (f & 15u) == ((f >> 2) & 15u); }

// The integer value of a double holding an integer 0 <= v < 2^52: 2^52 + v
// is exact and its 52 mantissa bits are v (one add and one mask instead of
// the general conversion's truncation, scaling and two converts).
__device__ __forceinline__ uint64_t int_of_double(double v) {
  return (uint64_t)__double_as_longlong(v + 0x1p52) & ((1ull << 52) - 1);
}

// Garner's algorithm on residues res[j * N] (|x| <= p_j/2 + 8) into the
// S-limb two's complement of the representative in (-P/2, P/2)
// (src/lib/crt.c:11-94 states the same CRT on 62-bit primes).
// Mixed-radix digits v_j in [0, p_j): v_j = (x_j - [v_0 + p_0 (v_1 + ...)]) /
// (p_0 ... p_{j-1}) mod p_j; all fp64 values stay below 2^50.
template <int K, int S>
__device__ __forceinline__ void garner(const double* __restrict__ res, uint32_t N,
                                       const RfConsts* __restrict__ C, uint64_t (&X)[S]) {
  double v[K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    const double p = C->p[j];
    const double x = res[(size_t)j * N];
    double r;
    if (j == 0) {
      r = x;
    } else {
      double acc = v[j - 1];
#pragma unroll
      for (int l = j - 2; l >= 0; l--) acc = rf_mulmod(acc, C->pm[j][l][0], C->pm[j][l][1], p) + v[l];
      r = rf_mulmod(x - acc, C->ci[j][0], C->ci[j][1], p);
    }
    // |r| < p: x is reduced (|x| <= p/2 + 8); acc, a product plus a
    // canonical digit, stays below 1.5 p + 8, so |x - acc| < 2^50 and the
    // product r satisfies |r| <= p (1/2 + 2^50 2^-52) = 0.75 p
    v[j] = rf_canon1(r, p, C->pinv[j]);
  }
  // Mixed radix to binary: X = v_{K-1}; X = X p_j + v_j for j = K-2 .. 0, in
  // 32-bit limbs.  p_j = phi 2^32 + plo with phi < 2^16 and, as p_j = 1 mod
  // 2^24, plo <= 2^32 - 2^24 + 1.  One pass per step, column by column:
  //   t = x_{l-1} phi + c;  t = x_l plo + t;  x_l = t mod 2^32;  c = t >> 32,
  // starting from c = v_j (< 2^48): t < (2^32 - 1)(2^32 - 2^24 + 1) + 2^49
  // < 2^64, so each column is two v_mad_u64_u32 whose 64-bit addend is the
  // previous one's result.  Before step j, X < 2^(48 (K-1-j)) has nl limbs;
  // after it, nl + 2 limbs hold X p_j + v_j.
  constexpr int W = 2 * S;
  uint32_t x[W];
#pragma unroll
  for (int l = 0; l < W; l++) x[l] = 0;
  {
    const uint64_t t = int_of_double(v[K - 1]);
    x[0] = (uint32_t)t;
    x[1] = (uint32_t)(t >> 32);
  }
#pragma unroll
  for (int j = K - 2; j >= 0; j--) {
    const int nl = (48 * (K - 1 - j) + 31) / 32;  // limbs of X before the step
    const uint64_t pj = C->pu[j];
    const uint32_t plo = (uint32_t)pj, phi = (uint32_t)(pj >> 32);
    uint64_t c = int_of_double(v[j]);
    uint32_t prev = 0;
#pragma unroll
    for (int l = 0; l < W; l++) {
      if (l <= nl) {
        const uint32_t xl = l < nl ? x[l] : 0u;
        uint64_t t = (uint64_t)prev * phi + c;
        t = (uint64_t)xl * plo + t;
        prev = xl;
        x[l] = (uint32_t)t;
        c = t >> 32;
      } else if (l == nl + 1) {
        x[l] = (uint32_t)c;
      }
    }
  }
  uint64_t X64[S];
#pragma unroll
  for (int l = 0; l < S; l++) X64[l] = (uint64_t)x[2 * l] | ((uint64_t)x[2 * l + 1] << 32);
#pragma unroll
  for (int l = 0; l < S; l++) X[l] = X64[l];
  // symmetric lift: X > floor(P / 2)  =>  X - P (two's complement in S limbs).
  // floor(P / 2) = (P - 1) / 2 has the mixed-radix digits (p_j - 1) / 2 (the
  // sum of (p_j - 1) / 2 p_0 .. p_{j-1} telescopes), and the mixed-radix
  // order is lexicographic from the top digit, which nearly always decides.
  const double htop = __builtin_fma(C->p[K - 1], 0.5, -0.5);
  bool gt = v[K - 1] > htop;
  if (v[K - 1] == htop) {
#pragma unroll
    for (int l = K - 2; l >= 0; l--) {
      const double h = __builtin_fma(C->p[l], 0.5, -0.5);
      if (v[l] != h) {
        gt = v[l] > h;
        break;
      }
    }
  }
  if (gt) {
    unsigned long long br = 0;
#pragma unroll
    for (int l = 0; l < S; l++) X[l] = __builtin_subcll(X[l], C->P[K][l], br, &br);
  }
}

// Sum of the slices of coefficients X_j (j0-based in LDS) falling on digit i
// (limbs [t i, t i + t)), plus the incoming carry cin; writes limb l through
// emit(l, value) and returns the signed carry out.  Coefficient j occupies
// limbs [t j, t j + S) unsigned, minus 2^64 at limb t j + S if negative.
// TT > 0 fixes t = TT (and M) at compile time so the loops unroll.
// Every coefficient read, i - m (m < M) and the sign term i - S / t, lies at
// or above j0 + 1 for the digits of a block (i >= digit0 - 1, halo M + 2),
// and the rows of coefficients below 0 hold zeros, so the reads need no
// tests and address one row with constant offsets.  A limb's sum is formed
// in 32-bit words with add-with-carry chains, its carries counted in one
// signed 32-bit word: at most M + 1 terms, so carries stay in [-1, M].
template <int S, int TT, typename Emit>
__device__ __forceinline__ int64_t digit_sum(const uint64_t* __restrict__ X, int j0, int i, int t_rt,
                                             int M_rt, int64_t cin, Emit emit) {
  const int t = TT > 0 ? TT : t_rt;
  const int M = TT > 0 ? (S + TT - 1) / TT : M_rt;
  const int lneg = S % t;
  const uint64_t* row = X + (size_t)(i - j0) * S;  // coefficient i; coefficient i - m at row - m S
  const uint32_t nb = (uint32_t)(X[(size_t)(i - S / t - j0) * S + (S - 1)] >> 63);
  uint32_t w0 = (uint32_t)cin, w1 = (uint32_t)((uint64_t)cin >> 32);
  int32_t h = (int32_t)(cin >> 63);
#pragma unroll
  for (int l = 0; l < t; l++) {
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int idx = m * t + l;
      if (idx < S) {
        const uint64_t x = row[idx - m * S];
        unsigned int cy;
        w0 = __builtin_addc(w0, (uint32_t)x, 0u, &cy);
        w1 = __builtin_addc(w1, (uint32_t)(x >> 32), cy, &cy);
        h += (int32_t)cy;
      }
    }
    if (l == lneg) {
      unsigned int br;
      w0 = __builtin_subc(w0, nb, 0u, &br);
      w1 = __builtin_subc(w1, 0u, br, &br);
      h -= (int32_t)br;
    }
    emit(l, ((uint64_t)w1 << 32) | w0);
    w0 = (uint32_t)h;
    w1 = (uint32_t)(h >> 31);
    h >>= 31;
  }
  return (int64_t)(((uint64_t)w1 << 32) | w0);
}

template <int K>
__global__ __launch_bounds__(RF_CRT_THREADS) void k_crt(RfDevTables tb,
                                                       const RfCrtTask* __restrict__ tasks,
                                                       const double* __restrict__ coef,
                                                       uint8_t* __restrict__ pre,
                                                       uint8_t* __restrict__ fun) {
  constexpr int S = (48 * K + 1 + 63) / 64;
  __shared__ uint64_t X[RF_CRT_THREADS * S];
  __shared__ int64_t E[RF_CRT_THREADS];
  __shared__ uint32_t code[RF_CRT_THREADS / 64];
  const int tid = threadIdx.x;
  const RfCrtTask tk = rf_wave_uniform_load(tasks + blockIdx.x);
  const int t = tk.t, M = (S + t - 1) / t, H = tk.halo;
  const int digit0 = (int)tk.digit0;
  const int j0 = digit0 - H;  // coefficient of thread 0
  {
    const int j = j0 + tid;
    uint64_t x[S];
    if (j >= 0 && (uint32_t)j < tk.nout) {
      garner<K, S>(coef + tk.coef_off + (uint32_t)j, tk.N, tb.consts, x);
    } else {
#pragma unroll
      for (int l = 0; l < S; l++) x[l] = 0;
    }
#pragma unroll
    for (int l = 0; l < S; l++) X[tid * S + l] = x[l];
  }
  __syncthreads();
  // Thread tid >= H - 1 owns digit i = digit0 + tid - H (digit0 - 1 for
  // tid = H - 1); its inputs X_{i-M+1..i} and the sign term X_{i - S/t}
  // all lie at or above j0 + 1 because H = M + 2.
  const int i = digit0 + tid - H;
  // pass 1: each digit on its own: its sum V_i (limbs 1.. written out now,
  // limb 0 kept) and carry out E_i
  const uint32_t ndig = (tk.rn + t - 1) / t;
  const bool own = tid >= H && (uint32_t)i < ndig;
  uint64_t RF_GLOBAL* dst = (uint64_t RF_GLOBAL*)(tk.out + (size_t)(own ? i : 0) * t);
  const uint32_t lim = own ? tk.rn - (uint32_t)i * t : 0u;  // limbs of the digit inside the output
  uint64_t v0 = 0;
  bool hz = true, ho = true;  // limbs 1 .. t-1 all zero / all ones
  {
    auto emit = [&](int l, uint64_t v) {
      if (l == 0) {
        v0 = v;
      } else {
        hz &= (v == 0);
        ho &= (v == ~(uint64_t)0);
        if ((uint32_t)l < lim) dst[l] = v;
      }
    };
    int64_t e = 0;
    if (tid >= H - 1 && i >= 0) {
      switch (t) {  // wave-uniform: one unrolled instance per chunk width
        case 1: e = digit_sum<S, 1>(X, j0, i, t, M, 0, emit); break;
        case 2: e = digit_sum<S, 2>(X, j0, i, t, M, 0, emit); break;
        case 3: e = digit_sum<S, 3>(X, j0, i, t, M, 0, emit); break;
        case 4: e = digit_sum<S, 4>(X, j0, i, t, M, 0, emit); break;
        case 5: e = digit_sum<S, 5>(X, j0, i, t, M, 0, emit); break;
        case 6: e = digit_sum<S, 6>(X, j0, i, t, M, 0, emit); break;
        case 7: e = digit_sum<S, 7>(X, j0, i, t, M, 0, emit); break;
        case 8: e = digit_sum<S, 8>(X, j0, i, t, M, 0, emit); break;
        default: e = digit_sum<S, 0>(X, j0, i, t, M, 0, emit); break;
      }
    }
    E[tid] = e;
  }
  __syncthreads();
  // pass 2: U_i = (V_i + E_{i-1}) mod 2^(64 t) and F_i = its carry in
  // {-1, 0, 1} (the carry out is E_i + F_i); the signed carry-in almost
  // always stops at limb 0
  uint32_t f = FN_ID;
  if (own) {
    int64_t c = E[tid - 1];
    const uint64_t r0 = v0 + (uint64_t)c;
    c = c >= 0 ? (int64_t)(r0 < v0) : -(int64_t)(r0 > v0);
    if (lim > 0) dst[0] = r0;
    bool all0 = r0 == 0 && hz, all1 = r0 == ~(uint64_t)0 && ho;
    if (c != 0) {
      // propagate through limbs 1 .. (this thread wrote them in pass 1); a
      // digit cut short by the output's end has no later digits, so its
      // carry function is never used
      all0 = r0 == 0;
      all1 = r0 == ~(uint64_t)0;
      for (int l = 1; l < t; l++) {
        if ((uint32_t)l >= lim) break;
        const uint64_t a = dst[l];
        uint64_t r = a;
        if (c != 0) {
          r = a + (uint64_t)c;
          c = c > 0 ? (int64_t)(r < a) : -(int64_t)(r > a);
          dst[l] = r;
        }
        all0 &= (r == 0);
        all1 &= (r == ~(uint64_t)0);
      }
    }
    const int F = (int)c;
    f = 0;
#pragma unroll
    for (int cc = -1; cc <= 1; cc++) {
      int o = F + ((cc == -1 && all0) ? -1 : 0) + ((cc == 1 && all1) ? 1 : 0);
      f |= (uint32_t)(o + 1) << (2 * (cc + 1));
    }
  }
  // Scan by composition: inclusive f_tid o ... o f_0 within each wave by
  // shuffles, then each wave composes the earlier waves' aggregates.  A
  // constant function composed after anything is itself, so in a wave whose
  // digits' functions are all constant (a digit that absorbs any incoming
  // carry: the usual case) every owning lane's inclusive prefix is its own
  // function and the shuffle scan is skipped.  The other lanes hold the
  // identity: the halo lanes below the block's first digit, whose prefix is
  // the identity as well, and lanes past the entry's last digit (only in its
  // last block), whose prefixes and aggregates nothing reads.
  const int lane = tid & 63, wave = tid >> 6;
  const bool slow = __builtin_amdgcn_ballot_w64(own && !fn_constant(f)) != 0;
  if (slow) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t g = __shfl_up(f, off);
      if (lane >= off) f = fn_compose(f, g);
    }
  }
  if (lane == 63) code[wave] = f;
  __syncthreads();
  // composition of the earlier waves' aggregates: the same for every lane,
  // formed in scalar registers
  uint32_t before = FN_ID;
  for (int w = 0; w < wave; w++) before = fn_compose(__builtin_amdgcn_readfirstlane(code[w]), before);
  // after the shortcut every lane that matters holds a constant (or, below
  // the first digit, the identity with the identity before it): its
  // inclusive prefix is its own function
  uint32_t incl = f;
  if (slow) incl = fn_compose(f, before);
  const uint32_t prev = __shfl_up(incl, 1);
  pre[tk.pre_off + tid] = (uint8_t)(lane == 0 ? before : prev);
  if (tid == RF_CRT_THREADS - 1) fun[tk.fun_idx] = (uint8_t)incl;
}

// Adds each digit's resolved carry into the output limbs.  The carry into
// the block is the composition, in order, of the aggregates of the entry's
// earlier blocks applied to 0: the first wave's lanes each compose a
// contiguous run of them, then an ordered shuffle tree combines the runs.
__global__ __launch_bounds__(64) void k_apply(const RfCrtTask* __restrict__ tasks,
                                           const uint8_t* __restrict__ pre,
                                           const uint8_t* __restrict__ fun) {
  // one wave per block: the composition of the earlier blocks' aggregates
  // (a shuffle tree leaves it in lane 0), then RF_CRT_THREADS / 64 digits per
  // lane (carries are rare, so most lanes only read their codes)
  const int lane = threadIdx.x;
  const RfCrtTask tk = rf_wave_uniform_load(tasks + blockIdx.x);
  const int t = tk.t, H = tk.halo;
  const uint32_t nprev = tk.digit0 / (uint32_t)(RF_CRT_THREADS - H);  // earlier blocks of the entry
  const uint32_t first = tk.fun_idx - nprev;
  const uint32_t cnt = (nprev + 63) / 64;
  uint32_t f = FN_ID;
  const uint32_t b1 = min(nprev, (uint32_t)(lane + 1) * cnt);
  for (uint32_t b = (uint32_t)lane * cnt; b < b1; b++) f = fn_compose(fun[first + b], f);
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t g = __shfl_down(f, off);
    if ((lane & (2 * off - 1)) == 0) f = fn_compose(g, f);
  }
  const int cin = __shfl(fn_apply(f, 0), 0);
  const uint32_t ndig = (tk.rn + t - 1) / t;
#pragma unroll
  for (int k = 0; k < RF_CRT_THREADS / 64; k++) {
    const int tid = lane + 64 * k;
    if (tid < H) continue;
    const uint32_t i = tk.digit0 + (uint32_t)(tid - H);
    if (i >= ndig) continue;
    const int c = fn_apply(pre[tk.pre_off + tid], cin);
    if (c == 0) continue;
    uint64_t RF_GLOBAL* dst = (uint64_t RF_GLOBAL*)(tk.out + (size_t)i * t);
    const uint32_t lim = min((uint32_t)t, tk.rn - i * t);
    for (uint32_t l = 0; l < lim; l++) {
      const uint64_t x = dst[l];
      if (c > 0) {
        dst[l] = x + 1;
        if (x != ~(uint64_t)0) break;
      } else {
        dst[l] = x - 1;
        if (x != 0) break;
      }
    }
  }
}

template <int K>
hipError_t launch_crt_t(const RfDevTables& tb, const RfCrtTask* tasks, int ntasks,
                        const double* coef, uint8_t* pre, uint8_t* fun, hipStream_t st) {
  hipLaunchKernelGGL(k_crt<K>, dim3(ntasks), dim3(RF_CRT_THREADS), 0, st, tb, tasks, coef,
                     pre, fun);
  return hipGetLastError();
}

}  // namespace

hipError_t rf_launch_crt(const RfDevTables& tb, int k, const RfCrtTask* tasks, int ntasks,
                         const double* coef, uint8_t* pre, uint8_t* fun, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  switch (k) {
#define RF_CASE(K) \
  case K:          \
    return launch_crt_t<K>(tb, tasks, ntasks, coef, pre, fun, st);
    RF_CASE(1) RF_CASE(2) RF_CASE(3) RF_CASE(4) RF_CASE(5) RF_CASE(6) RF_CASE(7) RF_CASE(8)
    RF_CASE(9) RF_CASE(10) RF_CASE(11) RF_CASE(12) RF_CASE(13) RF_CASE(14) RF_CASE(15)
    RF_CASE(16) RF_CASE(17) RF_CASE(18) RF_CASE(19) RF_CASE(20) RF_CASE(21) RF_CASE(22)
    RF_CASE(23) RF_CASE(24)
#undef RF_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t rf_launch_apply(const RfCrtTask* tasks, int ntasks, const uint8_t* pre, const uint8_t* fun,
                           hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply, dim3(ntasks), dim3(64), 0, st, tasks, pre, fun);
  return hipGetLastError();
}
