// rf_util.hip -- data kernels beside the product pipeline (gfx950).
//
//   k_gen       fills entries with the synthetic limb stream of a seeded
//               request (load generation for the request server)
//   k_checksum  64-bit checksums of result limbs, summed per request
//   k_fill      writes a constant over a u64 range (clears checksum slots
//               without a memset node, safe under graph capture)
//   k_sm_scan,  two's complement results -> sign and magnitude in place (a
//   k_sm_flip   product tree's level outputs become the next level's inputs)
//   k_mod_fix   the last step of a remainder X mod D: X - Q D and its
//               corrections by multi-limb subtraction
//
// The first three are flat, coalesced streams over lists of spans; a span is one
// entry (k_gen) or one C entry (k_checksum), split by the host into blocks
// of RF_UTIL_LIMBS limbs.
#include "rf_util.h"

namespace {

__global__ __launch_bounds__(RF_UTIL_THREADS) void k_gen(const RfSpan* __restrict__ spans,
                                                         const RfBlock* __restrict__ blocks, int nblocks) {
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  const RfBlock bk = blocks[b];
  const RfSpan sp = spans[bk.span];
  const uint64_t n = (sp.bits + 63) / 64;
  const uint32_t top_bits = (uint32_t)(sp.bits - 64 * (n - 1));  // 1 .. 64
#pragma unroll
  for (int u = 0; u < RF_UTIL_LIMBS / RF_UTIL_THREADS; u++) {
    const uint64_t i = bk.first + (uint64_t)u * RF_UTIL_THREADS + threadIdx.x;
    if (i >= n) break;
    uint64_t v = rf_gen_limb(sp.key, i);
    if (i == n - 1) {
      if (top_bits < 64) v &= (((uint64_t)1) << top_bits) - 1;
      v |= ((uint64_t)1) << (top_bits - 1);
    }
    sp.ptr[i] = v;
  }
}

__global__ __launch_bounds__(RF_UTIL_THREADS) void k_checksum(const RfSpan* __restrict__ spans,
                                                              const RfBlock* __restrict__ blocks, int nblocks,
                                                              unsigned long long* __restrict__ sums) {
  __shared__ uint64_t part[RF_UTIL_THREADS / 64];
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  const RfBlock bk = blocks[b];
  const RfSpan sp = spans[bk.span];
  const uint64_t n = sp.bits;  // limbs
  uint64_t h = 0;
#pragma unroll
  for (int u = 0; u < RF_UTIL_LIMBS / RF_UTIL_THREADS; u++) {
    const uint64_t i = bk.first + (uint64_t)u * RF_UTIL_THREADS + threadIdx.x;
    if (i < n) h += rf_sum_term(sp.ptr[i], sp.key, i);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
#pragma unroll
    for (int w = 0; w < RF_UTIL_THREADS / 64; w++) t += part[w];
    atomicAdd(sums + sp.slot, (unsigned long long)t);
  }
}

// The scan, one wave per entry (RF_UTIL_THREADS / 64 entries per
// workgroup): the sign is the top bit; a negative entry's magnitude ~x + 1 is
// 0 below its lowest nonzero limb l0, -x[l0] at l0 and ~x above, so its
// length is found without writing: limb i > l0 of the magnitude is zero
// exactly when x[i] is all ones.  Both ends are found by sweeps of 64 limbs
// decided by one ballot each, from the bottom for l0 and from the top for the
// length; the first sweep of each end is loaded together with the sign limb,
// and a typical entry needs no other.
__global__ __launch_bounds__(RF_UTIL_THREADS) void k_sm_scan(const RfConvTask* __restrict__ tasks, int ntasks,
                                                            RfEntryMeta* __restrict__ meta) {
  const int lane = threadIdx.x & 63;
  const int e = (int)blockIdx.x * (RF_UTIL_THREADS / 64) + (int)(threadIdx.x >> 6);
  if (e >= ntasks) return;  // the whole wave
  const RfConvTask tk = tasks[e];
  const uint64_t n = tk.n;
  const uint64_t* __restrict__ x = tk.limbs;
  const uint64_t ul = (uint64_t)lane;
  const uint64_t top = n ? x[n - 1] : 0;
  const uint64_t lo = ul < n ? x[ul] : 0;
  uint64_t v = ul < n ? x[n - 1 - ul] : 0;
  const bool neg = (top >> 63) != 0;
  uint64_t l0 = 0;
  if (neg) {  // a nonzero limb exists: the top one
    uint64_t b = 0;
    uint64_t m = __ballot(lo != 0);
    while (m == 0) {
      b += 64;
      m = __ballot(b + ul < n && x[b + ul] != 0);
    }
    l0 = b + (uint64_t)(__ffsll((unsigned long long)m) - 1);
  }
  const uint64_t fill = neg ? ~0ull : 0ull;
  uint64_t len = 0;
  // sweeps of [t - 64, t) down to l0 (lane i holds limb t - 1 - i); a
  // negative entry's sweep ends at l0 at the latest
  for (uint64_t t = n; t > l0;) {
    const uint64_t m = __ballot(ul < t - l0 && (v != fill || (neg && t - 1 - ul == l0)));
    if (m) {
      len = t - (uint64_t)(__ffsll((unsigned long long)m) - 1);
      break;
    }
    t = t - l0 > 64 ? t - 64 : l0;
    v = ul < t - l0 ? x[t - 1 - ul] : 0;
  }
  if (lane == 0) meta[e] = RfEntryMeta{len == 0 ? 0 : (neg ? -1 : 1), (uint32_t)len, (uint32_t)l0, 0};
}

// The negation of the negative entries: workgroup (x, y) turns limbs
// [x, x + 1) * RF_FLIP_LIMBS of entry y into the magnitude's, from l0 up
// (the limbs below l0 are zero in both), coalesced; loads first, then stores.
__global__ __launch_bounds__(RF_UTIL_THREADS) void k_sm_flip(const RfConvTask* __restrict__ tasks,
                                                            const RfEntryMeta* __restrict__ meta) {
  const RfEntryMeta m = meta[blockIdx.y];
  if (m.sign >= 0) return;
  const RfConvTask tk = tasks[blockIdx.y];
  const uint64_t t0 = (uint64_t)blockIdx.x * RF_FLIP_LIMBS;
  if (t0 >= tk.n) return;
  const uint64_t l0 = m.l0;
  constexpr int U = RF_FLIP_LIMBS / RF_UTIL_THREADS;
  uint64_t v[U];
#pragma unroll
  for (int k = 0; k < U; k++) {
    const uint64_t i = t0 + (uint64_t)k * RF_UTIL_THREADS + threadIdx.x;
    v[k] = i >= l0 && i < tk.n ? tk.limbs[i] : 0;
  }
#pragma unroll
  for (int k = 0; k < U; k++) {
    const uint64_t i = t0 + (uint64_t)k * RF_UTIL_THREADS + threadIdx.x;
    if (i >= l0 && i < tk.n) tk.limbs[i] = i == l0 ? 0 - v[k] : ~v[k];
  }
}

// The remainder fix-up runs one remainder per workgroup of RF_FIX_THREADS
// threads over tiles of as many limbs; the carry chain across its waves
// keeps NW = RF_FIX_THREADS / 64 entries per kind in shared memory.
#define RF_FIX_THREADS 1024
#define FIX_NW (RF_FIX_THREADS / 64)
#define FIX_C0 0
#define FIX_C1 FIX_NW
#define FIX_IN (2 * FIX_NW)
#define FIX_TILE (3 * FIX_NW)
#define FIX_FLAG (3 * FIX_NW + 1)
#define FIX_SH (3 * FIX_NW + 2)

// ---- multi-limb subtraction inside one workgroup ----
//
// dst = a - b over L limbs (limbs past an operand's length read 0), tiles of
// RF_FIX_THREADS limbs in order with the borrow carried between tiles.
// Within a wave, limb i generates a borrow (g: a_i < b_i) or propagates one
// (p: a_i == b_i); with the wave's masks G and P (disjoint) the borrows into
// its limbs are the carries of the addition x + y + c for x = G | P, y = G:
// carry into bit i = bit i of (x + y + c) ^ x ^ y, carry out of bit 63 =
// ((x & y) | ((x | y) & ~(x + y + c))) >> 63.  The waves' carries are
// chained by one thread (FIX_NW waves).  Returns the borrow out of limb L-1.  Thread j
// handles limbs j, j + RF_FIX_THREADS, ... in every call, so a thread reads
// back only what it wrote itself.
struct LimbSpan {
  const uint64_t* p;
  uint64_t n;
  __device__ uint64_t at(uint64_t i) const { return i < n ? p[i] : 0; }
};

__device__ int sub_limbs(LimbSpan a, LimbSpan b, uint64_t* __restrict__ dst, uint64_t L, int* sh) {
  // sh: [0, NW) carry out with c = 0, [NW, 2 NW) with c = 1, [2 NW, 3 NW) wave
  // carry-in, [3 NW] tile borrow (NW waves)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();  // every thread has read the previous call's result
  if (threadIdx.x == 0) sh[FIX_TILE] = 0;
  __syncthreads();
  for (uint64_t base = 0; base < L; base += RF_FIX_THREADS) {
    const uint64_t i = base + threadIdx.x;
    const bool in = i < L;
    const uint64_t x = in ? a.at(i) : 0, y = in ? b.at(i) : 0;
    const uint64_t G = __builtin_amdgcn_ballot_w64(in && x < y);
    const uint64_t P = __builtin_amdgcn_ballot_w64(!in || x == y);
    const uint64_t X = G | P, Y = G;
    const uint64_t S0 = X + Y, S1 = X + Y + 1;
    if (lane == 0) {
      sh[FIX_C0 + wave] = (int)(((X & Y) | ((X | Y) & ~S0)) >> 63);
      sh[FIX_C1 + wave] = (int)(((X & Y) | ((X | Y) & ~S1)) >> 63);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int c = sh[FIX_TILE];
      for (int w = 0; w < RF_FIX_THREADS / 64; w++) {
        sh[FIX_IN + w] = c;
        c = c ? sh[FIX_C1 + w] : sh[FIX_C0 + w];
      }
      sh[FIX_TILE] = c;
    }
    __syncthreads();
    const uint64_t S = sh[FIX_IN + wave] ? S1 : S0;
    const uint64_t bin = ((S ^ X ^ Y) >> lane) & 1;
    if (in) dst[i] = x - y - bin;
    __syncthreads();  // sh[] is rewritten by the next tile
  }
  return sh[FIX_TILE];
}

// a + b mod B^L into dst (dst may alias a or b); returns the carry out.
// The carries into the lanes come from the same generate / propagate ballots
// as sub_limbs' borrows: G = the lane's sum overflows, P = it is all ones.
__device__ int add_limbs(LimbSpan a, LimbSpan b, uint64_t* __restrict__ dst, uint64_t L, int* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();  // every thread has read the previous call's result
  if (threadIdx.x == 0) sh[FIX_TILE] = 0;
  __syncthreads();
  for (uint64_t base = 0; base < L; base += RF_FIX_THREADS) {
    const uint64_t i = base + threadIdx.x;
    const bool in = i < L;
    const uint64_t x = in ? a.at(i) : 0, y = in ? b.at(i) : 0;
    const uint64_t s = x + y;
    const uint64_t G = __builtin_amdgcn_ballot_w64(in && s < x);
    const uint64_t P = __builtin_amdgcn_ballot_w64(!in || s == ~(uint64_t)0);
    const uint64_t X = G | P, Y = G;
    const uint64_t S0 = X + Y, S1 = X + Y + 1;
    if (lane == 0) {
      sh[FIX_C0 + wave] = (int)(((X & Y) | ((X | Y) & ~S0)) >> 63);
      sh[FIX_C1 + wave] = (int)(((X & Y) | ((X | Y) & ~S1)) >> 63);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int c = sh[FIX_TILE];
      for (int w = 0; w < RF_FIX_THREADS / 64; w++) {
        sh[FIX_IN + w] = c;
        c = c ? sh[FIX_C1 + w] : sh[FIX_C0 + w];
      }
      sh[FIX_TILE] = c;
    }
    __syncthreads();
    const uint64_t S = sh[FIX_IN + wave] ? S1 : S0;
    const uint64_t cin = ((S ^ X ^ Y) >> lane) & 1;
    if (in) dst[i] = s + cin;
    __syncthreads();  // sh[] is rewritten by the next tile
  }
  return sh[FIX_TILE];
}

// a mod (B^W - 1) into dst (W limbs) for a of at most 2 W limbs: a_lo + a_hi,
// then the carry out (B^W = 1) added back.  a_lo + a_hi <= 2 (B^W - 1), so
// after a carry the low part is at most B^W - 2 and adding 1 cannot carry
// again.  The class of 0 may come out as B^W - 1.
__device__ void fold_limbs(LimbSpan a, uint64_t* __restrict__ dst, uint64_t W, int* sh) {
  const LimbSpan lo{a.p, a.n < W ? a.n : W};
  const LimbSpan hi{a.n > W ? a.p + W : a.p, a.n > W ? a.n - W : 0};
  if (add_limbs(lo, hi, dst, W, sh)) {
    const uint64_t one = 1;
    add_limbs(LimbSpan{dst, W}, LimbSpan{&one, 1}, dst, W, sh);
  }
}

__global__ __launch_bounds__(RF_FIX_THREADS) void k_mod_fix(const RfModTask* __restrict__ tasks, int ntasks) {
  __shared__ int sh[FIX_SH];
  if ((int)blockIdx.x >= ntasks) return;
  const RfModTask tk = tasks[blockIdx.x];
  const uint64_t L = tk.n + 1;
  uint64_t* R = tk.r;
  uint64_t* T = tk.t;
  const LimbSpan D{tk.d, tk.n};
  if (tk.sign == 0) {
    for (uint64_t i = threadIdx.x; i < tk.n; i += RF_FIX_THREADS) tk.out[i] = 0;
    return;
  }
  if (tk.wrap) {
    // R = (X - QD) mod (B^W - 1): exact, as 0 <= R < 3D < B^(n+1) - 1 <= B^W - 1
    const uint64_t W = tk.wrap;
    fold_limbs(LimbSpan{tk.x, tk.xn}, tk.xf, W, sh);
    fold_limbs(LimbSpan{tk.qd, tk.qdn}, T, W, sh);
    if (sub_limbs(LimbSpan{tk.xf, W}, LimbSpan{T, W}, R, W, sh)) {  // a borrow: add B^W - 1
      const uint64_t one = 1;
      sub_limbs(LimbSpan{R, W}, LimbSpan{&one, 1}, R, W, sh);
    }
    // B^W - 1 stands for 0; otherwise R's limbs from n + 1 on are zero
    if (threadIdx.x == 0) sh[FIX_FLAG] = 1;
    __syncthreads();
    for (uint64_t base = 0; base < W; base += RF_FIX_THREADS) {
      const uint64_t i = base + threadIdx.x;
      if (__builtin_amdgcn_ballot_w64(i < W && R[i] != ~(uint64_t)0) && (threadIdx.x & 63) == 0) sh[FIX_FLAG] = 0;
    }
    __syncthreads();
    if (sh[FIX_FLAG])
      for (uint64_t i = threadIdx.x; i < L; i += RF_FIX_THREADS) R[i] = 0;
    __syncthreads();
  } else {
    sub_limbs(LimbSpan{tk.x, tk.xn}, LimbSpan{tk.qd, tk.qdn}, R, L, sh);  // X - QD mod B^(n+1)
  }
  for (int k = 0; k < 2; k++) {
    const int borrow = sub_limbs(LimbSpan{R, L}, D, T, L, sh);  // R >= D iff no borrow
    if (!borrow) {
      uint64_t* u = R;
      R = T;
      T = u;
    }
  }
  if (tk.sign < 0) {
    if (threadIdx.x == 0) sh[FIX_FLAG] = 0;
    __syncthreads();
    for (uint64_t base = 0; base < L; base += RF_FIX_THREADS) {
      const uint64_t i = base + threadIdx.x;
      if (__builtin_amdgcn_ballot_w64(i < L && R[i] != 0) && (threadIdx.x & 63) == 0) atomicOr(&sh[FIX_FLAG], 1);
    }
    __syncthreads();
    if (sh[FIX_FLAG]) {
      sub_limbs(D, LimbSpan{R, L}, T, L, sh);  // D - R
      R = T;
    }
  }
  for (uint64_t i = threadIdx.x; i < tk.n; i += RF_FIX_THREADS) tk.out[i] = R[i];
}

// ---- the approximate-reciprocal recursion (one workgroup per modulus) ----
//
// With A_h X_h < B^(2h) <= A_h (X_h + 2) (the lower level's guarantee) and
// A = A_h B^l + A_l, T = A X_h < B^(n+h) + 2 B^n, so T >= B^(n+h) means limb
// n + h of T is 1 and at most a few subtractions of A (A >= B^n / 2) bring T
// below B^(n+h); the loop stops on the flag, bounded at 8 rounds.  Then
// 0 < B^(n+h) - T <= 2A < 2 B^n, so Tm = floor((B^(n+h) - T) / B^l) < 2 B^h
// fits h + 1 limbs.
__global__ __launch_bounds__(RF_FIX_THREADS) void k_recip_adjust(const RfRecipTask* __restrict__ tasks, int ntasks) {
  __shared__ int sh[FIX_SH];
  if ((int)blockIdx.x >= ntasks) return;
  const RfRecipTask tk = tasks[blockIdx.x];
  const uint64_t nh = tk.n + tk.h;
  const uint64_t one = 1;
  for (int round = 0; round < 8; round++) {
    __syncthreads();
    const bool over = nh < tk.tn && tk.t[nh] != 0;  // every thread reads the same limb
    if (!over) break;
    sub_limbs(LimbSpan{tk.t, tk.tn}, LimbSpan{tk.a, tk.n}, tk.t, tk.tn, sh);
    sub_limbs(LimbSpan{tk.xh, tk.h + 1}, LimbSpan{&one, 1}, tk.xh, tk.h + 1, sh);
    __syncthreads();
  }
  // B^(n+h) - T = -T mod B^(n+h) (T > 0), written over T
  const uint64_t zero = 0;
  sub_limbs(LimbSpan{&zero, 1}, LimbSpan{tk.t, nh}, tk.t, nh, sh);
  __syncthreads();
  for (uint64_t i = threadIdx.x; i <= tk.h; i += RF_FIX_THREADS) tk.tm[i] = tk.l + i < nh ? tk.t[tk.l + i] : 0;
}

// X = X_h B^l + floor(U / B^(2h-l)): the low l limbs are U's limbs
// [2h - l, 2h), the rest X_h plus U's limbs from 2h on (U < 4 B^(2h)).
__global__ __launch_bounds__(RF_FIX_THREADS) void k_recip_combine(const RfRecipTask* __restrict__ tasks, int ntasks) {
  __shared__ int sh[FIX_SH];
  if ((int)blockIdx.x >= ntasks) return;
  const RfRecipTask tk = tasks[blockIdx.x];
  const uint64_t lo = 2 * tk.h - tk.l;
  for (uint64_t i = threadIdx.x; i < tk.l; i += RF_FIX_THREADS) tk.x[i] = lo + i < tk.un ? tk.u[lo + i] : 0;
  const LimbSpan uh{tk.un > 2 * tk.h ? tk.u + 2 * tk.h : tk.u, tk.un > 2 * tk.h ? tk.un - 2 * tk.h : 0};
  add_limbs(LimbSpan{tk.xh, tk.h + 1}, uh, tk.x + tk.l, tk.h + 1, sh);
}

__global__ __launch_bounds__(256) void k_recip_final(const RfRecipFinal* __restrict__ tasks, int ntasks) {
  const RfRecipFinal tk = tasks[blockIdx.y];
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= tk.il) return;
  const uint64_t hi = j + tk.o < tk.xn ? tk.x[j + tk.o] : 0;
  const uint64_t lo = j + tk.o - 1 < tk.xn ? tk.x[j + tk.o - 1] : 0;
  tk.out[j] = tk.s ? (hi << tk.s) | (lo >> (64 - tk.s)) : hi;
}

__global__ __launch_bounds__(256) void k_mod_norm(const RfNormTask* __restrict__ tasks, int ntasks) {
  const RfNormTask tk = tasks[blockIdx.y];
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= tk.np) return;
  const uint64_t z = tk.np - tk.n;  // zero limbs below D 2^s
  uint64_t v = 0;
  if (j >= z) {
    const uint64_t i = j - z;
    v = tk.d[i] << tk.s;
    if (tk.s && i) v |= tk.d[i - 1] >> (64 - tk.s);
  }
  tk.out[j] = v;
}

__global__ void k_fill(uint64_t* __restrict__ dst, uint64_t n, uint64_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

__global__ __launch_bounds__(256) void k_sign_patch(char* __restrict__ img, const RfSignPatch* __restrict__ pt,
                                                      int n, const RfEntryMeta* __restrict__ meta) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) *(int32_t*)(img + pt[i].off) = meta[pt[i].src].sign;
}

}  // namespace

hipError_t rf_launch_sign_patch(char* image, const RfSignPatch* patches, int n, const RfEntryMeta* meta,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sign_patch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, image, patches, n, meta);
  return hipGetLastError();
}

hipError_t rf_launch_gen(const RfSpan* spans, const RfBlock* blocks, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen, dim3(nblocks), dim3(RF_UTIL_THREADS), 0, st, spans, blocks, nblocks);
  return hipGetLastError();
}

hipError_t rf_launch_checksum(const RfSpan* spans, const RfBlock* blocks, int nblocks, uint64_t* sums,
                              hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_checksum, dim3(nblocks), dim3(RF_UTIL_THREADS), 0, st, spans, blocks, nblocks,
                     (unsigned long long*)sums);
  return hipGetLastError();
}

hipError_t rf_launch_fill(uint64_t* dst, uint64_t n, uint64_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, n, v);
  return hipGetLastError();
}

hipError_t rf_launch_mod_fix(const RfModTask* tasks, int ntasks, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mod_fix, dim3(ntasks), dim3(RF_FIX_THREADS), 0, st, tasks, ntasks);
  return hipGetLastError();
}

hipError_t rf_launch_recip_adjust(const RfRecipTask* tasks, int ntasks, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_recip_adjust, dim3(ntasks), dim3(RF_FIX_THREADS), 0, st, tasks, ntasks);
  return hipGetLastError();
}

hipError_t rf_launch_recip_combine(const RfRecipTask* tasks, int ntasks, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_recip_combine, dim3(ntasks), dim3(RF_FIX_THREADS), 0, st, tasks, ntasks);
  return hipGetLastError();
}

hipError_t rf_launch_recip_final(const RfRecipFinal* tasks, int ntasks, uint64_t max_il, hipStream_t st) {
  if (ntasks <= 0 || max_il == 0) return hipSuccess;
  // the grid covers the longest output; shorter ones stop early; the grid's
  // y extent is at most 65535
  for (int t0 = 0; t0 < ntasks; t0 += 65535) {
    const int nt = ntasks - t0 < 65535 ? ntasks - t0 : 65535;
    hipLaunchKernelGGL(k_recip_final, dim3((unsigned)((max_il + 255) / 256), (unsigned)nt), dim3(256), 0, st,
                       tasks + t0, nt);
  }
  return hipGetLastError();
}

hipError_t rf_launch_mod_norm(const RfNormTask* tasks, int ntasks, uint64_t max_np, hipStream_t st) {
  if (ntasks <= 0 || max_np == 0) return hipSuccess;
  for (int t0 = 0; t0 < ntasks; t0 += 65535) {  // the grid's y extent is at most 65535
    const int nt = ntasks - t0 < 65535 ? ntasks - t0 : 65535;
    hipLaunchKernelGGL(k_mod_norm, dim3((unsigned)((max_np + 255) / 256), (unsigned)nt), dim3(256), 0, st,
                       tasks + t0, nt);
  }
  return hipGetLastError();
}

hipError_t rf_launch_signmag(const RfConvTask* tasks, int ntasks, RfEntryMeta* meta, uint64_t max_n,
                             hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  constexpr int EPB = RF_UTIL_THREADS / 64;  // entries per workgroup
  hipLaunchKernelGGL(k_sm_scan, dim3((unsigned)((ntasks + EPB - 1) / EPB)), dim3(RF_UTIL_THREADS), 0, st, tasks,
                     ntasks, meta);
  if (max_n == 0) return hipGetLastError();
  // every entry's tiles fit the grid's x extent (max_n bounds every n); the
  // y extent is at most 65535
  const unsigned gx = (unsigned)((max_n + RF_FLIP_LIMBS - 1) / RF_FLIP_LIMBS);
  for (int t0 = 0; t0 < ntasks; t0 += 65535) {
    const int nt = ntasks - t0 < 65535 ? ntasks - t0 : 65535;
    hipLaunchKernelGGL(k_sm_flip, dim3(gx, (unsigned)nt), dim3(RF_UTIL_THREADS), 0, st, tasks + t0, meta + t0);
  }
  return hipGetLastError();
}
