// rf_ntt.hip -- number-theoretic transforms of the product pipeline (gfx950).
//
//   k_fwd  pack (limb chunks -> residues, sign and 1/N folded into the pack
//          constants) + forward transform, natural order in, bit-reversed out
//   k_inv  pointwise r x d . d x d product per frequency + inverse transform,
//          bit-reversed in, natural order out (unscaled: the 1/N sits in the
//          B operands' pack constants)
//
// Transform definition (as src/include/fft62.h:35-54): with omega of order N,
// forward output position m holds sum_i a_i omega^(i rev(m)).
//
// Organisation.  Each thread keeps E = 16 residues in registers.  A pass runs
// R <= 4 butterfly stages on groups of 2^R residues whose indices differ only
// in R adjacent bits; between passes the workgroup exchanges residues through
// LDS (one write and one read of every residue, XOR-swizzled so that the
// strided accesses spread over the banks).  An N = 2^13 transform is
// four passes (4 + 4 + 4 + 1 stages) and three LDS exchanges.  Transforms
// smaller than a wave share a workgroup.
//
// Both directions use Cooley-Tukey butterflies (x, y) -> (x + w y, x - w y).
// A stage adds at most p (1/2 + |y| 2^-52) to a residue's bound, or doubles it
// where w = 1 is not multiplied (the first pass's block 0), and every eighth
// stage reduces, so all values stay below 12.2 p + 160 < 2^51.5 (primes below
// 2^51 / 9; the derivation is at RF_PMAX).
#include <atomic>

#include "rf_field.h"
#include "rf_kernels.h"

namespace {

__device__ __forceinline__ int xcd_task(int bid, int grid, int ntasks) {
  // Workgroup ids are dispatched round-robin over the 8 XCDs; remap so that
  // consecutive tasks (one entry's primes, or one prime's entries) share an
  // XCD and its L2.  grid is a multiple of 8.
  const int per = grid >> 3;
  const int t = (bid & 7) * per + (bid >> 3);
  return t < ntasks ? t : -1;
}

// LDS slot of residue i in an exchange.  Both sides of an exchange address
// residue i; the slot is chosen so that each 32-lane half of a
// ds_read_b64 / ds_write_b64 covers 32 distinct slots modulo 32 (all 64
// banks).
//   SW = 0: identity, enough when both sides' lanes run over 32 consecutive
//           residues.
//   SW = 2: i + 2 (i >> 5), rows of 32 residues padded by two slots: one
//           side's lanes run over 32 consecutive residues, the other's take
//           2^B consecutive residues (B = 1 .. 4) from each of 32 / 2^B rows.
//   SW = 3: i + (i >> 4), rows of 16 padded by one: one group of 16
//           consecutive residues per lane on one side against lanes 16 rows
//           apart on the other.
//   SW = 4: (i + 2 (i >> 5)) xor ((i >> 5) mod 2), SW = 2 with bit 0 flipped
//           in odd rows: lanes taking residue pairs from 16 rows (i =
//           32 (lane / 2) + lane mod 2 + c) need the pads of 16 consecutive
//           rows to share their parity, lanes two residues apart (i =
//           2 lane + c) need adjacent rows of opposite parity; the flip
//           gives both, which no plain padding does.
//   SW = 1: i xor h(i >> 5), h(m) = 2 (m mod 16) + (m mod 2), no padding:
//           it spreads both of the SW = 4 patterns as well, at one VALU
//           operation per access.
// Every map is injective with slots below N + N / 16 (below N for SW = 0, 1).
template <int SW>
__device__ __forceinline__ int lds_slot(int i) {
  if constexpr (SW == 0) {
    return i;
  } else if constexpr (SW == 2) {
    return i + ((i >> 5) << 1);
  } else if constexpr (SW == 3) {
    return i + (i >> 4);
  } else if constexpr (SW == 4) {
    return (i + ((i >> 5) << 1)) ^ ((i >> 5) & 1);
  } else {
    const int m = i >> 5;
    return i ^ (((m & 15) << 1) | (m & 1));
  }
}

// Address of residue t | o in an exchange, t holding the thread's index
// bits and o, a compile-time constant, the register position's.  The two
// never share a bit: every index below places the bits of the thread id, the
// group number u and the register index j at distinct positions, and the
// group of (tid, u) is tid + T u = tid | u T.  So
//   SW = 0, 2, 3: slot(t | o) = slot(t) + slot(o) (t | o = t + o, and
//                 (t | o) >> k = (t >> k) + (o >> k)), one base per thread
//                 and immediate offsets;
//   SW = 4:       the same sum with bit 0 set to (t0 | o0) xor row parity,
//                 at most four bases per thread;
//   SW = 1:       slot(t) xor slot(o) (h is linear over GF(2)).
// With constant offsets the compiler merges accesses into ds_read2_b64 /
// ds_write2_b64 pairs; pairs whose two slots lie in different banks conflict
// (64 conflict cycles per wave in the 2^13-point kernels, against 80 LDS
// instructions and no conflicts unmerged, which measured no faster).
template <int SW>
__device__ __forceinline__ double* lds_at(double* sh, int t, int o) {
  if constexpr (SW == 1) {
    return (double*)((char*)sh + ((lds_slot<1>(t) << 3) ^ (lds_slot<1>(o) << 3)));
  } else if constexpr (SW == 4) {
    const int pt = lds_slot<2>(t), po = lds_slot<2>(o);
    return sh + ((pt & ~1) + (po & ~1) + (((pt | po) ^ ((t ^ o) >> 5)) & 1));
  } else {
    return sh + lds_slot<SW>(t) + lds_slot<SW>(o);
  }
}

// Slot map of an exchange between a side whose lanes run over 2^A
// consecutive residues and one whose lanes run over 2^B (A, B: index bits
// below the stages of each pass), passes of R1 and R2 stages, in a
// transform of at least one wave (wide) or of less, with padded LDS or not.
constexpr int lds_map(bool pad, bool wide, int R1, int R2, int A, int B) {
  if (!wide) return 1;
  if (A > B) { const int t = A; A = B; B = t; }
  if (A >= 5) return 0;
  if (!pad) return 1;
  if (R1 >= 3 && R2 >= 3 && A >= 1 && B >= 5) return 2;
  if (R1 == 4 && R2 == 4 && A == 0 && B == 4) return 3;
  return 4;
}

// Slot for the pointwise-to-first-pass transposition (coalesced writes,
// groups of 16 consecutive residues read): the low four index bits XORed
// with the next four.
__device__ __forceinline__ int lds_pad(int i) { return i ^ ((i >> 4) & 15); }

template <int LGN>
struct Shape {
  static constexpr int N = 1 << LGN;
  static constexpr int E = LGN >= 4 ? 16 : N;          // residues per thread
  static constexpr int T = N / E;                      // threads per transform
  static constexpr int G = T >= 64 ? 1 : 64 / T;       // transforms per workgroup
  static constexpr int NP = (LGN + 3) / 4;             // forward passes
  // LDS per transform padded for the exchanges' slot maps from 2^13 on,
  // where two (2^13) or one (2^14) transforms per CU hold either way;
  // shorter transforms can run more workgroups per CU than 6 % padding
  // leaves room for, and keep N doubles
  static constexpr bool PAD = LGN >= 13;
  static constexpr int LDSN = PAD ? N + (N >> 4) : N;
  static constexpr int THREADS = T * G;
  // waves per SIMD the register allocation must allow (HIP's second
  // launch bound): four, as many as two 64 KiB transforms in LDS leave room
  // for, so at most 128 VGPRs
  static constexpr int MINB = 4;
};

// Pass P of a transform whose first pass runs R0 stages and every later one
// four (the last fewer).  Forward transforms use R0 = 4; inverse transforms
// start with the LGN mod 4 (or 4) leftover stages so that their groups of
// 2^R0 consecutive residues come straight from global memory.
template <int LGN, int P, int R0 = 4>
struct Pass {
  static constexpr int L0 = P == 0 ? 0 : R0 + 4 * (P - 1);   // first stage of the pass
  static constexpr int R = P == 0 ? (R0 < LGN ? R0 : LGN)
                                  : ((LGN - L0) < 4 ? (LGN - L0) : 4);  // stages in the pass
  static constexpr int NG = Shape<LGN>::E >> R;              // groups per thread
  static constexpr int GS = 1 << R;                          // group size
  // Reduce before this pass if the stages since the last reduction would
  // exceed eight.  Passes start after a reduction at P = 0.
  static constexpr int SINCE = (P == 0) ? 0 : (Pass<LGN, P - 1, R0>::SINCE_AFTER);
  static constexpr bool REDUCE_BEFORE = SINCE + R > 8;
  static constexpr int SINCE_AFTER = (REDUCE_BEFORE ? 0 : SINCE) + R;
};
template <int LGN, int R0>
struct Pass<LGN, -1, R0> {
  static constexpr int SINCE_AFTER = 0;
};
template <int LGN, int R0>
constexpr int npasses() {
  return LGN <= R0 ? 1 : 1 + (LGN - R0 + 3) / 4;
}
// First inverse pass: the leftover stages when they are one or two (groups
// of two or four residues loaded directly), otherwise four after an LDS
// transposition of the pointwise products.
template <int LGN>
constexpr int inv_r0() {
  return (Shape<LGN>::E == 16 && LGN >= 5 && ((LGN - 1) % 4) < 2) ? ((LGN - 1) % 4) + 1 : 4;
}

// Forward group indexing: stages L0 .. L0+R-1 act on index bits
// LGN-1-L0 .. LGN-L0-R.  Group g = (H, lo): i = H 2^(LGN-L0) + j 2^NLO + lo.
template <int LGN, int P>
__device__ __forceinline__ int fwd_index(int g, int j) {
  using Q = Pass<LGN, P>;
  constexpr int NLO = LGN - Q::L0 - Q::R;
  const int H = g >> NLO, lo = g & ((1 << NLO) - 1);
  return (H << (LGN - Q::L0)) | (j << NLO) | lo;
}

// Group of forward pass P held by thread tid as its u-th group.  Passes of
// four stages hold one group per thread: g = tid.  A last pass of fewer
// stages (several groups per thread) in a workgroup of several waves takes
// g = wave 2^(10-R) + u 64 + lane: the residues i = g 2^R + j of a wave are
// then the block [1024 wave, 1024 wave + 1024), the same block the wave held
// in the pass before (see fwd_local), and stores stay coalesced (consecutive
// lanes, consecutive groups).
template <int LGN, int P>
__device__ __forceinline__ int fwd_group(int tid, int u) {
  using Q = Pass<LGN, P>;
  if constexpr (P > 0 && P == Shape<LGN>::NP - 1 && Q::NG > 1 && Shape<LGN>::T > 64)
    return ((tid >> 6) << (10 - Q::R)) | (u << 6) | (tid & 63);
  else
    return tid + Shape<LGN>::T * u;
}

// Whether the exchange from forward pass P-1 into pass P stays inside each
// wave.  A transform of at most one wave trivially does.  Otherwise pass
// P-1 has four stages and one group per thread, i = H 2^(NLO+4) + j 2^NLO + lo
// with H = tid >> NLO, lo = tid mod 2^NLO; for NLO <= 6, i >> 10 = tid >> 6,
// so a wave holds exactly the block of 1024 residues numbered by its wave
// index, and so does pass P (NLO 4 lower, or the last-pass grouping above).
// The stages in between act on bits below 10 only, within those blocks.
template <int LGN, int P>
constexpr bool fwd_local() {
  return Shape<LGN>::T <= 64 || (LGN - Pass<LGN, P - 1>::L0 - 4) <= 6;
}

// Synchronisation around an exchange through LDS: the whole workgroup, or
// (LOCAL) only the wave, whose LDS accesses execute in order; the fences keep
// the compiler from moving this lane's accesses across the point.
template <bool LOCAL>
__device__ __forceinline__ void xsync() {
  if constexpr (LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Inverse group indexing: stages L0 .. L0+R-1 act on index bits L0 .. L0+R-1.
template <int LGN, int P, int R0 = inv_r0<LGN>()>
__device__ __forceinline__ int inv_index(int g, int j) {
  using Q = Pass<LGN, P, R0>;
  const int hi = g >> Q::L0, lo = g & ((1 << Q::L0) - 1);
  return (hi << (Q::L0 + Q::R)) | (j << Q::L0) | lo;
}

// Twiddles derived instead of loaded.  In a pass of R stages a group needs
// 2^s twiddles at stage s; all but one of them are products of the group's
// first one and a constant shared by every lane:
//   forward  T[H 2^s + b] = T[H 2^s] T[b]           (b < 2^s: the reversal of
//            H 2^s + b is the sum of the reversals of its two disjoint parts)
//   inverse  V[2^L + jj 2^L0 + lo] = V[2^L + lo] V[2^s + jj]
//            (omega_{2^(L+1)}^-(jj 2^L0 + lo), L = L0 + s, jj < 2^s)
// so a group loads R twiddles per pass and forms the other 2^R - 1 - R as
// exact modular products.  A derived twiddle w = rf_mulmod(base, c) is not
// canonical but |w| <= p (1/2 + p 2^-52) < 0.56 p, and its quotient
// wq = fl(w pinv) (two roundings) is within 0.6 2^-52 (1 + 2^-53) of w / p;
// fl(a wq) is then within |a| 0.9 2^-52 of a w / p, inside the |a| 2^-52
// that rf_mulmod's bound assumes for a canonical w, so every residue bound
// of the transforms holds unchanged.
#ifndef RF_DERIVE_FWD
#define RF_DERIVE_FWD 1
#endif
#ifndef RF_DERIVE_INV
#define RF_DERIVE_INV 1
#endif
#ifndef RF_DERIVE_INV_L0
#define RF_DERIVE_INV_L0 1  // inverse passes starting at stage >= this derive
#endif
#ifndef RF_FWD_BASES_R
#define RF_FWD_BASES_R 2  // forward passes of >= this many stages load bases early
#endif
// Shorter transforms keep loading every twiddle and storing through the
// caches: their tables stay in L1, and a batch of them often keeps its
// spectra in the Infinity Cache between the forward and inverse launches.
#ifndef RF_BIG_LGN
#define RF_BIG_LGN 13
#endif
__device__ __forceinline__ double2 derive_tw(double2 base, double2 c, double p, double pinv) {
  const double w = rf_mulmod(base.x, c.x, c.y, p);
  return make_double2(w, w * pinv);
}

// R forward stages on NG groups held in v[u * 2^R + j].  At stage L0 + s the
// pair (j, j + 2^(R-1-s)) of group (H, lo) uses twiddle T[H 2^s + (j >> (R-s))]
// (T[m] = omega^(rev(m)), shared prefix for all lengths).  A row of a longer
// transform (ROW) prefixes its row index hb to the block index:
// T[(hb 2^L0 + H) 2^s + (j >> (R-s))].  Passes after the first (and rows)
// derive the twiddles with b > 0 from T[H 2^s] and the uniform T[b].
template <int LGN, int P, bool ROW = false>
constexpr bool fwd_derive() {
  return RF_DERIVE_FWD && LGN >= RF_BIG_LGN && (ROW || P > 0) && Pass<LGN, P>::R >= RF_FWD_BASES_R;
}
// The R base twiddles T[H 2^s] of each of the thread's groups in forward
// pass P (loaded before the exchange into the pass, so that they arrive
// while it runs).
template <int LGN, int P, bool ROW = false>
__device__ __forceinline__ void fwd_bases(double2* bs, int tid, rf_gtw __restrict__ tw, int hb = 0) {
  using Q = Pass<LGN, P>;
  constexpr int NLO = LGN - Q::L0 - Q::R;
  if constexpr (fwd_derive<LGN, P, ROW>()) {
#pragma unroll
    for (int u = 0; u < Q::NG; u++) {
      const int H = (ROW ? (hb << Q::L0) : 0) + (fwd_group<LGN, P>(tid, u) >> NLO);
#pragma unroll
      for (int s = 0; s < Q::R; s++) bs[u * Q::R + s] = rf_ld(tw + (H << s));
    }
  }
}

template <int LGN, int P, int S0 = 0, bool ROW = false>
__device__ __forceinline__ void fwd_pass(double* v, int tid, rf_gtw __restrict__ tw,
                                         double p, double pinv, const double2* bs, int hb = 0) {
  using Q = Pass<LGN, P>;
  constexpr int NLO = LGN - Q::L0 - Q::R;
  constexpr bool DERIVE = fwd_derive<LGN, P, ROW>();
#pragma unroll
  for (int s = S0; s < Q::R; s++) {
    const int half = 1 << (Q::R - 1 - s);
#pragma unroll
    for (int u = 0; u < Q::NG; u++) {
      const int g = fwd_group<LGN, P>(tid, u);
      const int H = (ROW ? (hb << Q::L0) : 0) + (g >> NLO);
#pragma unroll
      for (int b = 0; b < (1 << s); b++) {
        // in pass 0 of a whole transform (H = 0) block 0 uses T[0] = 1: no
        // multiplication
        const bool one = !ROW && (P == 0) && (b == 0);
        double2 w;
        if (one)
          w = make_double2(1.0, 0.0);
        else if constexpr (DERIVE)
          w = b == 0 ? bs[u * Q::R + s] : derive_tw(bs[u * Q::R + s], rf_ld(tw + b), p, pinv);
        else
          w = rf_ld(tw + ((H << s) + b));
#pragma unroll
        for (int jj = 0; jj < half; jj++) {
          const int j0 = u * Q::GS + b * 2 * half + jj, j1 = j0 + half;
          const double z = one ? v[j1] : rf_mulmod(v[j1], w.x, w.y, p);
          const double x = v[j0];
          v[j0] = x + z;
          v[j1] = x - z;
        }
      }
    }
  }
}

// R inverse (decimation-in-time) stages: at stage L = L0 + s the pair
// (j, j + 2^s) uses V[2^L + (j mod 2^s) 2^L0 + lo] = omega_{2^(L+1)}^-(i mod 2^L).
// Passes whose groups span at least 32 distinct lo per wave derive the
// twiddles with jj > 0 from V[2^L + lo] and the uniform V[2^s + jj].
template <int LGN, int P, int R0 = inv_r0<LGN>()>
constexpr bool inv_derive() {
  return RF_DERIVE_INV && LGN >= RF_BIG_LGN && P > 0 && Pass<LGN, P, R0>::L0 >= RF_DERIVE_INV_L0 &&
         Pass<LGN, P, R0>::R >= 2;
}
// The R base twiddles V[2^(L0+s) + lo] of each of the thread's groups in
// inverse pass P (loaded before the exchange into the pass).
template <int LGN, int P, int R0 = inv_r0<LGN>()>
__device__ __forceinline__ void inv_bases(double2* bs, int tid, rf_gtw __restrict__ tw) {
  using Q = Pass<LGN, P, R0>;
  if constexpr (inv_derive<LGN, P, R0>()) {
#pragma unroll
    for (int u = 0; u < Q::NG; u++) {
      const int lo = (tid + Shape<LGN>::T * u) & ((1 << Q::L0) - 1);
#pragma unroll
      for (int s = 0; s < Q::R; s++) bs[u * Q::R + s] = rf_ld(tw + ((1 << (Q::L0 + s)) + lo));
    }
  }
}

template <int LGN, int P, int R0 = inv_r0<LGN>()>
__device__ __forceinline__ void inv_pass(double* v, int tid, rf_gtw __restrict__ tw,
                                         double p, double pinv, const double2* bs) {
  using Q = Pass<LGN, P, R0>;
  constexpr bool DERIVE = inv_derive<LGN, P, R0>();
#pragma unroll
  for (int s = 0; s < Q::R; s++) {
    const int half = 1 << s;
    const int len = 1 << (Q::L0 + s);
#pragma unroll
    for (int u = 0; u < Q::NG; u++) {
      const int g = tid + Shape<LGN>::T * u;
      const int lo = g & ((1 << Q::L0) - 1);
#pragma unroll
      for (int jj = 0; jj < half; jj++) {
        // in pass 0 (lo = 0) the pair with jj = 0 uses V[len] = 1
        const bool one = (P == 0) && (jj == 0);
        double2 w;
        if (one)
          w = make_double2(1.0, 0.0);
        else if constexpr (DERIVE)
          w = jj == 0 ? bs[u * Q::R + s] : derive_tw(bs[u * Q::R + s], rf_ld(tw + (half + jj)), p, pinv);
        else
          w = rf_ld(tw + (len + (jj << Q::L0) + lo));
#pragma unroll
        for (int b = 0; b < (Q::GS >> (s + 1)); b++) {
          const int j0 = u * Q::GS + b * 2 * half + jj, j1 = j0 + half;
          const double z = one ? v[j1] : rf_mulmod(v[j1], w.x, w.y, p);
          const double x = v[j0];
          v[j0] = x + z;
          v[j1] = x - z;
        }
      }
    }
  }
}

template <int E>
__device__ __forceinline__ void reduce_all(double* v, double p, double pinv) {
#pragma unroll
  for (int e = 0; e < E; e++) v[e] = rf_reduce(v[e], p, pinv);
}

// Stores of a transform's results, non-temporal when the launch writes more
// than the Infinity Cache holds (nt is uniform: the launcher sets it from
// the launch's output bytes).  Smaller launches keep their results cached
// for the next launch (the inverse, or a column pass that reads a row
// pass's output straight back).
#define RF_NT_BYTES ((size_t)256 << 20)
__device__ __forceinline__ void st1(bool nt, double* p, double x) {
  if (nt)
    rf_st_nt(p, x);
  else
    *p = x;
}
__device__ __forceinline__ void st2(bool nt, double* p, double x, double y) {
  if (nt)
    rf_st_nt2(p, x, y);
  else
    *reinterpret_cast<double2*>(p) = make_double2(x, y);
}

// Reduced residues of the last forward pass to out (in place order).  When
// its groups are consecutive residues (no index bits below its stages) they
// go out as 16-byte pairs.
template <int LGN>
__device__ __forceinline__ void fwd_store(double* __restrict__ out, const double* v, int tid, double p,
                                          double pinv, bool nt) {
  using S = Shape<LGN>;
  using QL = Pass<LGN, S::NP - 1>;
  constexpr int NLO = LGN - QL::L0 - QL::R;
#pragma unroll
  for (int u = 0; u < QL::NG; u++) {
    if constexpr (NLO == 0 && QL::GS >= 2) {
#pragma unroll
      for (int j = 0; j < QL::GS; j += 2)
        st2(nt, out + fwd_index<LGN, S::NP - 1>(fwd_group<LGN, S::NP - 1>(tid, u), j),
                               rf_reduce(v[u * QL::GS + j], p, pinv), rf_reduce(v[u * QL::GS + j + 1], p, pinv));
    } else {
#pragma unroll
      for (int j = 0; j < QL::GS; j++)
        st1(nt, out + fwd_index<LGN, S::NP - 1>(fwd_group<LGN, S::NP - 1>(tid, u), j), rf_reduce(v[u * QL::GS + j], p, pinv));
    }
  }
}

// Passes 1 .. NP-1 of the forward transform: LDS exchange, optional
// reduction, stages.
template <int LGN, int P, bool ROW = false>
__device__ __forceinline__ void fwd_rest(double* v, double* sh, int tid, rf_gtw tw,
                                         double p, double pinv, int hb = 0) {
  if constexpr (P < Shape<LGN>::NP) {
    using Q = Pass<LGN, P>;
    using Qp = Pass<LGN, P - 1>;
    double2 bs[fwd_derive<LGN, P, ROW>() ? Q::NG * Q::R : 1];
    fwd_bases<LGN, P, ROW>(bs, tid, tw, hb);
    if constexpr (Q::REDUCE_BEFORE) reduce_all<Shape<LGN>::E>(v, p, pinv);
    // index bits below the stages of each side: lanes run over that many
    // consecutive residues
    constexpr int SW = lds_map(Shape<LGN>::PAD, Shape<LGN>::T >= 64, Qp::R, Q::R, LGN - Qp::L0 - Qp::R,
                               LGN - Q::L0 - Q::R);
    const int tw0 = fwd_index<LGN, P - 1>(fwd_group<LGN, P - 1>(tid, 0), 0);
#pragma unroll
    for (int u = 0; u < Qp::NG; u++)
#pragma unroll
      for (int j = 0; j < Qp::GS; j++)
        *lds_at<SW>(sh, tw0, fwd_index<LGN, P - 1>(fwd_group<LGN, P - 1>(0, u), j)) = v[u * Qp::GS + j];
    // the exchanges after a wave-local one are wave-local too, so a wave
    // barrier after the reads suffices there as well
    xsync<fwd_local<LGN, P>()>();
    const int tr0 = fwd_index<LGN, P>(fwd_group<LGN, P>(tid, 0), 0);
#pragma unroll
    for (int u = 0; u < Q::NG; u++)
#pragma unroll
      for (int j = 0; j < Q::GS; j++)
        v[u * Q::GS + j] = *lds_at<SW>(sh, tr0, fwd_index<LGN, P>(fwd_group<LGN, P>(0, u), j));
    xsync<fwd_local<LGN, P>()>();
    fwd_pass<LGN, P, 0, ROW>(v, tid, tw, p, pinv, bs, hb);
    fwd_rest<LGN, P + 1, ROW>(v, sh, tid, tw, p, pinv, hb);
  }
}

template <int LGN, int P>
__device__ __forceinline__ void inv_rest(double* v, double* sh, int tid, rf_gtw tw,
                                         double p, double pinv) {
  constexpr int R0 = inv_r0<LGN>();
  if constexpr (P < npasses<LGN, R0>()) {
    using Q = Pass<LGN, P, R0>;
    using Qp = Pass<LGN, P - 1, R0>;
    double2 bs[inv_derive<LGN, P>() ? Q::NG * Q::R : 1];
    inv_bases<LGN, P>(bs, tid, tw);
    if constexpr (Q::REDUCE_BEFORE) reduce_all<Shape<LGN>::E>(v, p, pinv);
    constexpr int SW = lds_map(Shape<LGN>::PAD, Shape<LGN>::T >= 64, Qp::R, Q::R, Qp::L0, Q::L0);
    const int tw0 = inv_index<LGN, P - 1>(tid, 0);
#pragma unroll
    for (int u = 0; u < Qp::NG; u++)
#pragma unroll
      for (int j = 0; j < Qp::GS; j++)
        *lds_at<SW>(sh, tw0, inv_index<LGN, P - 1>(Shape<LGN>::T * u, j)) = v[u * Qp::GS + j];
    __syncthreads();
    const int tr0 = inv_index<LGN, P>(tid, 0);
#pragma unroll
    for (int u = 0; u < Q::NG; u++)
#pragma unroll
      for (int j = 0; j < Q::GS; j++)
        v[u * Q::GS + j] = *lds_at<SW>(sh, tr0, inv_index<LGN, P>(Shape<LGN>::T * u, j));
    __syncthreads();
    inv_pass<LGN, P>(v, tid, tw, p, pinv, bs);
    inv_rest<LGN, P + 1>(v, sh, tid, tw, p, pinv);
  }
}

// Residues of NCH chunks ci[c] (t limbs = 2t 32-bit words each) of an entry
// modulo p, times the scale folded into the pack constants:
// sum_w word_w K_w with K_w = 2^(32 w) S mod p split in 16-bit pieces
// K_w = k0 + k1 2^16 + k2 2^32, the constants holding k0, k1 2^16 and k2 2^32.
// At most 16 words are summed: s0 < 2^52; s1 is a multiple of 2^16 below
// 2^68 and s2 a multiple of 2^32 below 2^84, so all three accumulate
// exactly.  s1 and s2 are reduced on their own (rf_reduce_wide, results
// <= 0.51 p), and s0 + both < 2^53 once more by rf_reduce.  Limb l of all NCH
// chunks is loaded before any is used, so NCH loads are in flight.  When
// every chunk of the wave lies inside the entry the loads need no bounds
// checks; otherwise loads past the entry's end read limb 0 and are
// discarded.  Control flow depends only on the uniform t and that test.
template <int NCH>
__device__ __forceinline__ void pack_chunks(rf_glimbs __restrict__ limbs, uint32_t nlimbs,
                                            const uint32_t (&ci)[NCH], int t, const double (*pk)[3],
                                            double p, double pinv,
                                            double (&out)[NCH]) {
  uint32_t cmax = 0;
#pragma unroll
  for (int c = 0; c < NCH; c++) cmax = cmax > ci[c] ? cmax : ci[c];
  const bool full = __builtin_amdgcn_ballot_w64((uint64_t)(cmax + 1) * (uint32_t)t > nlimbs) == 0;
  double s0[NCH], s1[NCH], s2[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) s0[c] = s1[c] = s2[c] = 0.0;
#pragma unroll
  for (int l = 0; l < RF_TMAX; l++) {
    if (l < t) {
      uint64_t x[NCH];
      if (full) {
#pragma unroll
        for (int c = 0; c < NCH; c++) x[c] = limbs[ci[c] * (uint32_t)t + (uint32_t)l];
      } else {
#pragma unroll
        for (int c = 0; c < NCH; c++) {
          const uint32_t idx = ci[c] * (uint32_t)t + (uint32_t)l;
          const uint64_t y = limbs[idx < nlimbs ? idx : 0];
          x[c] = idx < nlimbs ? y : 0;
        }
      }
      const double a0 = pk[2 * l][0], a1 = pk[2 * l][1], a2 = pk[2 * l][2];
      const double b0 = pk[2 * l + 1][0], b1 = pk[2 * l + 1][1], b2 = pk[2 * l + 1][2];
#pragma unroll
      for (int c = 0; c < NCH; c++) {
        const double lo = (double)(uint32_t)x[c];
        const double hi = (double)(uint32_t)(x[c] >> 32);
        s0[c] = __builtin_fma(hi, b0, __builtin_fma(lo, a0, s0[c]));
        s1[c] = __builtin_fma(hi, b1, __builtin_fma(lo, a1, s1[c]));
        s2[c] = __builtin_fma(hi, b2, __builtin_fma(lo, a2, s2[c]));
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const double r = s0[c] + rf_reduce_wide(s1[c], p, pinv) + rf_reduce_wide(s2[c], p, pinv);
    out[c] = rf_reduce(r, p, pinv);
  }
}

// Packs with a compile-time chunk width T: every limb of a sub-batch of SB
// chunks (at most 16 limbs, or LIMBS on the aligned path) is loaded before
// any is used.  pack_full16: every
// chunk inside the entry, T even and the entry 16-byte aligned, so each chunk
// is T/2 aligned 16-byte loads.  pack_any8: 8-byte loads with the bounds
// checks of pack_chunks.  Same arithmetic and bounds as pack_chunks.
template <int T>
__device__ __forceinline__ double pack_combine(const uint64_t (&x)[T], const double (*pk)[3], double p,
                                               double pinv) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int l = 0; l < T; l++) {
    const double lo = (double)(uint32_t)x[l];
    const double hi = (double)(uint32_t)(x[l] >> 32);
    s0 = __builtin_fma(hi, pk[2 * l + 1][0], __builtin_fma(lo, pk[2 * l][0], s0));
    s1 = __builtin_fma(hi, pk[2 * l + 1][1], __builtin_fma(lo, pk[2 * l][1], s1));
    s2 = __builtin_fma(hi, pk[2 * l + 1][2], __builtin_fma(lo, pk[2 * l][2], s2));
  }
  const double r = s0 + rf_reduce_wide(s1, p, pinv) + rf_reduce_wide(s2, p, pinv);
  return rf_reduce(r, p, pinv);
}

// Chunks per sub-batch: the largest power of two with at most LIMBS limbs
// (at least one chunk), never more than NCH; NCH is a power of two, so
// sub-batches tile it exactly.
template <int NCH, int T, int LIMBS>
constexpr int pack_sb() {
  int sb = 1;
  while (sb * 2 <= NCH && sb * 2 * T <= LIMBS) sb *= 2;
  return sb;
}

template <int NCH, int T, int LIMBS>
__device__ __forceinline__ void pack_full16(rf_glimbs __restrict__ limbs, const uint32_t (&ci)[NCH],
                                            const double (*pk)[3], double p, double pinv,
                                            double (&out)[NCH]) {
  constexpr int SB = pack_sb<NCH, T, LIMBS>();
  static_assert(NCH % SB == 0, "sub-batches must tile the chunks");
#pragma unroll
  for (int c0 = 0; c0 < NCH; c0 += SB) {
    uint64_t x[SB][T];
#pragma unroll
    for (int c = 0; c < SB; c++) {
      rf_glimbs2 src = reinterpret_cast<rf_glimbs2>(limbs + (size_t)ci[c0 + c] * T);
#pragma unroll
      for (int l = 0; l < T / 2; l++) {
        const ulonglong2 y = src[l];
        x[c][2 * l] = y.x;
        x[c][2 * l + 1] = y.y;
      }
    }
#pragma unroll
    for (int c = 0; c < SB; c++) out[c0 + c] = pack_combine<T>(x[c], pk, p, pinv);
  }
}

template <int NCH, int T>
__device__ __forceinline__ void pack_any8(rf_glimbs __restrict__ limbs, uint32_t nlimbs,
                                          const uint32_t (&ci)[NCH], const double (*pk)[3],
                                          double p, double pinv, double (&out)[NCH]) {
  constexpr int SB = pack_sb<NCH, T, 16>();
  static_assert(NCH % SB == 0, "sub-batches must tile the chunks");
#pragma unroll
  for (int c0 = 0; c0 < NCH; c0 += SB) {
    uint64_t x[SB][T];
#pragma unroll
    for (int c = 0; c < SB; c++)
#pragma unroll
      for (int l = 0; l < T; l++) {
        const uint32_t idx = ci[c0 + c] * (uint32_t)T + (uint32_t)l;
        const uint64_t y = limbs[idx < nlimbs ? idx : 0];
        x[c][l] = idx < nlimbs ? y : 0;
      }
#pragma unroll
    for (int c = 0; c < SB; c++) out[c0 + c] = pack_combine<T>(x[c], pk, p, pinv);
  }
}

// T > 0: chunk width fixed at compile time (t == T); T == 0: run-time t.
// FULL_LIMBS: the limbs per sub-batch of the aligned, bounds-free path.
template <int NCH, int T, int FULL_LIMBS = 16>
__device__ __forceinline__ void pack_dispatch(rf_glimbs __restrict__ limbs, uint32_t nlimbs,
                                              const uint32_t (&ci)[NCH], int t, const double (*pk)[3],
                                              double p, double pinv, double (&out)[NCH]) {
  if constexpr (T == 0) {
    pack_chunks<NCH>(limbs, nlimbs, ci, t, pk, p, pinv, out);
  } else {
    uint32_t cmax = 0;
#pragma unroll
    for (int c = 0; c < NCH; c++) cmax = cmax > ci[c] ? cmax : ci[c];
    const bool full = __builtin_amdgcn_ballot_w64((uint64_t)(cmax + 1) * T > nlimbs) == 0;
    if (T % 2 == 0 && full && ((uintptr_t)limbs & 15) == 0)
      pack_full16<NCH, T, FULL_LIMBS>(limbs, ci, pk, p, pinv, out);
    else
      pack_any8<NCH, T>(limbs, nlimbs, ci, pk, p, pinv, out);
  }
}

template <int LGN, int T>
__global__ __launch_bounds__(Shape<LGN>::THREADS, Shape<LGN>::MINB) void k_fwd(RfDevTables tb,
                                                             const RfFwdTask* __restrict__ tasks,
                                                             int ntasks, double* __restrict__ spec, int nt) {
  using S = Shape<LGN>;
  using Q0 = Pass<LGN, 0>;
  extern __shared__ double sh_all[];
  const int sub = threadIdx.x / S::T, tid = threadIdx.x % S::T;
  const int wg = xcd_task(blockIdx.x, gridDim.x, (ntasks + S::G - 1) / S::G);
  if (wg < 0) return;
  const int ti = wg * S::G + sub;
  if (ti >= ntasks) return;  // only whole transforms drop out; no barrier spans them
  const RfFwdTask tk = S::G == 1 ? rf_wave_uniform_load(tasks + ti) : tasks[ti];
  double* sh = sh_all + sub * S::LDSN;
  const RfConsts* C = tb.consts;
  const double p = C->p[tk.q], pinv = C->pinv[tk.q];
  const double(*pk)[3] = C->pk[tk.q][tk.scale];
  rf_gtw tw = (rf_gtw)(tb.tw_fwd + (size_t)tk.q * RF_FWD_TW * 2);
  const double sg = tk.sign < 0 ? -1.0 : 1.0;
  const uint32_t nchunk = tk.sign != 0 ? tk.nchunk : 0u;
  double v[S::E];
  // Pass 0 holds, per group, residues i = j 2^(LGN-R) + g: the upper half of
  // j covers positions >= N/2.  When the entry has at most N/2 chunks those
  // are zero and the first stage (twiddle 1) just copies (x, 0) -> (x, x).
  // chunks per load batch (four with a run-time chunk width, whose pack
  // keeps three accumulators per chunk across all t limbs)
  constexpr int CB = S::E / 2 < 8 ? (S::E / 2 > 0 ? S::E / 2 : 1) : (T == 0 ? 4 : 8);
  // chunks of at most four limbs load a whole batch (up to 32 limbs) before
  // any is used: one memory round trip instead of two, still without spills
  // (k_fwd<13, 4>: 121 VGPRs; 2^13 points, 1-Mbit entries: k_fwd -1.6 %)
  constexpr int PACK_LIMBS = (T >= 1 && T <= 4) ? 32 : 16;
  if (nchunk <= (uint32_t)(S::N / 2) && Q0::GS > 1) {
#pragma unroll
    for (int u = 0; u < Q0::NG; u++)
#pragma unroll
      for (int j0 = 0; j0 < Q0::GS / 2; j0 += CB) {
        uint32_t ci[CB];
        double x[CB];
#pragma unroll
        for (int c = 0; c < CB; c++) ci[c] = (uint32_t)fwd_index<LGN, 0>(tid + S::T * u, j0 + c);
        if (nchunk) {  // a zero entry reads no limbs at all
          pack_dispatch<CB, T, PACK_LIMBS>((rf_glimbs)tk.limbs, tk.nlimbs, ci, tk.t, pk, p, pinv, x);
        } else {
#pragma unroll
          for (int c = 0; c < CB; c++) x[c] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < CB; c++) {
          const double y = ci[c] < nchunk ? sg * x[c] : 0.0;
          v[u * Q0::GS + j0 + c] = y;
          v[u * Q0::GS + j0 + c + Q0::GS / 2] = y;
        }
      }
    fwd_pass<LGN, 0, 1>(v, tid, tw, p, pinv, nullptr);
  } else {
#pragma unroll
    for (int u = 0; u < Q0::NG; u++)
#pragma unroll
      for (int j0 = 0; j0 < Q0::GS; j0 += CB) {
        uint32_t ci[CB];
        double x[CB];
#pragma unroll
        for (int c = 0; c < CB; c++) ci[c] = (uint32_t)fwd_index<LGN, 0>(tid + S::T * u, j0 + c);
        if (nchunk) {
          pack_dispatch<CB, T, PACK_LIMBS>((rf_glimbs)tk.limbs, tk.nlimbs, ci, tk.t, pk, p, pinv, x);
        } else {
#pragma unroll
          for (int c = 0; c < CB; c++) x[c] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < CB; c++) v[u * Q0::GS + j0 + c] = ci[c] < nchunk ? sg * x[c] : 0.0;
      }
    fwd_pass<LGN, 0, 0>(v, tid, tw, p, pinv, nullptr);
  }
  fwd_rest<LGN, 1>(v, sh, tid, tw, p, pinv);
  fwd_store<LGN>(spec + tk.spec_off, v, tid, p, pinv, nt != 0);
}

// Pointwise product sum_l a_l b_l of the thread's E frequencies
// tid + T e (coalesced: consecutive lanes, consecutive frequencies).  D > 0
// fixes the number of terms so that all loads of a batch are issued before
// any is used (|sum| <= 2.4 p needs no intermediate reduction); D == 0 takes
// d at run time and reduces every eighth term (|term| <= 0.6 p); D == -1
// (a task with d = 0) copies one stored spectrum, already a product, in.
template <int LGN, int D>
__device__ __forceinline__ void pointwise(double* v, int tid, rf_gres __restrict__ pa,
                                          rf_gres __restrict__ pb, uint32_t as, uint32_t bs, int d,
                                          double p, double pinv) {
  using S = Shape<LGN>;
  if constexpr (D < 0) {
#pragma unroll
    for (int e = 0; e < S::E; e++) v[e] = pa[tid + S::T * e];
  } else if constexpr (D > 0) {
    constexpr int EB0 = D == 1 ? 16 : (D == 2 ? 8 : 4);
    constexpr int EB = S::E < EB0 ? S::E : EB0;
#pragma unroll
    for (int e0 = 0; e0 < S::E; e0 += EB) {
      double a[D][EB], b[D][EB];
#pragma unroll
      for (int l = 0; l < D; l++)
#pragma unroll
        for (int e = 0; e < EB; e++) {
          a[l][e] = pa[(size_t)l * as + tid + S::T * (e0 + e)];
          b[l][e] = pb[(size_t)l * bs + tid + S::T * (e0 + e)];
        }
#pragma unroll
      for (int e = 0; e < EB; e++) {
        double acc = rf_mulmod_var(a[0][e], b[0][e], p, pinv);
#pragma unroll
        for (int l = 1; l < D; l++) acc += rf_mulmod_var(a[l][e], b[l][e], p, pinv);
        v[e0 + e] = rf_reduce(acc, p, pinv);
      }
    }
  } else {
    constexpr int EB = S::E < 8 ? S::E : 8;
#pragma unroll
    for (int e0 = 0; e0 < S::E; e0 += EB) {
      double acc[EB];
#pragma unroll
      for (int e = 0; e < EB; e++) acc[e] = 0.0;
      rf_gres qa = pa;
      rf_gres qb = pb;
      for (int l = 0; l < d; l++) {
        double a[EB], b[EB];
#pragma unroll
        for (int e = 0; e < EB; e++) {
          a[e] = qa[tid + S::T * (e0 + e)];
          b[e] = qb[tid + S::T * (e0 + e)];
        }
#pragma unroll
        for (int e = 0; e < EB; e++) {
          acc[e] += rf_mulmod_var(a[e], b[e], p, pinv);  // |term| <= 0.6 p
          if ((l & 7) == 7) acc[e] = rf_reduce(acc[e], p, pinv);
        }
        qa += as;
        qb += bs;
      }
#pragma unroll
      for (int e = 0; e < EB; e++) v[e0 + e] = rf_reduce(acc[e], p, pinv);
    }
  }
}

// Pointwise product straight into the first inverse pass's layout: group u
// of the thread holds the 2^R0 consecutive residues (g << R0) + j,
// g = tid + T u, loaded as 16-byte pairs.  D as in pointwise.
template <int LGN, int D, int R0>
__device__ __forceinline__ void pointwise_grp(double* v, int tid, rf_gres __restrict__ pa,
                                              rf_gres __restrict__ pb, uint32_t as, uint32_t bs, int d,
                                              double p, double pinv) {
  using S = Shape<LGN>;
  constexpr int GS = 1 << R0, NG = S::E >> R0, PR = GS / 2;  // 16-byte pairs per group
  constexpr int EB0 = (D == 1 || D == 0) ? 16 : (D == 2 ? 8 : 4);
  constexpr int GB = (EB0 >> R0) > 0 ? (EB0 >> R0) : 1;      // groups per batch
#pragma unroll
  for (int u0 = 0; u0 < NG; u0 += GB) {
    if constexpr (D < 0) {
#pragma unroll
      for (int u = 0; u < GB; u++)
#pragma unroll
        for (int k = 0; k < PR; k++) {
          const double2 a =
              rf_ld(reinterpret_cast<rf_gtw>(pa + (size_t)((tid + S::T * (u0 + u)) << R0) + 2 * k));
          v[(u0 + u) * GS + 2 * k] = a.x;
          v[(u0 + u) * GS + 2 * k + 1] = a.y;
        }
    } else if constexpr (D > 0) {
      double2 a[D][GB][PR], b[D][GB][PR];
#pragma unroll
      for (int l = 0; l < D; l++)
#pragma unroll
        for (int u = 0; u < GB; u++)
#pragma unroll
          for (int k = 0; k < PR; k++) {
            const size_t i = (size_t)((tid + S::T * (u0 + u)) << R0) + 2 * k;
            a[l][u][k] = rf_ld(reinterpret_cast<rf_gtw>(pa + (size_t)l * as + i));
            b[l][u][k] = rf_ld(reinterpret_cast<rf_gtw>(pb + (size_t)l * bs + i));
          }
#pragma unroll
      for (int u = 0; u < GB; u++)
#pragma unroll
        for (int k = 0; k < PR; k++) {
          double x = rf_mulmod_var(a[0][u][k].x, b[0][u][k].x, p, pinv);
          double y = rf_mulmod_var(a[0][u][k].y, b[0][u][k].y, p, pinv);
#pragma unroll
          for (int l = 1; l < D; l++) {
            x += rf_mulmod_var(a[l][u][k].x, b[l][u][k].x, p, pinv);
            y += rf_mulmod_var(a[l][u][k].y, b[l][u][k].y, p, pinv);
          }
          v[(u0 + u) * GS + 2 * k] = rf_reduce(x, p, pinv);
          v[(u0 + u) * GS + 2 * k + 1] = rf_reduce(y, p, pinv);
        }
    } else {
      double acc[GB][GS];
#pragma unroll
      for (int u = 0; u < GB; u++)
#pragma unroll
        for (int j = 0; j < GS; j++) acc[u][j] = 0.0;
      for (int l = 0; l < d; l++) {
        double2 a[GB][PR], b[GB][PR];
#pragma unroll
        for (int u = 0; u < GB; u++)
#pragma unroll
          for (int k = 0; k < PR; k++) {
            const size_t i = (size_t)((tid + S::T * (u0 + u)) << R0) + 2 * k;
            a[u][k] = rf_ld(reinterpret_cast<rf_gtw>(pa + (size_t)l * as + i));
            b[u][k] = rf_ld(reinterpret_cast<rf_gtw>(pb + (size_t)l * bs + i));
          }
#pragma unroll
        for (int u = 0; u < GB; u++)
#pragma unroll
          for (int k = 0; k < PR; k++) {
            acc[u][2 * k] += rf_mulmod_var(a[u][k].x, b[u][k].x, p, pinv);  // |term| <= 0.6 p
            acc[u][2 * k + 1] += rf_mulmod_var(a[u][k].y, b[u][k].y, p, pinv);
            if ((l & 7) == 7) {
              acc[u][2 * k] = rf_reduce(acc[u][2 * k], p, pinv);
              acc[u][2 * k + 1] = rf_reduce(acc[u][2 * k + 1], p, pinv);
            }
          }
      }
#pragma unroll
      for (int u = 0; u < GB; u++)
#pragma unroll
        for (int j = 0; j < GS; j++) v[(u0 + u) * GS + j] = rf_reduce(acc[u][j], p, pinv);
    }
  }
}

template <int LGN>
__global__ __launch_bounds__(Shape<LGN>::THREADS, Shape<LGN>::MINB) void k_inv(RfDevTables tb,
                                                             const RfInvTask* __restrict__ tasks,
                                                             int ntasks, const double* __restrict__ aspec,
                                                             const double* __restrict__ bspec,
                                                             double* __restrict__ coef, int nt) {
  using S = Shape<LGN>;
  extern __shared__ double sh_all[];
  const int sub = threadIdx.x / S::T, tid = threadIdx.x % S::T;
  const int wg = xcd_task(blockIdx.x, gridDim.x, (ntasks + S::G - 1) / S::G);
  if (wg < 0) return;
  const int ti = wg * S::G + sub;
  if (ti >= ntasks) return;
  const RfInvTask tk = S::G == 1 ? rf_wave_uniform_load(tasks + ti) : tasks[ti];
  double* sh = sh_all + sub * S::LDSN;
  const RfConsts* C = tb.consts;
  const double p = C->p[tk.q], pinv = C->pinv[tk.q];
  rf_gtw tw = (rf_gtw)(tb.tw_inv + (size_t)tk.q * RF_INV_TW * 2);
  double v[S::E];
  const rf_gres pa = (rf_gres)((tk.a_base ? tk.a_base : aspec) + tk.a_off);
  const rf_gres pb = (rf_gres)(bspec + tk.b_off);
  constexpr int R0 = inv_r0<LGN>();
  if constexpr (R0 <= 2) {
    // the first pass's groups of 2 or 4 consecutive frequencies come
    // straight from global memory as 16-byte pairs
    switch (tk.d) {
      case 0: pointwise_grp<LGN, -1, R0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 0, p, pinv); break;
      case 1: pointwise_grp<LGN, 1, R0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 1, p, pinv); break;
      case 2: pointwise_grp<LGN, 2, R0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 2, p, pinv); break;
      case 3: pointwise_grp<LGN, 3, R0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 3, p, pinv); break;
      case 4: pointwise_grp<LGN, 4, R0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 4, p, pinv); break;
      default: pointwise_grp<LGN, 0, R0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, tk.d, p, pinv); break;
    }
  } else {
    // Pointwise product in coalesced order (consecutive lanes, consecutive
    // frequencies) into LDS; pass 0 then reads its groups of 16 consecutive
    // frequencies from there.
    switch (tk.d) {
      case 0: pointwise<LGN, -1>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 0, p, pinv); break;
      case 1: pointwise<LGN, 1>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 1, p, pinv); break;
      case 2: pointwise<LGN, 2>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 2, p, pinv); break;
      case 3: pointwise<LGN, 3>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 3, p, pinv); break;
      case 4: pointwise<LGN, 4>(v, tid, pa, pb, tk.a_stride, tk.b_stride, 4, p, pinv); break;
      default: pointwise<LGN, 0>(v, tid, pa, pb, tk.a_stride, tk.b_stride, tk.d, p, pinv); break;
    }
    if constexpr (S::T > 1) {
      using Q0 = Pass<LGN, 0, R0>;
#pragma unroll
      for (int e = 0; e < S::E; e++) sh[lds_pad(tid + S::T * e)] = v[e];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < Q0::NG; u++)
#pragma unroll
        for (int j = 0; j < Q0::GS; j++) v[u * Q0::GS + j] = sh[lds_pad(inv_index<LGN, 0>(tid + S::T * u, j))];
      __syncthreads();
    }
  }
  inv_pass<LGN, 0>(v, tid, tw, p, pinv, nullptr);
  inv_rest<LGN, 1>(v, sh, tid, tw, p, pinv);
  constexpr int NPI = npasses<LGN, R0>();
  using QL = Pass<LGN, NPI - 1, R0>;
  double* out = coef + tk.out_off;
#pragma unroll
  for (int u = 0; u < QL::NG; u++)
#pragma unroll
    for (int j = 0; j < QL::GS; j++)
      st1(nt, out + inv_index<LGN, NPI - 1>(tid + S::T * u, j), rf_reduce(v[u * QL::GS + j], p, pinv));
}

// ---- twiddle tables, generated on the device at set-up ----
//
// Forward T[m] = omega^rev(m) (rev over RF_LGN_MAX - 1 bits), inverse
// V[len + j] = omega_{2 len}^-j = omi^(j 2^(RF_LGN_MAX - 1 - log2 len)), with
// omega of order 2^RF_LGN_MAX and omi its inverse; each entry is (w, fl(w / p))
// with w in [0, p), computed by square-and-multiply with exact modular
// products of canonical operands.
__device__ __forceinline__ double tw_pow(double b, uint32_t e, double p, double pinv) {
  double r = 1.0;
  while (e) {
    if (e & 1) r = rf_canon(rf_mulmod_var(r, b, p, pinv), p);
    b = rf_canon(rf_mulmod_var(b, b, p, pinv), p);
    e >>= 1;
  }
  return r;
}

__global__ __launch_bounds__(256) void k_twiddles(const RfConsts* __restrict__ C, const double* __restrict__ om,
                                                   const double* __restrict__ omi, double* __restrict__ fwd,
                                                   double* __restrict__ inv, int lgmax) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const int q = blockIdx.y;
  const double p = C->p[q], pinv = C->pinv[q];
  const uint32_t nf = 1u << (lgmax - 1), ni = 1u << lgmax;
  if (i < nf) {
    const uint32_t rev = __builtin_bitreverse32(i) >> (32 - (lgmax - 1));
    const double w = tw_pow(om[q], rev, p, pinv);
    double* F = fwd + ((size_t)q * nf + i) * 2;
    F[0] = w;
    F[1] = w / p;
  }
  if (i < ni) {
    double w = 0.0;
    if (i) {
      const int e = 31 - __builtin_clz(i);  // len = 2^e
      const uint32_t j = i - (1u << e);
      w = tw_pow(omi[q], j << (lgmax - 1 - e), p, pinv);
    }
    double* V = inv + ((size_t)q * ni + i) * 2;
    V[0] = w;
    V[1] = i ? w / p : 0.0;
  }
}

// The twiddle table of prime q for a transform of 2^lgN points: the set-up
// table up to 2^RF_LGN_TW, the middle one up to 2^RF_LGN_MID, the long one
// above (they agree on every index they share).
__device__ __forceinline__ rf_gtw fwd_table(const RfDevTables& tb, int q, int lgN) {
  return lgN > RF_LGN_MID  ? (rf_gtw)(tb.tw_fwd_long + (size_t)q * RF_FWD_TW_LONG * 2)
         : lgN > RF_LGN_TW ? (rf_gtw)(tb.tw_fwd_mid + (size_t)q * RF_FWD_TW_MID * 2)
                           : (rf_gtw)(tb.tw_fwd + (size_t)q * RF_FWD_TW * 2);
}
__device__ __forceinline__ rf_gtw inv_table(const RfDevTables& tb, int q, int lgN) {
  return lgN > RF_LGN_MID  ? (rf_gtw)(tb.tw_inv_long + (size_t)q * RF_INV_TW_LONG * 2)
         : lgN > RF_LGN_TW ? (rf_gtw)(tb.tw_inv_mid + (size_t)q * RF_INV_TW_MID * 2)
                           : (rf_gtw)(tb.tw_inv + (size_t)q * RF_INV_TW * 2);
}
// Whether the tables a transform of 2^lgN points reads have been made.
static bool tables_made(const RfDevTables& tb, int lgN) {
  if (lgN > RF_LGN_MID) return tb.tw_fwd_long && tb.tw_inv_long;
  if (lgN > RF_LGN_TW) return tb.tw_fwd_mid && tb.tw_inv_mid;
  return tb.tw_fwd && tb.tw_inv;
}

// ---- transforms longer than one workgroup: N = 2^(RF_LGN_ROW + N1) ----

// Forward column pass: thread `col` packs chunks row 2^lgc + col (rows
// 0 .. 2^N1 - 1, lgc = lgN - N1) and runs the first N1 stages, whose pairs
// differ in the row bits only; at stage s the pair (hi, hi + 2^(N1-1-s)) uses
// T[hi >> (N1-s)], the same for every column, so the whole pass is one
// 2^N1-point transform per column.  Output in place, reduced.
template <int N1, int T>
__global__ __launch_bounds__(RF_COL_THREADS, 4) void k_fwd_col(RfDevTables tb,
                                                            const RfFwdTask* __restrict__ tasks,
                                                            int ntasks, double* __restrict__ spec, int lgc) {
  constexpr int ROWS = 1 << N1;
  const int COLS = 1 << lgc, BPT = COLS / RF_COL_THREADS;
  const uint32_t N = (uint32_t)ROWS * (uint32_t)COLS;
  const int w = xcd_task(blockIdx.x, gridDim.x, ntasks * BPT);
  if (w < 0) return;
  const RfFwdTask tk = rf_wave_uniform_load(tasks + w / BPT);
  const int col = (w % BPT) * RF_COL_THREADS + (int)threadIdx.x;
  const RfConsts* C = tb.consts;
  const double p = C->p[tk.q], pinv = C->pinv[tk.q];
  const double(*pk)[3] = C->pk[tk.q][tk.scale];
  rf_gtw tw = (rf_gtw)(tb.tw_fwd + (size_t)tk.q * RF_FWD_TW * 2);
  const double sg = tk.sign < 0 ? -1.0 : 1.0;
  const uint32_t nchunk = tk.sign != 0 ? tk.nchunk : 0u;
  const bool half = nchunk <= N / 2;  // rows >= ROWS/2 are zero: stage 0 copies
  constexpr int CB = ROWS / 2 < 8 ? ROWS / 2 : 8;
  double v[ROWS];
#pragma unroll
  for (int r0 = 0; r0 < ROWS; r0 += CB) {
    uint32_t ci[CB];
    double x[CB];
#pragma unroll
    for (int c = 0; c < CB; c++) ci[c] = (uint32_t)(r0 + c) * COLS + (uint32_t)col;
    if (nchunk && !(half && r0 >= ROWS / 2)) {
      pack_dispatch<CB, T>((rf_glimbs)tk.limbs, tk.nlimbs, ci, tk.t, pk, p, pinv, x);
    } else {
#pragma unroll
      for (int c = 0; c < CB; c++) x[c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < CB; c++) v[r0 + c] = ci[c] < nchunk ? sg * x[c] : 0.0;
  }
  if (half) {
#pragma unroll
    for (int j = 0; j < ROWS / 2; j++) v[j + ROWS / 2] = v[j];
  } else {
#pragma unroll
    for (int j = 0; j < ROWS / 2; j++) {  // stage 0: T[0] = 1
      const double x = v[j], y = v[j + ROWS / 2];
      v[j] = x + y;
      v[j + ROWS / 2] = x - y;
    }
  }
#pragma unroll
  for (int st = 1; st < N1; st++) {
    const int hf = ROWS >> (st + 1);
#pragma unroll
    for (int b = 0; b < (1 << st); b++) {
      const double2 wt = rf_ld(tw + (b));
#pragma unroll
      for (int jj = 0; jj < hf; jj++) {
        const int j0 = b * 2 * hf + jj, j1 = j0 + hf;
        const double z = b == 0 ? v[j1] : rf_mulmod(v[j1], wt.x, wt.y, p);
        const double x = v[j0];
        v[j0] = x + z;
        v[j1] = x - z;
      }
    }
  }
  double* out = spec + tk.spec_off;
#pragma unroll
  for (int j = 0; j < ROWS; j++) out[(size_t)j * COLS + col] = rf_reduce(v[j], p, pinv);
}

// Forward row pass: the last 13 stages of one row, in LDS, with the row index
// prefixed to every twiddle index.  Output in place (bit-reversed overall).
__global__ __launch_bounds__(Shape<RF_LGN_ROW>::THREADS) void k_fwd_row(RfDevTables tb,
                                                                         const RfRowTask* __restrict__ tasks,
                                                                         int ntasks, double* __restrict__ spec,
                                                                         int nt) {
  constexpr int LGN = RF_LGN_ROW;
  using S = Shape<LGN>;
  using Q0 = Pass<LGN, 0>;
  static_assert(S::G == 1, "one row per workgroup");
  extern __shared__ double sh_all[];
  const int tid = threadIdx.x;
  const int wg = xcd_task(blockIdx.x, gridDim.x, ntasks);
  if (wg < 0) return;
  const RfRowTask tk = rf_wave_uniform_load(tasks + wg);
  const RfConsts* C = tb.consts;
  const double p = C->p[tk.q], pinv = C->pinv[tk.q];
  rf_gtw tw = fwd_table(tb, tk.q, tk.lg);
  double* a = spec + tk.off;
  double v[S::E];
  double2 bs[fwd_derive<LGN, 0, true>() ? Q0::NG * Q0::R : 1];
  fwd_bases<LGN, 0, true>(bs, tid, tw, (int)tk.row);
#pragma unroll
  for (int u = 0; u < Q0::NG; u++)
#pragma unroll
    for (int j = 0; j < Q0::GS; j++) v[u * Q0::GS + j] = a[fwd_index<LGN, 0>(tid + S::T * u, j)];
  fwd_pass<LGN, 0, 0, true>(v, tid, tw, p, pinv, bs, (int)tk.row);
  fwd_rest<LGN, 1, true>(v, sh_all, tid, tw, p, pinv, (int)tk.row);
  fwd_store<LGN>(a, v, tid, p, pinv, nt != 0);
}

// Forward middle pass of transforms above 2^(13+4) points: after the passes
// over the row bits above b0 + NB, the next NB stages act on row bits
// b0 + NB - 1 .. b0.  A thread holds rows hi 2^(b0+NB) + j 2^b0 + lo of one
// column for j < 2^NB, hi being the processed row bits and lo the ones still
// to come; at stage s of the pass the pair (j, j + 2^(NB-1-s)) uses
// T[(hi << s) + (j >> (NB - s))] (the top stage-count bits of the index; lo
// does not enter).  Output in place, reduced.
template <int NB>
__global__ __launch_bounds__(RF_COL_THREADS, 4) void k_fwd_mid(RfDevTables tb, const RfFwdTask* __restrict__ tasks,
                                                            int ntasks, double* __restrict__ spec, int lgN, int b0) {
  constexpr int J = 1 << NB;
  const int lgt = lgN - NB;                       // threads per transform: 2^lgt
  const int BPT = (1 << lgt) / RF_COL_THREADS;
  const int w = xcd_task(blockIdx.x, gridDim.x, ntasks * BPT);
  if (w < 0) return;
  const RfFwdTask tk = rf_wave_uniform_load(tasks + w / BPT);
  const int th = (w % BPT) * RF_COL_THREADS + (int)threadIdx.x;
  const int col = th & ((1 << RF_LGN_ROW) - 1);
  const int lo = (th >> RF_LGN_ROW) & ((1 << b0) - 1), hi = th >> (RF_LGN_ROW + b0);
  const RfConsts* C = tb.consts;
  const double p = C->p[tk.q], pinv = C->pinv[tk.q];
  rf_gtw tw = fwd_table(tb, tk.q, lgN);
  const size_t stride = (size_t)1 << (RF_LGN_ROW + b0);
  double* a = spec + tk.spec_off + ((((size_t)hi << (b0 + NB)) + lo) << RF_LGN_ROW) + col;
  double v[J];
#pragma unroll
  for (int j = 0; j < J; j++) v[j] = a[(size_t)j * stride];
#pragma unroll
  for (int s = 0; s < NB; s++) {
    const int hf = J >> (s + 1);
#pragma unroll
    for (int b = 0; b < (1 << s); b++) {
      const double2 wt = rf_ld(tw + ((hi << s) + b));
#pragma unroll
      for (int jj = 0; jj < hf; jj++) {
        const int j0 = b * 2 * hf + jj, j1 = j0 + hf;
        const double z = rf_mulmod(v[j1], wt.x, wt.y, p);
        const double x = v[j0];
        v[j0] = x + z;
        v[j1] = x - z;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < J; j++) a[(size_t)j * stride] = rf_reduce(v[j], p, pinv);
}

// Inverse column pass over row bits b0 .. b0+NB-1: after the row pass (k_inv
// on each row) and the passes over the row bits below b0, stage L = 13 + b0 + s
// pairs rows (R, R + 2^(b0+s)); a thread holds rows hi 2^(b0+NB) + j 2^b0 + lo
// of one column for j < 2^NB, and the pair (j, j + 2^s) uses
// V[2^L + ((j mod 2^s) 2^b0 + lo) 2^13 + col] (omega_{2^(L+1)} to minus the
// index mod 2^L).  Output in place, natural order, reduced.
template <int NB>
__global__ __launch_bounds__(RF_COL_THREADS, 4) void k_inv_col(RfDevTables tb,
                                                            const RfRowTask* __restrict__ tasks,
                                                            int ntasks, double* __restrict__ coef, int lgN, int b0) {
  constexpr int J = 1 << NB;
  const int lgt = lgN - NB;
  const int BPT = (1 << lgt) / RF_COL_THREADS;
  const int w = xcd_task(blockIdx.x, gridDim.x, ntasks * BPT);
  if (w < 0) return;
  const RfRowTask tk = rf_wave_uniform_load(tasks + w / BPT);
  const int th = (w % BPT) * RF_COL_THREADS + (int)threadIdx.x;
  const int col = th & ((1 << RF_LGN_ROW) - 1);
  const int lo = (th >> RF_LGN_ROW) & ((1 << b0) - 1), hi = th >> (RF_LGN_ROW + b0);
  const RfConsts* C = tb.consts;
  const double p = C->p[tk.q], pinv = C->pinv[tk.q];
  rf_gtw tw = inv_table(tb, tk.q, lgN);
  double RF_GLOBAL* a = (double RF_GLOBAL*)(coef + tk.off + ((((size_t)hi << (b0 + NB)) + lo) << RF_LGN_ROW) + col);
  const size_t stride = (size_t)1 << (RF_LGN_ROW + b0);
  const int low = (lo << RF_LGN_ROW) + col;
  double v[J];
#pragma unroll
  for (int j = 0; j < J; j++) v[j] = a[(size_t)j * stride];
#pragma unroll
  for (int st = 0; st < NB; st++) {
    const int hf = 1 << st;
    const int L = RF_LGN_ROW + b0 + st;
#pragma unroll
    for (int jj = 0; jj < hf; jj++) {
      const double2 wt = rf_ld(tw + ((1 << L) + (jj << (RF_LGN_ROW + b0)) + low));
#pragma unroll
      for (int b = 0; b < (J >> (st + 1)); b++) {
        const int j0 = b * 2 * hf + jj, j1 = j0 + hf;
        const double z = rf_mulmod(v[j1], wt.x, wt.y, p);
        const double x = v[j0];
        v[j0] = x + z;
        v[j1] = x - z;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < J; j++) a[(size_t)j * stride] = rf_reduce(v[j], p, pinv);
}

// Opts kernel fn into `lds` bytes of dynamic LDS when that exceeds 64 KiB,
// once per device (`done` holds one bit per device id below 64).
void lds_optin(const void* fn, size_t lds, std::atomic<uint64_t>& done) {
  if (lds <= 65536) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint64_t bit = 1ull << (dev & 63);
  if (done.load(std::memory_order_relaxed) & bit) return;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  done.fetch_or(bit, std::memory_order_relaxed);
}

template <int LGN, int T>
hipError_t launch_fwd_t(const RfDevTables& tb, const RfFwdTask* tasks, int ntasks, double* spec,
                        hipStream_t st) {
  using S = Shape<LGN>;
  const int wgs = (ntasks + S::G - 1) / S::G;
  const int grid = (wgs + 7) & ~7;
  const size_t lds = (size_t)S::G * S::LDSN * sizeof(double);
  static std::atomic<uint64_t> optin{0};
  lds_optin((const void*)k_fwd<LGN, T>, lds, optin);
  const int nt = (size_t)ntasks * S::N * sizeof(double) > RF_NT_BYTES;
  hipLaunchKernelGGL((k_fwd<LGN, T>), dim3(grid), dim3(S::THREADS), lds, st, tb, tasks, ntasks, spec, nt);
  return hipGetLastError();
}

// Transform lengths from 2^RF_FWD_TSPEC on get one kernel per chunk width.
template <int LGN>
hipError_t launch_fwd_l(const RfDevTables& tb, int t, const RfFwdTask* tasks, int ntasks, double* spec,
                        hipStream_t st) {
  if constexpr (LGN < RF_FWD_TSPEC) {
    return launch_fwd_t<LGN, 0>(tb, tasks, ntasks, spec, st);
  } else {
    switch (t) {
      case 1: return launch_fwd_t<LGN, 1>(tb, tasks, ntasks, spec, st);
      case 2: return launch_fwd_t<LGN, 2>(tb, tasks, ntasks, spec, st);
      case 3: return launch_fwd_t<LGN, 3>(tb, tasks, ntasks, spec, st);
      case 4: return launch_fwd_t<LGN, 4>(tb, tasks, ntasks, spec, st);
      case 5: return launch_fwd_t<LGN, 5>(tb, tasks, ntasks, spec, st);
      case 6: return launch_fwd_t<LGN, 6>(tb, tasks, ntasks, spec, st);
      case 7: return launch_fwd_t<LGN, 7>(tb, tasks, ntasks, spec, st);
      case 8: return launch_fwd_t<LGN, 8>(tb, tasks, ntasks, spec, st);
      default: return hipErrorInvalidValue;
    }
  }
}

template <int LGN>
hipError_t launch_inv_t(const RfDevTables& tb, const RfInvTask* tasks, int ntasks,
                        const double* aspec, const double* bspec, double* coef, hipStream_t st) {
  using S = Shape<LGN>;
  const int wgs = (ntasks + S::G - 1) / S::G;
  const int grid = (wgs + 7) & ~7;
  const size_t lds = (size_t)S::G * S::LDSN * sizeof(double);
  static std::atomic<uint64_t> optin{0};
  lds_optin((const void*)k_inv<LGN>, lds, optin);
  const int nt = (size_t)ntasks * S::N * sizeof(double) > RF_NT_BYTES;
  hipLaunchKernelGGL(k_inv<LGN>, dim3(grid), dim3(S::THREADS), lds, st, tb, tasks, ntasks, aspec, bspec,
                     coef, nt);
  return hipGetLastError();
}

}  // namespace

hipError_t rf_launch_fwd(const RfDevTables& tb, int lgN, int t, const RfFwdTask* tasks, int ntasks,
                         double* spec, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  switch (lgN) {
#define RF_CASE(L) \
  case L:          \
    return launch_fwd_l<L>(tb, t, tasks, ntasks, spec, st);
    RF_CASE(1) RF_CASE(2) RF_CASE(3) RF_CASE(4) RF_CASE(5) RF_CASE(6) RF_CASE(7)
    RF_CASE(8) RF_CASE(9) RF_CASE(10) RF_CASE(11) RF_CASE(12) RF_CASE(13) RF_CASE(14)
#undef RF_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t rf_launch_inv(const RfDevTables& tb, int lgN, const RfInvTask* tasks, int ntasks,
                         const double* aspec, const double* bspec, double* coef, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  switch (lgN) {
#define RF_CASE(L) \
  case L:          \
    return launch_inv_t<L>(tb, tasks, ntasks, aspec, bspec, coef, st);
    RF_CASE(1) RF_CASE(2) RF_CASE(3) RF_CASE(4) RF_CASE(5) RF_CASE(6) RF_CASE(7)
    RF_CASE(8) RF_CASE(9) RF_CASE(10) RF_CASE(11) RF_CASE(12) RF_CASE(13) RF_CASE(14)
#undef RF_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template <int N1>
static hipError_t launch_fwd_col_n(int t, dim3 grid, dim3 blk, const RfDevTables& tb, const RfFwdTask* tasks, int ntasks,
                            double* spec, int lgc, hipStream_t st) {
  switch (t) {
#define RF_TC(TT)                                                                                  \
  case TT:                                                                                         \
    hipLaunchKernelGGL((k_fwd_col<N1, TT>), grid, blk, 0, st, tb, tasks, ntasks, spec, lgc); \
    return hipGetLastError();
    RF_TC(1) RF_TC(2) RF_TC(3) RF_TC(4) RF_TC(5) RF_TC(6) RF_TC(7) RF_TC(8)
#undef RF_TC
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t rf_launch_twiddles(const RfConsts* consts, const double* om, const double* omi, int nprimes, int lgmax,
                              double* fwd, double* inv, hipStream_t st) {
  if (lgmax < 2 || lgmax > RF_LGN_MAX) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(((1u << lgmax) + 255) / 256), (unsigned)nprimes);
  hipLaunchKernelGGL(k_twiddles, grid, dim3(256), 0, st, consts, om, omi, fwd, inv, lgmax);
  return hipGetLastError();
}

// n1 = lgN - 13 row bits: the pack-fused column pass takes the top
// na = n1 - 4 floor((n1 - 1) / 4) (1 .. 4) of them, then middle passes take
// 4 at a time down to bit 0.
hipError_t rf_launch_fwd_col(const RfDevTables& tb, int lgN, int t, const RfFwdTask* tasks, int ntasks,
                             double* spec, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  const int n1 = lgN - RF_LGN_ROW;
  if (n1 < 1 || lgN > RF_LGN_MAX) return hipErrorInvalidValue;
  if (!tables_made(tb, lgN)) return hipErrorInvalidValue;
  const int na = n1 - 4 * ((n1 - 1) / 4);  // bits of the pack-fused column pass
  const int lgc = lgN - na;
  const dim3 grid(ntasks * ((1 << lgc) / RF_COL_THREADS)), blk(RF_COL_THREADS);
  hipError_t e;
  switch (na) {
    case 1: e = launch_fwd_col_n<1>(t, grid, blk, tb, tasks, ntasks, spec, lgc, st); break;
    case 2: e = launch_fwd_col_n<2>(t, grid, blk, tb, tasks, ntasks, spec, lgc, st); break;
    case 3: e = launch_fwd_col_n<3>(t, grid, blk, tb, tasks, ntasks, spec, lgc, st); break;
    case 4: e = launch_fwd_col_n<4>(t, grid, blk, tb, tasks, ntasks, spec, lgc, st); break;
    default: return hipErrorInvalidValue;
  }
  const dim3 grid2(ntasks * ((1 << (lgN - 4)) / RF_COL_THREADS));
  for (int b0 = n1 - na - 4; e == hipSuccess && b0 >= 0; b0 -= 4) {
    hipLaunchKernelGGL(k_fwd_mid<4>, grid2, blk, 0, st, tb, tasks, ntasks, spec, lgN, b0);
    e = hipGetLastError();
  }
  return e;
}

hipError_t rf_launch_fwd_row(const RfDevTables& tb, const RfRowTask* tasks, int ntasks, double* spec,
                             hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  using S = Shape<RF_LGN_ROW>;
  const int grid = (ntasks + 7) & ~7;
  const int nt = (size_t)ntasks * S::N * sizeof(double) > RF_NT_BYTES;
  const size_t lds = (size_t)S::LDSN * sizeof(double);
  static std::atomic<uint64_t> optin{0};
  lds_optin((const void*)k_fwd_row, lds, optin);
  hipLaunchKernelGGL(k_fwd_row, dim3(grid), dim3(S::THREADS), lds, st, tb, tasks,
                     ntasks, spec, nt);
  return hipGetLastError();
}

// One inverse column pass over row bits b0 .. b0 + nb - 1.
static hipError_t launch_inv_col_pass(const RfDevTables& tb, int lgN, int b0, int nb, const RfRowTask* tasks,
                                      int ntasks, double* coef, hipStream_t st) {
  const dim3 grid(ntasks * ((1 << (lgN - nb)) / RF_COL_THREADS)), blk(RF_COL_THREADS);
  switch (nb) {
#define RF_CASE(NB) \
  case NB:          \
    hipLaunchKernelGGL(k_inv_col<NB>, grid, blk, 0, st, tb, tasks, ntasks, coef, lgN, b0); \
    return hipGetLastError();
    RF_CASE(1) RF_CASE(2) RF_CASE(3) RF_CASE(4)
#undef RF_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t rf_launch_inv_col(const RfDevTables& tb, int lgN, const RfRowTask* tasks, int ntasks,
                             double* coef, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  const int n1 = lgN - RF_LGN_ROW;
  if (n1 < 1 || lgN > RF_LGN_MAX) return hipErrorInvalidValue;
  if (!tables_made(tb, lgN)) return hipErrorInvalidValue;
  hipError_t e = hipSuccess;
  for (int b0 = 0; e == hipSuccess && b0 < n1; b0 += 4)
    e = launch_inv_col_pass(tb, lgN, b0, n1 - b0 < 4 ? n1 - b0 : 4, tasks, ntasks, coef, st);
  return e;
}
