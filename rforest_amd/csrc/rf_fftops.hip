// rf_fftops.hip -- element-wise arithmetic on stored spectra (gfx950), the
// device side of the general transform API (add, sub, neg, set, products and
// matrix products of values held in the transform domain).
//
// Residue i of a value lies at [q][n] with q = i >> lgN; every stored residue
// satisfies |x| <= p/2 + 8.  A value holds 2^-lgN times the transform of its
// integer (the forward pass folds the factor into the pack constants), so the
// inverse transform needs no scaling, sums stay in that form, and a product
// X Y (two factors 2^-lgN) is multiplied by 2^lgN mod p once.  All kernels are
// flat coalesced streams, bound by HBM bandwidth.
#include "rf_field.h"
#include "rf_kernels.h"

namespace {

constexpr int EW_THREADS = 256;

// x 2^lgN mod p, then reduced: 2^lgN < p, wq = fl(2^lgN / p).
__device__ __forceinline__ double rescale(double x, double n, double nq, double p, double pinv) {
  return rf_reduce(rf_mulmod(x, n, nq, p), p, pinv);
}

template <int OP>
__global__ __launch_bounds__(EW_THREADS) void k_ew(RfDevTables tb, int lgN, uint64_t total,
                                                   const double* __restrict__ a, const double* __restrict__ b,
                                                   double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * EW_THREADS + threadIdx.x;
  if (i >= total) return;
  const int q = (int)(i >> lgN);
  const double p = tb.consts->p[q], pinv = tb.consts->pinv[q];
  const double x = a[i];
  double r;
  if constexpr (OP == RF_EW_SET) {
    r = x;
  } else if constexpr (OP == RF_EW_NEG) {
    r = -x;
  } else if constexpr (OP == RF_EW_ADD) {
    r = rf_reduce(x + b[i], p, pinv);
  } else if constexpr (OP == RF_EW_SUB) {
    r = rf_reduce(x - b[i], p, pinv);
  } else {
    const double n = (double)((uint64_t)1 << lgN);
    r = rescale(rf_mulmod_var(x, b[i], p, pinv), n, n / p, p, pinv);
  }
  out[i] = r;
}

// One output matrix entry per blockIdx.y; the sum over l of n2 products
// (each |.| <= 0.6 p) is reduced every eighth term.
__global__ __launch_bounds__(EW_THREADS) void k_ew_matmul(RfDevTables tb, int lgN, uint64_t total,
                                                          const double* const* __restrict__ a,
                                                          const double* const* __restrict__ b,
                                                          double* const* __restrict__ out, int n2, int n3) {
  const uint64_t i = (uint64_t)blockIdx.x * EW_THREADS + threadIdx.x;
  if (i >= total) return;
  const int row = (int)blockIdx.y / n3, col = (int)blockIdx.y % n3;
  const int q = (int)(i >> lgN);
  const double p = tb.consts->p[q], pinv = tb.consts->pinv[q];
  double acc = 0.0;
  for (int l = 0; l < n2; l++) {
    acc += rf_mulmod_var(a[row * n2 + l][i], b[l * n3 + col][i], p, pinv);
    if ((l & 7) == 7) acc = rf_reduce(acc, p, pinv);
  }
  const double n = (double)((uint64_t)1 << lgN);
  out[blockIdx.y][i] = rescale(rf_reduce(acc, p, pinv), n, n / p, p, pinv);
}

}  // namespace

hipError_t rf_launch_ew(const RfDevTables& tb, int op, int k, int lgN, const double* a, const double* b,
                        double* out, hipStream_t st) {
  const uint64_t total = (uint64_t)k << lgN;
  if (total == 0) return hipSuccess;
  const dim3 grid((unsigned)((total + EW_THREADS - 1) / EW_THREADS));
  switch (op) {
    case RF_EW_SET: hipLaunchKernelGGL(k_ew<RF_EW_SET>, grid, dim3(EW_THREADS), 0, st, tb, lgN, total, a, b, out); break;
    case RF_EW_ADD: hipLaunchKernelGGL(k_ew<RF_EW_ADD>, grid, dim3(EW_THREADS), 0, st, tb, lgN, total, a, b, out); break;
    case RF_EW_SUB: hipLaunchKernelGGL(k_ew<RF_EW_SUB>, grid, dim3(EW_THREADS), 0, st, tb, lgN, total, a, b, out); break;
    case RF_EW_NEG: hipLaunchKernelGGL(k_ew<RF_EW_NEG>, grid, dim3(EW_THREADS), 0, st, tb, lgN, total, a, b, out); break;
    case RF_EW_MUL: hipLaunchKernelGGL(k_ew<RF_EW_MUL>, grid, dim3(EW_THREADS), 0, st, tb, lgN, total, a, b, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t rf_launch_ew_matmul(const RfDevTables& tb, int k, int lgN, const double* const* a,
                               const double* const* b, double* const* out, int n1, int n2, int n3,
                               hipStream_t st) {
  const uint64_t total = (uint64_t)k << lgN;
  if (total == 0 || n1 <= 0 || n3 <= 0) return hipSuccess;
  const dim3 grid((unsigned)((total + EW_THREADS - 1) / EW_THREADS), (unsigned)(n1 * n3));
  hipLaunchKernelGGL(k_ew_matmul, grid, dim3(EW_THREADS), 0, st, tb, lgN, total, a, b, out, n2, n3);
  return hipGetLastError();
}
