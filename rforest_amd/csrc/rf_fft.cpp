// rf_fft.cpp -- integers held in the transform domain (the general transform
// API of include/rforest_hip.h, after the reference's mpzfft interface,
// src/include/mpzfft.h:13-120 and src/lib/mpzfft.c:10-55).
//
// A value is one integer's spectra, [prime][N] residues in device memory,
// always holding 2^-lgN times the transform (the forward pass folds the
// factor into its pack constants, like a product's right operand), so that
// sums need nothing, products rescale once (rf_fftops.hip) and the inverse is
// the plain inverse transform followed by the product pipeline's CRT and
// carry kernels.  Every entry point holds the engine lock, uses the engine's
// stream, arena and staging buffer, and returns after the stream is idle.
#include <gmp.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "rf_internal.h"
#include "rf_util.h"

struct rf_fft_params {
  RfPlan plan;
  uint64_t bits;
  unsigned terms;
};

struct rf_fft {
  RfPlan plan;
  uint64_t bits = 0;  // the params' output bound
  double* spec = nullptr;
};

namespace {
using namespace rfi;

uint64_t spec_len(const RfPlan& pl) { return (uint64_t)pl.k << pl.lgN; }

// Arena bytes of the forward task tables, and of the inverse's coefficient
// residues, task tables and CRT scratch: computed ahead so that the arena
// grows (and moves) before anything is placed in it.
size_t forward_scratch(const RfPlan& pl) {
  const size_t rows = plan_split(pl) ? (size_t)1 << (pl.lgN - RF_LGN_ROW) : 0;
  return align256(pl.k * sizeof(RfFwdTask)) + align256(pl.k * rows * sizeof(RfRowTask));
}
size_t inverse_scratch(const RfPlan& pl) {
  const size_t rows = plan_split(pl) ? (size_t)1 << (pl.lgN - RF_LGN_ROW) : 1;
  const size_t nblk = crt_blocks(pl);
  return align256(spec_len(pl) * 8) + align256(pl.k * rows * sizeof(RfInvTask)) +
         align256(pl.k * sizeof(RfRowTask)) + align256(nblk * sizeof(RfCrtTask)) +
         align256(nblk * RF_CRT_THREADS) + align256(nblk);
}

bool same_plan(const rf_fft* a, const rf_fft* b) {
  return a && b && a->plan.t == b->plan.t && a->plan.lgN == b->plan.lgN && a->plan.k == b->plan.k;
}

#define RF_FFT_ENTER()                                                                  \
  std::lock_guard<std::mutex> lk(engine_mutex());                                       \
  if (!engine_ready()) return fail(RF_ERR_NOT_INIT, "hw_mpz_setup() has not been called")

// Forward transforms (one per prime) of the entry at dptr into f->spec; the
// arena from byte `base` on holds the task tables.  Input limbs, if staged,
// occupy the staging buffer before byte `hbase`.
int forward_locked(rf_fft* f, const uint64_t* dptr, uint64_t nlimbs, int sign, size_t base, size_t hbase) {
  const RfPlan& pl = f->plan;
  const int k = pl.k, lgN = pl.lgN;
  const uint64_t N = (uint64_t)1 << lgN;
  hipStream_t st = engine_stream();
  if (sign == 0 || nlimbs == 0) {
    HIPCHK(rf_launch_fill((uint64_t*)f->spec, spec_len(pl), 0, st));
    HIPCHK(hipStreamSynchronize(st));
    return RF_OK;
  }
  const bool split = plan_split(pl);
  const uint32_t rows = split ? 1u << (lgN - RF_LGN_ROW) : 0u;
  std::vector<RfFwdTask> ft;
  std::vector<RfRowTask> rt;
  for (int q = 0; q < k; q++) {
    RfFwdTask t{};
    t.limbs = dptr;
    t.spec_off = (uint64_t)q * N;
    t.nlimbs = (uint32_t)nlimbs;
    t.nchunk = (uint32_t)std::min<uint64_t>((nlimbs + pl.t - 1) / pl.t, N);
    t.sign = sign;
    t.q = (uint16_t)q;
    t.t = (uint8_t)pl.t;
    t.scale = (uint8_t)lgN;  // the value holds 2^-lgN times its transform
    ft.push_back(t);
    for (uint32_t h = 0; h < rows; h++)
      rt.push_back(RfRowTask{t.spec_off + ((uint64_t)h << RF_LGN_ROW), h, (uint16_t)q, (uint16_t)pl.lgN});
  }
  const size_t fb = align256(ft.size() * sizeof(RfFwdTask)), rb = align256(rt.size() * sizeof(RfRowTask));
  char* d_base = nullptr;
  char* h_base = nullptr;
  int rc = engine_arena(base + fb + rb, &d_base);
  if (!rc) rc = engine_stage(hbase + fb + rb, &h_base);
  if (rc) return rc;
  memcpy(h_base + hbase, ft.data(), ft.size() * sizeof(RfFwdTask));
  if (!rt.empty()) memcpy(h_base + hbase + fb, rt.data(), rt.size() * sizeof(RfRowTask));
  RfFwdTask* d_ft = (RfFwdTask*)(d_base + base);
  RfRowTask* d_rt = (RfRowTask*)(d_base + base + fb);
  HIPCHK(hipMemcpyAsync(d_ft, h_base + hbase, fb + rb, hipMemcpyHostToDevice, st));
  const RfDevTables tb = engine_tables();
  if (split) {
    HIPCHK(rf_launch_fwd_col(tb, lgN, pl.t, d_ft, k, f->spec, st));
    HIPCHK(rf_launch_fwd_row(tb, d_rt, (int)rt.size(), f->spec, st));
  } else {
    HIPCHK(rf_launch_fwd(tb, lgN, pl.t, d_ft, k, f->spec, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return RF_OK;
}

// Inverse transform + CRT + carries of op into `out` (plan.rn limbs, two's
// complement); scratch from arena byte `base` on.  Leaves the stream busy.
int inverse_locked(const rf_fft* op, uint64_t* out, size_t base) {
  const RfPlan& pl = op->plan;
  const int k = pl.k, lgN = pl.lgN;
  const uint64_t N = (uint64_t)1 << lgN;
  const bool split = plan_split(pl);
  const uint32_t rows = split ? 1u << (lgN - RF_LGN_ROW) : 1u;
  const uint64_t rowlen = split ? (uint64_t)1 << RF_LGN_ROW : N;
  std::vector<RfInvTask> it;
  std::vector<RfRowTask> ct;
  for (int q = 0; q < k; q++) {
    for (uint32_t h = 0; h < rows; h++) {
      RfInvTask t{};
      t.a_off = (uint64_t)q * N + h * rowlen;
      t.out_off = t.a_off;
      t.q = (uint16_t)q;
      t.d = 0;  // the one stored spectrum, no product
      it.push_back(t);
    }
    if (split) ct.push_back(RfRowTask{(uint64_t)q * N, 0, (uint16_t)q, (uint16_t)pl.lgN});
  }
  std::vector<RfCrtTask> crt;
  uint32_t npre = 0, nblk = 0;
  crt_tasks(pl, 0, out, npre, nblk, crt);
  const size_t coef_b = align256(spec_len(pl) * 8);
  const size_t ib = align256(it.size() * sizeof(RfInvTask)), cb = align256(ct.size() * sizeof(RfRowTask));
  const size_t kb = align256(crt.size() * sizeof(RfCrtTask));
  const size_t pre_b = align256(npre), fun_b = align256(nblk);
  const size_t tasks_b = ib + cb + kb;
  char* d_base = nullptr;
  char* h_base = nullptr;
  int rc = engine_arena(base + coef_b + tasks_b + pre_b + fun_b, &d_base);
  if (!rc) rc = engine_stage(tasks_b, &h_base);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));  // the staging buffer is free
  memcpy(h_base, it.data(), it.size() * sizeof(RfInvTask));
  if (!ct.empty()) memcpy(h_base + ib, ct.data(), ct.size() * sizeof(RfRowTask));
  memcpy(h_base + ib + cb, crt.data(), crt.size() * sizeof(RfCrtTask));
  double* coef = (double*)(d_base + base);
  char* d_tasks = d_base + base + coef_b;
  uint8_t* pre = (uint8_t*)(d_tasks + tasks_b);
  uint8_t* fun = pre + pre_b;
  HIPCHK(hipMemcpyAsync(d_tasks, h_base, tasks_b, hipMemcpyHostToDevice, st));
  const RfDevTables tb = engine_tables();
  HIPCHK(rf_launch_inv(tb, split ? RF_LGN_ROW : lgN, (const RfInvTask*)d_tasks, (int)it.size(), op->spec, op->spec,
                       coef, st));
  if (split) HIPCHK(rf_launch_inv_col(tb, lgN, (const RfRowTask*)(d_tasks + ib), (int)ct.size(), coef, st));
  const RfCrtTask* d_crt = (const RfCrtTask*)(d_tasks + ib + cb);
  HIPCHK(rf_launch_crt(tb, k, d_crt, (int)crt.size(), coef, pre, fun, st));
  HIPCHK(rf_launch_apply(d_crt, (int)crt.size(), pre, fun, st));
  return RF_OK;
}

int ew_locked(int opc, rf_fft* rop, const rf_fft* a, const rf_fft* b) {
  if (!rop || !a || !same_plan(rop, a) || (b && !same_plan(rop, b)))
    return fail(RF_ERR_ARG, "transform-domain operands of different plans (or null)");
  hipStream_t st = engine_stream();
  HIPCHK(rf_launch_ew(engine_tables(), opc, rop->plan.k, rop->plan.lgN, a->spec, b ? b->spec : nullptr, rop->spec,
                      st));
  HIPCHK(hipStreamSynchronize(st));
  return RF_OK;
}

}  // namespace

extern "C" {

int rf_fft_params_create(uint64_t bits, unsigned terms, rf_fft_params** out) {
  std::lock_guard<std::mutex> lk(engine_mutex());
  int rc = engine_host_init();  // planning needs the primes only
  if (rc) return rc;
  if (!out) return fail(RF_ERR_ARG, "rf_fft_params_create: null output");
  *out = nullptr;
  if (bits == 0 || terms == 0) return fail(RF_ERR_ARG, "rf_fft_params_create: bits and terms must be positive");
  RfPlan pl;
  if (rf_plan_general(&engine_primes(), bits, terms, &pl) != 0)
    return fail(RF_ERR_UNSUPPORTED, "no transform plan for %llu-bit sums of %u products",
                (unsigned long long)bits, terms);
  rf_fft_params* p = new rf_fft_params();
  p->plan = pl;
  p->bits = bits;
  p->terms = terms;
  *out = p;
  return RF_OK;
}

int rf_fft_params_plan(const rf_fft_params* params, int* t, int* lgN, int* k) {
  if (!params || !t || !lgN || !k) return fail(RF_ERR_ARG, "rf_fft_params_plan: null argument");
  *t = params->plan.t;
  *lgN = params->plan.lgN;
  *k = params->plan.k;
  return RF_OK;
}

void rf_fft_params_destroy(rf_fft_params* params) { delete params; }

int rf_fft_create(const rf_fft_params* params, rf_fft** out) {
  RF_FFT_ENTER();
  if (!params || !out) return fail(RF_ERR_ARG, "rf_fft_create: null argument");
  *out = nullptr;
  const int rc = engine_long_tables(params->plan.lgN);
  if (rc) return rc;
  rf_fft* f = new rf_fft();
  f->plan = params->plan;
  f->bits = params->bits;
  const size_t bytes = spec_len(f->plan) * 8;
  hipError_t e = hipMalloc((void**)&f->spec, bytes);
  if (e != hipSuccess) {
    delete f;
    return fail(RF_ERR_NOMEM, "transform-domain value of %zu bytes: %s", bytes, hipGetErrorString(e));
  }
  hipStream_t st = engine_stream();
  hipError_t e2 = rf_launch_fill((uint64_t*)f->spec, spec_len(f->plan), 0, st);
  if (e2 == hipSuccess) e2 = hipStreamSynchronize(st);
  if (e2 != hipSuccess) {
    (void)hipFree(f->spec);
    delete f;
    return fail(RF_ERR_HIP, "clearing a transform-domain value: %s", hipGetErrorString(e2));
  }
  *out = f;
  return RF_OK;
}

void rf_fft_destroy(rf_fft* f) {
  if (!f) return;
  std::lock_guard<std::mutex> lk(engine_mutex());
  if (engine_ready()) (void)hipStreamSynchronize(engine_stream());
  if (f->spec) (void)hipFree(f->spec);
  delete f;
}

uint64_t rf_fft_out_limbs(const rf_fft* f) { return f ? (uint64_t)f->plan.rn : 0; }

int rf_fft_forward(rf_fft* rop, const mpz_t op) {
  RF_FFT_ENTER();
  if (!rop || !op) return fail(RF_ERR_ARG, "rf_fft_forward: null argument");
  const int sign = mpz_sgn(op);
  const uint64_t n = mpz_size(op);
  if (sign && mpz_sizeinbase(op, 2) > rop->bits)
    return fail(RF_ERR_ARG, "operand of %llu bits exceeds the %llu-bit bound of its plan",
                (unsigned long long)mpz_sizeinbase(op, 2), (unsigned long long)rop->bits);
  const size_t in_b = align256(n * 8 + 8), fs = forward_scratch(rop->plan);
  char* d_base = nullptr;
  char* h_base = nullptr;
  int rc = engine_arena(in_b + fs, &d_base);
  if (!rc) rc = engine_stage(in_b + fs, &h_base);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  if (n) {
    memcpy(h_base, op->_mp_d, n * 8);
    HIPCHK(hipMemcpyAsync(d_base, h_base, n * 8, hipMemcpyHostToDevice, st));
  }
  return forward_locked(rop, (const uint64_t*)d_base, n, sign, in_b, in_b);
}

int rf_fft_forward_device(rf_fft* rop, const rf_dev_entry* op) {
  RF_FFT_ENTER();
  if (!rop || !op) return fail(RF_ERR_ARG, "rf_fft_forward_device: null argument");
  const int sign = op->nlimbs ? (op->sign > 0 ? 1 : (op->sign < 0 ? -1 : 0)) : 0;
  if (sign && op->nlimbs > (rop->bits + 63) / 64)
    return fail(RF_ERR_ARG, "operand of %llu limbs exceeds the %llu-bit bound of its plan",
                (unsigned long long)op->nlimbs, (unsigned long long)rop->bits);
  HIPCHK(hipStreamSynchronize(engine_stream()));
  return forward_locked(rop, op->limbs, op->nlimbs, sign, 0, 0);
}

int rf_fft_inverse(mpz_t rop, const rf_fft* op) {
  RF_FFT_ENTER();
  if (!rop || !op) return fail(RF_ERR_ARG, "rf_fft_inverse: null argument");
  const size_t rn = op->plan.rn;
  const size_t out_b = align256(rn * 8), is = inverse_scratch(op->plan);
  char* d_base = nullptr;
  char* h_base = nullptr;
  int rc = engine_arena(out_b + is, &d_base);
  if (!rc) rc = engine_stage(std::max(out_b, is), &h_base);
  if (rc) return rc;
  rc = inverse_locked(op, (uint64_t*)d_base, out_b);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpyAsync(h_base, d_base, rn * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<uint64_t> x((uint64_t*)h_base, (uint64_t*)h_base + rn);
  bool neg = false;
  tc_to_signmag(x, neg);  // src/lib/mpzfft.c:45-54
  if (x.empty()) {
    mpz_set_ui(rop, 0);
    return RF_OK;
  }
  _mpz_realloc(rop, (mp_size_t)x.size());
  memcpy(rop->_mp_d, x.data(), x.size() * 8);
  rop->_mp_size = neg ? -(int)x.size() : (int)x.size();
  return RF_OK;
}

int rf_fft_inverse_device(const rf_fft* op, uint64_t* out, uint64_t cap, int* sign, uint64_t* nlimbs) {
  RF_FFT_ENTER();
  if (!op || !out || !sign || !nlimbs) return fail(RF_ERR_ARG, "rf_fft_inverse_device: null argument");
  const uint64_t rn = op->plan.rn;
  if (cap < rn) return fail(RF_ERR_CAPACITY, "output of %llu limbs, %llu needed", (unsigned long long)cap,
                            (unsigned long long)rn);
  const size_t conv_b = 256, meta_b = 256, is = inverse_scratch(op->plan);
  char* d_base = nullptr;
  char* h_base = nullptr;
  int rc = engine_arena(conv_b + meta_b + is, &d_base);
  if (!rc) rc = engine_stage(std::max(conv_b + meta_b, is), &h_base);
  if (rc) return rc;
  rc = inverse_locked(op, out, conv_b + meta_b);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  RfConvTask ct{out, rn};
  memcpy(h_base, &ct, sizeof ct);
  HIPCHK(hipMemcpyAsync(d_base, h_base, sizeof ct, hipMemcpyHostToDevice, st));
  RfEntryMeta* d_meta = (RfEntryMeta*)(d_base + conv_b);
  HIPCHK(rf_launch_signmag((const RfConvTask*)d_base, 1, d_meta, rn, st));
  HIPCHK(hipMemcpyAsync(h_base + conv_b, d_meta, sizeof(RfEntryMeta), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  RfEntryMeta m;
  memcpy(&m, h_base + conv_b, sizeof m);
  *sign = m.sign;
  *nlimbs = m.nlimbs;
  return RF_OK;
}

int rf_fft_set(rf_fft* rop, const rf_fft* op) {
  RF_FFT_ENTER();
  return ew_locked(RF_EW_SET, rop, op, nullptr);
}
int rf_fft_add(rf_fft* rop, const rf_fft* op1, const rf_fft* op2) {
  RF_FFT_ENTER();
  return ew_locked(RF_EW_ADD, rop, op1, op2);
}
int rf_fft_sub(rf_fft* rop, const rf_fft* op1, const rf_fft* op2) {
  RF_FFT_ENTER();
  return ew_locked(RF_EW_SUB, rop, op1, op2);
}
int rf_fft_neg(rf_fft* rop, const rf_fft* op) {
  RF_FFT_ENTER();
  return ew_locked(RF_EW_NEG, rop, op, nullptr);
}
int rf_fft_mul(rf_fft* rop, const rf_fft* op1, const rf_fft* op2) {
  RF_FFT_ENTER();
  return ew_locked(RF_EW_MUL, rop, op1, op2);
}

int rf_fft_matrix_mul(rf_fft* const* rop, rf_fft* const* op1, rf_fft* const* op2, unsigned dim1, unsigned dim2,
                      unsigned dim3) {
  RF_FFT_ENTER();
  if (!rop || !op1 || !op2 || dim1 == 0 || dim2 == 0 || dim3 == 0)
    return fail(RF_ERR_ARG, "rf_fft_matrix_mul: null argument or zero dimension");
  const size_t n1 = (size_t)dim1 * dim2, n2 = (size_t)dim2 * dim3, n3 = (size_t)dim1 * dim3;
  const rf_fft* ref = rop[0];
  for (size_t i = 0; i < n1; i++)
    if (!same_plan(ref, op1[i])) return fail(RF_ERR_ARG, "rf_fft_matrix_mul: operands of different plans");
  for (size_t i = 0; i < n2; i++)
    if (!same_plan(ref, op2[i])) return fail(RF_ERR_ARG, "rf_fft_matrix_mul: operands of different plans");
  for (size_t i = 0; i < n3; i++)
    if (!same_plan(ref, rop[i])) return fail(RF_ERR_ARG, "rf_fft_matrix_mul: operands of different plans");
  // outputs that alias an input are computed into the arena and copied
  bool alias = false;
  for (size_t i = 0; i < n3 && !alias; i++) {
    for (size_t j = 0; j < n1; j++) alias |= rop[i]->spec == op1[j]->spec;
    for (size_t j = 0; j < n2; j++) alias |= rop[i]->spec == op2[j]->spec;
  }
  const uint64_t len = spec_len(ref->plan);
  const size_t ptr_b = align256((n1 + n2 + n3) * sizeof(double*));
  const size_t tmp_b = alias ? align256(n3 * len * 8) : 0;
  char* d_base = nullptr;
  char* h_base = nullptr;
  int rc = engine_arena(ptr_b + tmp_b, &d_base);
  if (!rc) rc = engine_stage(ptr_b, &h_base);
  if (rc) return rc;
  hipStream_t st = engine_stream();
  HIPCHK(hipStreamSynchronize(st));
  double** hp = (double**)h_base;
  for (size_t i = 0; i < n1; i++) hp[i] = op1[i]->spec;
  for (size_t i = 0; i < n2; i++) hp[n1 + i] = op2[i]->spec;
  for (size_t i = 0; i < n3; i++) hp[n1 + n2 + i] = alias ? (double*)(d_base + ptr_b) + i * len : rop[i]->spec;
  HIPCHK(hipMemcpyAsync(d_base, h_base, (n1 + n2 + n3) * sizeof(double*), hipMemcpyHostToDevice, st));
  double** dp = (double**)d_base;
  HIPCHK(rf_launch_ew_matmul(engine_tables(), ref->plan.k, ref->plan.lgN, (const double* const*)dp,
                             (const double* const*)(dp + n1), dp + n1 + n2, (int)dim1, (int)dim2, (int)dim3, st));
  if (alias)
    for (size_t i = 0; i < n3; i++)
      HIPCHK(hipMemcpyAsync(rop[i]->spec, (double*)(d_base + ptr_b) + i * len, len * 8, hipMemcpyDeviceToDevice,
                            st));
  HIPCHK(hipStreamSynchronize(st));
  return RF_OK;
}

}  // extern "C"
