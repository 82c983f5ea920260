// rf_noalloc.hip -- the fixed device scratch of rf_batch_mult_device_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) run products with
// the pipeline's scratch -- spectra, coefficient residues, task tables and
// carry codes -- in this file-scope array instead of the engine's arena.
// The code object carries it, so every GPU the library runs on holds its own
// copy, at its own address, from module load until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address is taken once per device with hipGetSymbolAddress (at
// hw_mpz_setup, with that device current) and only that pointer is used.
#include <hip/hip_runtime.h>

#include "rf_noalloc.h"

namespace {

static __device__ __align__(256) char scratch[RF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
char* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device (engine lock held)

}  // namespace

hipError_t rf_noalloc_init() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (g_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[dev] = static_cast<char*>(p);
  return hipSuccess;
}

hipError_t rf_noalloc_scratch(char** ptr) {
  *ptr = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_ptr[dev]) {
    e = rf_noalloc_init();  // first call on a device set up elsewhere
    if (e != hipSuccess) return e;
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}
