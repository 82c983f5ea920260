// rf_noalloc.h -- the fixed device scratch behind rf_batch_mult_device_noalloc.
#pragma once
#include <hip/hip_runtime.h>

// Bytes of the static device array (per GPU, held until the process ends):
// room for the pipeline scratch of seven 2x2 products of 1,048,576-bit
// entries (8.4 MiB each), or many smaller ones.
#define RF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

// Takes the current device's address of the array (hipGetSymbolAddress),
// once per device; called by hw_mpz_setup with the device current.
hipError_t rf_noalloc_init();
// The current device's address of the array (initialising it on the first
// call on a device that hw_mpz_setup did not see).
hipError_t rf_noalloc_scratch(char** ptr);
