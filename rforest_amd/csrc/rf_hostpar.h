// Host-side parallel copies for the paths that move operands and results
// through the pinned staging buffer (the drop-in's limbs in, results out).
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace rfi {

// Runs fn(ctx, i) for every i in [0, n) on the copy pool's worker threads
// and the calling thread; returns when all n calls have returned.  Calls
// from several threads are serialised.  Workers spin for a short while
// after a job (so back-to-back calls find them ready) and then sleep.
void par_for(int n, void (*fn)(void* ctx, int i), void* ctx);

// Wakes sleeping workers ahead of a job that will follow shortly.
void par_wake();

// dst[0, n) = src[0, n) in pieces on the pool (n in bytes).
void par_memcpy(void* dst, const void* src, size_t n);

// Every segment's copy, in pieces of the segments' concatenation over the
// pool (small segments share a piece, large ones span several).
struct CopySeg {
  void* dst;
  const void* src;
  size_t n;  // bytes
};
void par_copy(const CopySeg* segs, int nseg);

}  // namespace rfi
