// launch.h -- flat-index launcher and device-memory shims.
//
// A functor F with `operator()(long t)` is run for t in [0, n) by k_for<F> on the context
// stream.  Block size 256 (four 64-wide wavefronts); the grid covers n exactly once, so
// every workgroup index maps to real work except the last partial block.
//
// LMFHD_HOST_DEBUG builds the same functors into a host-only debug library that runs the
// loops sequentially on the CPU; it exists only to check index arithmetic without a GPU
// and is never built into the product library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace lmfhd {

#ifndef LMFHD_HOST_DEBUG
template <class F>
__global__ void __launch_bounds__(256) k_for(F f, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f(t);
}
#endif

template <class F>
inline void launch(hipStream_t s, long n, const F& f) {
  if (n <= 0) return;
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (long t = 0; t < n; ++t) f(t);
#else
  const int bs = 256;
  const long nb = (n + bs - 1) / bs;
  hipLaunchKernelGGL(k_for<F>, dim3((unsigned)nb), dim3(bs), 0, s, f, n);
#endif
}

}  // namespace lmfhd
