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
#include <vector>

namespace lmfhd {

#ifndef LMFHD_HOST_DEBUG
template <class F>
__global__ void __launch_bounds__(256) k_for(F f, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f(t);
}
// the same with XCD-contiguous block ranges (hardware block b -> logical (b % 8) per + b / 8),
// so that the rows and planes a stencil shares with its neighbours stay in one XCD's L2
template <class F>
__global__ void __launch_bounds__(256) k_for_xcd(F f, long n, int nb, int per) {
  const int b = blockIdx.x;
  const int lb = (b & 7) * per + (b >> 3);
  if (lb >= nb) return;
  const long t = (long)lb * 256 + threadIdx.x;
  if (t < n) f(t);
}
#endif

// flat launches use XCD-contiguous block ranges unless LMFHD_XCD=0 (measured at 128^3 on
// MI355X: the finest Gauss-Seidel colour launch 17.5 -> 15.5 us, the FP64 solve 32.5 -> 31.4 ms)
inline bool xcd_order() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("LMFHD_XCD");
    v = (e && atoi(e) == 0) ? 0 : 1;
  }
  return v == 1;
}

// Column launcher: F::operator()(col, k0, k1) handles point `col` of an (nx*ny) plane for
// k in [k0, k1).  Logical block lb covers 256 columns of z-run lb / nbx.  Workgroups are handed
// to the eight XCDs round-robin (hardware block b runs on XCD b % 8), so hardware block b takes
// logical block (b % 8) * per + b / 8: each XCD walks a contiguous range of z-runs and the
// neighbouring rows and planes a stencil re-reads stay in that XCD's L2.
constexpr int COL_BS = 256;
#ifndef LMFHD_HOST_DEBUG
template <class F>
__global__ void __launch_bounds__(COL_BS) k_cols(F f, int ncols, int nz, int kc, int nbx, int nblog, int per) {
  const int b = blockIdx.x;
  const int lb = (b & 7) * per + (b >> 3);
  if (lb >= nblog) return;
  const int bz = lb / nbx, bx = lb - bz * nbx;
  const int col = bx * COL_BS + threadIdx.x;
  if (col >= ncols) return;
  const int k0 = bz * kc;
  const int k1 = k0 + kc < nz ? k0 + kc : nz;
  f(col, k0, k1);
}
#endif

// Functors travel by value in the kernel-argument segment; the runtime copies that segment
// when the launch is queued, so a functor may live on the caller's stack.  Argument blocks too
// large for it (the coarse multigrid kernels' level tables) are copied to device memory with a
// completed host-to-device copy and passed by pointer instead.
constexpr size_t MAX_KERNARG = 4096;

template <class F>
inline void launch_cols(hipStream_t s, int ncols, int nz, int kc, const F& f) {
  static_assert(sizeof(F) + 32 <= MAX_KERNARG, "functor too large for the kernel-argument segment");
  if (ncols <= 0 || nz <= 0) return;
  if (kc < 1) kc = 1;
  const int nbx = (ncols + COL_BS - 1) / COL_BS;
  const int nbz = (nz + kc - 1) / kc;
  const int nblog = nbx * nbz;
  const int per = (nblog + 7) / 8;
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  (void)per;
  for (int lb = 0; lb < nblog; ++lb) {
    const int bz = lb / nbx, bx = lb - bz * nbx;
    for (int tx = 0; tx < COL_BS; ++tx) {
      const int col = bx * COL_BS + tx;
      if (col >= ncols) continue;
      const int k0 = bz * kc;
      f(col, k0, k0 + kc < nz ? k0 + kc : nz);
    }
  }
#else
  hipLaunchKernelGGL(k_cols<F>, dim3((unsigned)(8 * per)), dim3(COL_BS), 0, s, f, ncols, nz, kc, nbx, nblog, per);
#endif
}

template <class F>
inline void launch(hipStream_t s, long n, const F& f) {
  static_assert(sizeof(F) + 32 <= MAX_KERNARG, "functor too large for the kernel-argument segment");
  if (n <= 0) return;
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (long t = 0; t < n; ++t) f(t);
#else
  const int bs = 256;
  const long nb = (n + bs - 1) / bs;
  if (xcd_order() && nb >= 64) {
    const int per = (int)((nb + 7) / 8);
    hipLaunchKernelGGL(k_for_xcd<F>, dim3((unsigned)(8 * per)), dim3(bs), 0, s, f, n, (int)nb, per);
  } else {
    hipLaunchKernelGGL(k_for<F>, dim3((unsigned)nb), dim3(bs), 0, s, f, n);
  }
#endif
}

}  // namespace lmfhd
