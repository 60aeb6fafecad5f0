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

// ---- recorded launches: lock-step batches of contexts -----------------------------------------
// While a Recorder is installed on the calling thread (rec_slot), launch / launch_cols /
// launch_rows / launch_blocks append a record -- the functor's bytes, the launch shape and the
// launcher of the batched form -- instead of launching.  Contexts of one size class record the
// same sequence of functor types and shapes; a batch replays record k of all of them in one
// launch whose grid has one y-slice per context, slice q running the functor copy q read from
// device memory.  Each context's items get the arithmetic of its own launch, so a batched run
// gives the bits of the separate runs.
struct Rec;
using MultiFn = void (*)(hipStream_t s, const unsigned char* fs, int nent, const Rec& r);
struct Rec {
  int kind = 0;                   // 0 flat, 1 cols, 2 rows, 3 blocks
  std::vector<unsigned char> f;   // the functor's bytes
  long n = 0;                     // flat: items; blocks: workgroups
  int ncols = 0, nz = 0, kc = 0;  // cols
  int nw = 0;                     // rows: waves
  MultiFn multi = nullptr;
};
struct Recorder {
  std::vector<Rec> recs;
};
inline Recorder*& rec_slot() {
  static thread_local Recorder* r = nullptr;
  return r;
}
template <class F>
inline void record(int kind, const F& f, MultiFn m, long n, int ncols = 0, int nz = 0, int kc = 0, int nw = 0) {
  Rec r;
  r.kind = kind;
  r.f.assign(reinterpret_cast<const unsigned char*>(&f), reinterpret_cast<const unsigned char*>(&f) + sizeof(F));
  r.n = n;
  r.ncols = ncols;
  r.nz = nz;
  r.kc = kc;
  r.nw = nw;
  r.multi = m;
  rec_slot()->recs.push_back(std::move(r));
}

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
void multi_cols(hipStream_t s, const unsigned char* fs, int nent, const Rec& r);
template <class F>
inline void launch_cols(hipStream_t s, int ncols, int nz, int kc, const F& f) {
  static_assert(sizeof(F) + 32 <= MAX_KERNARG, "functor too large for the kernel-argument segment");
  if (ncols <= 0 || nz <= 0) return;
  if (kc < 1) kc = 1;
  if (rec_slot()) {
    record(1, f, &multi_cols<F>, 0, ncols, nz, kc);
    return;
  }
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

// Row launcher: F::operator()(w, lane) for logical waves w in [0, nw), four 64-lane waves per
// 256-thread workgroup.  w is read from the first lane, so it and whatever the functor derives
// from it alone (the j and k of the row a wave walks) are wave-uniform: the compiler keeps them,
// and the wrap / reflect arithmetic on them, in scalar registers, and only the lane's own
// coordinate along the row costs vector instructions.  XCD-contiguous block order as in k_cols.
constexpr int ROW_WAVES = 4;
#ifndef LMFHD_HOST_DEBUG
template <class F>
__global__ void __launch_bounds__(256) k_rows(F f, int nw, int nblog, int per) {
  const int b = blockIdx.x;
  const int lb = (b & 7) * per + (b >> 3);
  if (lb >= nblog) return;
  const int w = __builtin_amdgcn_readfirstlane(lb * ROW_WAVES + (int)(threadIdx.x >> 6));
  if (w >= nw) return;
  f(w, (int)(threadIdx.x & 63));
}
#endif

template <class F>
void multi_rows(hipStream_t s, const unsigned char* fs, int nent, const Rec& r);
template <class F>
inline void launch_rows(hipStream_t s, int nw, const F& f) {
  static_assert(sizeof(F) + 32 <= MAX_KERNARG, "functor too large for the kernel-argument segment");
  if (nw <= 0) return;
  if (rec_slot()) {
    record(2, f, &multi_rows<F>, 0, 0, 0, 0, nw);
    return;
  }
  const int nblog = (nw + ROW_WAVES - 1) / ROW_WAVES;
  const int per = (nblog + 7) / 8;
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  (void)per;
  for (int w = 0; w < nw; ++w)
    for (int lane = 0; lane < 64; ++lane) f(w, lane);
#else
  hipLaunchKernelGGL(k_rows<F>, dim3((unsigned)(8 * per)), dim3(64 * ROW_WAVES), 0, s, f, nw, nblog, per);
#endif
}

// batched forms: slice blockIdx.y runs functor copy fs[blockIdx.y] over the shape of its record
#ifndef LMFHD_HOST_DEBUG
template <class F>
__global__ void __launch_bounds__(256) k_for_multi(const F* __restrict__ fs, long n) {
  const F& f = fs[blockIdx.y];
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) f(t);
}
template <class F>
__global__ void __launch_bounds__(COL_BS) k_cols_multi(const F* __restrict__ fs, int ncols, int nz, int kc, int nbx) {
  const F& f = fs[blockIdx.y];
  const int lb = blockIdx.x;
  const int bz = lb / nbx, bx = lb - bz * nbx;
  const int col = bx * COL_BS + threadIdx.x;
  if (col >= ncols) return;
  const int k0 = bz * kc;
  f(col, k0, k0 + kc < nz ? k0 + kc : nz);
}
template <class F>
__global__ void __launch_bounds__(256) k_rows_multi(const F* __restrict__ fs, int nw) {
  const F& f = fs[blockIdx.y];
  const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ROW_WAVES + (int)(threadIdx.x >> 6));
  if (w >= nw) return;
  f(w, (int)(threadIdx.x & 63));
}
#endif
template <class F>
void multi_flat(hipStream_t s, const unsigned char* fs, int nent, const Rec& r) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (int q = 0; q < nent; ++q)
    for (long t = 0; t < r.n; ++t) reinterpret_cast<const F*>(fs)[q](t);
#else
  hipLaunchKernelGGL(k_for_multi<F>, dim3((unsigned)((r.n + 255) / 256), (unsigned)nent), dim3(256), 0, s,
                     reinterpret_cast<const F*>(fs), r.n);
#endif
}
template <class F>
void multi_cols(hipStream_t s, const unsigned char* fs, int nent, const Rec& r) {
  const int nbx = (r.ncols + COL_BS - 1) / COL_BS, nbz = (r.nz + r.kc - 1) / r.kc;
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (int q = 0; q < nent; ++q)
    for (int lb = 0; lb < nbx * nbz; ++lb) {
      const int bz = lb / nbx, bx = lb - bz * nbx;
      for (int tx = 0; tx < COL_BS; ++tx) {
        const int col = bx * COL_BS + tx;
        if (col >= r.ncols) continue;
        const int k0 = bz * r.kc;
        reinterpret_cast<const F*>(fs)[q](col, k0, k0 + r.kc < r.nz ? k0 + r.kc : r.nz);
      }
    }
#else
  hipLaunchKernelGGL(k_cols_multi<F>, dim3((unsigned)(nbx * nbz), (unsigned)nent), dim3(COL_BS), 0, s,
                     reinterpret_cast<const F*>(fs), r.ncols, r.nz, r.kc, nbx);
#endif
}
template <class F>
void multi_rows(hipStream_t s, const unsigned char* fs, int nent, const Rec& r) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (int q = 0; q < nent; ++q)
    for (int w = 0; w < r.nw; ++w)
      for (int lane = 0; lane < 64; ++lane) reinterpret_cast<const F*>(fs)[q](w, lane);
#else
  hipLaunchKernelGGL(k_rows_multi<F>, dim3((unsigned)((r.nw + ROW_WAVES - 1) / ROW_WAVES), (unsigned)nent),
                     dim3(64 * ROW_WAVES), 0, s, reinterpret_cast<const F*>(fs), r.nw);
#endif
}

// Workgroup functors: F::block(b) runs workgroup b of a launch of nb workgroups of F::BS threads
// (it may use __shared__ memory and barriers); launch_blocks runs them, recordable like the rest.
// In the host debug build a workgroup functor provides block_host(b) instead.
#ifndef LMFHD_HOST_DEBUG
template <class F>
__global__ void __launch_bounds__(F::BS) k_blocks(F f) {
  f.block((int)blockIdx.x);
}
template <class F>
__global__ void __launch_bounds__(F::BS) k_blocks_multi(const F* __restrict__ fs) {
  fs[blockIdx.y].block((int)blockIdx.x);
}
#endif
template <class F>
void multi_blocks(hipStream_t s, const unsigned char* fs, int nent, const Rec& r) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (int q = 0; q < nent; ++q)
    for (long b = 0; b < r.n; ++b) reinterpret_cast<const F*>(fs)[q].block_host((int)b);
#else
  hipLaunchKernelGGL(k_blocks_multi<F>, dim3((unsigned)r.n, (unsigned)nent), dim3(F::BS), 0, s,
                     reinterpret_cast<const F*>(fs));
#endif
}
template <class F>
inline void launch_blocks(hipStream_t s, int nb, const F& f) {
  static_assert(sizeof(F) + 32 <= MAX_KERNARG, "functor too large for the kernel-argument segment");
  if (nb <= 0) return;
  if (rec_slot()) {
    record(3, f, &multi_blocks<F>, nb);
    return;
  }
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (int b = 0; b < nb; ++b) f.block_host(b);
#else
  hipLaunchKernelGGL(k_blocks<F>, dim3((unsigned)nb), dim3(F::BS), 0, s, f);
#endif
}

template <class F>
inline void launch(hipStream_t s, long n, const F& f) {
  static_assert(sizeof(F) + 32 <= MAX_KERNARG, "functor too large for the kernel-argument segment");
  if (n <= 0) return;
  if (rec_slot()) {
    record(0, f, &multi_flat<F>, n);
    return;
  }
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
