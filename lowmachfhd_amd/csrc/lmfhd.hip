// lmfhd.hip -- MI355X-native implicit staggered-grid Stokes solver: host-side sequencing and C ABI.
//
// Product path of the Stokes solve of LowMachFHD (staggered_grid/src_gmres).  Every stencil,
// smoother, transfer, reduction and Gram-Schmidt step runs as a HIP kernel on gfx950; the host
// only sequences launches and runs the O(m^2) Givens/least-squares scalar work of GMRES.
//
//   gmres          gmres.f90:32-621          -> gmres_left() / fgmres()
//   apply_matrix   apply_matrix.f90:24-157   -> apply_matrix()
//   apply_precon   apply_precon.f90:29-536   -> precon<T>()  (precon_type 1)
//   stag_mg_solver stag_mg_solver.f90:37-864 -> StagMG<T>::solve()
//   macproject     macproject.f90:34-123     -> MacMG<T>::solve() (own cell-centred MG)
//   norms          norm_inner_product.f90    -> mdot() weighted reductions
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstddef>
#include <type_traits>
#include <chrono>
#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lmfhd.h"
#include "launch.h"
#include "stencils.h"

namespace lmfhd {

// ------------------------------------------------------------------------------------------
// errors
static thread_local char g_err[512] = "";
static void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
struct Fail {
  int code;
};
#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      set_err("HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__);       \
      throw Fail{e_ == hipErrorOutOfMemory ? LMFHD_E_NOMEM : LMFHD_E_HIP};               \
    }                                                                                    \
  } while (0)

// ------------------------------------------------------------------------------------------
// device memory
// Owns its allocations: release() (or the destructor, e.g. when a Fail unwinds a temporary)
// frees them.  hipFree waits for the device, so no kernel still uses a buffer it frees.
struct DevMem {
  std::vector<void*> ptrs;
  long long bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  template <typename T>
  T* alloc(long n) {
    if (n <= 0) n = 1;
    void* p = nullptr;
#ifdef LMFHD_HOST_DEBUG
    p = calloc((size_t)n, sizeof(T));
    if (!p) throw Fail{LMFHD_E_NOMEM};
#else
    HIPCHK(hipMalloc(&p, (size_t)n * sizeof(T)));
#endif
    ptrs.push_back(p);
    bytes += (long long)n * sizeof(T);
    return (T*)p;
  }
  // the same, cleared to zero bits (completed before it returns).  On the calling thread's
  // own stream, not the legacy one: a legacy-stream operation would wait on every blocking
  // stream, and fails while another thread captures a graph
  template <typename T>
  T* alloc_zero(long n) {
    T* p = alloc<T>(n);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipMemsetAsync(p, 0, (size_t)(n <= 0 ? 1 : n) * sizeof(T), hipStreamPerThread));
    HIPCHK(hipStreamSynchronize(hipStreamPerThread));
#endif
    return p;
  }
  void release() {
    for (void* p : ptrs) {
#ifdef LMFHD_HOST_DEBUG
      free(p);
#else
      (void)hipFree(p);
#endif
    }
    ptrs.clear();
    bytes = 0;
  }
};

// host -> device, complete on return: every source here is pageable host memory that the
// caller reuses or frees right away (stack structs, scalars, reused staging vectors), and
// hipMemcpyAsync only promises to have read it when the copy has executed on the stream
static void h2d(void* dst, const void* src, size_t b, hipStream_t s) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  memcpy(dst, src, b);
#else
  HIPCHK(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
#endif
}
static void d2h(void* dst, const void* src, size_t b, hipStream_t s) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  memcpy(dst, src, b);
#else
  HIPCHK(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
#endif
}
static void d2d(void* dst, const void* src, size_t b, hipStream_t s) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  memmove(dst, src, b);
#else
  HIPCHK(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, s));
#endif
}
#ifdef LMFHD_HOST_DEBUG
#define HIPCHK_NOTHROW(x) ((void)0)
#else
#define HIPCHK_NOTHROW(x) ((void)(x))
#endif
static void sync_stream(hipStream_t s) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
#else
  HIPCHK(hipStreamSynchronize(s));
#endif
}
static void check_launch() {
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipGetLastError());
#endif
}

template <typename T>
static void fill(hipStream_t s, T* p, long n, T v) {
  launch(s, n, FillFn<T>{p, v});
}

// ------------------------------------------------------------------------------------------
// reductions (deterministic: fixed grid, per-block partials, one final block)
static const int RED_BLOCKS = 4096;  // partial-sum slots (the grid of a reduction is at most this)
static const int RED_NV = 16;
// partial sum of workgroup b for product q: q-major, so that the final reduction's lanes (one
// per b) read consecutive addresses
#define PART(b, q) ((long)(q) * RED_BLOCKS + (b))

#ifndef LMFHD_HOST_DEBUG
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[PART(b, q)] = sum over workgroup b's elements (of nb) of weight(e) w[e] V_q[e]
template <int NV>
__device__ void mdot_block(int b, int nb, const double* __restrict__ w, const double* __restrict__ V, long vstride,
                           int nv, const VecLayout& L, long n, int weighted, long e0, double* __restrict__ part,
                           const double* __restrict__ skip) {
  if (skip && skip[0] != 0.0) return;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (long e = (long)b * 256 + threadIdx.x; e < n; e += (long)nb * 256) {
    const double we = weighted ? w[e] * L.weight(e0 + e) : w[e];
    if (V == nullptr) {
      acc[0] += we;
      continue;
    }
    // every load unconditional (q >= nv re-reads vector nv-1 from cache; its sums are not
    // stored), so the NV loads of an element are in flight together
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += we * V[(q < nv ? q : nv - 1) * vstride + e];
  }
  __shared__ double sh[NV][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) sh[q][wid] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    const int q = threadIdx.x;
    part[PART(b, q)] = ((sh[q][0] + sh[q][1]) + sh[q][2]) + sh[q][3];
  }
}

// out[q] = sum_b part[PART(b, q)] (optionally sqrt); one workgroup
__device__ void red_final_block(const double* __restrict__ part, int nb, int nv, int do_sqrt, double* __restrict__ out,
                                const double* __restrict__ skip) {
  if (skip && skip[0] != 0.0) {
    for (int q = threadIdx.x; q < nv; q += 256) out[q] = 0.0;
    return;
  }
  // wave w sums the partials of q = w, w + 4, ...: no barriers, a fixed order per q
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = wid; q < nv; q += 4) {
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += part[PART(b, q)];
    s = wave_sum(s);
    if (lane == 0) out[q] = do_sqrt ? sqrt(s) : s;
  }
}

template <int NV>
__global__ void __launch_bounds__(256) k_mdot(const double* __restrict__ w, const double* __restrict__ V, long vstride,
                                              int nv, VecLayout L, long n, int weighted, long e0,
                                              double* __restrict__ part, const double* __restrict__ skip) {
  mdot_block<NV>((int)blockIdx.x, (int)gridDim.x, w, V, vstride, nv, L, n, weighted, e0, part, skip);
}
__global__ void __launch_bounds__(256) k_red_final(const double* __restrict__ part, int nb, int nv, int do_sqrt,
                                                   double* __restrict__ out, const double* __restrict__ skip) {
  red_final_block(part, nb, nv, do_sqrt, out, skip);
}
// the same two steps as workgroup functors (launch_blocks: recordable)
template <int NV>
struct MdotBlk {
  static constexpr int BS = 256;
  const double* w;
  const double* V;
  long vstride;
  int nv;
  VecLayout L;
  long n;
  int weighted;
  long e0;
  double* part;
  const double* skip;
  int nb;
  __device__ void block(int b) const { mdot_block<NV>(b, nb, w, V, vstride, nv, L, n, weighted, e0, part, skip); }
};
struct RedFinalBlk {
  static constexpr int BS = 256;
  const double* part;
  int nb, nv, do_sqrt;
  double* out;
  const double* skip;
  __device__ void block(int) const { red_final_block(part, nb, nv, do_sqrt, out, skip); }
};

// max |x| of a float or double array (stag_mg_solver residual norms)
template <typename T>
__global__ void __launch_bounds__(256) k_absmax(const T* __restrict__ x, long n, double* __restrict__ part) {
  double m = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) m = fmax(m, fabs((double)x[e]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// y += sgn * sum_q coef[q] V_q
template <int NV>
__global__ void __launch_bounds__(256) k_maxpy(double* __restrict__ y, const double* __restrict__ V, long vstride,
                                               int nv, const double* __restrict__ coef, double sgn, long n,
                                               const double* __restrict__ skip) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n || (skip && skip[0] != 0.0)) return;
  // every load unconditional (q >= nv re-reads vector nv-1 from cache) so that they are in
  // flight together; the sum takes the q < nv terms in order
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = V[(q < nv ? q : nv - 1) * vstride + e];
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < NV; ++q)
    if (q < nv) s += coef[q] * v[q];
  y[e] = y[e] + sgn * s;
}

// The fused Gram-Schmidt passes of the Arnoldi step (grid-stride over the nb partial blocks of
// k_mdot, partial sums part[PART(b, q)] reduced by k_red_final in a fixed order):
//   k_mdot_self:  part_q = <w, V_q> for q < nv and part_nv = <w, w> (nv + 1 <= RED_NV);
//   k_maxpy_dot:  w -= sum_q coef[q] V_q, then part_q = <w_new, V_q> for q < nv;
//   k_maxpy_norm: unless *skip, w -= sum_q coef[q] V_q; part_0 = <w, w> of the result.
// Inner products carry the Stokes-vector weights of element e0 + e.
template <int NV>
__global__ void __launch_bounds__(256) k_mdot_self(const double* __restrict__ w, const double* __restrict__ V,
                                                   long vstride, int nv, VecLayout L, long n, long e0,
                                                   double* __restrict__ part) {
  double acc[NV + 1];
#pragma unroll
  for (int q = 0; q <= NV; ++q) acc[q] = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double x = w[e];
    const double we = x * L.weight(e0 + e);
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += we * V[(q < nv ? q : nv - 1) * vstride + e];
    acc[NV] += we * x;
  }
  __shared__ double sh[NV + 1][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q <= NV; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) sh[q][wid] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x <= nv) {
    const int q = threadIdx.x == nv ? NV : threadIdx.x;
    part[PART(blockIdx.x, threadIdx.x)] = ((sh[q][0] + sh[q][1]) + sh[q][2]) + sh[q][3];
  }
}
template <int NV>
__global__ void __launch_bounds__(256) k_maxpy_dot(double* __restrict__ w, const double* __restrict__ V, long vstride,
                                                   int nv, const double* __restrict__ coef, VecLayout L, long n,
                                                   long e0, double* __restrict__ part) {
  double c[NV], acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    c[q] = q < nv ? coef[q] : 0.0;
    acc[q] = 0.0;
  }
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    double v[NV];
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      v[q] = V[(q < nv ? q : nv - 1) * vstride + e];
      if (q < nv) sum += c[q] * v[q];
    }
    const double x = w[e] + (-1.0) * sum;
    w[e] = x;
    const double we = x * L.weight(e0 + e);
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += we * v[q];
  }
  __shared__ double sh[NV][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double r = wave_sum(acc[q]);
    if (lane == 0) sh[q][wid] = r;
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    const int q = threadIdx.x;
    part[PART(blockIdx.x, q)] = ((sh[q][0] + sh[q][1]) + sh[q][2]) + sh[q][3];
  }
}
template <int NV>
__global__ void __launch_bounds__(256) k_maxpy_norm(double* __restrict__ w, const double* __restrict__ V, long vstride,
                                                    int nv, const double* __restrict__ coef, VecLayout L, long n,
                                                    long e0, double* __restrict__ part,
                                                    const double* __restrict__ skip) {
  const bool upd = !(skip && skip[0] != 0.0);
  double c[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) c[q] = q < nv ? coef[q] : 0.0;
  double acc = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    double x = w[e];
    if (upd) {
      // every load unconditional (q >= nv re-reads vector nv-1 from cache), so the loads of an
      // element are in flight together; the sum takes the q < nv terms as before
      double v[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] = V[(q < nv ? q : nv - 1) * vstride + e];
      double sum = 0.0;
#pragma unroll
      for (int q = 0; q < NV; ++q)
        if (q < nv) sum += c[q] * v[q];
      x = x + (-1.0) * sum;
      w[e] = x;
    }
    acc += (x * L.weight(e0 + e)) * x;
  }
  __shared__ double sh[4];
  const double r = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) part[PART(blockIdx.x, 0)] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// skip = 1 when |w - V h|^2 = wn2 - |h|^2 >= wn2 / 2 (no second Gram-Schmidt pass), else 0
__global__ void k_dgks(const double* __restrict__ h, int nh, const double* __restrict__ wn2, double* __restrict__ skip) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double hh = 0.0;
  for (int q = 0; q < nh; ++q) hh += h[q] * h[q];
  skip[0] = (wn2[0] - hh >= 0.5 * wn2[0]) ? 1.0 : 0.0;
}

// dst = src / *den
__global__ void __launch_bounds__(256) k_scale_div(double* __restrict__ dst, const double* __restrict__ src,
                                                   const double* __restrict__ den, long n) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) dst[e] = src[e] / den[0];
}
// dst = a - dst
__global__ void __launch_bounds__(256) k_rsub(double* __restrict__ dst, const double* __restrict__ a, long n) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) dst[e] = -(dst[e] - a[e]);
}
// single-workgroup mean removal for small arrays (MAC multigrid bottom)
template <typename T>
struct RemoveMeanBlk {
  static constexpr int BS = 256;
  T* x;
  long n;
  __device__ void block(int) const {
    __shared__ double sh[4];
    double s = 0.0;
    for (long e = threadIdx.x; e < n; e += 256) s += (double)x[e];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    const double mean = (((sh[0] + sh[1]) + sh[2]) + sh[3]) / (double)n;
    for (long e = threadIdx.x; e < n; e += 256) x[e] = T((double)x[e] - mean);
  }
};
#endif

#ifdef LMFHD_HOST_DEBUG
// host form of mdot (one "workgroup" forms every sum), recordable like the device form
struct MdotHostBlk {
  static constexpr int BS = 1;
  const double* w;
  const double* V;
  long vstride;
  int nv;
  VecLayout L;
  long n;
  bool weighted;
  long e0;
  const double* skip;
  bool do_sqrt;
  double* out_dev;
  void block_host(int) const {
    if (skip && skip[0] != 0.0) {
      for (int q = 0; q < nv; ++q) out_dev[q] = 0.0;
      return;
    }
    for (int q = 0; q < nv; ++q) {
      double acc = 0.0;
      for (long e = 0; e < n; ++e)
        acc += (weighted ? w[e] * L.weight(e0 + e) : w[e]) * (V ? V[q * vstride + e] : 1.0);
      out_dev[q] = do_sqrt ? std::sqrt(acc) : acc;
    }
  }
};
#endif

struct Reducer {
  double* part = nullptr;  // RED_BLOCKS * RED_NV
  double* out = nullptr;   // 4 * RED_NV results
  hipStream_t s;
  VecLayout L;

  int max_blocks = 1024;  // workgroups of a reduction launch (<= RED_BLOCKS; a tuning parameter)
  int nblocks(long n) const { return (int)std::min<long>(max_blocks, std::max<long>(1, (n + 255) / 256)); }

  // out_dev[q] = <w, V_q> (weighted with the Stokes-vector weights of element e0+e, or plain)
  // for q < nv <= RED_NV; V == nullptr gives the (weighted) sum of w.
  // skip (device scalar, may be null): when nonzero the products are not formed and out = 0
  void mdot(const double* w, const double* V, long vstride, int nv, long n, bool weighted, double* out_dev,
            bool do_sqrt = false, long e0 = 0, const double* skip = nullptr) {
    if (nv < 1 || nv > RED_NV) {  // k_mdot reads vector min(q, nv - 1)
      set_err("internal: mdot over %d vectors", nv);
      throw Fail{LMFHD_E_ARG};
    }
#ifdef LMFHD_HOST_DEBUG
    launch_blocks(s, 1, MdotHostBlk{w, V, vstride, nv, L, n, weighted, e0, skip, do_sqrt, out_dev});
#else
    const int nb = nblocks(n);
    if (nv <= 1) launch_blocks(s, nb, MdotBlk<1>{w, V, vstride, nv, L, n, (int)weighted, e0, part, skip, nb});
    else if (nv <= 4) launch_blocks(s, nb, MdotBlk<4>{w, V, vstride, nv, L, n, (int)weighted, e0, part, skip, nb});
    else if (nv <= 8) launch_blocks(s, nb, MdotBlk<8>{w, V, vstride, nv, L, n, (int)weighted, e0, part, skip, nb});
    else launch_blocks(s, nb, MdotBlk<16>{w, V, vstride, nv, L, n, (int)weighted, e0, part, skip, nb});
    launch_blocks(s, 1, RedFinalBlk{part, nb, nv, (int)do_sqrt, out_dev, skip});
    check_launch();
#endif
  }
  // Fused Gram-Schmidt passes of the Arnoldi step, weighted inner products, nv + 1 <= RED_NV:
  // out[q] = <w, V_q> (q < nv), out[nv] = <w, w>
  void mdot_self(const double* w, const double* V, long vstride, int nv, long n, double* out) {
#ifdef LMFHD_HOST_DEBUG
    for (int q = 0; q <= nv; ++q) {
      double acc = 0.0;
      for (long e = 0; e < n; ++e) acc += (w[e] * L.weight(e)) * (q < nv ? V[q * vstride + e] : w[e]);
      out[q] = acc;
    }
#else
    const int nb = nblocks(n);
    if (nv <= 1) hipLaunchKernelGGL(k_mdot_self<1>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, L, n, 0L, part);
    else if (nv <= 4) hipLaunchKernelGGL(k_mdot_self<4>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, L, n, 0L, part);
    else if (nv <= 8) hipLaunchKernelGGL(k_mdot_self<8>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, L, n, 0L, part);
    else hipLaunchKernelGGL(k_mdot_self<16>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, L, n, 0L, part);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(256), 0, s, part, nb, nv + 1, 0, out, nullptr);
    check_launch();
#endif
  }
  // w -= sum_q coef[q] V_q; out[q] = <w_new, V_q>, or 0 when *skip_out != 0
  void maxpy_dot(double* w, const double* V, long vstride, int nv, const double* coef, long n, double* out,
                 const double* skip_out) {
#ifdef LMFHD_HOST_DEBUG
    for (long e = 0; e < n; ++e) {
      double sum = 0.0;
      for (int q = 0; q < nv; ++q) sum += coef[q] * V[q * vstride + e];
      w[e] = w[e] + (-1.0) * sum;
    }
    for (int q = 0; q < nv; ++q) {
      double acc = 0.0;
      for (long e = 0; e < n; ++e) acc += (w[e] * L.weight(e)) * V[q * vstride + e];
      out[q] = (skip_out && skip_out[0] != 0.0) ? 0.0 : acc;
    }
#else
    const int nb = nblocks(n);
    if (nv <= 1) hipLaunchKernelGGL(k_maxpy_dot<1>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part);
    else if (nv <= 4) hipLaunchKernelGGL(k_maxpy_dot<4>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part);
    else if (nv <= 8) hipLaunchKernelGGL(k_maxpy_dot<8>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part);
    else hipLaunchKernelGGL(k_maxpy_dot<16>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(256), 0, s, part, nb, nv, 0, out, skip_out);
    check_launch();
#endif
  }
  // unless *skip: w -= sum_q coef[q] V_q; *nrm = sqrt <w, w>
  void maxpy_norm(double* w, const double* V, long vstride, int nv, const double* coef, long n, double* nrm,
                  const double* skip) {
#ifdef LMFHD_HOST_DEBUG
    const bool upd = !(skip && skip[0] != 0.0);
    double acc = 0.0;
    for (long e = 0; e < n; ++e) {
      if (upd) {
        double sum = 0.0;
        for (int q = 0; q < nv; ++q) sum += coef[q] * V[q * vstride + e];
        w[e] = w[e] + (-1.0) * sum;
      }
      acc += (w[e] * L.weight(e)) * w[e];
    }
    nrm[0] = std::sqrt(acc);
#else
    const int nb = nblocks(n);
    if (nv <= 1) hipLaunchKernelGGL(k_maxpy_norm<1>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part, skip);
    else if (nv <= 4) hipLaunchKernelGGL(k_maxpy_norm<4>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part, skip);
    else if (nv <= 8) hipLaunchKernelGGL(k_maxpy_norm<8>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part, skip);
    else hipLaunchKernelGGL(k_maxpy_norm<16>, dim3(nb), dim3(256), 0, s, w, V, vstride, nv, coef, L, n, 0L, part, skip);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(256), 0, s, part, nb, 1, 1, nrm, nullptr);
    check_launch();
#endif
  }
  template <typename T>
  void absmax(const T* x, long n, double* out_dev) {
#ifdef LMFHD_HOST_DEBUG
    double m = 0.0;
    for (long e = 0; e < n; ++e) m = std::max(m, std::fabs((double)x[e]));
    out_dev[0] = m;
#else
    const int nb = nblocks(n);
    hipLaunchKernelGGL(k_absmax<T>, dim3(nb), dim3(256), 0, s, x, n, part);
    std::vector<double> h(nb);
    d2h(h.data(), part, nb * sizeof(double), s);
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    h2d(out_dev, &m, sizeof(double), s);
#endif
  }
};

static void dgks(hipStream_t s, const double* h, int nh, const double* wn2, double* skip) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  double hh = 0.0;
  for (int q = 0; q < nh; ++q) hh += h[q] * h[q];
  skip[0] = (wn2[0] - hh >= 0.5 * wn2[0]) ? 1.0 : 0.0;
#else
  hipLaunchKernelGGL(k_dgks, dim3(1), dim3(64), 0, s, h, nh, wn2, skip);
  check_launch();
#endif
}

static void maxpy(hipStream_t s, double* y, const double* V, long vstride, int nv, const double* coef, double sgn,
                  long n, const double* skip = nullptr) {
  for (int q0 = 0; q0 < nv; q0 += RED_NV) {
    const int c = std::min(RED_NV, nv - q0);
#ifdef LMFHD_HOST_DEBUG
    (void)s;
    if (skip && skip[0] != 0.0) return;
    for (long e = 0; e < n; ++e) {
      double acc = 0.0;
      for (int q = 0; q < c; ++q) acc += coef[q0 + q] * V[(q0 + q) * vstride + e];
      y[e] = y[e] + sgn * acc;
    }
#else
    const long nb = (n + 255) / 256;
    const double* Vq = V + q0 * vstride;
    if (c <= 4)
      hipLaunchKernelGGL(k_maxpy<4>, dim3(nb), dim3(256), 0, s, y, Vq, vstride, c, coef + q0, sgn, n, skip);
    else if (c <= 8)
      hipLaunchKernelGGL(k_maxpy<8>, dim3(nb), dim3(256), 0, s, y, Vq, vstride, c, coef + q0, sgn, n, skip);
    else
      hipLaunchKernelGGL(k_maxpy<16>, dim3(nb), dim3(256), 0, s, y, Vq, vstride, c, coef + q0, sgn, n, skip);
    check_launch();
#endif
  }
}
static void scale_div(hipStream_t s, double* dst, const double* src, const double* den, long n) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (long e = 0; e < n; ++e) dst[e] = src[e] / den[0];
#else
  hipLaunchKernelGGL(k_scale_div, dim3((n + 255) / 256), dim3(256), 0, s, dst, src, den, n);
  check_launch();
#endif
}
static void rsub(hipStream_t s, double* dst, const double* a, long n) {
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  for (long e = 0; e < n; ++e) dst[e] = -(dst[e] - a[e]);
#else
  hipLaunchKernelGGL(k_rsub, dim3((n + 255) / 256), dim3(256), 0, s, dst, a, n);
  check_launch();
#endif
}
struct SubMeanFn {  // x -= *sum / cnt
  double* x;
  const double* sum;
  double cnt;
  HD void operator()(long e) const { x[e] -= sum[0] / cnt; }
};
static void sub_mean(hipStream_t s, double* x, const double* sum, double cnt, long n) {
  launch(s, n, SubMeanFn{x, sum, cnt});
  check_launch();
}
#ifdef LMFHD_HOST_DEBUG
template <typename T>
struct RemoveMeanHostBlk {
  static constexpr int BS = 1;
  T* x;
  long n;
  void block_host(int) const {
    double acc = 0.0;
    for (long e = 0; e < n; ++e) acc += (double)x[e];
    const double m = acc / (double)n;
    for (long e = 0; e < n; ++e) x[e] = T((double)x[e] - m);
  }
};
#endif
template <typename T>
static void remove_mean_small(hipStream_t s, T* x, long n) {
#ifdef LMFHD_HOST_DEBUG
  launch_blocks(s, 1, RemoveMeanHostBlk<T>{x, n});
#else
  launch_blocks(s, 1, RemoveMeanBlk<T>{x, n});
  check_launch();
#endif
}

// ------------------------------------------------------------------------------------------
static inline bool noslip_code(int c) { return c >= 100 && c <= 103; }
static inline bool slip_code(int c) { return c >= 200 && c <= 203; }
static inline bool wall_code(int c) { return noslip_code(c) || slip_code(c); }

static Geo make_geo(int dm, const int* n, const int* bclo, const int* bchi) {
  Geo g;
  g.dm = dm;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = d < dm ? n[d] : 1;
    const int lo = d < dm ? bclo[d] : -1, hi = d < dm ? bchi[d] : -1;
    g.per[d] = lo == -1 ? 1 : 0;
    g.wlo[d] = wall_code(lo) ? 1 : 0;
    g.whi[d] = wall_code(hi) ? 1 : 0;
    g.slo[d] = noslip_code(lo) ? -1 : (slip_code(lo) ? 1 : 0);
    g.shi[d] = noslip_code(hi) ? -1 : (slip_code(hi) ? 1 : 0);
  }
  return g;
}
static Geo coarsen_geo(const Geo& g) {
  Geo c = g;
  for (int d = 0; d < g.dm; ++d) c.n[d] = g.n[d] / 2;
  return c;
}
static int nlevs_mg(const Geo& g, int minwidth) {
  int nl = -1;
  for (int d = 0; d < g.dm; ++d) {
    int t = g.n[d], r = 1;
    while (t % 2 == 0 && t / minwidth != 1) { t /= 2; ++r; }
    nl = nl < 0 ? r : std::min(nl, r);
  }
  return nl;
}
static bool has_wall(const Geo& g) {
  for (int d = 0; d < g.dm; ++d)
    if (g.wlo[d] || g.whi[d]) return true;
  return false;
}

// column launches over the nodal (n+1)-box, one plane per thread: measured on MI355X (128^3
// FP64 stag_applyop 75 us with runs of 1 plane, 78 with 2, 90 with 4, 118 with 8) the
// register-heavier marching threads lose more occupancy than the plane reuse gains
struct ColShape {
  int ncols, nz, kc;
  FastDiv cx;
};
static ColShape col_shape(const Geo& g) {
  ColShape c;
  const int nx = g.n[0] + 1, ny = g.n[1] + 1;
  c.ncols = nx * ny;
  c.nz = g.dm == 3 ? g.n[2] + 1 : 1;
  c.cx = make_fastdiv((unsigned)nx);
  c.kc = 1;
  return c;
}
// row mapping (launch_rows) of a level: lanes along x in chunks of 64, the x faces i = n_x as
// tail waves with lanes along y
struct RowShape {
  int nw, wt;
  FastDiv cw, nyr, ct;
};
static RowShape row_shape(const Geo& g) {
  RowShape r;
  const int cw = (g.n[0] + 63) / 64, nyr = g.n[1] + 1, nzr = g.dm == 3 ? g.n[2] + 1 : 1;
  const int ct = (g.n[1] + 63) / 64, nzt = g.dm == 3 ? g.n[2] : 1;
  r.cw = make_fastdiv((unsigned)cw);
  r.nyr = make_fastdiv((unsigned)nyr);
  r.ct = make_fastdiv((unsigned)ct);
  r.wt = cw * nyr * nzr;
  r.nw = r.wt + ct * nzt;
  return r;
}
// positions of one parity of a level's nodal box ((n_x + 1) (n_y + 1) (n_z + 1), split order)
static unsigned pos3_half(const Geo& g) {
  return (unsigned)((g.n[0] + 2) / 2) * (unsigned)(g.n[1] + 1) * (unsigned)(g.dm == 3 ? g.n[2] + 1 : 1);
}
template <typename T, int VT, int DM>
static void apply_all_cols(hipStream_t s, const Geo& g, const Coefs<T>& C, T* const* phi, T* const* rhs, T* const* out,
                           int mode) {
  const ColShape cs = col_shape(g);
  ApplyAllFn<T, VT, DM> f;
  f.g = g;
  f.C = C;
  for (int a = 0; a < 3; ++a) {
    f.phi[a] = phi[a < DM ? a : DM - 1];
    f.rhs[a] = rhs ? rhs[a < DM ? a : DM - 1] : nullptr;
    f.out[a] = out[a < DM ? a : DM - 1];
  }
  f.cx = cs.cx;
  f.mode = mode;
  launch_cols(s, cs.ncols, cs.nz, cs.kc, f);
}
// sp: phi, rhs, out and alpha in parity-split storage (|visc_type| 1 levels only)
template <typename T, int VT, int DM>
static void apply_all(hipStream_t s, const Geo& g, const Coefs<T>& C, T* const* phi, T* const* rhs, T* const* out,
                      int mode, bool rows = false, int sp = 0) {
  if (sp) {
    if constexpr ((VT == 1 || VT == -1) && sizeof(T) == 4) {
      SplitPos3Fn<T, VT, DM, 0> f;
      f.g = g;
      f.C = C;
      for (int a = 0; a < 3; ++a) {
        const int q = a < DM ? a : DM - 1;
        f.phi[a] = phi[q];
        f.rhs[a] = rhs ? rhs[q] : nullptr;
        f.out[a] = out[q];
      }
      f.omega = T(0);
      f.dn = make_dec_color(g.n[0] + 1, g.n[1] + 1);
      f.Hn = pos3_half(g);
      f.par = 0;
      f.mode = mode;
      launch(s, 2L * f.Hn, f);
      return;
    } else if constexpr (VT == 1 || VT == -1) {
      ResidSplitFn<T, VT, DM> f;
      f.g = g;
      f.C = C;
      long mx = 0;
      for (int a = 0; a < 3; ++a) {
        const int q = a < DM ? a : DM - 1;
        f.phi[a] = phi[q];
        f.rhs[a] = rhs ? rhs[q] : nullptr;
        f.out[a] = out[q];
        f.dc[a] = make_dec_color(fext(g, q, 0), fext(g, q, 1));
        f.H[a] = a < DM ? sH(g, q) : 0;
        mx = std::max(mx, 2L * f.H[a]);
      }
      f.mode = mode;
      launch(s, (long)DM * ((mx + 255) / 256) * 256, f);
      return;
    }
    set_err("internal: split storage with |visc_type| %d", VT < 0 ? -VT : VT);
    throw Fail{LMFHD_E_ARG};
  }
  if (rows) {
    const RowShape rs = row_shape(g);
    ApplyRowsFn<T, VT, DM> f;
    f.g = g;
    f.C = C;
    for (int a = 0; a < 3; ++a) {
      f.phi[a] = phi[a < DM ? a : DM - 1];
      f.rhs[a] = rhs ? rhs[a < DM ? a : DM - 1] : nullptr;
      f.out[a] = out[a < DM ? a : DM - 1];
    }
    f.mode = mode;
    f.wt = rs.wt;
    f.cw = rs.cw;
    f.nyr = rs.nyr;
    f.ct = rs.ct;
    launch_rows(s, rs.nw, f);
    return;
  }
  apply_all_cols<T, VT, DM>(s, g, C, phi, rhs, out, mode);
}

static bool even_level(const Geo& g) {
  for (int d = 0; d < g.dm; ++d)
    if (g.n[d] % 2) return false;
  return true;
}

// compile-time dispatch on (visc_type, dm)
template <class F>
static void dispatch_vd(int vt, int dm, F&& f) {
#define LMFHD_VD(V, D) f(std::integral_constant<int, V>{}, std::integral_constant<int, D>{})
  if (dm == 3) {
    switch (vt) {
      case -1: LMFHD_VD(-1, 3); break;
      case -2: LMFHD_VD(-2, 3); break;
      case -3: LMFHD_VD(-3, 3); break;
      case 1: LMFHD_VD(1, 3); break;
      case 2: LMFHD_VD(2, 3); break;
      default: LMFHD_VD(3, 3); break;
    }
  } else {
    switch (vt) {
      case -1: LMFHD_VD(-1, 2); break;
      case -2: LMFHD_VD(-2, 2); break;
      case -3: LMFHD_VD(-3, 2); break;
      case 1: LMFHD_VD(1, 2); break;
      case 2: LMFHD_VD(2, 2); break;
      default: LMFHD_VD(3, 2); break;
    }
  }
#undef LMFHD_VD
}


// ------------------------------------------------------------------------------------------
// Coarse levels in one workgroup.  Every level whose cell counts are all <= COARSE_N is
// smoothed, restricted and prolongated by a single 1024-thread workgroup with a barrier
// between phases, replacing ~10^2 tiny launches (each ~5 us of GPU time whatever its size) per
// V-cycle by one.  Phase order and arithmetic are those of the per-level kernels.
static const int COARSE_MAXL = 16;
static const int COARSE_N = 16;
static const int COARSE_THREADS = 1024;

// LDS staging.  The arrays of the levels whose extents are all <= COARSE_N come from one block
// per multigrid (the arena), coarsest level first, so the levels of a one-workgroup V-cycle
// (level l0 and every coarser one) are one prefix of it.  When that prefix fits STAGE_BYTES
// the kernel copies it into LDS, runs every phase on the LDS copies (a phase then waits on
// LDS, not on L2, for its operands) and copies level l0's phi, the only result read later,
// back.  The phases and their arithmetic are unchanged.
static const unsigned STAGE_BYTES = 144u * 1024u;
struct Arena {  // byte offsets of the small levels' arrays in the arena, 16-byte aligned
  long bytes = 0;
  long take(long nbytes) {
    const long o = (bytes + 15) / 16 * 16;
    bytes = o + nbytes;
    return o;
  }
  long end() const { return (bytes + 15) / 16 * 16; }
};
struct StageInfo {
  const unsigned char* arena;  // the arena in device memory
  unsigned bytes;              // prefix staged in LDS (0: the V-cycle runs on device memory)
  unsigned out_off, out_bytes; // level l0's phi within it
};

template <typename T>
struct StagLevArgs {
  Geo g;
  T dxsqinv;
  T* alpha[3];
  T* beta;
  T* gamma;
  T* bed[3];
  T* phi[3];
  T* rhs[3];
  T* res[3];
  Dec dface[3], dcolor[3];  // decoders of the level's face arrays (all faces / one colour)
};
template <typename T>
struct StagCoarseArgs {
  int l0, nl, nu_down, nu_up, nu_bottom;
  int fuse;  // one phase per parity for all components where they decouple (|visc_type| 1)
  T omega;
  StageInfo st;
  StagLevArgs<T> L[COARSE_MAXL];
};
// The level table (several KB) exceeds the kernel-argument segment (launch.h), so it lives in
// device memory, written by a completed copy from the host (h2d) before any kernel reads it.

// The same per-component functor over the three components' item ranges in one phase
// ([0, n0) component 0, [n0, n01) component 1, the rest component 2): for steps whose
// components write disjoint arrays from inputs the phase does not change (residual,
// restriction, prolongation).
template <class F>
struct Comp3Fn {
  F f[3];
  long n0, n01;
  HD void operator()(long t) const {
    if (t < n0) f[0](t);
    else if (t < n01) f[1](t - n0);
    else f[2](t - n01);
  }
};
// The same functor over the components' item ranges in one flat launch, the components
// alternating by workgroup (block b runs items 256 (b / DM) + [0, 256) of component b % DM): the
// branch on the component is uniform in a workgroup
template <class F, int DM>
struct Comp3BlkFn {
  F f[3];
  unsigned n[3];
  HD void operator()(long t) const {
    const unsigned blk = (unsigned)(t >> 8), q = blk / DM;
    const int a = (int)(blk - q * DM);
    const unsigned u = (q << 8) + ((unsigned)t & 255u);
    if (a == 0) {
      if (u < n[0]) f[0](u);
    } else if (DM == 2 || a == 1) {
      if (u < n[1]) f[1](u);
    } else if (u < n[2]) {
      f[2](u);
    }
  }
};
template <class F, int DM, class Make, class Count>
HD Comp3Fn<F> comp3(const Make& make, const Count& count, long& total) {
  Comp3Fn<F> c{{make(0), make(DM > 1 ? 1 : 0), make(DM > 2 ? 2 : DM - 1)}, 0, 0};
  const long c0 = count(0), c1 = DM > 1 ? count(1) : 0, c2 = DM > 2 ? count(2) : 0;
  c.n0 = c0;
  c.n01 = c0 + c1;
  total = c0 + c1 + c2;
  return c;
}
template <typename T, int VT, int DM, class Run, class Map>
HD void stag_coarse_body(const StagCoarseArgs<T>& A, Run& run, const Map& M) {
  // phi of level l0 > 0 arrives zeroed (the restriction launch before this kernel), and each
  // restriction below zeroes phi of the level it writes; level 0's phi (l0 = 0) is the initial
  // guess
  auto coefs = [&](int n) __attribute__((always_inline)) {
    Coefs<T> C;
    for (int d = 0; d < 3; ++d) { C.alpha[d] = M(A.L[n].alpha[d]); C.bed[d] = M(A.L[n].bed[d]); }
    C.beta = M(A.L[n].beta);
    C.gamma = M(A.L[n].gamma);
    C.dxsqinv = A.L[n].dxsqinv;
    C.theta = T(1);
    return C;
  };
  auto smooth = [&](int n, int nsweeps) __attribute__((always_inline)) {
    const StagLevArgs<T>& l = A.L[n];
    const Coefs<T> C = coefs(n);
    bool inplace = true;
    for (int d = 0; d < DM; ++d) if (l.g.n[d] % 2) inplace = false;
    for (int m = 0; m < nsweeps; ++m) {
      if constexpr (VT == 1 || VT == -1) {
        if (inplace && A.fuse) {  // one phase per parity for all components (GsColorAllFn)
          GsColorAllFn<T, VT, DM> F;
          F.g = l.g;
          F.C = C;
          long cum = 0, ncs[3] = {0, 0, 0};
          for (int a = 0; a < DM; ++a) {
            F.phi[a] = M(l.phi[a]);
            F.rhs[a] = M(l.rhs[a]);
            F.dc[a] = l.dcolor[a];
            ncs[a] = (long)((fext(l.g, a, 0) + 1) / 2) * fext(l.g, a, 1) * fext(l.g, a, 2);
          }
          for (int a = DM; a < 3; ++a) { F.phi[a] = M(l.phi[DM - 1]); F.rhs[a] = M(l.rhs[DM - 1]); F.dc[a] = l.dcolor[DM - 1]; }
          // each component's items from a wave boundary: a wave then holds one component and
          // takes one branch of the component dispatch (on the smallest levels a wave would
          // otherwise run the three components' stencils one after the other)
          for (int a = 0; a < 3; ++a) F.nc[a] = a < DM ? (unsigned)ncs[a] : 0u;
          F.n0 = (ncs[0] + 63) / 64 * 64;
          F.n01 = F.n0 + (ncs[1] + 63) / 64 * 64;
          F.ilv = 0;
          cum = DM == 3 ? F.n01 + ncs[2] : F.n0 + ncs[1];
          F.omega = A.omega;
          for (int par = 0; par < 2; ++par) {
            F.par = par;
            run(F, cum);
          }
          continue;
        }
      }
      for (int color = 1; color <= 2 * DM; ++color) {
        const int a = (color - 1) / 2, par = (color + 1) % 2;
        const int nx = fext(l.g, a, 0), ny = fext(l.g, a, 1);
        const long nc = (long)((nx + 1) / 2) * ny * fext(l.g, a, 2);
        if (inplace) {
          run(GsColorFn<T, VT, DM>{l.g, C, {M(l.phi[0]), M(l.phi[1]), M(l.phi[2])}, M(l.rhs[a]), A.omega, a, l.dcolor[a], par}, nc);
        } else {
          run(ApplyColorFn<T, VT, DM>{l.g, C, {M(l.phi[0]), M(l.phi[1]), M(l.phi[2])}, M(l.res[a]), a, l.dcolor[a], par}, nc);
          run(UpdateColorFn<T, VT, DM>{l.g, C, M(l.phi[a]), M(l.rhs[a]), M(l.res[a]), A.omega, a, l.dcolor[a], par}, nc);
        }
      }
    }
  };
  for (int n = A.l0; n < A.nl - 1; ++n) {
    smooth(n, A.nu_down);
    const StagLevArgs<T>& f = A.L[n];
    const StagLevArgs<T>& c = A.L[n + 1];
    const Coefs<T> C = coefs(n);
    long cnt = 0;
    const auto R = comp3<ApplyFn<T, VT, DM>, DM>(
        [&](int a) { return ApplyFn<T, VT, DM>{f.g, C, {M(f.phi[0]), M(f.phi[1]), M(f.phi[2])}, M(f.rhs[a]), M(f.res[a]), a, f.dface[a], 1}; },
        [&](int a) { return fsize(f.g, a); }, cnt);
    run(R, cnt);
    const auto Q = comp3<RestrictFn<T, DM>, DM>(
        [&](int a) { return RestrictFn<T, DM>{f.g, c.g, M(f.res[a]), M(c.rhs[a]), a, c.dface[a], 0, 0, 0, M(c.phi[a])}; },
        [&](int a) { return fsize(c.g, a); }, cnt);
    run(Q, cnt);
  }
  smooth(A.nl - 1, A.nu_bottom);
  for (int n = A.nl - 2; n >= A.l0; --n) {
    const StagLevArgs<T>& f = A.L[n];
    const StagLevArgs<T>& c = A.L[n + 1];
    long cnt = 0;
    const auto P = comp3<ProlongFn<T, DM>, DM>(
        [&](int a) { return ProlongFn<T, DM>{f.g, c.g, M(f.phi[a]), M(c.phi[a]), a, f.dface[a]}; },
        [&](int a) { return fsize(f.g, a); }, cnt);
    run(P, cnt);
    smooth(n, A.nu_up);
  }
}

template <typename T>
struct MacLevArgs {
  Geo g;
  T h2i;
  T* b[3];
  T* phi;
  T* rhs;
  T* tmp;
  Dec dcell, dcolor;
};
template <typename T>
struct MacCoarseArgs {
  int l0, nl, nu_down, nu_up, nu_bottom;
  StageInfo st;
  MacLevArgs<T> L[COARSE_MAXL];
};

// the LDS staging of a coarse V-cycle from level l0 (StageInfo): the arena prefix of levels
// l0 .. nl-1 when it fits STAGE_BYTES; level l0's phi block goes back.  LMFHD_STAGE=0 when the
// context is built turns it off (the V-cycle then runs on device memory, with the same bits)
static StageInfo stage_info(const unsigned char* arena, int arena_l, const std::vector<long>& lev_end,
                            long phi_off, long phi_bytes, int l0) {
  const char* e = getenv("LMFHD_STAGE");
  const bool off = e && e[0] == '0';
  StageInfo st{arena, 0u, 0u, 0u};
  if (off || !arena || l0 < arena_l || l0 >= (int)lev_end.size() || lev_end[l0] > (long)STAGE_BYTES) return st;
  st.bytes = (unsigned)lev_end[l0];
  st.out_off = (unsigned)phi_off;
  st.out_bytes = (unsigned)phi_bytes;
  return st;
}

// The coarse bodies take every array through a map: the identity on device memory; for a
// staged V-cycle, an arena address to the same offset in the LDS block (formed from the block
// itself, so the compiler addresses LDS directly: 32-bit offsets and LDS instructions)
struct NoMap {
  template <typename P>
  HD P operator()(P p) const { return p; }
};
struct StageMap {
  const unsigned char* arena;
  unsigned char* buf;
  template <typename P>
  HD P operator()(P p) const { return (P)(buf + (unsigned)((const unsigned char*)p - arena)); }
};

template <typename T, int DM, class Run, class Mean, class Map>
HD void mac_coarse_body(const MacCoarseArgs<T>& A, Run& run, Mean& remove_mean, const Map& M) {
  // phi of level l0 > 0 arrives zeroed, and each restriction below zeroes the level it writes
  auto gs = [&](int n, int nsweeps) __attribute__((always_inline)) {
    const MacLevArgs<T>& l = A.L[n];
    bool inplace = true;
    for (int d = 0; d < DM; ++d) if (l.g.n[d] % 2) inplace = false;
    const long nc = (long)((l.g.n[0] + 1) / 2) * l.g.n[1] * l.g.n[2];
    for (int m = 0; m < nsweeps; ++m)
      for (int par = 0; par < 2; ++par) {
        if (inplace) {
          run(MacGsFn<T, DM>{l.g, {M(l.b[0]), M(l.b[1]), M(l.b[2])}, l.h2i, M(l.phi), M(l.rhs), par, l.dcolor}, nc);
        } else {
          run(MacGsNewFn<T, DM>{l.g, {M(l.b[0]), M(l.b[1]), M(l.b[2])}, l.h2i, M(l.phi), M(l.rhs), M(l.tmp), par, l.dcolor}, nc);
          run(CopyColorCellFn<T>{l.g, M(l.phi), M(l.tmp), par, l.dcolor}, nc);
        }
      }
  };
  for (int n = A.l0; n < A.nl - 1; ++n) {
    gs(n, A.nu_down);
    const MacLevArgs<T>& f = A.L[n];
    const MacLevArgs<T>& c = A.L[n + 1];
    run(MacResidRestrictFn<T, DM>{f.g, c.g, {M(f.b[0]), M(f.b[1]), M(f.b[2])}, f.h2i, M(f.phi), M(f.rhs), M(c.rhs), c.dcell, M(c.phi)}, csize(c.g));
  }
  remove_mean(M(A.L[A.nl - 1].rhs), csize(A.L[A.nl - 1].g));
  gs(A.nl - 1, A.nu_bottom);
  for (int n = A.nl - 2; n >= A.l0; --n) {
    run(MacProlongFn<T>{A.L[n].g, A.L[n + 1].g, M(A.L[n].phi), M(A.L[n + 1].phi), A.L[n].dcell}, csize(A.L[n].g));
    gs(n, A.nu_up);
  }
}

#ifndef LMFHD_HOST_DEBUG
struct BlockRun {
  template <class F>
  __device__ __forceinline__ void operator()(const F& f, long n) const {
    for (long t = threadIdx.x; t < n; t += blockDim.x) f(t);
    __syncthreads();
  }
  // each component's items as a loop of its own: choosing the functor per lane made the
  // compiler index the functor array with a lane-varying offset, which kept it in scratch
  // Component a's loop starts a ring of a third of the workgroup's waves further on, so that on
  // the small levels, where each component has fewer items than threads, the three run on
  // different waves at once instead of one after the other (the items of a phase are
  // independent, so which thread takes which does not change the result)
  template <class F>
  __device__ __forceinline__ void operator()(const Comp3Fn<F>& c, long n) const {
    const long n1 = c.n01 - c.n0, n2 = n - c.n01;
    const unsigned nt = blockDim.x, sh = ((nt >> 6) / 3) * 64;  // 2 sh < nt
    const long t1 = (threadIdx.x + nt - sh) % nt, t2 = (threadIdx.x + nt - 2 * sh) % nt;
    for (long t = threadIdx.x; t < c.n0; t += nt) c.f[0](t);
    for (long t = t1; t < n1; t += nt) c.f[1](t);
    for (long t = t2; t < n2; t += nt) c.f[2](t);
    __syncthreads();
  }
};


template <typename T>
struct BlockMean {  // subtract the mean of x[0, n) (workgroup reduction in FP64)
  __device__ void operator()(T* x, long n) const {
    __shared__ double sh[COARSE_THREADS / 64];
    double s = 0.0;
    for (long e = threadIdx.x; e < n; e += blockDim.x) s += (double)x[e];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += sh[w];
    const double mean = tot / (double)n;
    __syncthreads();
    for (long e = threadIdx.x; e < n; e += blockDim.x) x[e] = T((double)x[e] - mean);
    __syncthreads();
  }
};

// bytes (a multiple of 4) from src to dst by the workgroup: 16-byte words (both ends are
// 16-byte aligned: arena offsets and the LDS block), up to 8 loads of a thread in flight (past
// the end they reread the last word, and are not stored), then the 4-byte tail.  A native
// vector type: the HIP uint4 struct kept the staging array in scratch
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void blk_copy(void* dst, const void* src, unsigned bytes) {
  const unsigned n16 = bytes / 16, nt = blockDim.x, tid = threadIdx.x;
  const u32x4* s4 = (const u32x4*)src;
  u32x4* d4 = (u32x4*)dst;
  for (unsigned b = 0; b < n16; b += 8 * nt) {
    u32x4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const unsigned e = b + k * nt + tid;
      v[k] = s4[e < n16 ? e : n16 - 1];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const unsigned e = b + k * nt + tid;
      if (e < n16) d4[e] = v[k];
    }
  }
  const unsigned* s1 = (const unsigned*)src;
  unsigned* d1 = (unsigned*)dst;
  for (unsigned e = n16 * 4 + tid; e < bytes / 4; e += nt) d1[e] = s1[e];
}
// a coarse V-cycle on LDS copies of its levels (StageInfo): arena prefix in, body through
// StageMap, level l0's phi out
template <class Args, class Body>
__device__ __forceinline__ void staged_block(const Args& A, const Body& body) {
  __shared__ __align__(16) unsigned char buf[STAGE_BYTES];
  blk_copy(buf, A.st.arena, A.st.bytes);
  __syncthreads();
  body(StageMap{A.st.arena, buf});
  __syncthreads();
  blk_copy((void*)(A.st.arena + A.st.out_off), buf + A.st.out_off, A.st.out_bytes);
}

#else
struct HostRun {
  template <class F>
  void operator()(const F& f, long n) const {
    for (long t = 0; t < n; ++t) f(t);
  }
};
template <typename T>
struct HostMean {
  void operator()(T* x, long n) const {
    double acc = 0.0;
    for (long e = 0; e < n; ++e) acc += (double)x[e];
    const double m = acc / (double)n;
    for (long e = 0; e < n; ++e) x[e] = T((double)x[e] - m);
  }
};
// the LDS staging of staged_block with a host buffer in place of LDS (checks the arena layout
// and the pointer map in a CPU build)
template <class Args, class Body>
void staged_block_host(const Args& A, const Body& body) {
  std::vector<unsigned char> buf(STAGE_BYTES);
  memcpy(buf.data(), A.st.arena, A.st.bytes);
  body(StageMap{A.st.arena, buf.data()});
  memcpy((void*)(A.st.arena + A.st.out_off), buf.data() + A.st.out_off, A.st.out_bytes);
}

#endif

// The coarse V-cycle kernels as workgroup functors (launch_blocks, recordable): one workgroup
// of NT threads runs the body on its argument block.  NT threads: 1024 caps a thread at 128
// VGPRs (the stencils then spill to scratch); 256 and 512 leave room at the price of fewer
// lanes per colour step.
template <typename T, int VT, int DM, int NT>
struct StagCoarseBlk {
  static constexpr int BS = NT;
  const StagCoarseArgs<T>* A;
#ifdef LMFHD_HOST_DEBUG
  void block_host(int) const {
    const StagCoarseArgs<T>& a = *A;
    auto body = [&](const auto& M) {
      HostRun run;
      stag_coarse_body<T, VT, DM>(a, run, M);
    };
    if (a.st.bytes) staged_block_host(a, body);
    else body(NoMap{});
  }
#else
  // inlined into the kernel: as a called function the body kept the callee-saved registers of
  // the calling convention in scratch and spilled (248 VGPRs, 840 B of scratch per lane)
  __device__ __forceinline__ void block(int) const {
    const StagCoarseArgs<T>& a = *A;
    auto body = [&](const auto& M) __attribute__((always_inline)) {
      BlockRun run;
      stag_coarse_body<T, VT, DM>(a, run, M);
    };
    if (a.st.bytes) staged_block(a, body);
    else body(NoMap{});
  }
#endif
};
template <typename T, int DM, int NT>
struct MacCoarseBlk {
  static constexpr int BS = NT;
  const MacCoarseArgs<T>* A;
#ifdef LMFHD_HOST_DEBUG
  void block_host(int) const {
    const MacCoarseArgs<T>& a = *A;
    auto body = [&](const auto& M) {
      HostRun run;
      HostMean<T> mean;
      mac_coarse_body<T, DM>(a, run, mean, M);
    };
    if (a.st.bytes) staged_block_host(a, body);
    else body(NoMap{});
  }
#else
  __device__ __forceinline__ void block(int) const {
    const MacCoarseArgs<T>& a = *A;
    auto body = [&](const auto& M) __attribute__((always_inline)) {
      BlockRun run;
      BlockMean<T> mean;
      mac_coarse_body<T, DM>(a, run, mean, M);
    };
    if (a.st.bytes) staged_block(a, body);
    else body(NoMap{});
  }
#endif
};

static int coarse_level(const Geo& g0, int nl, int cn) {  // first level with all n_d <= cn
  Geo g = g0;
  for (int n = 0; n < nl; ++n) {
    bool small = true;
    for (int d = 0; d < g.dm; ++d) if (g.n[d] > cn) small = false;
    if (small) return (nl <= COARSE_MAXL) ? n : nl;
    g = coarsen_geo(g);
  }
  return nl;
}

// ------------------------------------------------------------------------------------------
// colour launches of at least this many items interleave the components by workgroup
constexpr long GS_ILV_MIN = 1L << 16;
// levels whose every extent is at least this store phi, rhs, residual and alpha split by
// parity (|visc_type| 1 with the Gauss-Seidel smoother on even sizes; far above the coarse
// kernel's levels, which stay in natural storage)
constexpr int SPLIT_MIN_N = 32;

// Staggered velocity multigrid for (theta alpha - A0) phi = rhs in precision T.

template <typename T>
struct StagMG {
  int dm = 3, nl = 1, vt = 1;
  int rows_n = 0;  // levels with n_x >= rows_n: row-mapped smoothing and operator launches
  int nu_down = 2, nu_up = 2, nu_bottom = 8, smoother = 1, max_vcycles = 1;
  T omega = T(1);
  double rel_tol = 1e-9;
  struct Lev {
    Geo g;
    int sp = 0;         // phi, rhs, res and alpha in parity-split storage
    long phi_span = 0;  // the components' phi arrays lie in one block of this many elements
    T dxsqinv;
    T* alpha[3] = {nullptr, nullptr, nullptr};
    T* beta = nullptr;
    T* gamma = nullptr;
    T* bed[3] = {nullptr, nullptr, nullptr};
    T* phi[3] = {nullptr, nullptr, nullptr};
    T* rhs[3] = {nullptr, nullptr, nullptr};
    T* res[3] = {nullptr, nullptr, nullptr};
    Coefs<T> C() const {
      Coefs<T> c;
      for (int d = 0; d < 3; ++d) { c.alpha[d] = alpha[d]; c.bed[d] = bed[d]; }
      c.beta = beta;
      c.gamma = gamma;
      c.dxsqinv = dxsqinv;
      c.theta = T(1);
      return c;
    }
  };
  std::vector<Lev> L;

  void build(const Geo& g0, double dx, int visc, const lmfhd_desc& d, DevMem& mem) {
    dm = g0.dm;
    vt = visc;
    nl = nlevs_mg(g0, d.stag_mg_minwidth);
    nu_down = d.stag_mg_nsmooths_down;
    nu_up = d.stag_mg_nsmooths_up;
    nu_bottom = d.stag_mg_nsmooths_bottom;
    smoother = d.stag_mg_smoother;
    max_vcycles = d.stag_mg_max_vcycles;
    omega = T(d.stag_mg_omega);
    rel_tol = d.stag_mg_rel_tol;
    L.resize(nl);
    Geo g = g0;
    double h = dx;
    for (int n = 0; n < nl; ++n) {
      Lev& l = L[n];
      l.g = g;
      l.dxsqinv = T(1.0 / (h * h));
      l.sp = split_level(g) ? 1 : 0;
      // phi of every component in one block (offsets rounded up to 64 elements), so that one
      // launch clears it
      l.phi_span = 0;
      for (int a = 0; a < dm; ++a) l.phi_span += (fstore(g, a, l.sp) + 63) / 64 * 64;
      g = coarsen_geo(g);
      h *= 2.0;
    }
    // the small levels' arrays in the arena, coarsest level first (the coarse kernel's LDS
    // staging); the rest in allocations of their own.  Everything starts zeroed (split storage
    // has slack slots, odd extents, that are never written)
    arena_l = nl;
    while (arena_l > 0) {
      bool small = true;
      for (int d = 0; d < dm; ++d) if (L[arena_l - 1].g.n[d] > COARSE_N) small = false;
      if (!small) break;
      --arena_l;
    }
    Arena ar;
    lev_end.assign(nl, 0);
    std::vector<std::array<long, 17>> offs(nl);
    for (int n = nl - 1; n >= arena_l; --n) {
      const Geo& gl = L[n].g;
      auto& o = offs[n];
      for (int a = 0; a < dm; ++a) o[a] = ar.take(fstore(gl, a, 0) * (long)sizeof(T));
      o[3] = ar.take(csize(gl) * (long)sizeof(T));
      o[4] = ar.take(csize(gl) * (long)sizeof(T));
      for (int e = 0; e < (dm == 3 ? 3 : 1); ++e) o[5 + e] = ar.take(esize(gl, e) * (long)sizeof(T));
      o[8] = ar.take(L[n].phi_span * (long)sizeof(T));
      for (int a = 0; a < dm; ++a) o[11 + a] = ar.take(fstore(gl, a, 0) * (long)sizeof(T));
      for (int a = 0; a < dm; ++a) o[14 + a] = ar.take(fstore(gl, a, 0) * (long)sizeof(T));
      lev_end[n] = ar.end();
    }
    arena = arena_l < nl ? mem.alloc_zero<unsigned char>(ar.end()) : nullptr;
    phi_off_l.assign(nl, 0);
    for (int n = 0; n < nl; ++n) {
      Lev& l = L[n];
      const Geo& gl = l.g;
      const bool in_arena = n >= arena_l;
      auto get = [&](int slot, long count) {
        return in_arena ? (T*)(arena + offs[n][slot]) : mem.alloc_zero<T>(count);
      };
      T* phis = get(8, l.phi_span);
      phi_off_l[n] = in_arena ? offs[n][8] : 0;
      long off = 0;
      for (int a = 0; a < dm; ++a) {
        const long ns = fstore(gl, a, l.sp);
        l.alpha[a] = get(a, ns);
        l.phi[a] = phis + off;
        off += (ns + 63) / 64 * 64;
        l.rhs[a] = get(11 + a, ns);
        l.res[a] = get(14 + a, ns);
      }
      l.beta = get(3, csize(gl));
      l.gamma = get(4, csize(gl));
      for (int e = 0; e < (dm == 3 ? 3 : 1); ++e) l.bed[e] = get(5 + e, esize(gl, e));
    }
  }
  // the arena: levels arena_l .. nl-1; lev_end[n] = end of level n's part (levels n .. nl-1 are
  // its prefix); phi_off_l[n] = offset of level n's phi block
  unsigned char* arena = nullptr;
  int arena_l = 0;
  std::vector<long> lev_end, phi_off_l;

  bool split_level(const Geo& g) const {
    if (smoother != 1 || (vt != 1 && vt != -1) || !even_level(g)) return false;
    for (int d = 0; d < dm; ++d)
      if (g.n[d] < SPLIT_MIN_N) return false;
    return true;
  }
  // natural-storage face array of level 0, component a -> the level's storage (and back)
  template <typename Ti>
  void to_level0(hipStream_t s, const Ti* src, T* dst, int a, Ti scale) {
    const Lev& l0 = L[0];
    if (l0.sp) launch(s, fsize(l0.g, a), SplitCopyFn<Ti, T>{l0.g, src, dst, scale, a, make_dec(fext(l0.g, a, 0), fext(l0.g, a, 1)), 1});
    else launch(s, fsize(l0.g, a), ConvertFn<T, Ti>{src, dst, scale});
  }
  template <typename To>
  void from_level0(hipStream_t s, const T* src, To* dst, int a) {
    const Lev& l0 = L[0];
    if (l0.sp) launch(s, fsize(l0.g, a), SplitCopyFn<T, To>{l0.g, src, dst, T(1), a, make_dec(fext(l0.g, a, 0), fext(l0.g, a, 1)), 0});
    else launch(s, fsize(l0.g, a), ConvertFn<To, T>{src, dst, T(1)});
  }

  // level-0 coefficients from FP64 arrays (alpha scaled by theta), then coarsen
  void set_coeffs(hipStream_t s, double* const* alpha64, const double* beta64, const double* gamma64,
                  double* const* bed64, double theta) {
    Lev& l0 = L[0];
    for (int a = 0; a < dm; ++a) to_level0<double>(s, alpha64[a], l0.alpha[a], a, theta);
    launch(s, csize(l0.g), ConvertFn<T, double>{beta64, l0.beta, 1.0});
    launch(s, csize(l0.g), ConvertFn<T, double>{gamma64, l0.gamma, 1.0});
    for (int e = 0; e < (dm == 3 ? 3 : 1); ++e) launch(s, esize(l0.g, e), ConvertFn<T, double>{bed64[e], l0.bed[e], 1.0});
    for (int n = 1; n < nl; ++n) {
      const Lev& f = L[n - 1];
      Lev& c = L[n];
      if (dm == 3) {
        launch(s, csize(c.g), CcRestrictFn<T, 3>{f.g, c.g, f.beta, c.beta, make_dec(c.g.n[0], c.g.n[1])});
        launch(s, csize(c.g), CcRestrictFn<T, 3>{f.g, c.g, f.gamma, c.gamma, make_dec(c.g.n[0], c.g.n[1])});
      } else {
        launch(s, csize(c.g), CcRestrictFn<T, 2>{f.g, c.g, f.beta, c.beta, make_dec(c.g.n[0], c.g.n[1])});
        launch(s, csize(c.g), CcRestrictFn<T, 2>{f.g, c.g, f.gamma, c.gamma, make_dec(c.g.n[0], c.g.n[1])});
      }
      for (int a = 0; a < dm; ++a) {
        const int nx = fext(c.g, a, 0), ny = fext(c.g, a, 1);
        if (dm == 3) launch(s, fsize(c.g, a), RestrictFn<T, 3>{f.g, c.g, f.alpha[a], c.alpha[a], a, make_dec(nx, ny), 1, f.sp, c.sp});
        else launch(s, fsize(c.g, a), RestrictFn<T, 2>{f.g, c.g, f.alpha[a], c.alpha[a], a, make_dec(nx, ny), 1, f.sp, c.sp});
      }
      for (int e = 0; e < (dm == 3 ? 3 : 1); ++e) launch(s, esize(c.g, e), EdgeRestrictFn<T>{f.g, c.g, f.bed[e], c.bed[e], e, make_dec(eext(c.g, e, 0), eext(c.g, e, 1))});
    }
    check_launch();
  }

  bool rows(const Geo& g) const { return rows_n > 0 && g.n[0] >= rows_n && !split_level(g); }

  // one smoothing sweep set at level n
  void smooth(hipStream_t s, int n, int nsweeps) {
    Lev& l = L[n];
    const bool inplace = even_level(l.g);
    const bool by_rows = rows(l.g);
    dispatch_vd(vt, dm, [&](auto V, auto D) {
      constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
      const Coefs<T> C = l.C();
      for (int m = 0; m < nsweeps; ++m) {
        if (smoother == 0) {  // weighted Jacobi: L on all faces, then update all faces
          apply_all<T, VT, DM>(s, l.g, C, l.phi, nullptr, l.res, 0, by_rows);
          for (int a = 0; a < dm; ++a) {
            const int nx = fext(l.g, a, 0), ny = fext(l.g, a, 1);
            const long nc = (long)((nx + 1) / 2) * ny * fext(l.g, a, 2);
            for (int par = 0; par < 2; ++par)
              launch(s, nc, UpdateColorFn<T, VT, DM>{l.g, C, l.phi[a], l.rhs[a], l.res[a], omega, a, make_dec_color(nx, ny), par});
          }
          continue;
        }
        if constexpr (VT == 1 || VT == -1) {
          if (inplace && by_rows) {  // the same, one (j, k) row of one component per wave
            GsRowsFn<T, VT, DM> F;
            F.g = l.g;
            F.C = C;
            int w = 0;
            for (int a = 0; a < 3; ++a) {
              const int q = a < dm ? a : dm - 1;
              F.phi[a] = l.phi[q];
              F.rhs[a] = l.rhs[q];
              const int nxm = fext(l.g, q, 0) - (q == 0 ? 1 : 0);
              const int cw = ((nxm + 1) / 2 + 63) / 64, ny = fext(l.g, q, 1);
              F.cw[a] = make_fastdiv((unsigned)cw);
              F.ny[a] = make_fastdiv((unsigned)ny);
              if (a < dm) w += cw * ny * fext(l.g, q, 2);
              F.wb[a] = w;
            }
            const int ct = ((l.g.n[1] + 1) / 2 + 63) / 64;
            F.ct = make_fastdiv((unsigned)ct);
            const int nw = w + ct * (dm == 3 ? l.g.n[2] : 1);
            F.omega = omega;
            for (int par = 0; par < 2; ++par) {
              F.par = par;
              launch_rows(s, nw, F);
            }
            continue;
          }
          // one launch per parity for all components (|visc_type| 1), natural or split storage
          auto fused = [&](auto SPc) {
            constexpr int SP = decltype(SPc)::value;
            GsColorAllFn<T, VT, DM, SP> F;
            F.g = l.g;
            F.C = C;
            long ncs[3] = {0, 0, 0}, cum = 0;
            for (int a = 0; a < 3; ++a) {
              const int q = a < dm ? a : dm - 1;
              const int nx = fext(l.g, q, 0), ny = fext(l.g, q, 1);
              F.phi[a] = l.phi[q];
              F.rhs[a] = l.rhs[q];
              F.dc[a] = make_dec_color(nx, ny);
              if (a < dm) ncs[a] = (long)((nx + 1) / 2) * ny * fext(l.g, q, 2);
              cum += ncs[a];
            }
            F.n0 = ncs[0];
            F.n01 = ncs[0] + ncs[1];
            F.omega = omega;
            // components interleaved by workgroup on levels of many workgroups
            long mx = 0;
            for (int a = 0; a < dm; ++a) {
              F.nc[a] = (unsigned)ncs[a];
              mx = std::max(mx, ncs[a]);
            }
            for (int a = dm; a < 3; ++a) F.nc[a] = 0;
            F.ilv = cum >= GS_ILV_MIN ? 1 : 0;
            const long nl = F.ilv ? (long)dm * ((mx + 255) / 256) * 256 : cum;
            for (int par = 0; par < 2; ++par) {
              F.par = par;
              launch(s, nl, F);
            }
          };
          if (inplace && l.sp && sizeof(T) == 4) {  // three components per position
            SplitPos3Fn<T, VT, DM, 1> F;
            F.g = l.g;
            F.C = C;
            for (int a = 0; a < 3; ++a) {
              const int q = a < dm ? a : dm - 1;
              F.phi[a] = l.phi[q];
              F.rhs[a] = l.rhs[q];
              F.out[a] = nullptr;
            }
            F.omega = omega;
            F.dn = make_dec_color(l.g.n[0] + 1, l.g.n[1] + 1);
            F.Hn = pos3_half(l.g);
            F.mode = 0;
            for (int par = 0; par < 2; ++par) {
              F.par = par;
              launch(s, (long)F.Hn, F);
            }
            continue;
          }
          if (inplace && l.sp) {
            fused(std::integral_constant<int, 1>{});
            continue;
          }
          if (inplace) {
            fused(std::integral_constant<int, 0>{});
            continue;
          }
        }
        for (int color = 1; color <= 2 * dm; ++color) {
          const int a = (color - 1) / 2, par = (color + 1) % 2;
          const int nx = fext(l.g, a, 0), ny = fext(l.g, a, 1);
          const long nc = (long)((nx + 1) / 2) * ny * fext(l.g, a, 2);
          if (inplace) {
            launch(s, nc, GsColorFn<T, VT, DM>{l.g, C, {l.phi[0], l.phi[1], l.phi[2]}, l.rhs[a], omega, a, make_dec_color(nx, ny), par});
          } else {
            launch(s, nc, ApplyColorFn<T, VT, DM>{l.g, C, {l.phi[0], l.phi[1], l.phi[2]}, l.res[a], a, make_dec_color(nx, ny), par});
            launch(s, nc, UpdateColorFn<T, VT, DM>{l.g, C, l.phi[a], l.rhs[a], l.res[a], omega, a, make_dec_color(nx, ny), par});
          }
        }
      }
    });
    check_launch();
  }

  // res = rhs - L phi at level n (zero on wall faces)
  void residual(hipStream_t s, int n) {
    Lev& l = L[n];
    dispatch_vd(vt, dm, [&](auto V, auto D) {
      constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
      apply_all<T, VT, DM>(s, l.g, l.C(), l.phi, l.rhs, l.res, 1, rows(l.g), l.sp);
    });
    check_launch();
  }
  // L phi on all faces of level 0 into out (no rhs); phi and out in natural storage.  A split
  // level 0 runs it on its own storage: phi into res, L phi into rhs (both are scratch between
  // solves), then back
  void apply(hipStream_t s, T* const* phi, T* const* out) {
    Lev& l = L[0];
    if (l.sp) {
      for (int a = 0; a < dm; ++a) to_level0<T>(s, phi[a], l.res[a], a, T(1));
      dispatch_vd(vt, dm, [&](auto V, auto D) {
        constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
        apply_all<T, VT, DM>(s, l.g, l.C(), l.res, nullptr, l.rhs, 0, false, 1);
      });
      for (int a = 0; a < dm; ++a) from_level0<T>(s, l.rhs[a], out[a], a);
      check_launch();
      return;
    }
    dispatch_vd(vt, dm, [&](auto V, auto D) {
      constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
      apply_all<T, VT, DM>(s, l.g, l.C(), phi, nullptr, out, 0, rows(l.g));
    });
    check_launch();
  }

  // residual restriction, all components in one column launch; zeroes the coarse level's phi
  template <int DM>
  void restrict_res_dm(hipStream_t s, int n) {
    const Lev& f = L[n];
    Lev& c = L[n + 1];
    AllFacesFn<RestrictFn<T, DM>, DM> F;
    F.g = c.g;
    for (int a = 0; a < 3; ++a) {
      const int b = a < DM ? a : DM - 1;
      F.f[a] = RestrictFn<T, DM>{f.g, c.g, f.res[b], c.rhs[b], b, Dec{}, 0, f.sp, c.sp, c.phi[b]};
    }
    const ColShape cs = col_shape(c.g);
    F.cx = cs.cx;
    launch_cols(s, cs.ncols, cs.nz, cs.kc, F);
  }
  void restrict_res(hipStream_t s, int n) {
    if (dm == 3) restrict_res_dm<3>(s, n);
    else restrict_res_dm<2>(s, n);
    check_launch();
  }
  // prolongation-add into level n from the coarse side (ProlongCFn), all components in one
  // launch
  void prolong(hipStream_t s, int n) {
    Lev& f = L[n];
    const Lev& c = L[n + 1];
    if (dm == 3) prolong_all<3>(s, f, c);
    else prolong_all<2>(s, f, c);
    check_launch();
  }
  template <int DM>
  void prolong_all(hipStream_t s, Lev& f, const Lev& c) {
    Comp3BlkFn<ProlongCFn<T, DM>, DM> F;
    long mx = 0;
    for (int a = 0; a < 3; ++a) {
      const int b = a < DM ? a : DM - 1;
      F.f[a] = ProlongCFn<T, DM>{f.g, c.g, f.phi[b], c.phi[b], b, make_dec(fext(c.g, b, 0), fext(c.g, b, 1)), f.sp, c.sp};
      F.n[a] = a < DM ? (unsigned)fsize(c.g, b) : 0u;
      mx = std::max(mx, (long)F.n[a]);
    }
    launch(s, (long)DM * ((mx + 255) / 256) * 256, F);
  }

  // levels l0..nl-1 of one V-cycle in one workgroup (phi of levels >= max(l0,1) zeroed by the
  // restrictions into them); the argument block (pointers and decoders of every level) lives in
  // device memory
  StagCoarseArgs<T>* coarse_args = nullptr;
  int coarse_l0 = -1;
  int coarse_nt = 512;  // threads of the one-workgroup kernel: 256, 512 or 1024
  void build_coarse(int l0, int nt, hipStream_t s, DevMem& mem) {
    coarse_l0 = l0;
    coarse_nt = nt;
    if (l0 >= nl) return;
    StagCoarseArgs<T> A;
    memset(&A, 0, sizeof(A));
    A.l0 = l0;
    A.nl = nl;
    A.nu_down = nu_down;
    A.nu_up = nu_up;
    A.nu_bottom = nu_bottom;
    A.omega = omega;
    A.fuse = 1;
    for (int n = l0; n < nl; ++n) {
      StagLevArgs<T>& a = A.L[n];
      const Lev& l = L[n];
      a.g = l.g;
      a.dxsqinv = l.dxsqinv;
      for (int d = 0; d < 3; ++d) {
        a.alpha[d] = l.alpha[d];
        a.bed[d] = l.bed[d];
        a.phi[d] = l.phi[d];
        a.rhs[d] = l.rhs[d];
        a.res[d] = l.res[d];
        if (d < dm) {
          a.dface[d] = make_dec(fext(l.g, d, 0), fext(l.g, d, 1));
          a.dcolor[d] = make_dec_color(fext(l.g, d, 0), fext(l.g, d, 1));
        }
      }
      a.beta = l.beta;
      a.gamma = l.gamma;
    }
    A.st = stage_info(arena, arena_l, lev_end, phi_off_l[l0], L[l0].phi_span * (long)sizeof(T), l0);
    if (!coarse_args) coarse_args = mem.alloc<StagCoarseArgs<T>>(1);
    h2d(coarse_args, &A, sizeof(A), s);
  }
  void coarse(hipStream_t s, int l0) {
    (void)l0;
    dispatch_vd(vt, dm, [&](auto V, auto D) {
      constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
      const int nt = coarse_nt;
      if (nt == 256) launch_blocks(s, 1, StagCoarseBlk<T, VT, DM, 256>{coarse_args});
      else if (nt == 512) launch_blocks(s, 1, StagCoarseBlk<T, VT, DM, 512>{coarse_args});
      else launch_blocks(s, 1, StagCoarseBlk<T, VT, DM, 1024>{coarse_args});
    });
    check_launch();
  }

  // V-cycles on L[0].rhs with L[0].phi as the initial guess (stag_mg_solver.f90:323-807)
  void solve(hipStream_t s, Reducer& R) {
    double resid0[3] = {0, 0, 0};
    if (max_vcycles > 1) {
      residual(s, 0);
      for (int a = 0; a < dm; ++a) {
        R.absmax(L[0].res[a], fstore(L[0].g, a, L[0].sp), R.out);  // slack slots hold zeros
        d2h(&resid0[a], R.out, sizeof(double), s);
      }
      bool allzero = true, anyzero = false;
      double mx = 0.0;
      for (int a = 0; a < dm; ++a) {
        if (resid0[a] != 0.0) allzero = false; else anyzero = true;
        mx = std::max(mx, resid0[a]);
      }
      if (allzero) return;
      if (anyzero) for (int a = 0; a < dm; ++a) resid0[a] = mx;
    }
    const int l0 = (smoother == 1) ? coarse_l0 : nl;
    for (int vc = 0; vc < max_vcycles; ++vc) {
      // the restriction into level n + 1 zeroes that level's phi (every face; the slack slots of
      // split storage are zero from allocation and never written)
      for (int n = 0; n < std::min(l0, nl - 1); ++n) {
        smooth(s, n, nu_down);
        residual(s, n);
        restrict_res(s, n);
      }
      if (l0 < nl) {
        coarse(s, l0);
      } else {
        smooth(s, nl - 1, nu_bottom);
      }
      for (int n = std::min(l0, nl - 1) - 1; n >= 0; --n) {
        prolong(s, n);
        smooth(s, n, nu_up);
      }
      if (max_vcycles > 1) {
        residual(s, 0);
        bool done = true;
        for (int a = 0; a < dm; ++a) {
          double r;
          R.absmax(L[0].res[a], fstore(L[0].g, a, L[0].sp), R.out);
          d2h(&r, R.out, sizeof(double), s);
          if (!(r / resid0[a] < rel_tol)) done = false;
        }
        if (done) break;
      }
    }
  }
};

// ------------------------------------------------------------------------------------------
// Cell-centred multigrid for D(b G) phi = rhs, b = 1/alpha_fc (0 on wall faces).
template <typename T>
struct MacMG {
  int dm = 3, nl = 1, nu_down = 2, nu_up = 2, nu_bottom = 8, max_vcycles = 1;
  struct Lev {
    Geo g;
    T h2i;
    T* b[3] = {nullptr, nullptr, nullptr};
    T* phi = nullptr;
    T* rhs = nullptr;
    T* tmp = nullptr;
  };
  std::vector<Lev> L;

  void build(const Geo& g0, double dx, const lmfhd_desc& d, DevMem& mem) {
    dm = g0.dm;
    nl = nlevs_mg(g0, d.mg_minwidth);
    nu_down = d.mg_nsmooths_down;
    nu_up = d.mg_nsmooths_up;
    nu_bottom = d.mg_nsmooths_bottom;
    max_vcycles = d.mg_max_vcycles;
    L.resize(nl);
    Geo g = g0;
    double h = dx;
    for (int n = 0; n < nl; ++n) {
      L[n].g = g;
      L[n].h2i = T(1.0 / (h * h));
      g = coarsen_geo(g);
      h *= 2.0;
    }
    // the small levels in the arena, coarsest first (the coarse kernel's LDS staging)
    arena_l = nl;
    while (arena_l > 0) {
      bool small = true;
      for (int d = 0; d < dm; ++d) if (L[arena_l - 1].g.n[d] > COARSE_N) small = false;
      if (!small) break;
      --arena_l;
    }
    Arena ar;
    lev_end.assign(nl, 0);
    std::vector<std::array<long, 6>> offs(nl);
    for (int n = nl - 1; n >= arena_l; --n) {
      const Geo& gl = L[n].g;
      for (int a = 0; a < dm; ++a) offs[n][a] = ar.take(fsize(gl, a) * (long)sizeof(T));
      for (int q = 3; q < 6; ++q) offs[n][q] = ar.take(csize(gl) * (long)sizeof(T));
      lev_end[n] = ar.end();
    }
    arena = arena_l < nl ? mem.alloc_zero<unsigned char>(ar.end()) : nullptr;
    phi_off_l.assign(nl, 0);
    for (int n = 0; n < nl; ++n) {
      Lev& l = L[n];
      const bool in_arena = n >= arena_l;
      auto get = [&](int slot, long count) { return in_arena ? (T*)(arena + offs[n][slot]) : mem.alloc<T>(count); };
      for (int a = 0; a < dm; ++a) l.b[a] = get(a, fsize(l.g, a));
      l.phi = get(3, csize(l.g));
      l.rhs = get(4, csize(l.g));
      l.tmp = get(5, csize(l.g));
      phi_off_l[n] = in_arena ? offs[n][3] : 0;
    }
  }
  unsigned char* arena = nullptr;  // levels arena_l .. nl-1, as in StagMG
  int arena_l = 0;
  std::vector<long> lev_end, phi_off_l;
  void set_coeffs(hipStream_t s, double* const* alpha64) {
    Lev& l0 = L[0];
    for (int a = 0; a < dm; ++a) {
      const int nx = fext(l0.g, a, 0), ny = fext(l0.g, a, 1);
      launch(s, fsize(l0.g, a), MacCoef0Fn<T>{l0.g, alpha64[a], l0.b[a], a, make_dec(nx, ny)});
    }
    for (int n = 1; n < nl; ++n) {
      const Lev& f = L[n - 1];
      Lev& c = L[n];
      for (int a = 0; a < dm; ++a) {
        const int nx = fext(c.g, a, 0), ny = fext(c.g, a, 1);
        if (dm == 3) launch(s, fsize(c.g, a), RestrictFn<T, 3>{f.g, c.g, f.b[a], c.b[a], a, make_dec(nx, ny), 1});
        else launch(s, fsize(c.g, a), RestrictFn<T, 2>{f.g, c.g, f.b[a], c.b[a], a, make_dec(nx, ny), 1});
      }
    }
    check_launch();
  }
  template <int DM>
  void gs(hipStream_t s, int n, int nsweeps) {
    Lev& l = L[n];
    const long nc = (long)((l.g.n[0] + 1) / 2) * l.g.n[1] * l.g.n[2];
    const bool inplace = even_level(l.g);
    for (int m = 0; m < nsweeps; ++m)
      for (int par = 0; par < 2; ++par) {
        if (inplace) {
          launch(s, nc, MacGsFn<T, DM>{l.g, {l.b[0], l.b[1], l.b[2]}, l.h2i, l.phi, l.rhs, par, make_dec_color(l.g.n[0], l.g.n[1])});
        } else {
          launch(s, nc, MacGsNewFn<T, DM>{l.g, {l.b[0], l.b[1], l.b[2]}, l.h2i, l.phi, l.rhs, l.tmp, par, make_dec_color(l.g.n[0], l.g.n[1])});
          launch(s, nc, CopyColorCellFn<T>{l.g, l.phi, l.tmp, par, make_dec_color(l.g.n[0], l.g.n[1])});
        }
      }
  }
  MacCoarseArgs<T>* coarse_args = nullptr;
  int coarse_l0 = -1;
  int coarse_nt = 512;
  void build_coarse(int l0, int nt, hipStream_t s, DevMem& mem) {
    coarse_l0 = l0;
    coarse_nt = nt;
    if (l0 >= nl) return;
    MacCoarseArgs<T> A;
    memset(&A, 0, sizeof(A));
    A.l0 = l0;
    A.nl = nl;
    A.nu_down = nu_down;
    A.nu_up = nu_up;
    A.nu_bottom = nu_bottom;
    for (int n = l0; n < nl; ++n) {
      MacLevArgs<T>& a = A.L[n];
      a.g = L[n].g;
      a.h2i = L[n].h2i;
      for (int d = 0; d < 3; ++d) a.b[d] = L[n].b[d];
      a.phi = L[n].phi;
      a.rhs = L[n].rhs;
      a.tmp = L[n].tmp;
      a.dcell = make_dec(L[n].g.n[0], L[n].g.n[1]);
      a.dcolor = make_dec_color(L[n].g.n[0], L[n].g.n[1]);
    }
    A.st = stage_info(arena, arena_l, lev_end, phi_off_l[l0], csize(L[l0].g) * (long)sizeof(T), l0);
    if (!coarse_args) coarse_args = mem.alloc<MacCoarseArgs<T>>(1);
    h2d(coarse_args, &A, sizeof(A), s);
  }
  template <int DM>
  void coarse(hipStream_t s, int l0) {
    (void)l0;
    const int nt = coarse_nt;
    if (nt == 256) launch_blocks(s, 1, MacCoarseBlk<T, DM, 256>{coarse_args});
    else if (nt == 512) launch_blocks(s, 1, MacCoarseBlk<T, DM, 512>{coarse_args});
    else launch_blocks(s, 1, MacCoarseBlk<T, DM, 1024>{coarse_args});
    check_launch();
  }
  template <int DM>
  void solve_dm(hipStream_t s) {
    const int l0 = coarse_l0;
    if (l0 < nl) {
      for (int vc = 0; vc < max_vcycles; ++vc) {
        for (int n = 0; n < l0; ++n) {  // the restriction zeroes phi of level n + 1
          gs<DM>(s, n, nu_down);
          const Lev& f = L[n];
          Lev& c = L[n + 1];
          launch(s, csize(c.g), MacResidRestrictFn<T, DM>{f.g, c.g, {f.b[0], f.b[1], f.b[2]}, f.h2i, f.phi, f.rhs, c.rhs, make_dec(c.g.n[0], c.g.n[1]), c.phi});
        }
        coarse<DM>(s, l0);
        for (int n = l0 - 1; n >= 0; --n) {
          launch(s, csize(L[n].g), MacProlongFn<T>{L[n].g, L[n + 1].g, L[n].phi, L[n + 1].phi, make_dec(L[n].g.n[0], L[n].g.n[1])});
          gs<DM>(s, n, nu_up);
        }
      }
      check_launch();
      return;
    }
    for (int vc = 0; vc < max_vcycles; ++vc) {
      for (int n = 0; n < nl - 1; ++n) {  // the restriction zeroes phi of level n + 1
        gs<DM>(s, n, nu_down);
        const Lev& f = L[n];
        Lev& c = L[n + 1];
        launch(s, csize(c.g), MacResidRestrictFn<T, DM>{f.g, c.g, {f.b[0], f.b[1], f.b[2]}, f.h2i, f.phi, f.rhs, c.rhs, make_dec(c.g.n[0], c.g.n[1]), c.phi});
      }
      remove_mean_small<T>(s, L[nl - 1].rhs, csize(L[nl - 1].g));
      gs<DM>(s, nl - 1, nu_bottom);
      for (int n = nl - 2; n >= 0; --n) {
        launch(s, csize(L[n].g), MacProlongFn<T>{L[n].g, L[n + 1].g, L[n].phi, L[n + 1].phi, make_dec(L[n].g.n[0], L[n].g.n[1])});
        gs<DM>(s, n, nu_up);
      }
    }
    check_launch();
  }
  // phi (level 0) from zero for L[0].rhs; phi_zeroed: the caller's last kernel zeroed it
  void solve(hipStream_t s, bool phi_zeroed = false) {
    if (!phi_zeroed) fill<T>(s, L[0].phi, csize(L[0].g), T(0));
    if (dm == 3) solve_dm<3>(s);
    else solve_dm<2>(s);
  }
};

// ------------------------------------------------------------------------------------------
// GMRES least squares on the device in FP64 (least_squares, rotmat, SolveUTriangular,
// gmres.f90:531-619).  A restart cycle's Givens state lives in device memory -- H, (m+2) x (m+1)
// and 1-based as in the reference, the rotations cs, sn, the rotated right-hand side s and the
// solution y -- and one thread does the O(i) work of a column.  The host reads only the status
// {|s(i+1)|, stop flag, H(i+1, i)} of each step.
struct LsqDev {
  double* H;
  double* cs;
  double* sn;
  double* sv;
  double* y;
  double* stat;  // [0] |s(i+1)|, the residual estimate; [1] 1 when it passes the stop test; [2] H(i+1, i)
  double* thr;   // stop thresholds: [0] relative (rel_tol times the reference norm), [1] absolute
  int m;
  HD double& h(int r, int col) const { return H[(long)r * (m + 1) + col]; }
};
HD void rotmat_hd(double a, double b, double& cs, double& sn) {
  if (b == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (fabs(b) > fabs(a)) {
    const double t = a / b;
    sn = 1.0 / sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = b / a;
    cs = 1.0 / sqrt(1.0 + t * t);
    sn = t * cs;
  }
}
// least_squares (gmres.f90:531-566) for column i: H(1..i, i) = h1 + h2 (the two Gram-Schmidt
// passes), H(i+1, i) = |w|; earlier rotations, the new rotation, s, the estimate and the test
HD void lsq_column_hd(const LsqDev& G, int i, const double* h1, const double* h2, const double* nrm) {
  for (int r = 1; r <= i; ++r) G.h(r, i) = h1[r - 1] + h2[r - 1];
  G.h(i + 1, i) = nrm[0];
  for (int k = 1; k <= i - 1; ++k) {
    const double t = G.cs[k] * G.h(k, i) + G.sn[k] * G.h(k + 1, i);
    G.h(k + 1, i) = -G.sn[k] * G.h(k, i) + G.cs[k] * G.h(k + 1, i);
    G.h(k, i) = t;
  }
  rotmat_hd(G.h(i, i), G.h(i + 1, i), G.cs[i], G.sn[i]);
  const double t = G.cs[i] * G.sv[i];
  G.sv[i + 1] = -G.sn[i] * G.sv[i];
  G.sv[i] = t;
  G.h(i, i) = G.cs[i] * G.h(i, i) + G.sn[i] * G.h(i + 1, i);
  G.h(i + 1, i) = 0.0;
  const double est = fabs(G.sv[i + 1]);
  G.stat[0] = est;
  G.stat[1] = (est <= G.thr[0] || est <= G.thr[1]) ? 1.0 : 0.0;
  G.stat[2] = nrm[0];
}
HD void lsq_init_hd(const LsqDev& G, const double* nrm, double thr_rel, double thr_abs) {
  G.sv[1] = nrm[0];
  G.thr[0] = thr_rel;
  G.thr[1] = thr_abs;
}
// SolveUTriangular (gmres.f90:599-619) for k = i - 1, then coef[q] = y(q + 1) for update_sol
HD void lsq_solve_hd(const LsqDev& G, int i, double* coef) {
  const int k = i - 1;
  G.y[k + 1] = G.sv[k + 1] / G.h(k + 1, k + 1);
  for (int r = k; r >= 1; --r) {
    double dot = 0.0;
    for (int col = r + 1; col <= k + 1; ++col) dot += G.h(r, col) * G.y[col];
    G.y[r] = (G.sv[r] - dot) / G.h(r, r);
  }
  for (int q = 0; q < i; ++q) coef[q] = G.y[q + 1];
}
#ifndef LMFHD_HOST_DEBUG
__global__ void k_lsq_column(LsqDev G, int i, const double* __restrict__ h1, const double* __restrict__ h2,
                             const double* __restrict__ nrm) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lsq_column_hd(G, i, h1, h2, nrm);
}
__global__ void k_lsq_init(LsqDev G, const double* __restrict__ nrm, double thr_rel, double thr_abs) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lsq_init_hd(G, nrm, thr_rel, thr_abs);
}
__global__ void k_lsq_solve(LsqDev G, int i, double* __restrict__ coef) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lsq_solve_hd(G, i, coef);
}
#endif

// ------------------------------------------------------------------------------------------
// Launch parameters of a size class (the autotuner's choices; lmfhd_tune in the C ABI).
struct Tune {
  int coarse_n = 8;         // staggered levels with every n_d <= coarse_n run in one workgroup
  int mac_coarse_n = 8;     // the same for the MAC hierarchy
  int coarse_threads = 512; // threads of that workgroup
  int red_blocks = 1024;    // workgroups of an inner-product launch
  int graphs = 0;           // HIP-graph replay of Arnoldi steps and preconditioner applications
  int rows_n = 0;           // staggered levels with n_x >= rows_n smooth and apply by rows (0: never)
  double ms = 0.0;          // measured time per solve of the winner (autotuner), 0 when set by hand
};

// ------------------------------------------------------------------------------------------
struct Ctx {
  lmfhd_desc d;
  Geo g;
  int dm = 3;
  double dx = 1.0;
  // the caller's boxes (fab views per box); labels: no box list, one fab whose lo / hi are labels
  int nb = 1;
  bool labels = true;
  std::vector<int> blo, bhi;  // 3 per box
  hipStream_t s = nullptr;
  DevMem mem;
  Reducer R;
  VecLayout VL;
  long nf[3] = {0, 0, 0}, nc = 0, vlen = 0, off[4] = {0, 0, 0, 0};
  // FP64 coefficients
  double* alpha64[3] = {nullptr, nullptr, nullptr};
  double* beta64 = nullptr;
  double* gamma64 = nullptr;
  double* bed64[3] = {nullptr, nullptr, nullptr};
  long ne[3] = {0, 0, 0};
  double theta = 0.0;
  bool have_coef = false;
  StagMG<double> mg64;
  StagMG<float> mg32;
  MacMG<double> mac64;
  MacMG<float> mac32;
  // visc_schur_approx = 1: the MAC hierarchy of face coefficient Minv = 1/diag(A0), and D =
  // h^2 diag(A0) on the faces (Stokes-vector sized)
  MacMG<double> macmu64;
  MacMG<float> macmu32;
  double* diaglap = nullptr;
  // filtering_width > 0: dm cell stresses, 2 x (3-D: 3, 2-D: 1) edge stresses and one spare
  // array for the filter passes, each stress_slot doubles (allocated on first use)
  double* stress = nullptr;
  long stress_slot = 0;
  // inhomogeneous walls: phi_bc, then staging for vel_bc_n and vel_bc_t (allocated on first use)
  double* bc_buf = nullptr;
  // MAC right-hand side / stag MG output staging per precision
  double* macrhs64 = nullptr;
  float* macrhs32 = nullptr;
  // Stokes vectors (FP64): Krylov basis and work
  int m = 5;
  double* V = nullptr;  // (m+1) * vlen
  double* Z = nullptr;  // m * vlen (FGMRES)
  double* r = nullptr;
  double* w = nullptr;
  double* tmp = nullptr;
  double* io_a = nullptr;  // staging for the fab entry points
  double* io_b = nullptr;
  // momentum time steps (allocated on first use, Stokes-vector sized): v^n, A^n, L(v^n), the
  // right-hand side and the increment of each solve
  double* st_vn = nullptr;
  double* st_adv = nullptr;
  double* st_dif = nullptr;
  double* st_b = nullptr;
  double* st_dx = nullptr;
  double* scal = nullptr;  // device scalars: h1 [0, m+1), h2 [m+1, 2m+2), norm, temp
  double* s_h1() const { return scal; }
  double* s_h2() const { return scal + (m + 1); }
  double* s_nrm() const { return scal + 2 * (m + 1); }
  double* s_tmp() const { return scal + 2 * (m + 1) + 1; }
  double* s_wn2() const { return scal + 2 * (m + 1) + 3; }
  double* s_skip() const { return scal + 2 * (m + 1) + 4; }
  long nscal() const { return 2 * (m + 1) + 8; }
  double* coef = nullptr;  // device coefficients for maxpy (m+1)
  LsqDev G;                // device Givens state of the restart cycle
  Tune tune;               // launch parameters in use
  // lmfhd_desc.batch: the size class's lock-step group, the recorded preconditioner sequence
  // (pin -> pout) and its staging vectors, the event marking pin ready
  std::shared_ptr<struct Batch> batch;
  Recorder rec;
  bool rec_ok = false;
  double* pin = nullptr;
  double* pout = nullptr;
  hipEvent_t ev_in = nullptr;
  // launch sequences captured as HIP graphs (run_graphed): kind 0 a preconditioner
  // application between buffers (a, b), kind 1 / 2 the Arnoldi step idx of GMRES / FGMRES
  struct GraphEntry {
    int kind, idx;
    const void* a;
    const void* b;
#ifndef LMFHD_HOST_DEBUG
    hipGraphExec_t exec;
#endif
  };
  std::vector<GraphEntry> graphs;
  bool use_graphs = false;
  bool graph_failed = false;
  void clear_graphs() {
#ifndef LMFHD_HOST_DEBUG
    if (!graphs.empty()) (void)hipStreamSynchronize(s);  // none may still be in flight
    for (auto& e : graphs) (void)hipGraphExecDestroy(e.exec);
#endif
    graphs.clear();
  }

  double* u(double* v, int a) const { return v + off[a]; }
  double* p(double* v) const { return v + off[3]; }
};

// ------------------------------------------------------------------------------------------
// apply_matrix: b = A x ; x wall faces zeroed in place first (apply_matrix.f90:71-115)
static void apply_matrix(Ctx& c, double* x, double* b) {
  hipStream_t s = c.s;
  if (has_wall(c.g))
    for (int a = 0; a < c.dm; ++a) {
      const int nx = fext(c.g, a, 0), ny = fext(c.g, a, 1);
      launch(s, c.nf[a], ZeroWallFacesFn<double>{c.g, c.u(x, a), a, make_dec(nx, ny)});
    }
  Coefs<double> C;
  for (int a = 0; a < 3; ++a) { C.alpha[a] = c.alpha64[a]; C.bed[a] = c.bed64[a]; }
  C.beta = c.beta64;
  C.gamma = c.gamma64;
  C.dxsqinv = 1.0 / (c.dx * c.dx);
  C.theta = c.theta;
  const double dxinv = 1.0 / c.dx;
  if (c.d.gmres_spatial_order == 4) {
    auto run4 = [&](auto D) {
      constexpr int DM = decltype(D)::value;
      for (int a = 0; a < c.dm; ++a) {
        AmFace4Fn<DM> F;
        F.g = c.g;
        for (int d = 0; d < 3; ++d) {
          const int q = d < c.dm ? d : c.dm - 1;
          F.alpha[d] = c.alpha64[q];
          F.xu[d] = c.u(x, q);
        }
        F.theta = c.theta;
        F.beta = c.beta64;
        F.xp = c.p(x);
        F.out = c.u(b, a);
        F.t12sq = 1.0 / (12.0 * (c.dx * c.dx));
        F.t12 = 1.0 / (12.0 * c.dx);
        F.a = a;
        F.dc = make_dec(fext(c.g, a, 0), fext(c.g, a, 1));
        launch(s, c.nf[a], F);
      }
      launch(s, c.nc, NegDivFn<DM>{c.g, {c.u(x, 0), c.u(x, 1), c.dm == 3 ? c.u(x, 2) : c.u(x, 1)}, c.p(b), dxinv, make_dec(c.g.n[0], c.g.n[1])});
    };
    if (c.dm == 3) run4(std::integral_constant<int, 3>{});
    else run4(std::integral_constant<int, 2>{});
    check_launch();
    return;
  }
  dispatch_vd(c.d.visc_type, c.dm, [&](auto V, auto D) {
    constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
    // one launch per component and one for the cells: measured faster than one column
    // launch for everything (108 us against 137 us at 128^3)
    for (int a = 0; a < c.dm; ++a) {
      const int nx = fext(c.g, a, 0), ny = fext(c.g, a, 1);
      launch(s, c.nf[a], AmFaceFn<VT, DM>{c.g, C, {c.u(x, 0), c.u(x, 1), c.dm == 3 ? c.u(x, 2) : c.u(x, 1)}, c.p(x), c.u(b, a), dxinv, a, make_dec(nx, ny)});
    }
    launch(s, c.nc, NegDivFn<DM>{c.g, {c.u(x, 0), c.u(x, 1), c.dm == 3 ? c.u(x, 2) : c.u(x, 1)}, c.p(b), dxinv, make_dec(c.g.n[0], c.g.n[1])});
  });
  check_launch();
}

// FP64 Lphi = (theta alpha - A0) phi on all faces (packed face arrays)
static void stag_applyop64(Ctx& c, double* phi, double* out, int color) {
  Coefs<double> C;
  for (int a = 0; a < 3; ++a) { C.alpha[a] = c.alpha64[a]; C.bed[a] = c.bed64[a]; }
  C.beta = c.beta64;
  C.gamma = c.gamma64;
  C.dxsqinv = 1.0 / (c.dx * c.dx);
  C.theta = c.theta;
  dispatch_vd(c.d.visc_type, c.dm, [&](auto V, auto D) {
    constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
    double* ph[3] = {c.u(phi, 0), c.u(phi, 1), c.dm == 3 ? c.u(phi, 2) : c.u(phi, 1)};
    if (color == 0) {
      double* po[3] = {c.u(out, 0), c.u(out, 1), c.dm == 3 ? c.u(out, 2) : c.u(out, 1)};
      apply_all<double, VT, DM>(c.s, c.g, C, ph, nullptr, po, 0, c.tune.rows_n > 0 && c.g.n[0] >= c.tune.rows_n);
      return;
    }
    for (int a = 0; a < c.dm; ++a) {
      const int nx = fext(c.g, a, 0), ny = fext(c.g, a, 1);
      if ((color - 1) / 2 == a) {
        const int par = (color + 1) % 2;
        const long ncol = (long)((nx + 1) / 2) * ny * fext(c.g, a, 2);
        launch(c.s, ncol, ApplyColorFn<double, VT, DM>{c.g, C, {ph[0], ph[1], ph[2]}, c.u(out, a), a, make_dec_color(nx, ny), par});
      }
    }
  });
  // stag_applyop_level ends with multifab_physbc_domainvel on every component of Lphi
  if (color != 0)
    for (int a = 0; a < c.dm; ++a)
      launch(c.s, c.nf[a], ZeroWallFacesFn<double>{c.g, c.u(out, a), a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  check_launch();
}

// null-space handling at the end of apply_precon (apply_precon.f90:433-458): subtract the
// mean pressure; with theta = 0 and no no-slip wall also the velocity means
static void precon_tail(Ctx& c, double* x) {
  hipStream_t s = c.s;
  const int dm = c.dm;
  double* xp = x + c.off[3];
  c.R.mdot(xp, nullptr, 0, 1, c.nc, false, c.s_tmp());
  sub_mean(s, xp, c.s_tmp(), (double)c.nc, c.nc);
  if (c.theta == 0.0) {
    bool no_wall_is_no_slip = true;
    for (int d = 0; d < dm; ++d)
      if (noslip_code(c.d.bc_lo[d]) || noslip_code(c.d.bc_hi[d])) no_wall_is_no_slip = false;
    if (no_wall_is_no_slip)
      for (int a = 0; a < dm; ++a) {
        if (!c.g.per[a]) continue;
        c.R.mdot(x + c.off[a], nullptr, 0, 1, c.nf[a], true, c.s_tmp(), false, c.off[a]);
        sub_mean(s, x + c.off[a], c.s_tmp(), (double)c.nc, c.nf[a]);
      }
  }
  check_launch();
}

// apply_precon, precon_type 1 (apply_precon.f90:103-195,433-460): b -> x (FP64 vectors)
template <typename T>
static void precon(Ctx& c, StagMG<T>& mg, MacMG<T>& mac, T* macrhs, const double* b, double* x) {
  hipStream_t s = c.s;
  const int dm = c.dm;
  auto& l0 = mg.L[0];
  const ColShape cs = col_shape(c.g);
  auto enter = [&](auto D) {
    constexpr int DM = decltype(D)::value;
    AllFacesFn<EnterFn<T>, DM> F;
    F.g = c.g;
    for (int a = 0; a < 3; ++a) {
      const int q = a < DM ? a : DM - 1;
      F.f[a] = EnterFn<T>{b + c.off[q], l0.rhs[q], l0.phi[q], l0.g, q, l0.sp};
    }
    F.cx = cs.cx;
    launch_cols(s, cs.ncols, cs.nz, cs.kc, F);
  };
  if (dm == 3) enter(std::integral_constant<int, 3>{});
  else enter(std::integral_constant<int, 2>{});
  mg.solve(s, c.R);
  // mac_rhs = D x_u + b_p (kept twice: the MAC solve consumes its copy)
  const T* ph[3] = {l0.phi[0], l0.phi[1], dm == 3 ? l0.phi[2] : l0.phi[1]};
  if (dm == 3) {
    DivPlusFn<T, 3> f{c.g, {ph[0], ph[1], ph[2]}, b + c.off[3], mac.L[0].rhs, T(1.0 / c.dx), make_dec(c.g.n[0], c.g.n[1])};
    f.out2 = macrhs;
    f.sp = l0.sp;
    f.z = mac.L[0].phi;
    launch(s, c.nc, f);
  } else {
    DivPlusFn<T, 2> f{c.g, {ph[0], ph[1], ph[2]}, b + c.off[3], mac.L[0].rhs, T(1.0 / c.dx), make_dec(c.g.n[0], c.g.n[1])};
    f.out2 = macrhs;
    f.sp = l0.sp;
    f.z = mac.L[0].phi;
    launch(s, c.nc, f);
  }
  mac.solve(s, true);
  auto leave = [&](auto D) {
    constexpr int DM = decltype(D)::value;
    FacesCellFn<PreconFaceFn<T>, PreconCellFn<T>, DM> F;
    F.g = c.g;
    for (int a = 0; a < 3; ++a) {
      const int q = a < DM ? a : DM - 1;
      F.f[a] = PreconFaceFn<T>{c.g, l0.phi[q], mac.L[0].phi, c.alpha64[q], x + c.off[q], 1.0 / c.dx, q, Dec{}, l0.sp};
    }
    F.c = PreconCellFn<T>{macrhs, mac.L[0].phi, c.beta64, c.gamma64, x + c.off[3], c.theta, std::abs(c.d.visc_type)};
    F.cx = cs.cx;
    launch_cols(s, cs.ncols, cs.nz, cs.kc, F);
  };
  if (dm == 3) leave(std::integral_constant<int, 3>{});
  else leave(std::integral_constant<int, 2>{});
  // subtract the mean pressure (sum_umac_press, apply_precon.f90:433-437)
  precon_tail(c, x);
}

// visc_schur_complement (apply_precon.f90:538-685): x_p = L_mu^-1 D Minv (0 - A0) Minv G
// L_mu^-1 r (+ theta phi when the L_alpha solve ran), L_mu the MAC operator of face coefficient
// Minv (mmu; its level-0 coefficients are 1/D, zero on wall faces).  Level-0 staggered buffers
// serve as scratch: res holds Minv G phi_mu, rhs the stencil result (the velocity in phi is
// left alone).  A0 runs with alpha = 0 (theta 0) and the wrap / reflect ghost rules.
template <typename T>
static void visc_schur(Ctx& c, StagMG<T>& mg, MacMG<T>& mac, MacMG<T>& mmu, const T* r, bool add_phi, double* xp) {
  hipStream_t s = c.s;
  const int dm = c.dm;
  auto& l0 = mg.L[0];
  auto& u0 = mmu.L[0];
  const T dxinv = T(1.0 / c.dx);
  const Dec dcell = make_dec(c.g.n[0], c.g.n[1]);
  launch(s, c.nc, ConvertFn<T, T>{r, u0.rhs, T(1)});
  mmu.solve(s);
  for (int a = 0; a < dm; ++a)
    launch(s, c.nf[a], WGradFn<T>{c.g, u0.phi, u0.b[a], l0.res[a], dxinv, a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  Coefs<T> C = l0.C();
  C.theta = T(0);
  dispatch_vd(c.d.visc_type, dm, [&](auto V, auto D) {
    constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
    apply_all<T, VT, DM>(s, l0.g, C, l0.res, nullptr, l0.rhs, 0, mg.rows(l0.g));
    for (int a = 0; a < dm; ++a) launch(s, c.nf[a], MulFn<T>{l0.rhs[a], u0.b[a]});
    const T* lg[3] = {l0.rhs[0], l0.rhs[1], dm == 3 ? l0.rhs[2] : l0.rhs[1]};
    launch(s, c.nc, DivPlusFn<T, DM>{c.g, {lg[0], lg[1], lg[2]}, nullptr, u0.rhs, dxinv, dcell});
  });
  mmu.solve(s);
  launch(s, c.nc, ViscOutFn<T>{u0.phi, mac.L[0].phi, c.theta, add_phi ? 1 : 0, xp});
  check_launch();
}

// apply_precon for precon_type -1 and +-2 .. +-5 (apply_precon.f90:101-426, visc_schur_approx
// = 0): the same building blocks as type 1 in plain launches
template <typename T>
static void precon_general(Ctx& c, StagMG<T>& mg, MacMG<T>& mac, T* macrhs, const double* b, double* x,
                           MacMG<T>* mmu) {
  hipStream_t s = c.s;
  const int dm = c.dm, pt = c.d.precon_type, apt = std::abs(pt), avt = std::abs(c.d.visc_type);
  auto& l0 = mg.L[0];
  auto& m0 = mac.L[0];
  const double dxinv = 1.0 / c.dx;
  const Dec dcell = make_dec(c.g.n[0], c.g.n[1]);
  // stag MG from phi = 0 (or from the current phi) on rhs b_u - G x_p (x_p may be null)
  auto stag_solve = [&](bool zero_phi, const double* xp_grad) {
    for (int a = 0; a < dm; ++a) {
      launch(s, c.nf[a], RhsGradFn<T>{c.g, b + c.off[a], xp_grad, l0.rhs[a], dxinv, a,
                                      make_dec(fext(c.g, a, 0), fext(c.g, a, 1)), l0.sp});
    }
    if (zero_phi) fill<T>(s, l0.phi[0], l0.phi_span, T(0));
    mg.solve(s, c.R);
  };
  auto u_out = [&]() {
    for (int a = 0; a < dm; ++a) mg.template from_level0<double>(s, l0.phi[a], x + c.off[a], a);
  };
  // mac_rhs = D x_u + b_p into the MAC rhs and its copy; the MAC phi zeroed
  auto div_plus_bp = [&]() {
    const T* ph[3] = {l0.phi[0], l0.phi[1], dm == 3 ? l0.phi[2] : l0.phi[1]};
    if (dm == 3) {
      DivPlusFn<T, 3> f{c.g, {ph[0], ph[1], ph[2]}, b + c.off[3], m0.rhs, T(dxinv), dcell};
      f.out2 = macrhs;
      f.sp = l0.sp;
      f.z = m0.phi;
      launch(s, c.nc, f);
    } else {
      DivPlusFn<T, 2> f{c.g, {ph[0], ph[1], ph[2]}, b + c.off[3], m0.rhs, T(dxinv), dcell};
      f.out2 = macrhs;
      f.sp = l0.sp;
      f.z = m0.phi;
      launch(s, c.nc, f);
    }
  };
  const bool mac_solve = apt == 1 || std::fabs(c.theta) > 0.0;
  auto schur_simple = [&](const T* r) {
    if (!mac_solve) fill<T>(s, m0.phi, c.nc, T(0));
    launch(s, c.nc, PreconCellFn<T>{r, m0.phi, c.beta64, c.gamma64, x + c.off[3], c.theta, avt});
  };
  double* xp = x + c.off[3];
  if (apt == 1) {  // -1
    stag_solve(true, nullptr);
    div_plus_bp();
    mac.solve(s, true);
    for (int a = 0; a < dm; ++a)
      launch(s, c.nf[a], PreconFaceFn<T>{c.g, l0.phi[a], m0.phi, c.alpha64[a], x + c.off[a], dxinv, a,
                                          make_dec(fext(c.g, a, 0), fext(c.g, a, 1)), l0.sp});
    if (mmu) {
      visc_schur<T>(c, mg, mac, *mmu, macrhs, std::fabs(c.theta) > 0.0, xp);
    } else if (c.theta == 0.0) {
      schur_simple(macrhs);
    } else {  // -chi L_alpha Phi in place of -chi mac_rhs (cc_applyop, apply_precon.f90:150-152)
      if (dm == 3) launch(s, c.nc, MacApplyFn<T, 3>{c.g, {m0.b[0], m0.b[1], m0.b[2]}, m0.h2i, m0.phi, m0.tmp, dcell});
      else launch(s, c.nc, MacApplyFn<T, 2>{c.g, {m0.b[0], m0.b[1], m0.b[2]}, m0.h2i, m0.phi, m0.tmp, dcell});
      schur_simple(m0.tmp);
    }
  } else if (apt == 2 || apt == 5) {  // lower triangular (:196-276)
    stag_solve(true, nullptr);
    div_plus_bp();
    if (mac_solve) mac.solve(s, true);
    if (mmu) visc_schur<T>(c, mg, mac, *mmu, macrhs, mac_solve, xp);
    else schur_simple(macrhs);
    if (pt == -2 || pt == -5) launch(s, c.nc, ConvertFn<double, double>{xp, xp, -1.0});
    if (apt == 5) stag_solve(false, xp);  // from the first solve's result
    u_out();
  } else if (apt == 3) {  // upper triangular (:278-344)
    const double sg = pt == -3 ? -1.0 : 1.0;
    launch(s, c.nc, ConvertFn<T, double>{b + c.off[3], m0.rhs, sg});
    launch(s, c.nc, ConvertFn<T, double>{b + c.off[3], macrhs, sg});
    if (mac_solve) mac.solve(s);
    if (mmu) visc_schur<T>(c, mg, mac, *mmu, macrhs, mac_solve, xp);
    else schur_simple(macrhs);
    stag_solve(true, xp);
    u_out();
  } else {  // 4: block diagonal (:346-422)
    stag_solve(true, nullptr);
    if (mac_solve) {
      launch(s, c.nc, ConvertFn<T, double>{b + c.off[3], m0.rhs, 1.0});
      mac.solve(s);
    } else {
      fill<T>(s, m0.phi, c.nc, T(0));
    }
    if (mmu) {  // the MAC right-hand side is b_p when the L_alpha solve ran (apply_precon.f90:358-365)
      if (mac_solve) launch(s, c.nc, ConvertFn<T, double>{b + c.off[3], macrhs, 1.0});
      else fill<T>(s, macrhs, c.nc, T(0));
      visc_schur<T>(c, mg, mac, *mmu, macrhs, mac_solve, xp);
    } else {
      launch(s, c.nc, PreconCell4Fn<T>{b + c.off[3], m0.phi, c.beta64, c.gamma64, xp, c.theta, mac_solve ? 1 : 0, avt, pt});
    }
    u_out();
  }
  precon_tail(c, x);
}

static void precon_eager(Ctx& c, const double* b, double* x) {
  const bool visc = c.d.visc_schur_approx != 0;
  const bool general = c.d.precon_type != 1 || visc;
  if (c.d.precision == LMFHD_MIXED) {
    if (general) precon_general<float>(c, c.mg32, c.mac32, c.macrhs32, b, x, visc ? &c.macmu32 : nullptr);
    else precon<float>(c, c.mg32, c.mac32, c.macrhs32, b, x);
  } else {
    if (general) precon_general<double>(c, c.mg64, c.mac64, c.macrhs64, b, x, visc ? &c.macmu64 : nullptr);
    else precon<double>(c, c.mg64, c.mac64, c.macrhs64, b, x);
  }
}

// HIP graphs.  With graphs on (lmfhd_desc.graphs), a fixed launch sequence -- one Arnoldi step
// of GMRES (operator, preconditioner, Gram-Schmidt, Givens update), or a preconditioner
// application between two fixed buffers -- is captured as a graph on its first use and replayed
// afterwards instead of launching its ~100 kernels one by one.  A captured sequence holds kernel
// arguments (theta, the buffers), so new coefficients drop the graphs; the sequences contain no
// host synchronisation and no memset nodes (fills are kernels).  Replays run the same kernels
// with the same arguments as the launch-by-launch path, so the results are the same bits.
template <class Body>
static void run_graphed(Ctx& c, int kind, int idx, const void* pa, const void* pb, Body&& body) {
#ifndef LMFHD_HOST_DEBUG
  if (c.use_graphs && !c.graph_failed) {
    for (auto& e : c.graphs)
      if (e.kind == kind && e.idx == idx && e.a == pa && e.b == pb) {
        HIPCHK(hipGraphLaunch(e.exec, c.s));
        return;
      }
    if (c.graphs.size() < 256) {
      hipGraph_t gr = nullptr;
      if (hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        bool ok = true;
        try {
          body();
        } catch (const Fail&) {
          ok = false;
        }
        const hipError_t ec = hipStreamEndCapture(c.s, &gr);
        hipGraphExec_t ex = nullptr;
        if (ok && ec == hipSuccess && gr && hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) == hipSuccess) {
          (void)hipGraphDestroy(gr);
          c.graphs.push_back(Ctx::GraphEntry{kind, idx, pa, pb, ex});
          HIPCHK(hipGraphLaunch(ex, c.s));
          return;
        }
        if (gr) (void)hipGraphDestroy(gr);
        (void)hipGetLastError();
      }
      c.graph_failed = true;  // fall back to launches for this context
    }
  }
#else
  (void)kind;
  (void)idx;
  (void)pa;
  (void)pb;
#endif
  body();
}

// apply_precon b -> x; captured as a graph when graphs are on and one staggered V-cycle (no
// residual-norm test with its host read-back) is asked for
// The size class of a descriptor (grid, boundary conditions, solver form): the key of its
// tuning record and of its lock-step batch.
struct TuneKey {
  int v[16];
  bool operator==(const TuneKey& o) const { return memcmp(v, o.v, sizeof(v)) == 0; }
};
static TuneKey tune_key(const lmfhd_desc& d) {
  TuneKey k;
  memset(&k, 0, sizeof(k));
  int q = 0;
  k.v[q++] = d.dm;
  for (int j = 0; j < 3; ++j) k.v[q++] = j < d.dm ? d.n_cells[j] : 1;
  for (int j = 0; j < 3; ++j) k.v[q++] = j < d.dm ? d.bc_lo[j] : 0;
  for (int j = 0; j < 3; ++j) k.v[q++] = j < d.dm ? d.bc_hi[j] : 0;
  k.v[q++] = d.visc_type;
  k.v[q++] = d.precision;
  k.v[q++] = d.krylov;
  k.v[q++] = d.precon_type;
  k.v[q++] = d.visc_schur_approx;
  return k;
}
// ---- lock-step batches of preconditioner applications (lmfhd_desc.batch) ----------------------
// Contexts of one size class created with batch = 1 share a Batch.  A context records its
// preconditioner once (pin -> pout: every launch of precon_eager, none executed) and, while in a
// solve, takes a seat.  An application copies its input to pin, marks it with an event and
// arrives; the last of the seated contexts to arrive replays the recordings of everyone who has
// arrived on the batch stream: record k of every arrived context in one launch, one grid slice
// each (the batched forms of launch.h), after waiting on their events.  Each context's stream
// then waits for the batch's completion event and copies pout to its output.  The functors are
// the ones the context would have launched, so the results are its own bits.
struct Batch {
  std::mutex mu;
  std::condition_variable cv;
  int seated = 0;
  std::vector<Ctx*> arrived;
  long round = 0;
  int err = LMFHD_OK;  // the error of the last round, reported to everyone in it
  long err_round = -1;
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  DevMem mem;  // functor tables of the member set `set`
  std::vector<Ctx*> set;
  std::vector<const unsigned char*> tab;  // per record: nent functor copies
  ~Batch() {
#ifndef LMFHD_HOST_DEBUG
    if (s) (void)hipStreamSynchronize(s);
    if (done) (void)hipEventDestroy(done);
    if (s) (void)hipStreamDestroy(s);
#endif
  }
};
static std::mutex g_batch_mu;
static std::vector<std::pair<std::pair<TuneKey, int>, std::weak_ptr<Batch>>> g_batches;

static bool batchable(const lmfhd_desc& d) {
  return d.batch == 1 && d.precon_type == 1 && d.visc_schur_approx == 0 && d.stag_mg_max_vcycles == 1 &&
         d.mg_max_vcycles == 1;
}
static std::shared_ptr<Batch> batch_for(const lmfhd_desc& d) {
  const auto key = std::make_pair(tune_key(d), d.device);
  std::lock_guard<std::mutex> lk(g_batch_mu);
  for (auto& e : g_batches)
    if (e.first.first == key.first && e.first.second == key.second)
      if (auto p = e.second.lock()) return p;
  auto p = std::make_shared<Batch>();
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipStreamCreateWithFlags(&p->s, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
#endif
  for (auto& e : g_batches)
    if (e.second.expired()) {
      e = {key, p};
      return p;
    }
  g_batches.push_back({key, p});
  return p;
}

static void record_precon(Ctx& c) {
  Recorder r;
  rec_slot() = &r;
  try {
    precon_eager(c, c.pin, c.pout);
  } catch (...) {
    rec_slot() = nullptr;
    throw;
  }
  rec_slot() = nullptr;
  c.rec = std::move(r);
  c.rec_ok = true;
}
static bool same_shape(const Rec& a, const Rec& b) {
  return a.kind == b.kind && a.multi == b.multi && a.f.size() == b.f.size() && a.n == b.n && a.ncols == b.ncols &&
         a.nz == b.nz && a.kc == b.kc && a.nw == b.nw;
}
// one round (B.mu held): replay the recordings of the arrived contexts on the batch stream
static void batch_run(Batch& B) {
  std::vector<Ctx*> set = B.arrived;
  std::sort(set.begin(), set.end());
  bool same = true;
  for (Ctx* c : set) {
    same = same && c->rec.recs.size() == set[0]->rec.recs.size();
    for (size_t k = 0; same && k < c->rec.recs.size(); ++k) same = same_shape(c->rec.recs[k], set[0]->rec.recs[k]);
  }
  if (!same) {  // not one size class after all: one context per slice group
    std::vector<Ctx*> all = set;
    for (Ctx* c : all) {
      B.arrived.assign(1, c);
      batch_run(B);
    }
    return;
  }
  const size_t nrec = set[0]->rec.recs.size();
  const int nent = (int)set.size();
  if (set != B.set) {  // functor tables of this member set
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(B.s));  // no launch of the old tables may still run
#endif
    B.mem.release();
    B.tab.assign(nrec, nullptr);
    size_t total = 0;
    for (size_t k = 0; k < nrec; ++k) total += (set[0]->rec.recs[k].f.size() * nent + 255) / 256 * 256;
    std::vector<unsigned char> h(total);
    unsigned char* d = B.mem.alloc<unsigned char>((long)total);
    size_t o = 0;
    for (size_t k = 0; k < nrec; ++k) {
      const size_t fs = set[0]->rec.recs[k].f.size();
      for (int q = 0; q < nent; ++q) memcpy(h.data() + o + q * fs, set[q]->rec.recs[k].f.data(), fs);
      B.tab[k] = d + o;
      o += (fs * nent + 255) / 256 * 256;
    }
    h2d(d, h.data(), total, B.s);  // completed before it returns
    B.set = set;
  }
#ifndef LMFHD_HOST_DEBUG
  for (Ctx* c : set) HIPCHK(hipStreamWaitEvent(B.s, c->ev_in, 0));
#endif
  for (size_t k = 0; k < nrec; ++k) set[0]->rec.recs[k].multi(B.s, B.tab[k], nent, set[0]->rec.recs[k]);
  check_launch();
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipEventRecord(B.done, B.s));
#endif
}
// the last arrival (or a departure that completes the round) runs it; errors reach everyone
static void batch_round(Batch& B) {
  try {
    batch_run(B);
  } catch (const Fail& f) {
    B.err = f.code;
    B.err_round = B.round;
  }
  B.arrived.clear();
  ++B.round;
  B.cv.notify_all();
}
// a seat for the duration of a solve
struct BatchSeat {
  Batch* B;
  explicit BatchSeat(Ctx& c) : B(c.batch.get()) {
    if (!B) return;
    std::lock_guard<std::mutex> lk(B->mu);
    ++B->seated;
  }
  ~BatchSeat() {
    if (!B) return;
    std::lock_guard<std::mutex> lk(B->mu);
    --B->seated;
    if (!B->arrived.empty() && (int)B->arrived.size() >= B->seated) batch_round(*B);
  }
};
static void copy_vec(hipStream_t s, double* dst, const double* src, long n) {
  if (dst == src) return;
#ifdef LMFHD_HOST_DEBUG
  (void)s;
  memcpy(dst, src, n * sizeof(double));
#else
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
#endif
}
static void precon_batched(Ctx& c, const double* b, double* x) {
  Batch& B = *c.batch;
  if (!c.rec_ok) record_precon(c);
  copy_vec(c.s, c.pin, b, c.vlen);
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipEventRecord(c.ev_in, c.s));
#endif
  {
    std::unique_lock<std::mutex> lk(B.mu);
    const long my = B.round;
    B.arrived.push_back(&c);
    if ((int)B.arrived.size() >= B.seated) batch_round(B);
    else B.cv.wait(lk, [&]() { return B.round > my; });
    if (B.err_round == my) {
      set_err("batched preconditioner round failed");
      throw Fail{B.err};
    }
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamWaitEvent(c.s, B.done, 0));  // under the lock: no later round has re-recorded it
#endif
  }
  copy_vec(c.s, x, c.pout, c.vlen);
}

// the preconditioner inside an Arnoldi step (captured with the step when graphs are on)
static void precon_step(Ctx& c, const double* b, double* x) {
  if (c.batch) precon_batched(c, b, x);
  else precon_eager(c, b, x);
}

static void precon_any(Ctx& c, const double* b, double* x) {
  if (c.batch) {
    precon_batched(c, b, x);
    return;
  }
  if (c.d.stag_mg_max_vcycles == 1) run_graphed(c, 0, 0, b, x, [&]() { precon_eager(c, b, x); });
  else precon_eager(c, b, x);
}

// weighted Stokes-vector norm (stag_l2_norm / cc_l2_norm with p_norm_weight), host value
static double vnorm(Ctx& c, const double* v) {
  c.R.mdot(v, v, 0, 1, c.vlen, true, c.s_tmp(), true);
  double h;
  d2h(&h, c.s_tmp(), sizeof(double), c.s);
  return h;
}

// the Givens column of step i and the stop test, on the device (one thread)
static void lsq_column(Ctx& c, int i) {
#ifdef LMFHD_HOST_DEBUG
  lsq_column_hd(c.G, i, c.s_h1(), c.s_h2(), c.s_nrm());
#else
  hipLaunchKernelGGL(k_lsq_column, dim3(1), dim3(64), 0, c.s, c.G, i, (const double*)c.s_h1(), (const double*)c.s_h2(),
                     (const double*)c.s_nrm());
  check_launch();
#endif
}
// a new restart cycle: s(1) = *nrm, stop thresholds (rel, abs)
static void lsq_init(Ctx& c, const double* nrm, double thr_rel, double thr_abs) {
#ifdef LMFHD_HOST_DEBUG
  lsq_init_hd(c.G, nrm, thr_rel, thr_abs);
#else
  hipLaunchKernelGGL(k_lsq_init, dim3(1), dim3(64), 0, c.s, c.G, nrm, thr_rel, thr_abs);
  check_launch();
#endif
}
// y of the cycle's i columns into c.coef
static void lsq_solve(Ctx& c, int i) {
#ifdef LMFHD_HOST_DEBUG
  lsq_solve_hd(c.G, i, c.coef);
#else
  hipLaunchKernelGGL(k_lsq_solve, dim3(1), dim3(64), 0, c.s, c.G, i, c.coef);
  check_launch();
#endif
}

// Arnoldi step: orthogonalise w against V_0..V_{i-1} by classical Gram-Schmidt with one
// reorthogonalisation (CGS2), V_i = w/|w|, then the Givens update of column i on the device.
// No host synchronisation: the step can be captured as a graph.
static void arnoldi(Ctx& c, double* w, int i) {
  hipStream_t s = c.s;
  double* h1 = c.s_h1();
  double* h2 = c.s_h2();
  double* wn2 = c.s_wn2();
  double* skip = c.s_skip();
  // classical Gram-Schmidt, repeated only when the first pass lost more than half of |w|^2
  // (Daniel, Gragg, Kaufman & Stewart 1976 with eta = 1/sqrt 2; |w - V h|^2 = |w|^2 - |h|^2
  // for the orthonormal basis).  The reference orthogonalises by modified Gram-Schmidt
  // (gmres.f90:352-372).
  double* nrm = c.s_nrm();
  if (i + 1 <= RED_NV) {
    // the same passes with the dot products folded into the sweeps that already read the
    // vectors: <w, V>, |w|^2 together; the first update with the second pass's dot products;
    // the second update (when needed) with the norm. Same sums in the same order.
    c.R.mdot_self(w, c.V, c.vlen, i, c.vlen, h1);  // h1[i] = |w|^2
    dgks(s, h1, i, h1 + i, skip);
    c.R.maxpy_dot(w, c.V, c.vlen, i, h1, c.vlen, h2, skip);
    c.R.maxpy_norm(w, c.V, c.vlen, i, h2, c.vlen, nrm, skip);
  } else {
    c.R.mdot(w, w, 0, 1, c.vlen, true, wn2);
    for (int pass = 0; pass < 2; ++pass) {
      double* h = pass == 0 ? h1 : h2;
      const double* sk = pass == 0 ? nullptr : skip;
      for (int q0 = 0; q0 < i; q0 += RED_NV) {
        const int nv = std::min(RED_NV, i - q0);
        c.R.mdot(w, c.V + (long)q0 * c.vlen, c.vlen, nv, c.vlen, true, h + q0, false, 0, sk);
      }
      maxpy(s, w, c.V, c.vlen, i, h, -1.0, c.vlen, sk);
      if (pass == 0) dgks(s, h1, i, wn2, skip);
    }
    c.R.mdot(w, w, 0, 1, c.vlen, true, nrm, true);
  }
  scale_div(s, c.V + (long)i * c.vlen, w, nrm, c.vlen);
  lsq_column(c, i);
}

static bool finite(double v) { return std::isfinite(v); }

// status of the last Givens column on the host: {|s(i+1)|, stop flag, H(i+1, i)}
static void lsq_status(Ctx& c, double* st) { d2h(st, c.G.stat, 3 * sizeof(double), c.s); }

// gmres (gmres.f90:32-621), left preconditioning, scale_factor = 1
static int gmres_left(Ctx& c, const double* b, double* x, lmfhd_stats* st) {
  const lmfhd_desc& d = c.d;
  hipStream_t s = c.s;
  const int m = c.m;
  lmfhd_stats S;
  memset(&S, 0, sizeof(S));
  int total_iter = 0, iter = 0, nh = 0;
  double norm_init_resid = 0.0, norm_resid = 0.0, norm_resid_stokes = 0.0, norm_b = 0.0;
  auto finish = [&](int rc) {
    S.iters = total_iter;
    S.outer_iters = iter;
    S.norm_init_resid = norm_init_resid;
    S.norm_resid = norm_resid;
    S.norm_resid_stokes = norm_resid_stokes;
    S.true_rel_resid = norm_b > 0.0 ? norm_resid_stokes / norm_b : 0.0;
    if (st) *st = S;
    return rc;
  };
  precon_any(c, b, c.tmp);
  S.n_precon++;
  const double norm_pre_b = vnorm(c, c.tmp);
  norm_b = vnorm(c, b);
  S.norm_b = norm_b;
  S.norm_pre_b = norm_pre_b;
  if (!finite(norm_b) || !finite(norm_pre_b)) { set_err("non-finite norm of the right-hand side"); return finish(LMFHD_E_NAN); }
  if (norm_b <= d.gmres_abs_tol) {
    fill<double>(s, x, c.vlen, 0.0);
    S.converged = 1;
    return finish(LMFHD_OK);
  }
  double stat[3];
  while (true) {
    apply_matrix(c, x, c.tmp);
    S.n_matrix++;
    rsub(s, c.tmp, b, c.vlen);  // tmp = b - A x
    norm_resid_stokes = vnorm(c, c.tmp);
    precon_any(c, c.tmp, c.r);
    S.n_precon++;
    norm_resid = vnorm(c, c.r);  // also left in s_tmp on the device
    if (!finite(norm_resid) || !finite(norm_resid_stokes)) { set_err("non-finite residual norm"); return finish(LMFHD_E_NAN); }
    if (iter == 0) norm_init_resid = norm_resid;
    if (total_iter >= d.gmres_max_iter) break;
    if (total_iter >= d.gmres_min_iter) {
      if (norm_resid <= d.gmres_rel_tol * std::min(norm_pre_b, norm_init_resid)) { S.converged = 1; break; }
      if (norm_resid <= d.gmres_abs_tol) { S.converged = 1; break; }
    }
    if (iter >= d.gmres_max_outer) break;
    iter++;
    // V_0 = r / |r|; s(1) = |r| (the norm of r is on the device from vnorm)
    scale_div(s, c.V, c.r, c.s_tmp(), c.vlen);
    lsq_init(c, c.s_tmp(), d.gmres_rel_tol * std::min(norm_pre_b, norm_init_resid), d.gmres_abs_tol);
    int i_copy = 0;
    for (int i = 1; i <= m; ++i) {
      total_iter++;
      i_copy = i;
      double* Vi = c.V + (long)(i - 1) * c.vlen;
      run_graphed(c, 1, i, Vi, nullptr, [&]() {
        apply_matrix(c, Vi, c.tmp);
        precon_step(c, c.tmp, c.w);
        arnoldi(c, c.w, i);
      });
      S.n_matrix++;
      S.n_precon++;
      lsq_status(c, stat);
      if (!finite(stat[2])) { set_err("non-finite Arnoldi norm"); return finish(LMFHD_E_NAN); }
      if (stat[2] == 0.0) { set_err("gmres: error in orthogonalization"); return finish(LMFHD_E_BREAKDOWN); }
      if (nh < 256) S.resid_hist[nh++] = stat[0];
      if (total_iter >= d.gmres_max_iter) break;
      if (total_iter >= d.gmres_min_iter && stat[1] != 0.0) break;
    }
    lsq_solve(c, i_copy);
    maxpy(s, x, c.V, c.vlen, i_copy, c.coef, 1.0, c.vlen);
  }
  return finish(LMFHD_OK);
}

// right-preconditioned flexible GMRES: x = x0 + Z y with Z_i = M^-1 V_i; the restart test is
// on the true residual |b - A x| <= rel_tol |b| (modelled on the reference's standalone
// exec/stokes_preconditioner/test_gmres/gmres/fgmres.f90).  |M^-1 b| (gmres.f90 norm_pre_rhs,
// the time-step callers' tolerance hand-off) is formed only when want_pre is set: it costs one
// preconditioner application the flexible iteration does not otherwise need.
static int fgmres(Ctx& c, const double* b, double* x, lmfhd_stats* st, bool want_pre) {
  const lmfhd_desc& d = c.d;
  hipStream_t s = c.s;
  const int m = c.m;
  lmfhd_stats S;
  memset(&S, 0, sizeof(S));
  int total_iter = 0, iter = 0, nh = 0;
  double beta = 0.0, beta0 = 0.0, norm_b = 0.0;
  auto finish = [&](int rc) {
    S.iters = total_iter;
    S.outer_iters = iter;
    S.norm_init_resid = beta0;
    S.norm_resid = beta;
    S.norm_resid_stokes = beta;
    S.true_rel_resid = norm_b > 0.0 ? beta / norm_b : 0.0;
    if (st) *st = S;
    return rc;
  };
  if (want_pre) {
    precon_any(c, b, c.tmp);
    S.n_precon++;
    S.norm_pre_b = vnorm(c, c.tmp);
    if (!finite(S.norm_pre_b)) { set_err("non-finite norm of the preconditioned right-hand side"); return finish(LMFHD_E_NAN); }
  }
  norm_b = vnorm(c, b);
  S.norm_b = norm_b;
  if (!finite(norm_b)) { set_err("non-finite norm of the right-hand side"); return finish(LMFHD_E_NAN); }
  if (norm_b <= d.gmres_abs_tol) {
    fill<double>(s, x, c.vlen, 0.0);
    S.converged = 1;
    return finish(LMFHD_OK);
  }
  double stat[3];
  while (true) {
    apply_matrix(c, x, c.tmp);
    S.n_matrix++;
    rsub(s, c.tmp, b, c.vlen);
    beta = vnorm(c, c.tmp);  // also left in s_tmp on the device
    if (!finite(beta)) { set_err("non-finite residual norm"); return finish(LMFHD_E_NAN); }
    if (iter == 0) beta0 = beta;
    if (total_iter >= d.gmres_max_iter) break;
    if (total_iter >= d.gmres_min_iter) {
      if (beta <= d.gmres_rel_tol * norm_b || beta <= d.gmres_abs_tol) { S.converged = 1; break; }
    }
    if (iter >= d.gmres_max_outer) break;
    iter++;
    scale_div(s, c.V, c.tmp, c.s_tmp(), c.vlen);
    lsq_init(c, c.s_tmp(), d.gmres_rel_tol * norm_b, d.gmres_abs_tol);
    int i_copy = 0;
    for (int i = 1; i <= m; ++i) {
      total_iter++;
      i_copy = i;
      double* Vi = c.V + (long)(i - 1) * c.vlen;
      double* Zi = c.Z + (long)(i - 1) * c.vlen;
      run_graphed(c, 2, i, Vi, Zi, [&]() {
        precon_step(c, Vi, Zi);
        apply_matrix(c, Zi, c.w);
        arnoldi(c, c.w, i);
      });
      S.n_precon++;
      S.n_matrix++;
      lsq_status(c, stat);
      if (!finite(stat[2])) { set_err("non-finite Arnoldi norm"); return finish(LMFHD_E_NAN); }
      if (nh < 256) S.resid_hist[nh++] = stat[0];
      if (stat[2] == 0.0) break;  // exact solution in the Krylov space: finish this cycle
      if (total_iter >= d.gmres_max_iter) break;
      if (total_iter >= d.gmres_min_iter && stat[1] != 0.0) break;
    }
    lsq_solve(c, i_copy);
    maxpy(s, x, c.Z, c.vlen, i_copy, c.coef, 1.0, c.vlen);
  }
  return finish(LMFHD_OK);
}

static int gmres_any(Ctx& c, const double* b, double* x, lmfhd_stats* st, bool want_pre = false) {
  if (!c.have_coef) {
    set_err("coefficients not set");
    if (st) memset(st, 0, sizeof(*st));
    return LMFHD_E_STATE;
  }
  BatchSeat seat(c);
  return c.d.krylov == LMFHD_FGMRES ? fgmres(c, b, x, st, want_pre) : gmres_left(c, b, x, st);
}

// ------------------------------------------------------------------------------------------
// momentum time steps for one species (density and viscosities constant in time, homogeneous
// velocity boundaries): advance_timestep_inertial / advance_timestep_overdamped, momentum part
static void step_bufs(Ctx& c) {
  if (c.st_b) return;
  c.st_vn = c.mem.alloc<double>(c.vlen);
  c.st_adv = c.mem.alloc<double>(c.vlen);
  c.st_dif = c.mem.alloc<double>(c.vlen);
  c.st_b = c.mem.alloc<double>(c.vlen);
  c.st_dx = c.mem.alloc<double>(c.vlen);
  fill<double>(c.s, c.st_adv, c.vlen, 0.0);
  fill<double>(c.s, c.st_dif, c.vlen, 0.0);
}

// b = the right-hand side of one solve of the step (StepRhsFn kind) for the state x, and its
// cell part D v (compute_div with increment onto zero)
static void step_rhs(Ctx& c, int kind, double dt, double* x, const double* f1, const double* f2, double* b) {
  hipStream_t s = c.s;
  const int dm = c.dm;
  const int q2 = dm == 3 ? 2 : 1;
  Coefs<double> C;
  for (int a = 0; a < 3; ++a) { C.alpha[a] = c.alpha64[a]; C.bed[a] = c.bed64[a]; }
  C.beta = c.beta64;
  C.gamma = c.gamma64;
  C.dxsqinv = 1.0 / (c.dx * c.dx);
  C.theta = 0.0;
  const double dxinv = 1.0 / c.dx;
  const double* vn = kind == 1 ? c.st_vn : x;
  dispatch_vd(c.d.visc_type, dm, [&](auto V, auto D) {
    constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
    for (int a = 0; a < dm; ++a) {
      StepRhsFn<VT, DM> F;
      F.g = c.g;
      F.C = C;
      for (int d = 0; d < 3; ++d) {
        const int q = d < dm ? d : q2;
        F.v[d] = x + c.off[q];
        F.vn[d] = vn + c.off[q];
        F.rho[d] = c.alpha64[q];
        F.f1[d] = f1 ? f1 + c.off[q] : nullptr;
        F.f2[d] = f2 ? f2 + c.off[q] : nullptr;
        F.adv[d] = c.st_adv + c.off[q];
        F.dif[d] = c.st_dif + c.off[q];
        F.out[d] = b + c.off[q];
      }
      F.p = x + c.off[3];
      F.dxinv = dxinv;
      F.dtinv = kind == 2 ? 0.0 : 1.0 / dt;
      F.kind = kind;
      F.a = a;
      F.dc = make_dec(fext(c.g, a, 0), fext(c.g, a, 1));
      launch(s, c.nf[a], F);
    }
    launch(s, c.nc, DivPlusFn<double, DM>{c.g, {x + c.off[0], x + c.off[1], x + c.off[q2]}, nullptr, b + c.off[3], dxinv,
                                          make_dec(c.g.n[0], c.g.n[1])});
  });
  check_launch();
}

// the two solves of a step with the reference's absolute-tolerance hand-off: the corrector
// stops at max(gmres_abs_tol, |preconditioned rhs of the predictor| rtol)
// (advance_timestep_inertial.f90:493-496, 887)
static int advance_step(Ctx& c, bool inertial, double dt, double* x, const double* f1, const double* f2,
                        lmfhd_stats* st2) {
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  if (inertial && !(dt > 0.0 && c.theta == 1.0 / dt)) {
    set_err("inertial step: the context's theta_alpha must be 1/dt (%g), it is %g", 1.0 / dt, c.theta);
    return LMFHD_E_ARG;
  }
  if (!inertial && c.theta != 0.0) { set_err("overdamped step: theta_alpha must be 0"); return LMFHD_E_ARG; }
  hipStream_t s = c.s;
  step_bufs(c);
  // multifab_physbc_domainvel on v^n (homogeneous walls)
  if (has_wall(c.g))
    for (int a = 0; a < c.dm; ++a)
      launch(s, c.nf[a], ZeroWallFacesFn<double>{c.g, c.u(x, a), a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  if (inertial) d2d(c.st_vn, x, (size_t)c.off[3] * sizeof(double), s);
  lmfhd_stats S[2];
  memset(S, 0, sizeof(S));
  const double abs_in = c.d.gmres_abs_tol;
  int rc = LMFHD_OK;
  for (int stage = 0; stage < 2 && rc == LMFHD_OK; ++stage) {
    const int kind = inertial ? stage : 2;
    const double* fa = inertial ? f1 : (stage == 0 ? f1 : f2);
    const double* fb = inertial && stage == 1 ? f2 : nullptr;
    step_rhs(c, kind, dt, x, fa, fb, c.st_b);
    fill<double>(s, c.st_dx, c.vlen, 0.0);
    try {
      rc = gmres_any(c, c.st_b, c.st_dx, &S[stage], true);
    } catch (...) {
      c.d.gmres_abs_tol = abs_in;
      throw;
    }
    if (rc != LMFHD_OK) break;
    if (stage == 0) c.d.gmres_abs_tol = std::max(abs_in, S[0].norm_pre_b * c.d.gmres_rel_tol);
    launch(s, c.vlen, AddFn{x, c.st_dx});
    check_launch();
  }
  c.d.gmres_abs_tol = abs_in;
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipStreamSynchronize(s));
#endif
  if (st2) { st2[0] = S[0]; st2[1] = S[1]; }
  return rc;
}


// ------------------------------------------------------------------------------------------
// Tuning records per size class: the grid, boundary conditions and solver form of a
// descriptor.  Filled by the autotuner (lmfhd_autotune) or by hand (lmfhd_tune_set); a context
// takes its size class's record when it is created, else the defaults above.
static std::mutex g_tune_mu;
static std::vector<std::pair<TuneKey, Tune>> g_tunes;
static bool tune_lookup(const lmfhd_desc& d, Tune& t) {
  const TuneKey k = tune_key(d);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& e : g_tunes)
    if (e.first == k) { t = e.second; return true; }
  return false;
}
static void tune_store(const lmfhd_desc& d, const Tune& t) {
  const TuneKey k = tune_key(d);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& e : g_tunes)
    if (e.first == k) { e.second = t; return; }
  g_tunes.push_back({k, t});
}
static bool tune_valid(const Tune& t) {
  return t.coarse_n >= 0 && t.coarse_n <= COARSE_N && t.mac_coarse_n >= 0 && t.mac_coarse_n <= COARSE_N &&
         (t.coarse_threads == 256 || t.coarse_threads == 512 || t.coarse_threads == 1024) && t.red_blocks >= 64 &&
         t.red_blocks <= RED_BLOCKS && (t.graphs == 0 || t.graphs == 1) && t.rows_n >= 0;
}
// (re)apply launch parameters to a context: coarse-kernel argument blocks, reduction width,
// graph replay (captured graphs are dropped: they hold the old launch shapes)
static void apply_tune(Ctx& c, const Tune& t) {
  c.clear_graphs();
  c.rec_ok = false;
  c.tune = t;
  c.R.max_blocks = t.red_blocks;
  c.use_graphs = t.graphs == 1 && !c.batch;  // batched solves leave the host between steps
  c.graph_failed = false;
  auto stag = [&](auto& mg) {
    mg.build_coarse(coarse_level(mg.L[0].g, mg.nl, t.coarse_n), t.coarse_threads, c.s, c.mem);
    mg.rows_n = t.rows_n;
  };
  auto mac = [&](auto& mm) { mm.build_coarse(coarse_level(mm.L[0].g, mm.nl, t.mac_coarse_n), t.coarse_threads, c.s, c.mem); };
  if (c.d.precision == LMFHD_MIXED) {
    stag(c.mg32);
    mac(c.mac32);
    if (c.d.visc_schur_approx) mac(c.macmu32);
  } else {
    stag(c.mg64);
    mac(c.mac64);
    if (c.d.visc_schur_approx) mac(c.macmu64);
  }
}

// ------------------------------------------------------------------------------------------
// context
static void ctx_build(Ctx& c, const lmfhd_desc& d) {
  c.d = d;
  c.d.box_lo = c.d.box_hi = nullptr;  // the caller's arrays: copied below, not kept
  c.labels = d.nboxes == 0;
  c.nb = c.labels ? 1 : d.nboxes;
  if (!c.labels) {
    c.blo.assign(d.box_lo, d.box_lo + 3 * d.nboxes);
    c.bhi.assign(d.box_hi, d.box_hi + 3 * d.nboxes);
    if (d.dm == 2)
      for (int b = 0; b < d.nboxes; ++b) c.blo[3 * b + 2] = c.bhi[3 * b + 2] = 0;
  }
  c.dm = d.dm;
  c.dx = d.dx;
  c.g = make_geo(d.dm, d.n_cells, d.bc_lo, d.bc_hi);
  {  // kernels address one field through 32-bit byte offsets
    long big = csize(c.g);
    for (int a = 0; a < c.dm; ++a) big = std::max(big, fsize(c.g, a));
    for (int e = 0; e < (c.dm == 3 ? 3 : 1); ++e) big = std::max(big, esize(c.g, e));
    if (big * (long)sizeof(double) >= (1L << 32)) {
      set_err("grid too large: one field of %ld doubles exceeds 4 GiB", big);
      throw Fail{LMFHD_E_ARG};
    }
  }
  c.m = std::max(1, d.gmres_max_inner);
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipSetDevice(d.device));
  HIPCHK(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
#endif
  long o = 0;
  for (int a = 0; a < 3; ++a) {
    c.off[a] = o;
    c.nf[a] = a < c.dm ? fsize(c.g, a) : 0;
    o += c.nf[a];
  }
  c.off[3] = o;
  c.nc = csize(c.g);
  c.vlen = o + c.nc;
  c.VL.off[0] = c.off[0];
  c.VL.off[1] = c.off[1];
  c.VL.off[2] = c.off[2];
  c.VL.off[3] = c.off[3];
  c.VL.off[4] = c.vlen;
  for (int dd = 0; dd < 3; ++dd) c.VL.n[dd] = c.g.n[dd];
  c.VL.dm = c.dm;
  c.VL.pw2 = d.p_norm_weight * d.p_norm_weight;
  c.VL.init_div();
  for (int a = 0; a < c.dm; ++a) c.alpha64[a] = c.mem.alloc<double>(c.nf[a]);
  c.beta64 = c.mem.alloc<double>(c.nc);
  c.gamma64 = c.mem.alloc<double>(c.nc);
  for (int e = 0; e < (c.dm == 3 ? 3 : 1); ++e) {
    c.ne[e] = esize(c.g, e);
    c.bed64[e] = c.mem.alloc<double>(c.ne[e]);
  }
  if (d.precision == LMFHD_MIXED) {
    c.mg32.build(c.g, c.dx, d.visc_type, d, c.mem);
    c.mac32.build(c.g, c.dx, d, c.mem);
    c.macrhs32 = c.mem.alloc<float>(c.nc);
    if (d.visc_schur_approx) c.macmu32.build(c.g, c.dx, d, c.mem);
  } else {
    c.mg64.build(c.g, c.dx, d.visc_type, d, c.mem);
    c.mac64.build(c.g, c.dx, d, c.mem);
    c.macrhs64 = c.mem.alloc<double>(c.nc);
    if (d.visc_schur_approx) c.macmu64.build(c.g, c.dx, d, c.mem);
  }
  if (d.visc_schur_approx) c.diaglap = c.mem.alloc<double>(c.vlen);
  c.V = c.mem.alloc<double>((long)(c.m + 1) * c.vlen);
  if (d.krylov == LMFHD_FGMRES) c.Z = c.mem.alloc<double>((long)c.m * c.vlen);
  c.r = c.mem.alloc<double>(c.vlen);
  c.w = c.mem.alloc<double>(c.vlen);
  c.tmp = c.mem.alloc<double>(c.vlen);
  c.io_a = c.mem.alloc<double>(c.vlen);
  c.io_b = c.mem.alloc<double>(c.vlen);
  c.scal = c.mem.alloc<double>(c.nscal());
  c.coef = c.mem.alloc<double>(c.m + 2);
  {
    const long nH = (long)(c.m + 2) * (c.m + 1);
    double* q = c.mem.alloc<double>(nH + 4 * (long)(c.m + 2) + 8);
    fill<double>(c.s, q, nH + 4 * (long)(c.m + 2) + 8, 0.0);
    c.G.H = q;
    c.G.cs = q + nH;
    c.G.sn = c.G.cs + (c.m + 2);
    c.G.sv = c.G.sn + (c.m + 2);
    c.G.y = c.G.sv + (c.m + 2);
    c.G.stat = c.G.y + (c.m + 2);
    c.G.thr = c.G.stat + 4;
    c.G.m = c.m;
  }
  if (batchable(d)) {
    c.pin = c.mem.alloc<double>(c.vlen);
    c.pout = c.mem.alloc<double>(c.vlen);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipEventCreateWithFlags(&c.ev_in, hipEventDisableTiming));
#endif
    c.batch = batch_for(d);
  }
  c.R.part = c.mem.alloc<double>((long)RED_BLOCKS * RED_NV);
  c.R.out = c.mem.alloc<double>(4 * RED_NV);
  c.R.s = c.s;
  c.R.L = c.VL;
  Tune t;
  tune_lookup(d, t);
  if (d.graphs == 0 || d.graphs == 1) t.graphs = d.graphs;
  apply_tune(c, t);
}

// ------------------------------------------------------------------------------------------
// fab views: check extents, gather to / scatter from a packed host buffer
static bool fab_ok(const lmfhd_fab& f, int dm, const int* ext) {
  if (!f.data || f.ng < 0) return false;
  for (int d = 0; d < 3; ++d) {
    const int e = f.hi[d] - f.lo[d] + 1;
    if (d < dm) { if (e != ext[d]) return false; }
    else if (e != 1) return false;
  }
  return true;
}
static void fab_get(const lmfhd_fab& f, int dm, double* dst) {
  const int g[3] = {f.ng, f.ng, dm == 3 ? f.ng : 0};
  const long sx = f.hi[0] - f.lo[0] + 1 + 2 * g[0], sy = f.hi[1] - f.lo[1] + 1 + 2 * g[1];
  long p = 0;
  for (int k = 0; k <= f.hi[2] - f.lo[2]; ++k)
    for (int j = 0; j <= f.hi[1] - f.lo[1]; ++j) {
      const double* row = f.data + (g[0] + sx * ((j + g[1]) + sy * (long)(k + g[2])));
      const int nx = f.hi[0] - f.lo[0] + 1;
      memcpy(dst + p, row, nx * sizeof(double));
      p += nx;
    }
}
static void fab_put(const lmfhd_fab& f, int dm, const double* src) {
  const int g[3] = {f.ng, f.ng, dm == 3 ? f.ng : 0};
  const long sx = f.hi[0] - f.lo[0] + 1 + 2 * g[0], sy = f.hi[1] - f.lo[1] + 1 + 2 * g[1];
  long p = 0;
  for (int k = 0; k <= f.hi[2] - f.lo[2]; ++k)
    for (int j = 0; j <= f.hi[1] - f.lo[1]; ++j) {
      double* row = f.data + (g[0] + sx * ((j + g[1]) + sy * (long)(k + g[2])));
      const int nx = f.hi[0] - f.lo[0] + 1;
      memcpy(row, src + p, nx * sizeof(double));
      p += nx;
    }
}
static void face_ext(const Ctx& c, int a, int* ext) {
  for (int d = 0; d < 3; ++d) ext[d] = fext(c.g, a, d);
}
static void edge_ext(const Ctx& c, int e, int* ext) {
  for (int d = 0; d < 3; ++d) ext[d] = eext(c.g, e, d);
}
// ---- the caller's multifabs ------------------------------------------------------------------
// One field arrives as c.nb fab views, box b's at f[b] (a multi-component argument holds c.nb
// per component, component-major).  Box b's valid region is its cells extended by one point in
// each nodal direction of the field (nod[d] = 1).  The packed field is the valid region of the
// domain, x fastest.  A point that two boxes hold (nodal, on their common boundary) is read from
// the first box listed and written to every box.
struct Nod {
  int v[3];
};
static Nod nod_cell() { return Nod{{0, 0, 0}}; }
static Nod nod_face(int a) { return Nod{{a == 0, a == 1, a == 2}}; }
static Nod nod_edge(const Ctx& c, int e) { return Nod{{enodal(c.dm, e, 0), enodal(c.dm, e, 1), enodal(c.dm, e, 2)}}; }

static bool mf_ok(const Ctx& c, const lmfhd_fab* f, const Nod& nod) {
  if (!f) return false;
  if (c.labels) {
    int ext[3];
    for (int d = 0; d < 3; ++d) ext[d] = c.g.n[d] + nod.v[d];
    return fab_ok(f[0], c.dm, ext);
  }
  for (int b = 0; b < c.nb; ++b) {
    const lmfhd_fab& q = f[b];
    if (!q.data || q.ng < 0) return false;
    for (int d = 0; d < 3; ++d) {
      if (d < c.dm) {
        if (q.lo[d] != c.blo[3 * b + d] || q.hi[d] != c.bhi[3 * b + d] + nod.v[d]) return false;
      } else if (q.hi[d] != q.lo[d]) {
        return false;
      }
    }
  }
  return true;
}
// (valid-region) row of box fab q at global (j, k): pointer to its point x = q.lo[0]
static long fab_row(const lmfhd_fab& q, int dm, int j, int k) {
  const int g[3] = {q.ng, q.ng, dm == 3 ? q.ng : 0};
  const long sx = q.hi[0] - q.lo[0] + 1 + 2 * g[0], sy = q.hi[1] - q.lo[1] + 1 + 2 * g[1];
  return g[0] + sx * ((j - q.lo[1] + g[1]) + sy * (long)(k - q.lo[2] + g[2]));
}
static void mf_get(const Ctx& c, const lmfhd_fab* f, const Nod& nod, double* dst) {
  if (c.labels) {
    fab_get(f[0], c.dm, dst);
    return;
  }
  const long ex = c.g.n[0] + nod.v[0], ey = c.g.n[1] + nod.v[1];
  for (int b = c.nb - 1; b >= 0; --b) {  // the first box listed is copied last: its values stand
    const lmfhd_fab& q = f[b];
    const int nx = q.hi[0] - q.lo[0] + 1;
    for (int k = q.lo[2]; k <= q.hi[2]; ++k)
      for (int j = q.lo[1]; j <= q.hi[1]; ++j)
        memcpy(dst + q.lo[0] + ex * (j + ey * (long)(c.dm == 3 ? k : 0)), q.data + fab_row(q, c.dm, j, k),
               nx * sizeof(double));
  }
}
static void mf_put(const Ctx& c, const double* src, lmfhd_fab* f, const Nod& nod) {
  if (c.labels) {
    fab_put(f[0], c.dm, src);
    return;
  }
  const long ex = c.g.n[0] + nod.v[0], ey = c.g.n[1] + nod.v[1];
  for (int b = 0; b < c.nb; ++b) {
    const lmfhd_fab& q = f[b];
    const int nx = q.hi[0] - q.lo[0] + 1;
    for (int k = q.lo[2]; k <= q.hi[2]; ++k)
      for (int j = q.lo[1]; j <= q.hi[1]; ++j)
        memcpy(q.data + fab_row(q, c.dm, j, k), src + q.lo[0] + ex * (j + ey * (long)(c.dm == 3 ? k : 0)),
               nx * sizeof(double));
  }
}
// a cell field with one ghost layer around the domain, packed (n + 2)^dm (2-D: one z plane):
// each box's valid cells, and the ghost cells of the domain from the boxes whose fabs hold them
// (ng >= 1; the first box listed stands)
static bool mf_get_ghost1(const Ctx& c, const lmfhd_fab* f, std::vector<double>& h) {
  const int dm = c.dm;
  const long gx = c.g.n[0] + 2, gy = c.g.n[1] + 2, gz = dm == 3 ? c.g.n[2] + 2 : 1;
  h.assign((size_t)(gx * gy * gz), 0.0);
  if (!mf_ok(c, f, nod_cell())) return false;
  for (int b = c.nb - 1; b >= 0; --b) {
    const lmfhd_fab& q = f[b];
    if (q.ng < 1) return false;
    const int G = q.ng;
    const long sx = q.hi[0] - q.lo[0] + 1 + 2 * G, sy = q.hi[1] - q.lo[1] + 1 + 2 * G;
    // box-relative origin: with labels the fab covers the domain whatever its lo
    const int o[3] = {c.labels ? 0 : q.lo[0], c.labels ? 0 : q.lo[1], c.labels || dm == 2 ? 0 : q.lo[2]};
    const int e[3] = {q.hi[0] - q.lo[0] + 1, q.hi[1] - q.lo[1] + 1, dm == 3 ? q.hi[2] - q.lo[2] + 1 : 1};
    const int kg0 = dm == 3 ? -1 : 0, kg1 = dm == 3 ? e[2] : 0;
    for (int kk = kg0; kk <= kg1; ++kk)
      for (int jj = -1; jj <= e[1]; ++jj)
        for (int ii = -1; ii <= e[0]; ++ii) {
          const int I[3] = {o[0] + ii, o[1] + jj, o[2] + kk};
          bool outside = false, inbox = true;
          for (int d = 0; d < dm; ++d) {
            outside = outside || I[d] < 0 || I[d] >= c.g.n[d];
            const int r = d == 0 ? ii : (d == 1 ? jj : kk);
            inbox = inbox && r >= 0 && r < e[d];
          }
          if (!outside && !inbox) continue;  // interior ghost: another box's valid cell
          const long src = (ii + G) + sx * ((jj + G) + sy * (long)(dm == 3 ? kk + G : 0));
          h[(I[0] + 1) + gx * ((I[1] + 1) + gy * (long)(dm == 3 ? I[2] + 1 : 0))] = q.data[src];
        }
  }
  return true;
}

// upload a Stokes vector given as dm face fields (+ optional cell field) into device vector v
static bool upload_vec(Ctx& c, const lmfhd_fab* u, const lmfhd_fab* p, double* v, std::vector<double>& h) {
  h.assign(c.vlen, 0.0);
  for (int a = 0; a < c.dm; ++a) {
    if (!mf_ok(c, u + a * c.nb, nod_face(a))) return false;
    mf_get(c, u + a * c.nb, nod_face(a), h.data() + c.off[a]);
  }
  if (p) {
    if (!mf_ok(c, p, nod_cell())) return false;
    mf_get(c, p, nod_cell(), h.data() + c.off[3]);
  }
  h2d(v, h.data(), c.vlen * sizeof(double), c.s);
  return true;
}
static bool download_vec(Ctx& c, const double* v, lmfhd_fab* u, lmfhd_fab* p, std::vector<double>& h) {
  for (int a = 0; a < c.dm; ++a)
    if (!mf_ok(c, u + a * c.nb, nod_face(a))) return false;
  if (p && !mf_ok(c, p, nod_cell())) return false;
  h.assign(c.vlen, 0.0);
  d2h(h.data(), v, c.vlen * sizeof(double), c.s);
  for (int a = 0; a < c.dm; ++a) mf_put(c, h.data() + c.off[a], u + a * c.nb, nod_face(a));
  if (p) mf_put(c, h.data() + c.off[3], p, nod_cell());
  return true;
}

static void set_coeffs_dev(Ctx& c) {
  hipStream_t s = c.s;
  c.clear_graphs();
  if (c.d.precision == LMFHD_MIXED) {
    c.mg32.set_coeffs(s, c.alpha64, c.beta64, c.gamma64, c.bed64, c.theta);
    c.mac32.set_coeffs(s, c.alpha64);
  } else {
    c.mg64.set_coeffs(s, c.alpha64, c.beta64, c.gamma64, c.bed64, c.theta);
    c.mac64.set_coeffs(s, c.alpha64);
  }
  if (c.d.visc_schur_approx) {
    double* dg[3];
    for (int a = 0; a < 3; ++a) dg[a] = c.diaglap + c.off[a < c.dm ? a : c.dm - 1];
    dispatch_vd(c.d.visc_type, c.dm, [&](auto V, auto D) {
      constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
      if constexpr (VT == 3 || VT == -3) {
        return;  // rejected by the descriptor check (inverse_diag_lap has no |visc_type| 3)
      } else {
        for (int a = 0; a < c.dm; ++a)
          launch(s, c.nf[a], DiagSumFn<VT, DM>{c.g, c.beta64, {c.bed64[0], c.bed64[1], c.bed64[2]}, dg[a], a,
                                               make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
      }
    });
    check_launch();
    if (c.d.precision == LMFHD_MIXED) c.macmu32.set_coeffs(s, dg);
    else c.macmu64.set_coeffs(s, dg);
  }
  c.have_coef = true;
  c.rec_ok = false;  // a recorded preconditioner holds theta
}

static bool valid_desc(const lmfhd_desc* d) {
  if (!d || (d->dm != 2 && d->dm != 3)) return false;
  if (d->visc_type == 0 || d->visc_type < -3 || d->visc_type > 3) return false;
  if (d->dx <= 0.0) return false;
  for (int k = 0; k < d->dm; ++k) {
    if (d->n_cells[k] < 2) return false;
    const int lo = d->bc_lo[k], hi = d->bc_hi[k];
    if ((lo == -1) != (hi == -1)) return false;
    if (lo != -1 && !wall_code(lo)) return false;
    if (hi != -1 && !wall_code(hi)) return false;
  }
  if (d->precon_type == 0 || d->precon_type < -5 || d->precon_type > 5) return false;
  if (d->precision != LMFHD_FP64 && d->precision != LMFHD_MIXED) return false;
  if (d->krylov != LMFHD_GMRES_LEFT && d->krylov != LMFHD_FGMRES) return false;
  if (d->gmres_max_inner < 1 || d->stag_mg_minwidth < 1 || d->mg_minwidth < 1) return false;
  if (d->gmres_spatial_order != 2 && d->gmres_spatial_order != 4) return false;
  if (d->gmres_spatial_order == 4) {  // stag_applyop_4o: visc_type 1 only (stag_applyop.f90:1340)
    if (d->visc_type != 1) return false;
    for (int k = 0; k < d->dm; ++k)
      if (d->bc_lo[k] != -1) return false;  // its two ghost layers: periodic images here
  }
  if (d->visc_schur_approx != 0 && d->visc_schur_approx != 1) return false;
  if (d->graphs < -1 || d->graphs > 1) return false;
  if (d->nboxes < 0) return false;
  if (d->nboxes > 0) {  // the boxes tile the domain: each inside it, volumes summing to it, no overlap
    if (!d->box_lo || !d->box_hi) return false;
    const int nb = d->nboxes;
    long vol = 0, dom = 1;
    for (int k = 0; k < d->dm; ++k) dom *= d->n_cells[k];
    for (int b = 0; b < nb; ++b) {
      long v = 1;
      for (int k = 0; k < d->dm; ++k) {
        const int lo = d->box_lo[3 * b + k], hi = d->box_hi[3 * b + k];
        if (lo < 0 || hi < lo || hi >= d->n_cells[k]) return false;
        v *= hi - lo + 1;
      }
      vol += v;
    }
    if (vol != dom) return false;
    for (int b = 0; b < nb; ++b)
      for (int q = b + 1; q < nb; ++q) {
        bool meet = true;
        for (int k = 0; k < d->dm; ++k)
          meet = meet && d->box_lo[3 * b + k] <= d->box_hi[3 * q + k] && d->box_lo[3 * q + k] <= d->box_hi[3 * b + k];
        if (meet) return false;
      }
  }
  if (d->visc_schur_approx == 1 && (d->visc_type == 3 || d->visc_type == -3)) return false;  // inverse_diag_lap
  {  // filter_stoch_m_flux takes widths 0, 1, 2, 4 and reads periodic images as ghosts
    const int w = d->filtering_width;
    if (w != 0 && w != 1 && w != 2 && w != 4) return false;
    if (w > 0)
      for (int k = 0; k < d->dm; ++k)
        if (d->bc_lo[k] != -1 || d->bc_hi[k] != -1 || d->n_cells[k] < w) return false;
  }
  return true;
}

template <typename T>
static int stag_mg_api(Ctx& c, StagMG<T>& mg, const lmfhd_fab* rhs, lmfhd_fab* phi) {
  std::vector<double> h;
  if (!upload_vec(c, rhs, nullptr, c.io_a, h)) { set_err("rhs shape"); return LMFHD_E_ARG; }
  if (!upload_vec(c, phi, nullptr, c.io_b, h)) { set_err("phi shape"); return LMFHD_E_ARG; }
  for (int a = 0; a < c.dm; ++a) {
    mg.template to_level0<double>(c.s, c.io_a + c.off[a], mg.L[0].rhs[a], a, 1.0);
    mg.template to_level0<double>(c.s, c.io_b + c.off[a], mg.L[0].phi[a], a, 1.0);
    launch(c.s, c.nf[a], ZeroWallFacesFn<T>{c.g, mg.L[0].phi[a], a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1)), mg.L[0].sp});
  }
  mg.solve(c.s, c.R);
  for (int a = 0; a < c.dm; ++a) mg.template from_level0<double>(c.s, mg.L[0].phi[a], c.io_b + c.off[a], a);
  check_launch();
  if (!download_vec(c, c.io_b, phi, nullptr, h)) return LMFHD_E_ARG;
  return LMFHD_OK;
}

template <typename T>
static int mac_mg_api(Ctx& c, MacMG<T>& mac, const lmfhd_fab* rhs, lmfhd_fab* phi) {
  std::vector<double> h(c.nc);
  if (!mf_ok(c, rhs, nod_cell()) || !mf_ok(c, phi, nod_cell())) { set_err("cell shape"); return LMFHD_E_ARG; }
  mf_get(c, rhs, nod_cell(), h.data());
  h2d(c.io_a, h.data(), c.nc * sizeof(double), c.s);
  launch(c.s, c.nc, ConvertFn<T, double>{c.io_a, mac.L[0].rhs, 1.0});
  mac.solve(c.s);
  launch(c.s, c.nc, ConvertFn<double, T>{mac.L[0].phi, c.io_b, T(1)});
  check_launch();
  d2h(h.data(), c.io_b, c.nc * sizeof(double), c.s);
  mf_put(c, h.data(), phi, nod_cell());
  return LMFHD_OK;
}

static int avg_api(Ctx& c, const lmfhd_fab* cc, lmfhd_fab* out, bool edge) {
  const int dm = c.dm;
  // cc with one ghost layer (ng >= 1), repacked to an (n+2)^dm array
  std::vector<double> h;
  if (!mf_get_ghost1(c, cc, h)) { set_err("cc needs one ghost layer"); return LMFHD_E_ARG; }
  const int nout = edge ? (dm == 3 ? 3 : 1) : dm;
  for (int q = 0; q < nout; ++q)
    if (!mf_ok(c, out + q * c.nb, edge ? nod_edge(c, q) : nod_face(q))) { set_err("output shape"); return LMFHD_E_ARG; }
  DevMem tm;
  double* dcc = tm.alloc<double>((long)h.size());
  h2d(dcc, h.data(), h.size() * sizeof(double), c.s);
  std::vector<double> o;
  for (int q = 0; q < nout; ++q) {
    int ex[3];
    if (edge) edge_ext(c, q, ex); else face_ext(c, q, ex);
    const long n = (long)ex[0] * ex[1] * ex[2];
    if (edge) launch(c.s, n, AvgEdgeFn{c.g, dcc, c.io_b, q, make_dec(ex[0], ex[1])});
    else launch(c.s, n, AvgFaceFn{c.g, dcc, c.io_b, q, make_dec(ex[0], ex[1])});
    check_launch();
    o.resize(n);
    d2h(o.data(), c.io_b, n * sizeof(double), c.s);
    mf_put(c, o.data(), out + q * c.nb, edge ? nod_edge(c, q) : nod_face(q));
  }
  tm.release();
  return LMFHD_OK;
}

}  // namespace lmfhd

// ==========================================================================================
// C ABI
using namespace lmfhd;

struct lmfhd_ctx {
  Ctx c;
  // one asynchronous solve at a time, run on a host thread owned by the context
  std::thread job;
  std::atomic<int> job_done{1};
  int job_rc = LMFHD_OK;
  lmfhd_stats job_stats;
  char job_err[512] = "";
  void join() {
    if (job.joinable()) job.join();
  }
};

#define API_BEGIN try {
#define API_END                    \
  }                                \
  catch (const Fail& f) {          \
    return f.code;                 \
  }                                \
  catch (...) {                    \
    set_err("unexpected exception"); \
    return LMFHD_E_HIP;            \
  }

extern "C" {

const char* lmfhd_last_error(void) { return g_err; }

void lmfhd_desc_defaults(lmfhd_desc* d) {
  memset(d, 0, sizeof(*d));
  d->dm = 3;
  d->n_cells[0] = d->n_cells[1] = d->n_cells[2] = 32;
  d->dx = 1.0 / 32;
  d->dz2d = 1.0;
  for (int k = 0; k < 3; ++k) { d->bc_lo[k] = -1; d->bc_hi[k] = -1; }
  d->visc_type = 1;
  d->precision = LMFHD_FP64;
  d->krylov = LMFHD_GMRES_LEFT;
  d->precon_type = 1;
  d->gmres_max_inner = 5;
  d->gmres_max_outer = 20;
  d->gmres_max_iter = 100;
  d->gmres_min_iter = 1;
  d->gmres_rel_tol = 1e-9;
  d->gmres_abs_tol = 0.0;
  d->p_norm_weight = 1.0;
  d->stag_mg_nsmooths_down = 2;
  d->stag_mg_nsmooths_up = 2;
  d->stag_mg_nsmooths_bottom = 8;
  d->stag_mg_minwidth = 2;
  d->stag_mg_max_vcycles = 1;
  d->stag_mg_smoother = 1;
  d->stag_mg_omega = 1.0;
  d->stag_mg_rel_tol = 1e-9;
  d->mg_nsmooths_down = 2;
  d->mg_nsmooths_up = 2;
  d->mg_nsmooths_bottom = 8;
  d->mg_minwidth = 2;
  d->mg_max_vcycles = 1;
  d->gmres_spatial_order = 2;
  d->visc_schur_approx = 0;
  d->filtering_width = 0;
  d->graphs = -1;  // the size class's tuning record decides (no graphs without one)
  d->nboxes = 0;
  d->box_lo = nullptr;
  d->box_hi = nullptr;
  d->device = 0;
}

int lmfhd_ctx_create(const lmfhd_desc* d, lmfhd_ctx** out) {
  if (!out || !valid_desc(d)) { set_err("invalid lmfhd_desc"); return LMFHD_E_ARG; }
  lmfhd_ctx* x = new lmfhd_ctx;
  try {
    ctx_build(x->c, *d);
  } catch (const Fail& f) {
    x->c.mem.release();
    delete x;
    return f.code;
  }
  *out = x;
  return LMFHD_OK;
}

int lmfhd_ctx_destroy(lmfhd_ctx* x) {
  if (!x) return LMFHD_OK;
  x->join();
#ifndef LMFHD_HOST_DEBUG
  (void)hipStreamSynchronize(x->c.s);
#endif
  x->c.clear_graphs();
  if (x->c.batch) {  // the group's tables may hold this context's functors: drop them
    std::lock_guard<std::mutex> lk(x->c.batch->mu);
    auto& set = x->c.batch->set;
    if (std::find(set.begin(), set.end(), &x->c) != set.end()) {
#ifndef LMFHD_HOST_DEBUG
      (void)hipStreamSynchronize(x->c.batch->s);
#endif
      x->c.batch->mem.release();
      set.clear();
    }
  }
  x->c.batch.reset();
  x->c.mem.release();
#ifndef LMFHD_HOST_DEBUG
  if (x->c.ev_in) (void)hipEventDestroy(x->c.ev_in);
  if (x->c.s) (void)hipStreamDestroy(x->c.s);
#endif
  delete x;
  return LMFHD_OK;
}

void* lmfhd_ctx_stream(lmfhd_ctx* x) { return x ? (void*)x->c.s : nullptr; }
long long lmfhd_ctx_device_bytes(lmfhd_ctx* x) { return x ? x->c.mem.bytes : 0; }
long long lmfhd_dev_vec_len(lmfhd_ctx* x) { return x ? x->c.vlen : 0; }

int lmfhd_sync(lmfhd_ctx* x) {
  API_BEGIN
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipStreamSynchronize(x->c.s));
#endif
  return LMFHD_OK;
  API_END
}

int lmfhd_set_coeffs(lmfhd_ctx* x, const lmfhd_fab* alpha_fc, const lmfhd_fab* beta_cc, const lmfhd_fab* beta_ed,
                     const lmfhd_fab* gamma_cc, double theta_alpha) {
  API_BEGIN
  Ctx& c = x->c;
  if (!alpha_fc || !beta_cc || !beta_ed || !gamma_cc) { set_err("null coefficient"); return LMFHD_E_ARG; }
  std::vector<double> h;
  for (int a = 0; a < c.dm; ++a) {
    if (!mf_ok(c, alpha_fc + a * c.nb, nod_face(a))) { set_err("alpha_fc[%d] shape", a); return LMFHD_E_ARG; }
    h.resize(c.nf[a]);
    mf_get(c, alpha_fc + a * c.nb, nod_face(a), h.data());
    for (double v : h)
      if (!(v != 0.0) || !std::isfinite(v)) { set_err("alpha_fc must be finite and nonzero"); return LMFHD_E_ARG; }
    h2d(c.alpha64[a], h.data(), h.size() * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(c.s));
#endif
  }
  if (!mf_ok(c, beta_cc, nod_cell()) || !mf_ok(c, gamma_cc, nod_cell())) { set_err("beta/gamma shape"); return LMFHD_E_ARG; }
  h.resize(c.nc);
  mf_get(c, beta_cc, nod_cell(), h.data());
  h2d(c.beta64, h.data(), h.size() * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipStreamSynchronize(c.s));
#endif
  mf_get(c, gamma_cc, nod_cell(), h.data());
  h2d(c.gamma64, h.data(), h.size() * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipStreamSynchronize(c.s));
#endif
  for (int e = 0; e < (c.dm == 3 ? 3 : 1); ++e) {
    if (!mf_ok(c, beta_ed + e * c.nb, nod_edge(c, e))) { set_err("beta_ed[%d] shape", e); return LMFHD_E_ARG; }
    h.resize(c.ne[e]);
    mf_get(c, beta_ed + e * c.nb, nod_edge(c, e), h.data());
    h2d(c.bed64[e], h.data(), h.size() * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(c.s));
#endif
  }
  c.theta = theta_alpha;
  set_coeffs_dev(c);
#ifndef LMFHD_HOST_DEBUG
  HIPCHK(hipStreamSynchronize(c.s));
#endif
  return LMFHD_OK;
  API_END
}

int lmfhd_stag_applyop(lmfhd_ctx* x, const lmfhd_fab* phi, lmfhd_fab* Lphi, int color) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  if (color < 0 || color > 2 * c.dm) { set_err("bad color"); return LMFHD_E_ARG; }
  std::vector<double> h;
  if (!upload_vec(c, phi, nullptr, c.io_a, h)) { set_err("phi shape"); return LMFHD_E_ARG; }
  if (!upload_vec(c, Lphi, nullptr, c.io_b, h)) { set_err("Lphi shape"); return LMFHD_E_ARG; }
  // the callers hand phi with wall-normal values set by multifab_physbc_domainvel
  for (int a = 0; a < c.dm; ++a)
    launch(c.s, c.nf[a], ZeroWallFacesFn<double>{c.g, c.u(c.io_a, a), a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  stag_applyop64(c, c.io_a, c.io_b, color);
  if (!download_vec(c, c.io_b, Lphi, nullptr, h)) return LMFHD_E_ARG;
  return LMFHD_OK;
  API_END
}

// ---- inhomogeneous velocity boundary values -------------------------------------------------
// wall direction d and component b of vel_bc_t slot q (the reference's order)
static void vbt_pair(int dm, int q, int& d, int& b) {
  if (dm == 2) { d = q; b = 1 - q; return; }
  d = q / 2;
  b = q % 2 == 0 ? (d == 0 ? 1 : 0) : (d == 2 ? 1 : 2);
}
static long vbt_len(const Ctx& c, int q) {
  int d, b;
  vbt_pair(c.dm, q, d, b);
  return c.ne[c.dm == 2 ? 0 : edge_id(3, d, b)];
}
static double* bc_buffers(Ctx& c) {
  if (!c.bc_buf) {
    long m = 2 * c.vlen;
    for (int q = 0; q < (c.dm == 3 ? 6 : 2); ++q) m += vbt_len(c, q);
    c.bc_buf = c.mem.alloc<double>(m);
  }
  return c.bc_buf;
}

// The lift of convert_to_homogeneous (convert_to_homogeneous.f90:19-106): with phi_bc zero
// except for vel_bc_n on the wall faces and the inhomogeneous ghosts of vel_bc_t,
// b_u += su (theta alpha - A0) phi_bc and b_p += sp D phi_bc. apply_matrix with vel_bc_n /
// vel_bc_t (apply_matrix.f90:83-105) is the homogeneous product plus this lift (su = 1,
// sp = -1): the operators are linear and the wall fills are affine in the interior values.
// vbn: dm face arrays packed as the faces of a Stokes vector (read on wall faces); vbt: the
// 2 (2-D) / 6 (3-D) wall arrays back to back in slot order.
static void bc_lift(Ctx& c, const double* vbn, const double* vbt, double* b, double su, double sp) {
  const int dm = c.dm;
  double* phi = bc_buffers(c);
  for (int a = 0; a < dm; ++a)
    launch(c.s, c.nf[a], BcPhiFn{c.g, vbn + c.off[a], phi + c.off[a], a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  const double* vt[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const double* q = vbt;
  for (int k = 0; k < (dm == 3 ? 6 : 2); ++k) {
    vt[k] = q;
    q += vbt_len(c, k);
  }
  Coefs<double> C;
  for (int a = 0; a < 3; ++a) { C.alpha[a] = c.alpha64[a]; C.bed[a] = c.bed64[a]; }
  C.beta = c.beta64;
  C.gamma = c.gamma64;
  C.dxsqinv = 1.0 / (c.dx * c.dx);
  C.theta = c.theta;
  double* ph[3] = {phi + c.off[0], phi + c.off[1], phi + c.off[dm == 3 ? 2 : 1]};
  dispatch_vd(c.d.visc_type, dm, [&](auto V, auto D) {
    constexpr int VT = decltype(V)::value, DM = decltype(D)::value;
    for (int a = 0; a < dm; ++a) {
      BcLiftFn<VT, DM> F;
      F.g = c.g;
      F.C = C;
      for (int d = 0; d < 3; ++d) F.phi[d] = ph[d];
      for (int k = 0; k < 6; ++k) F.vbt[k] = vt[k];
      F.dx = c.dx;
      F.sgn = su;
      F.out = b + c.off[a];
      F.a = a;
      F.dc = make_dec(fext(c.g, a, 0), fext(c.g, a, 1));
      launch(c.s, c.nf[a], F);
    }
    launch(c.s, c.nc, DivAddFn<DM>{c.g, {ph[0], ph[1], ph[2]}, b + c.off[3], 1.0 / c.dx, sp,
                                    make_dec(c.g.n[0], c.g.n[1])});
  });
  check_launch();
}

// upload vel_bc_n (dm face fabs) and vel_bc_t (2 / 6 wall fabs) into the staging after phi_bc
static bool upload_bc(Ctx& c, const lmfhd_fab* vbn, const lmfhd_fab* vbt, const double*& dn, const double*& dt) {
  double* base = bc_buffers(c);
  std::vector<double> h;
  if (!upload_vec(c, vbn, nullptr, base + c.vlen, h)) return false;
  dn = base + c.vlen;
  double* q = base + 2 * c.vlen;
  dt = q;
  for (int k = 0; k < (c.dm == 3 ? 6 : 2); ++k) {
    int d, b;
    vbt_pair(c.dm, k, d, b);
    const Nod nod = nod_edge(c, c.dm == 2 ? 0 : edge_id(3, d, b));
    if (!mf_ok(c, vbt + k * c.nb, nod)) return false;
    const long m = vbt_len(c, k);
    std::vector<double> w(m);
    mf_get(c, vbt + k * c.nb, nod, w.data());
    h2d(q, w.data(), m * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(c.s));
#endif
    q += m;
  }
  return true;
}

int lmfhd_apply_matrix_bc(lmfhd_ctx* x, const lmfhd_fab* x_u, const lmfhd_fab* x_p, lmfhd_fab* b_u,
                          lmfhd_fab* b_p, const lmfhd_fab* vel_bc_n, const lmfhd_fab* vel_bc_t) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  if (!vel_bc_n || !vel_bc_t) { set_err("vel_bc_n and vel_bc_t are required"); return LMFHD_E_ARG; }
  std::vector<double> h;
  if (!upload_vec(c, x_u, x_p, c.io_a, h)) { set_err("x shape"); return LMFHD_E_ARG; }
  const double *dn, *dt;
  if (!upload_bc(c, vel_bc_n, vel_bc_t, dn, dt)) { set_err("vel_bc shape"); return LMFHD_E_ARG; }
  apply_matrix(c, c.io_a, c.io_b);
  bc_lift(c, dn, dt, c.io_b, 1.0, -1.0);
  if (!download_vec(c, c.io_b, b_u, b_p, h)) { set_err("b shape"); return LMFHD_E_ARG; }
  return LMFHD_OK;
  API_END
}

int lmfhd_convert_to_homogeneous(lmfhd_ctx* x, lmfhd_fab* b_u, lmfhd_fab* b_p, const lmfhd_fab* vel_bc_n,
                                 const lmfhd_fab* vel_bc_t) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  if (!vel_bc_n || !vel_bc_t) { set_err("vel_bc_n and vel_bc_t are required"); return LMFHD_E_ARG; }
  std::vector<double> h;
  if (!upload_vec(c, b_u, b_p, c.io_a, h)) { set_err("b shape"); return LMFHD_E_ARG; }
  const double *dn, *dt;
  if (!upload_bc(c, vel_bc_n, vel_bc_t, dn, dt)) { set_err("vel_bc shape"); return LMFHD_E_ARG; }
  bc_lift(c, dn, dt, c.io_a, -1.0, 1.0);
  if (!download_vec(c, c.io_a, b_u, b_p, h)) { set_err("b shape"); return LMFHD_E_ARG; }
  return LMFHD_OK;
  API_END
}

// ---- plotfile fields -------------------------------------------------------------------------
static int plot_fields(Ctx& c, const double* x, int flags, double* out) {
  PlotCcFn F;
  F.g = c.g;
  for (int d = 0; d < 3; ++d) F.u[d] = x + c.off[d < c.dm ? d : c.dm - 1];
  F.p = x + c.off[3];
  F.out = out;
  F.nc = c.nc;
  F.nq = 0;
  for (int k = 0; k < 2; ++k)
    if (flags & (1 << k))
      for (int a = 0; a < c.dm; ++a) { F.kind[F.nq] = k; F.comp[F.nq] = a; ++F.nq; }
  if (flags & 4) { F.kind[F.nq] = 2; F.comp[F.nq] = 0; ++F.nq; }
  F.dc = make_dec(c.g.n[0], c.g.n[1]);
  if (F.nq > 0) launch(c.s, c.nc, F);
  check_launch();
  return F.nq;
}

int lmfhd_plot_fields_dev(lmfhd_ctx* x, const double* xv, int flags, double* out) {
  API_BEGIN
  if (!x || !xv || flags < 0 || flags > 7 || (!out && flags)) { set_err("invalid argument"); return LMFHD_E_ARG; }
  if (flags) plot_fields(x->c, xv, flags, out);
  return LMFHD_OK;
  API_END
}

int lmfhd_plot_fields(lmfhd_ctx* x, const lmfhd_fab* x_u, const lmfhd_fab* x_p, int flags, lmfhd_fab* out) {
  API_BEGIN
  if (!x || !x_u || !x_p || flags < 0 || flags > 7 || (!out && flags)) { set_err("invalid argument"); return LMFHD_E_ARG; }
  Ctx& c = x->c;
  if (!flags) return LMFHD_OK;
  std::vector<double> h;
  if (!upload_vec(c, x_u, x_p, c.io_a, h)) { set_err("x shape"); return LMFHD_E_ARG; }
  const int nq = (flags & 1 ? c.dm : 0) + (flags & 2 ? c.dm : 0) + (flags & 4 ? 1 : 0);
  std::vector<double> w((size_t)nq * c.nc);
  if (nq) {
    DevMem tmpmem;
    double* d = tmpmem.alloc<double>((long)nq * c.nc);
    plot_fields(c, c.io_a, flags, d);
    d2h(w.data(), d, w.size() * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(c.s));
#endif
    tmpmem.release();
  }
  for (int q = 0; q < nq; ++q) {
    if (!mf_ok(c, out + q * c.nb, nod_cell())) { set_err("out shape"); return LMFHD_E_ARG; }
    mf_put(c, w.data() + (size_t)q * c.nc, out + q * c.nb, nod_cell());
  }
  return LMFHD_OK;
  API_END
}

int lmfhd_apply_matrix(lmfhd_ctx* x, const lmfhd_fab* x_u, const lmfhd_fab* x_p, lmfhd_fab* b_u, lmfhd_fab* b_p) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  std::vector<double> h;
  if (!upload_vec(c, x_u, x_p, c.io_a, h)) { set_err("x shape"); return LMFHD_E_ARG; }
  apply_matrix(c, c.io_a, c.io_b);
  if (!download_vec(c, c.io_b, b_u, b_p, h)) { set_err("b shape"); return LMFHD_E_ARG; }
  return LMFHD_OK;
  API_END
}

int lmfhd_apply_precon(lmfhd_ctx* x, const lmfhd_fab* b_u, const lmfhd_fab* b_p, lmfhd_fab* x_u, lmfhd_fab* x_p) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  std::vector<double> h;
  if (!upload_vec(c, b_u, b_p, c.io_a, h)) { set_err("b shape"); return LMFHD_E_ARG; }
  precon_any(c, c.io_a, c.io_b);
  if (!download_vec(c, c.io_b, x_u, x_p, h)) { set_err("x shape"); return LMFHD_E_ARG; }
  return LMFHD_OK;
  API_END
}

int lmfhd_stag_mg_solver(lmfhd_ctx* x, const lmfhd_fab* rhs, lmfhd_fab* phi) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  if (c.d.precision == LMFHD_MIXED) return stag_mg_api<float>(c, c.mg32, rhs, phi);
  return stag_mg_api<double>(c, c.mg64, rhs, phi);
  API_END
}

int lmfhd_mac_mg_solver(lmfhd_ctx* x, const lmfhd_fab* rhs, lmfhd_fab* phi) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) { set_err("coefficients not set"); return LMFHD_E_STATE; }
  if (c.d.precision == LMFHD_MIXED) return mac_mg_api<float>(c, c.mac32, rhs, phi);
  return mac_mg_api<double>(c, c.mac64, rhs, phi);
  API_END
}

int lmfhd_gmres(lmfhd_ctx* x, const lmfhd_fab* b_u, const lmfhd_fab* b_p, lmfhd_fab* x_u, lmfhd_fab* x_p,
                double* norm_pre_rhs, lmfhd_stats* stats) {
  API_BEGIN
  Ctx& c = x->c;
  std::vector<double> h;
  if (!upload_vec(c, b_u, b_p, c.io_a, h)) { set_err("b shape"); return LMFHD_E_ARG; }
  if (!upload_vec(c, x_u, x_p, c.io_b, h)) { set_err("x shape"); return LMFHD_E_ARG; }
  lmfhd_stats st{};
  const int rc = gmres_any(c, c.io_a, c.io_b, &st, norm_pre_rhs != nullptr);
  if (stats) *stats = st;
  if (norm_pre_rhs) *norm_pre_rhs = st.norm_pre_b;
  if (rc != LMFHD_OK && rc != LMFHD_E_BREAKDOWN) return rc;
  if (!download_vec(c, c.io_b, x_u, x_p, h)) return LMFHD_E_ARG;
  return rc;
  API_END
}

int lmfhd_gmres_dev(lmfhd_ctx* x, const double* b, double* xv, lmfhd_stats* stats) {
  API_BEGIN
  lmfhd_stats st{};
  const int rc = gmres_any(x->c, b, xv, &st);
  if (stats) *stats = st;
  return rc;
  API_END
}

int lmfhd_advance_inertial_dev(lmfhd_ctx* x, double dt, double* xv, const double* f1, const double* f2,
                               lmfhd_stats* stats) {
  API_BEGIN
  if (!x || !xv) { set_err("null argument"); return LMFHD_E_ARG; }
  if (!x->job_done.load()) { set_err("a solve is running on this context"); return LMFHD_E_STATE; }
  return advance_step(x->c, true, dt, xv, f1, f2, stats);
  API_END
}

int lmfhd_advance_overdamped_dev(lmfhd_ctx* x, double* xv, const double* f1, const double* f2, lmfhd_stats* stats) {
  API_BEGIN
  if (!x || !xv) { set_err("null argument"); return LMFHD_E_ARG; }
  if (!x->job_done.load()) { set_err("a solve is running on this context"); return LMFHD_E_STATE; }
  return advance_step(x->c, false, 0.0, xv, f1, f2, stats);
  API_END
}

int lmfhd_advective_m_fluxdiv(lmfhd_ctx* x, const lmfhd_fab* u, const lmfhd_fab* rho, lmfhd_fab* m_update,
                              int increment) {
  API_BEGIN
  if (!x || !u || !rho || !m_update) { set_err("null argument"); return LMFHD_E_ARG; }
  Ctx& c = x->c;
  step_bufs(c);
  std::vector<double> h;
  if (!upload_vec(c, u, nullptr, c.io_a, h)) { set_err("u shape"); return LMFHD_E_ARG; }
  if (!upload_vec(c, rho, nullptr, c.io_b, h)) { set_err("rho shape"); return LMFHD_E_ARG; }
  if (increment && !upload_vec(c, m_update, nullptr, c.st_b, h)) { set_err("m_update shape"); return LMFHD_E_ARG; }
  // the velocity as the reference's callers leave it: wall-normal faces zero (domainvel)
  if (has_wall(c.g))
    for (int a = 0; a < c.dm; ++a)
      launch(c.s, c.nf[a], ZeroWallFacesFn<double>{c.g, c.u(c.io_a, a), a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  const int q2 = c.dm == 3 ? 2 : 1;
  const double dxinv = 1.0 / c.dx;
  for (int a = 0; a < c.dm; ++a) {
    AdvMFn<3> F3;
    AdvMFn<2> F2;
    auto setup = [&](auto& F) {
      F.g = c.g;
      for (int d = 0; d < 3; ++d) {
        const int q = d < c.dm ? d : q2;
        F.u[d] = c.io_a + c.off[q];
        F.rho[d] = c.io_b + c.off[q];
      }
      F.out = c.st_b + c.off[a];
      F.dxinv = dxinv;
      F.a = a;
      F.increment = increment;
      F.dc = make_dec(fext(c.g, a, 0), fext(c.g, a, 1));
    };
    if (c.dm == 3) { setup(F3); launch(c.s, c.nf[a], F3); }
    else { setup(F2); launch(c.s, c.nf[a], F2); }
  }
  check_launch();
  if (!download_vec(c, c.st_b, m_update, nullptr, h)) { set_err("m_update shape"); return LMFHD_E_ARG; }
  return LMFHD_OK;
  API_END
}

long long lmfhd_dev_coef_len(lmfhd_ctx* x) {
  if (!x) return 0;
  const Ctx& c = x->c;
  long n = 2 * c.nc;
  for (int a = 0; a < c.dm; ++a) n += c.nf[a];
  for (int e = 0; e < (c.dm == 3 ? 3 : 1); ++e) n += c.ne[e];
  return n;
}

int lmfhd_set_coeffs_dev(lmfhd_ctx* x, const double* coef, double theta_alpha) {
  API_BEGIN
  if (!x || !coef) { set_err("null argument"); return LMFHD_E_ARG; }
  Ctx& c = x->c;
  const double* p = coef;
  for (int a = 0; a < c.dm; ++a) {
    d2d(c.alpha64[a], p, c.nf[a] * sizeof(double), c.s);
    p += c.nf[a];
  }
  d2d(c.beta64, p, c.nc * sizeof(double), c.s);
  p += c.nc;
  for (int e = 0; e < (c.dm == 3 ? 3 : 1); ++e) {
    d2d(c.bed64[e], p, c.ne[e] * sizeof(double), c.s);
    p += c.ne[e];
  }
  d2d(c.gamma64, p, c.nc * sizeof(double), c.s);
  c.theta = theta_alpha;
  set_coeffs_dev(c);
  return LMFHD_OK;
  API_END
}

// the body of an asynchronous or batched solve, on the calling host thread
static int gmres_job(lmfhd_ctx* x, const double* b, double* xv, lmfhd_stats* st) {
  try {
    int rc = LMFHD_OK;
    try {
      rc = gmres_any(x->c, b, xv, st);
    } catch (...) {
      // the caller may free b and x once the job ends: let queued kernels finish first
      HIPCHK_NOTHROW(hipStreamSynchronize(x->c.s));
      throw;
    }
    sync_stream(x->c.s);  // on every return, so that b and x are free for reuse afterwards
    return rc;
  } catch (const Fail& f) {
    return f.code;
  } catch (const std::exception& e) {
    set_err("solve failed: %s", e.what());
    return LMFHD_E_HIP;
  } catch (...) {
    set_err("solve failed: unexpected exception");
    return LMFHD_E_HIP;
  }
}

int lmfhd_gmres_dev_async(lmfhd_ctx* x, const double* b, double* xv) {
  if (!x || !b || !xv) { set_err("null argument"); return LMFHD_E_ARG; }
  if (!x->job_done.load()) { set_err("an asynchronous solve is already running on this context"); return LMFHD_E_STATE; }
  x->join();
  x->job_done.store(0);
  x->job_rc = LMFHD_OK;
  x->job_err[0] = 0;
  x->job = std::thread([x, b, xv]() {
    HIPCHK_NOTHROW(hipSetDevice(x->c.d.device));
    x->job_rc = gmres_job(x, b, xv, &x->job_stats);
    if (x->job_rc != LMFHD_OK) snprintf(x->job_err, sizeof(x->job_err), "%s", g_err);
    x->job_done.store(1);
  });
  return LMFHD_OK;
}

int lmfhd_poll(lmfhd_ctx* x) { return (!x || x->job_done.load()) ? 1 : 0; }

int lmfhd_wait(lmfhd_ctx* x, lmfhd_stats* stats) {
  if (!x) { set_err("null context"); return LMFHD_E_ARG; }
  x->join();
  if (stats) *stats = x->job_stats;
  if (x->job_rc != LMFHD_OK) set_err("%s", x->job_err);
  return x->job_rc;
}

int lmfhd_gmres_dev_batch(lmfhd_ctx** ctxs, int n, const double* const* b, double* const* xv, lmfhd_stats* stats,
                          int* rcs, int max_threads) {
  if (n < 0 || (n > 0 && (!ctxs || !b || !xv))) { set_err("invalid batch arguments"); return LMFHD_E_ARG; }
  for (int q = 0; q < n; ++q)
    if (!ctxs[q] || !b[q] || !xv[q] || !ctxs[q]->job_done.load()) {
      set_err("batch entry %d: null pointer or busy context", q);
      return LMFHD_E_ARG;
    }
  for (int q = 0; q < n; ++q)
    for (int r = q + 1; r < n; ++r)
      if (ctxs[q] == ctxs[r]) { set_err("batch entries %d and %d share a context", q, r); return LMFHD_E_ARG; }
  const int nt = std::max(1, std::min(n, max_threads > 0 ? max_threads : 8));
  std::atomic<int> next{0};
  std::vector<std::string> errs(n);
  std::vector<int> rc(n, LMFHD_OK);
  auto worker = [&]() {
    for (;;) {
      const int q = next.fetch_add(1);
      if (q >= n) break;
#ifndef LMFHD_HOST_DEBUG
      if (hipSetDevice(ctxs[q]->c.d.device) != hipSuccess) { rc[q] = LMFHD_E_HIP; errs[q] = "hipSetDevice failed"; continue; }
#endif
      lmfhd_stats st{};
      rc[q] = gmres_job(ctxs[q], b[q], xv[q], &st);
      if (stats) stats[q] = st;
      if (rc[q] != LMFHD_OK) errs[q] = g_err;
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
  int first = LMFHD_OK;
  for (int q = 0; q < n; ++q) {
    if (rcs) rcs[q] = rc[q];
    if (rc[q] != LMFHD_OK && first == LMFHD_OK) {
      first = rc[q];
      set_err("batch entry %d: %s", q, errs[q].c_str());
    }
  }
  return first;
}

int lmfhd_apply_matrix_dev(lmfhd_ctx* x, const double* xv, double* b) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  d2d(c.io_a, xv, c.vlen * sizeof(double), c.s);
  apply_matrix(c, c.io_a, b);
  return LMFHD_OK;
  API_END
}

int lmfhd_apply_matrix_bc_dev(lmfhd_ctx* x, const double* xv, double* b, const double* vel_bc_n,
                              const double* vel_bc_t) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  if (!xv || !b || !vel_bc_n || !vel_bc_t) { set_err("invalid argument"); return LMFHD_E_ARG; }
  d2d(c.io_a, xv, c.vlen * sizeof(double), c.s);
  apply_matrix(c, c.io_a, b);
  bc_lift(c, vel_bc_n, vel_bc_t, b, 1.0, -1.0);
  return LMFHD_OK;
  API_END
}

int lmfhd_convert_to_homogeneous_dev(lmfhd_ctx* x, double* b, const double* vel_bc_n, const double* vel_bc_t) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  if (!b || !vel_bc_n || !vel_bc_t) { set_err("invalid argument"); return LMFHD_E_ARG; }
  bc_lift(c, vel_bc_n, vel_bc_t, b, -1.0, 1.0);
  return LMFHD_OK;
  API_END
}

long long lmfhd_dev_vel_bc_t_len(lmfhd_ctx* x) {
  if (!x) return -1;
  long long m = 0;
  for (int q = 0; q < (x->c.dm == 3 ? 6 : 2); ++q) m += vbt_len(x->c, q);
  return m;
}

int lmfhd_apply_precon_dev(lmfhd_ctx* x, const double* b, double* xv) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  precon_any(c, b, xv);
  return LMFHD_OK;
  API_END
}

int lmfhd_stag_applyop_dev(lmfhd_ctx* x, const double* phi, double* Lphi) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  stag_applyop64(c, const_cast<double*>(phi), Lphi, 0);
  return LMFHD_OK;
  API_END
}

int lmfhd_stag_applyop_mg_dev(lmfhd_ctx* x, const void* phi, void* Lphi) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  if (c.d.precision == LMFHD_MIXED) {
    float* p = (float*)const_cast<void*>(phi);
    float* o = (float*)Lphi;
    float* ph[3] = {p + c.off[0], p + c.off[1], p + c.off[2]};
    float* ou[3] = {o + c.off[0], o + c.off[1], o + c.off[2]};
    c.mg32.apply(c.s, ph, ou);
  } else {
    double* p = (double*)const_cast<void*>(phi);
    double* o = (double*)Lphi;
    double* ph[3] = {p + c.off[0], p + c.off[1], p + c.off[2]};
    double* ou[3] = {o + c.off[0], o + c.off[1], o + c.off[2]};
    c.mg64.apply(c.s, ph, ou);
  }
  return LMFHD_OK;
  API_END
}

int lmfhd_stag_smooth_dev(lmfhd_ctx* x, int nsweeps) {
  API_BEGIN
  Ctx& c = x->c;
  if (!c.have_coef) return LMFHD_E_STATE;
  if (nsweeps < 0) return LMFHD_E_ARG;
  if (c.d.precision == LMFHD_MIXED) c.mg32.smooth(c.s, 0, nsweeps);
  else c.mg64.smooth(c.s, 0, nsweeps);
  return LMFHD_OK;
  API_END
}

int lmfhd_compute_div(lmfhd_ctx* x, const lmfhd_fab* u, lmfhd_fab* div) {
  API_BEGIN
  Ctx& c = x->c;
  std::vector<double> h;
  if (!upload_vec(c, u, nullptr, c.io_a, h)) { set_err("u shape"); return LMFHD_E_ARG; }
  if (!mf_ok(c, div, nod_cell())) { set_err("div shape"); return LMFHD_E_ARG; }
  const double* uu[3] = {c.io_a + c.off[0], c.io_a + c.off[1], c.io_a + (c.dm == 3 ? c.off[2] : c.off[1])};
  if (c.dm == 3) launch(c.s, c.nc, DivPlusFn<double, 3>{c.g, {uu[0], uu[1], uu[2]}, nullptr, c.io_b, 1.0 / c.dx, make_dec(c.g.n[0], c.g.n[1])});
  else launch(c.s, c.nc, DivPlusFn<double, 2>{c.g, {uu[0], uu[1], uu[2]}, nullptr, c.io_b, 1.0 / c.dx, make_dec(c.g.n[0], c.g.n[1])});
  check_launch();
  h.resize(c.nc);
  d2h(h.data(), c.io_b, c.nc * sizeof(double), c.s);
  mf_put(c, h.data(), div, nod_cell());
  return LMFHD_OK;
  API_END
}

int lmfhd_compute_grad(lmfhd_ctx* x, const lmfhd_fab* p, lmfhd_fab* gradp) {
  API_BEGIN
  Ctx& c = x->c;
  if (!mf_ok(c, p, nod_cell())) { set_err("p shape"); return LMFHD_E_ARG; }
  std::vector<double> h(c.nc);
  mf_get(c, p, nod_cell(), h.data());
  h2d(c.io_a, h.data(), c.nc * sizeof(double), c.s);
  for (int a = 0; a < c.dm; ++a)
    launch(c.s, c.nf[a], GradFn{c.g, c.io_a, c.io_b + c.off[a], 1.0 / c.dx, a, make_dec(fext(c.g, a, 0), fext(c.g, a, 1))});
  check_launch();
  if (!download_vec(c, c.io_b, gradp, nullptr, h)) { set_err("grad shape"); return LMFHD_E_ARG; }
  return LMFHD_OK;
  API_END
}

// ---- staging memory of the no-allocation entry point -------------------------------------------
// lmfhd_stochastic_m_fluxdiv_dev_noalloc stages the filtered stresses here instead of in memory
// the context allocates: a file-scope device array of 64 MiB, one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory): the device address comes from hipGetSymbolAddress, taken once
// per device with that device current, and kept in g_noalloc_ptr.  Every call on a GPU shares
// its copy, so calls that overlap on one GPU (contexts on different streams) race.  Each call
// writes every staged value before reading it, so nothing has to start at zero.
constexpr long NOALLOC_SCRATCH_DOUBLES = 8L << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
#ifdef LMFHD_HOST_DEBUG
static double scratch[NOALLOC_SCRATCH_DOUBLES];
#else
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
#endif
static std::mutex g_noalloc_mu;
static double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

static double* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_err("device %d out of range for the no-allocation staging memory", dev);
    throw Fail{LMFHD_E_ARG};
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
#ifdef LMFHD_HOST_DEBUG
    g_noalloc_ptr[dev] = scratch;
#else
    int prev = 0;
    HIPCHK(hipGetDevice(&prev));
    HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    HIPCHK(hipSetDevice(prev));
    if (e != hipSuccess || !p) {
      set_err("hipGetSymbolAddress of the staging memory failed: %s", hipGetErrorString(e));
      throw Fail{LMFHD_E_HIP};
    }
    g_noalloc_ptr[dev] = static_cast<double*>(p);
#endif
  }
  return g_noalloc_ptr[dev];
}

// The flux divergence of stochastic_m_fluxdiv from F[a], one configured StochFn per face
// component. With filtering_width > 0 the scaled stresses are first written out
// (StochStageFn), filtered along each axis in turn (StochFilterFn, ping-ponging with one spare
// array) and the divergence then reads the filtered arrays (:172-175, :226-231).
// Doubles of staging the filtered form needs: narr + 1 slots of the largest stress array.
static long stoch_stage_doubles(const Ctx& c, long* slot_out) {
  const int ne = c.dm == 3 ? 3 : 1, narr = c.dm + 2 * ne;
  long m = c.nc;
  for (int e = 0; e < ne; ++e) m = std::max(m, c.ne[e]);
  const long slot = (m + 63) / 64 * 64;
  if (slot_out) *slot_out = slot;
  return slot * (narr + 1);
}

// stage: staging memory of at least stoch_stage_doubles(c) doubles for the filtered form, or
// null to use the context's own (allocated on first use).
static void launch_stoch(Ctx& c, StochFn* F, double* stage = nullptr) {
  const int dm = c.dm, w = c.d.filtering_width;
  if (w == 0) {
    for (int a = 0; a < dm; ++a) launch(c.s, c.nf[a], F[a]);
    return;
  }
  const int ne = dm == 3 ? 3 : 1, narr = dm + 2 * ne;
  long stage_slot = 0;
  const long need = stoch_stage_doubles(c, &stage_slot);
  if (!stage) {
    if (!c.stress) {
      c.stress_slot = stage_slot;
      c.stress = c.mem.alloc<double>(need);
    }
    stage = c.stress;
  }
  double* slot[10];
  for (int r = 0; r <= narr; ++r) slot[r] = stage + stage_slot * r;
  double* spare = slot[narr];
  int ext[10][3];
  for (int a = 0; a < dm; ++a) {
    for (int d = 0; d < 3; ++d) ext[a][d] = c.g.n[d];
    launch(c.s, c.nc, StochStageFn{F[a], -1, slot[a], make_dec(ext[a][0], ext[a][1])});
  }
  for (int e = 0; e < ne; ++e) {
    int nd[2], q = 0;  // the two nodal directions of edge array e, ascending
    for (int d = 0; d < dm; ++d)
      if (eext(c.g, e, d) == c.g.n[d] + 1) nd[q++] = d;
    for (int cq = 0; cq < 2; ++cq) {
      // component 0 is read by the face component of the lower direction, 1 by the upper
      const int a = nd[cq], b = nd[1 - cq];
      int tb[2];
      transverse(a, tb[0], tb[1]);
      const int t = tb[0] == b ? 0 : 1;
      const int r = dm + 2 * e + cq;
      for (int d = 0; d < 3; ++d) ext[r][d] = eext(c.g, e, d);
      launch(c.s, c.ne[e], StochStageFn{F[a], t, slot[r], make_dec(ext[r][0], ext[r][1])});
    }
  }
  for (int r = 0; r < narr; ++r) {
    const long cnt = (long)ext[r][0] * ext[r][1] * ext[r][2];
    const Dec dc = make_dec(ext[r][0], ext[r][1]);
    for (int d = 0; d < dm; ++d) {
      const long stride = d == 0 ? 1 : (d == 1 ? (long)ext[r][0] : (long)ext[r][0] * ext[r][1]);
      launch(c.s, cnt, StochFilterFn{slot[r], spare, stride, d, c.g.n[d], w, dc});
      std::swap(slot[r], spare);
    }
  }
  for (int a = 0; a < dm; ++a) {
    StochFn G = F[a];
    G.pre = 1;
    G.gen = 0;
    G.Wcc = slot[a];
    int tb[2];
    transverse(a, tb[0], tb[1]);
    for (int t = 0; t < dm - 1; ++t) {
      const int b = tb[t];
      G.Wed[t] = slot[dm + 2 * G.eid[t] + (a < b ? 0 : 1)];
    }
    launch(c.s, c.nf[a], G);
  }
}

int lmfhd_stochastic_m_fluxdiv(lmfhd_ctx* x, const lmfhd_fab* W_cc, const lmfhd_fab* W_ed, const lmfhd_fab* eta_cc,
                               const lmfhd_fab* eta_ed, const lmfhd_fab* T_cc, const lmfhd_fab* T_ed, double dt,
                               double k_B, double variance_coef_mom, lmfhd_fab* f) {
  API_BEGIN
  Ctx& c = x->c;
  const int dm = c.dm, ne = dm == 3 ? 3 : 1;
  if (!(dt > 0.0)) { set_err("dt must be positive"); return LMFHD_E_ARG; }
  DevMem tmpmem;
  auto up_cell = [&](const lmfhd_fab* fb) -> double* {
    if (!mf_ok(c, fb, nod_cell())) throw Fail{LMFHD_E_ARG};
    std::vector<double> h(c.nc);
    mf_get(c, fb, nod_cell(), h.data());
    double* d = tmpmem.alloc<double>(c.nc);
    h2d(d, h.data(), c.nc * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(c.s));
#endif
    return d;
  };
  auto up_edge = [&](const lmfhd_fab* fb, int e) -> double* {
    if (!mf_ok(c, fb, nod_edge(c, e))) throw Fail{LMFHD_E_ARG};
    std::vector<double> h(c.ne[e]);
    mf_get(c, fb, nod_edge(c, e), h.data());
    double* d = tmpmem.alloc<double>(c.ne[e]);
    h2d(d, h.data(), c.ne[e] * sizeof(double), c.s);
#ifndef LMFHD_HOST_DEBUG
    HIPCHK(hipStreamSynchronize(c.s));
#endif
    return d;
  };
  int rc = LMFHD_OK;
  try {
    double* wcc[3] = {nullptr, nullptr, nullptr};
    for (int a = 0; a < dm; ++a) wcc[a] = up_cell(W_cc + a * c.nb);
    double* wed[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int e = 0; e < ne; ++e)
      for (int q = 0; q < 2; ++q) wed[2 * e + q] = up_edge(W_ed + (2 * e + q) * c.nb, e);
    double* ecc = up_cell(eta_cc);
    double* eed[3] = {nullptr, nullptr, nullptr};
    for (int e = 0; e < ne; ++e) eed[e] = up_edge(eta_ed + e * c.nb, e);
    double* tcc = T_cc ? up_cell(T_cc) : nullptr;
    double* ted[3] = {nullptr, nullptr, nullptr};
    if (T_ed)
      for (int e = 0; e < ne; ++e) ted[e] = up_edge(T_ed + e * c.nb, e);
    const double dz = dm == 3 ? c.dx : c.d.dz2d;
    const double var = std::sqrt(variance_coef_mom * 2.0 * k_B / (c.dx * c.dx * dz * dt));
    StochFn Fs[3];
    for (int a = 0; a < dm; ++a) {
      StochFn F;
      F.g = c.g;
      F.Wcc = wcc[a];
      F.eta_cc = ecc;
      F.T_cc = tcc;
      int tb[2];
      transverse(a, tb[0], tb[1]);
      for (int t = 0; t < 2; ++t) {
        F.Wed[t] = nullptr; F.eta_ed[t] = nullptr; F.T_ed[t] = nullptr; F.eid[t] = 0;
        if (t >= dm - 1) continue;
        const int b = tb[t];
        const int e = edge_id(dm, a, b);
        const int q = a < b ? 0 : 1;
        F.Wed[t] = wed[2 * e + q];
        F.eta_ed[t] = eed[e];
        F.T_ed[t] = ted[e];
        F.eid[t] = e;
      }
      F.var = var;
      F.dxinv = 1.0 / c.dx;
      F.sqrt2 = std::sqrt(2.0);
      F.out = c.io_b + c.off[a];
      F.a = a;
      F.dc = make_dec(fext(c.g, a, 0), fext(c.g, a, 1));
      Fs[a] = F;
    }
    launch_stoch(c, Fs);
    check_launch();
    std::vector<double> h;
    if (!download_vec(c, c.io_b, f, nullptr, h)) rc = LMFHD_E_ARG;
  } catch (const Fail& e) {
    rc = e.code;
    if (rc == LMFHD_E_ARG) set_err("stochastic_m_fluxdiv: array shape");
  }
#ifndef LMFHD_HOST_DEBUG
  (void)hipStreamSynchronize(c.s);
#endif
  tmpmem.release();
  return rc;
  API_END
}

// Philox key word 1 (the first hexadecimal digits of pi); key word 0 is the caller's seed
static const unsigned long long NOISE_KEY1 = 0x243F6A8885A308D3ull;

// The body of lmfhd_stochastic_m_fluxdiv_dev; stage_mem as in launch_stoch.
static int stoch_dev(lmfhd_ctx* x, unsigned long long seed, unsigned long long stage, int stoch_stress_form,
                     const double* eta_cc, const double* eta_ed, double dt, double k_B, double variance_coef_mom,
                     int increment, double* f, double* stage_mem) {
  Ctx& c = x->c;
  if (!c.have_coef && (!eta_cc || !eta_ed)) { set_err("no viscosity: pass eta or set the coefficients first"); return LMFHD_E_STATE; }
  const int dm = c.dm, ne = dm == 3 ? 3 : 1;
  const double* ecc = eta_cc ? eta_cc : c.beta64;
  const double* eed[3] = {nullptr, nullptr, nullptr};
  const double* pe = eta_ed;
  for (int e = 0; e < ne; ++e) {
    eed[e] = eta_ed ? pe : c.bed64[e];
    if (eta_ed) pe += c.ne[e];
  }
  const double dz = dm == 3 ? c.dx : c.d.dz2d;
  const double var = std::sqrt(variance_coef_mom * 2.0 * k_B / (c.dx * c.dx * dz * dt));
  StochFn Fs[3];
  for (int a = 0; a < dm; ++a) {
    StochFn F;
    F.g = c.g;
    F.Wcc = nullptr;
    F.eta_cc = ecc;
    F.T_cc = nullptr;
    int tb[2];
    transverse(a, tb[0], tb[1]);
    for (int t = 0; t < 2; ++t) {
      F.Wed[t] = nullptr; F.eta_ed[t] = nullptr; F.T_ed[t] = nullptr; F.eid[t] = 0; F.ecomp[t] = 0;
      if (t >= dm - 1) continue;
      const int b = tb[t];
      const int e = edge_id(dm, a, b);
      F.eta_ed[t] = eed[e];
      F.eid[t] = e;
      F.ecomp[t] = stoch_stress_form == 0 ? (a < b ? 0 : 1) : 0;
    }
    F.var = var;
    F.dxinv = 1.0 / c.dx;
    F.sqrt2 = std::sqrt(2.0);
    F.out = f + c.off[a];
    F.a = a;
    F.dc = make_dec(fext(c.g, a, 0), fext(c.g, a, 1));
    F.gen = 1;
    F.ng.ph.k0 = seed;
    F.ng.ph.k1 = NOISE_KEY1;
    F.ng.stage = stage;
    F.ccfac = stoch_stress_form == 0 ? 1.0 : std::sqrt(2.0);
    F.incr = increment ? 1 : 0;
    Fs[a] = F;
  }
  launch_stoch(c, Fs, stage_mem);
  if (!increment) fill<double>(c.s, f + c.off[3], c.nc, 0.0);
  check_launch();
  return LMFHD_OK;
}

int lmfhd_stochastic_m_fluxdiv_dev(lmfhd_ctx* x, unsigned long long seed, unsigned long long stage,
                                   int stoch_stress_form, const double* eta_cc, const double* eta_ed, double dt,
                                   double k_B, double variance_coef_mom, int increment, double* f) {
  API_BEGIN
  if (!x || !f || !(dt > 0.0)) { set_err("invalid argument"); return LMFHD_E_ARG; }
  return stoch_dev(x, seed, stage, stoch_stress_form, eta_cc, eta_ed, dt, k_B, variance_coef_mom, increment, f,
                   nullptr);
  API_END
}

int lmfhd_stochastic_m_fluxdiv_dev_noalloc(lmfhd_ctx* x, unsigned long long seed, unsigned long long stage,
                                           int stoch_stress_form, const double* eta_cc, const double* eta_ed,
                                           double dt, double k_B, double variance_coef_mom, int increment,
                                           double* f) {
  API_BEGIN
  if (!x || !f || !(dt > 0.0)) { set_err("invalid argument"); return LMFHD_E_ARG; }
  Ctx& c = x->c;
  double* stage_mem = nullptr;
  if (c.d.filtering_width > 0) {
    if (stoch_stage_doubles(c, nullptr) > NOALLOC_SCRATCH_DOUBLES) {
      set_err("grid too large for the no-allocation staging memory (64 MiB: 3-D grids up to about 90^3 "
              "with filtering)");
      return LMFHD_E_ARG;
    }
    stage_mem = noalloc_scratch(c.d.device);
  }
  return stoch_dev(x, seed, stage, stoch_stress_form, eta_cc, eta_ed, dt, k_B, variance_coef_mom, increment, f,
                   stage_mem);
  API_END
}

int lmfhd_average_cc_to_face(lmfhd_ctx* x, const lmfhd_fab* cc, lmfhd_fab* face) {
  API_BEGIN
  return avg_api(x->c, cc, face, false);
  API_END
}

int lmfhd_average_cc_to_edge(lmfhd_ctx* x, const lmfhd_fab* cc, lmfhd_fab* edge) {
  API_BEGIN
  return avg_api(x->c, cc, edge, true);
  API_END
}

// ---- launch-parameter autotuner ----------------------------------------------------------------
static lmfhd_tune to_c(const Tune& t) {
  lmfhd_tune o;
  o.coarse_n = t.coarse_n;
  o.mac_coarse_n = t.mac_coarse_n;
  o.coarse_threads = t.coarse_threads;
  o.red_blocks = t.red_blocks;
  o.graphs = t.graphs;
  o.rows_n = t.rows_n;
  o.ms_per_solve = t.ms;
  return o;
}
static Tune from_c(const lmfhd_tune& o) {
  Tune t;
  t.coarse_n = o.coarse_n;
  t.mac_coarse_n = o.mac_coarse_n;
  t.coarse_threads = o.coarse_threads;
  t.red_blocks = o.red_blocks;
  t.graphs = o.graphs;
  t.rows_n = o.rows_n;
  t.ms = o.ms_per_solve;
  return t;
}

int lmfhd_tune_set(const lmfhd_desc* d, const lmfhd_tune* t) {
  if (!d || !t || !valid_desc(d) || !tune_valid(from_c(*t))) { set_err("invalid tuning record"); return LMFHD_E_ARG; }
  tune_store(*d, from_c(*t));
  return LMFHD_OK;
}

int lmfhd_tune_get(const lmfhd_desc* d, lmfhd_tune* t) {
  if (!d || !t || !valid_desc(d)) { set_err("invalid argument"); return LMFHD_E_ARG; }
  Tune v;
  const bool found = tune_lookup(*d, v);
  *t = to_c(v);
  return found ? 1 : 0;
}

int lmfhd_ctx_tune(lmfhd_ctx* x, lmfhd_tune* t) {
  if (!x || !t) { set_err("invalid argument"); return LMFHD_E_ARG; }
  *t = to_c(x->c.tune);
  return LMFHD_OK;
}

// Coordinate search over the launch parameters of one size class: starting from the defaults
// (or the class's record), each parameter in turn takes the candidate value with the least
// median time per solve of a synthetic problem of the class (alpha = 1, beta = 1 on cells and
// edges, gamma = 0, theta_alpha = 10, device noise as the right-hand side, the descriptor's
// tolerances), the others held at their best so far.  The winner is recorded for the class.
int lmfhd_autotune(const lmfhd_desc* d, int reps, lmfhd_tune* out) {
  API_BEGIN
  if (!d || !valid_desc(d)) { set_err("invalid lmfhd_desc"); return LMFHD_E_ARG; }
  if (reps < 1) reps = 3;
  lmfhd_desc dd = *d;
  dd.graphs = 0;
  lmfhd_ctx* x = nullptr;
  int rc = lmfhd_ctx_create(&dd, &x);
  if (rc != LMFHD_OK) return rc;
  struct Guard {
    lmfhd_ctx* x;
    ~Guard() { lmfhd_ctx_destroy(x); }
  } guard{x};
  Ctx& c = x->c;
  DevMem tm;
  const long ncoef = lmfhd_dev_coef_len(x);
  double* coef = tm.alloc<double>(ncoef);
  double* b = tm.alloc<double>(c.vlen);
  double* xv = tm.alloc<double>(c.vlen);
  const long nalpha = c.off[3];
  fill<double>(c.s, coef, ncoef, 1.0);
  fill<double>(c.s, coef + ncoef - c.nc, c.nc, 0.0);  // gamma
  (void)nalpha;
  if ((rc = lmfhd_set_coeffs_dev(x, coef, 10.0)) != LMFHD_OK) return rc;
  if ((rc = lmfhd_stochastic_m_fluxdiv_dev(x, 12345ull, 0ull, 0, nullptr, nullptr, 0.1, 1.0, 1.0, 0, b)) != LMFHD_OK)
    return rc;
  sync_stream(c.s);
  auto time_solve = [&](const Tune& t) -> double {
    apply_tune(c, t);
    std::vector<double> ms;
    for (int r = 0; r <= reps; ++r) {  // the first solve warms up (and captures graphs)
      fill<double>(c.s, xv, c.vlen, 0.0);
      sync_stream(c.s);
      const auto t0 = std::chrono::steady_clock::now();
      lmfhd_stats st{};
      const int rr = gmres_any(c, b, xv, &st);
      sync_stream(c.s);
      const auto t1 = std::chrono::steady_clock::now();
      if (rr != LMFHD_OK) throw Fail{rr};
      if (r > 0) ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
  };
  Tune best;
  tune_lookup(*d, best);
  best.ms = time_solve(best);
  const int cn[] = {0, 4, 8, 16};
  const int nt[] = {256, 512, 1024};
  const int rb[] = {256, 1024, 4096};
  const int gr[] = {0, 1};
  const int rn[] = {0, 32, 64, 96};
  auto search = [&](int Tune::*field, const int* vals, int nv) {
    for (int q = 0; q < nv; ++q) {
      if (vals[q] == best.*field) continue;
      Tune t = best;
      t.*field = vals[q];
      t.ms = time_solve(t);
      if (t.ms < best.ms) best = t;
    }
  };
  search(&Tune::coarse_n, cn, 4);
  search(&Tune::mac_coarse_n, cn, 4);
  search(&Tune::coarse_threads, nt, 3);
  search(&Tune::red_blocks, rb, 3);
  search(&Tune::graphs, gr, 2);
  search(&Tune::rows_n, rn, 4);
  tune_store(*d, best);
  if (out) *out = to_c(best);
  return LMFHD_OK;
  API_END
}

}  // extern "C"
