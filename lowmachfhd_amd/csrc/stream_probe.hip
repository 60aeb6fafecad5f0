// stream_probe.hip -- achievable-bandwidth probe for the roofline of the stencil kernels.
//
// The stencil kernels of the solver read 11 to 15 arrays per point and write one to three.  A
// plain copy does not bound them: with many concurrent read streams the memory system delivers
// less.  This probe times, with HIP events, kernels that do nothing but stream:
//   * copy3: out = a + b + c, float4 per lane (three read streams, one write stream);
//   * sum13: out = sum of 13 arrays, one float per lane (thirteen read streams, one write);
// each at a working set that fits the 256 MB Infinity Cache (arrays of 128^3 floats) and at one
// that does not (256^3), and prints one JSON object of GB/s figures (bytes read + written per
// launch over the average launch time).  Built in-tree by lowmachfhd_amd/build.py; bench.py runs
// it as a child process.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                 \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "stream_probe: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      exit(1);                                                                \
    }                                                                         \
  } while (0)

__global__ void __launch_bounds__(256) k_copy3(const float4* __restrict__ a, const float4* __restrict__ b,
                                               const float4* __restrict__ c, float4* __restrict__ o, long n4) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n4) {
    const float4 x = a[t], y = b[t], z = c[t];
    o[t] = make_float4(x.x + y.x + z.x, x.y + y.y + z.y, x.z + y.z + z.z, x.w + y.w + z.w);
  }
}

__global__ void __launch_bounds__(256) k_sum13(const float* __restrict__ a, float* __restrict__ o, long n) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 13; ++q) s += a[q * n + t];
    o[t] = s;
  }
}

// average launch time (ms) of `reps` launches after one warm-up
template <class F>
static float time_ms(hipStream_t s, int reps, F&& go) {
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  go();
  CK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) go();
  CK(hipEventRecord(e1, s));
  CK(hipEventSynchronize(e1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, e0, e1));
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return ms / reps;
}

int main() {
  hipStream_t s;
  CK(hipStreamCreate(&s));
  const int reps = 30;
  const long sizes[2] = {128L * 128 * 128, 256L * 256 * 256};
  const char* names[2] = {"128^3", "256^3"};
  printf("{");
  for (int z = 0; z < 2; ++z) {
    const long n = sizes[z];
    // copy3 over 8 n floats per array (the same bytes as 2 + 2 of the 13-stream arrays)
    {
      const long m = 8 * n, n4 = m / 4;
      float *a, *b, *c, *o;
      CK(hipMalloc(&a, m * 4));
      CK(hipMalloc(&b, m * 4));
      CK(hipMalloc(&c, m * 4));
      CK(hipMalloc(&o, m * 4));
      CK(hipMemset(a, 0, m * 4));
      CK(hipMemset(b, 0, m * 4));
      CK(hipMemset(c, 0, m * 4));
      const unsigned nb = (unsigned)((n4 + 255) / 256);
      const float ms = time_ms(s, reps, [&]() {
        hipLaunchKernelGGL(k_copy3, dim3(nb), dim3(256), 0, s, (const float4*)a, (const float4*)b, (const float4*)c,
                           (float4*)o, n4);
      });
      printf("%s\"copy3_%s_GBps\": %.1f, \"copy3_%s_MB\": %.1f", z ? ", " : "", names[z], 16.0 * m / (1e-3 * ms) / 1e9,
             names[z], 16.0 * m / 1e6);
      CK(hipFree(a));
      CK(hipFree(b));
      CK(hipFree(c));
      CK(hipFree(o));
    }
    {
      float *a, *o;
      CK(hipMalloc(&a, 13 * n * 4));
      CK(hipMalloc(&o, n * 4));
      CK(hipMemset(a, 0, 13 * n * 4));
      const unsigned nb = (unsigned)((n + 255) / 256);
      const float ms = time_ms(s, reps, [&]() { hipLaunchKernelGGL(k_sum13, dim3(nb), dim3(256), 0, s, a, o, n); });
      printf(", \"sum13_%s_GBps\": %.1f, \"sum13_%s_MB\": %.1f", names[z], 14.0 * 4 * n / (1e-3 * ms) / 1e9, names[z],
             14.0 * 4 * n / 1e6);
      CK(hipFree(a));
      CK(hipFree(o));
    }
  }
  printf("}\n");
  CK(hipStreamDestroy(s));
  return 0;
}
