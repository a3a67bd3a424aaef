// Copy kernel of known size, the calibration point for HBM byte counters:
// a profile of FETCH_SIZE / WRITE_SIZE on this program reads N bytes fetched
// and N bytes written per dispatch, so the ratio of what the counters report
// to N converts counter readings of other kernels to bytes.
//
//   p2z_copy_calibrate [MiB=1024] [launches=5] [workgroups=8192; 0 = one element per thread] [reads=1]
//
// Without a profiler it also measures the achievable copy bandwidth.  With
// reads = k > 1 it runs a read-mostly stream instead, one float4 per thread:
// each thread sums k float4 from k slices of MiB / k MiB and writes one, so
// the kernel reads MiB and writes MiB / k (the fused pass reads 5.3 bytes
// per byte it writes).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// The GPU's clocks take tens of milliseconds of sustained load to rise after
// an idle spell: about 0.2 s of untimed launches (at about 6 TB/s) come first.
int warmup_launches(double bytes_per_launch) {
  const double n = 0.2 * 6e12 / bytes_per_launch;
  return n < 3 ? 3 : n > 100000 ? 100000 : (int)n;
}

__global__ void k_stream(const float4* __restrict__ a, float4* __restrict__ b, size_t nout, int k) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout) return;
  float4 s = a[i];
  for (int j = 1; j < k; ++j) {
    const float4 v = a[(size_t)j * nout + i];
    s.x += v.x;
    s.y += v.y;
    s.z += v.z;
    s.w += v.w;
  }
  b[i] = s;
}

int stream(size_t bytes, int reps, int k) {
  const size_t nout = bytes / sizeof(float4) / (size_t)k;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, nout * k * sizeof(float4)) != hipSuccess || hipMalloc(&b, nout * sizeof(float4)) != hipSuccess) {
    std::fprintf(stderr, "allocation failed\n");
    return 1;
  }
  if (hipMemset(a, 0, nout * k * sizeof(float4)) != hipSuccess) return 1;
  hipEvent_t t0, t1;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return 1;
  const unsigned grid = (unsigned)((nout + 255) / 256);
  for (int r = 0; r < warmup_launches((double)nout * (k + 1) * sizeof(float4)); ++r)
    hipLaunchKernelGGL(k_stream, dim3(grid), dim3(256), 0, 0, a, b, nout, k);
  if (hipEventRecord(t0) != hipSuccess) return 2;
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream, dim3(grid), dim3(256), 0, 0, a, b, nout, k);
  if (hipEventRecord(t1) != hipSuccess) return 2;
  if (hipEventSynchronize(t1) != hipSuccess) return 2;
  float ms = 0;
  if (hipEventElapsedTime(&ms, t0, t1) != hipSuccess) return 2;
  const double rd = (double)nout * k * sizeof(float4), wr = (double)nout * sizeof(float4);
  std::printf("stream: read %.1f MiB, write %.1f MiB (%d : 1), %u workgroups: %.4f ms per launch, %.1f GB/s (read + write)\n",
              rd / 1048576, wr / 1048576, k, grid, ms / reps, (rd + wr) / (ms / reps * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  const long wg = argc > 3 ? std::atol(argv[3]) : 8192;
  const int k = argc > 4 ? std::atoi(argv[4]) : 1;
  if (k < 1 || k > 1024 || reps < 1) {
    std::fprintf(stderr, "usage: p2z_copy_calibrate [MiB] [launches >= 1] [workgroups] [reads 1..1024]\n");
    return 1;
  }
  if (k > 1) return stream(mib << 20, reps, k);
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "allocation of 2 x %zu MiB failed\n", mib);
    return 1;
  }
  if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  hipEvent_t t0, t1;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return 1;
  const unsigned grid = wg > 0 ? (unsigned)wg : (unsigned)((n + 255) / 256);
  for (int r = 0; r < warmup_launches(2.0 * bytes); ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  if (hipEventRecord(t0) != hipSuccess) return 2;
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  if (hipEventRecord(t1) != hipSuccess) return 2;
  if (hipEventSynchronize(t1) != hipSuccess) return 2;
  float ms = 0;
  if (hipEventElapsedTime(&ms, t0, t1) != hipSuccess) return 2;
  std::printf("copy %zu MiB x %d, %u workgroups: %.4f ms per launch, %.1f GB/s (read + write)\n", mib, reps, grid,
              ms / reps, 2.0 * bytes / (ms / reps * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
