// Copy kernel of known size, the calibration point for HBM byte counters:
// a profile of FETCH_SIZE / WRITE_SIZE on this program reads N bytes fetched
// and N bytes written per dispatch, so the ratio of what the counters report
// to N converts counter readings of other kernels to bytes.
//
//   p2z_copy_calibrate [MiB=1024] [launches=5] [workgroups=8192; 0 = one element per thread]
//
// Without a profiler it also measures the achievable copy bandwidth.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  const long wg = argc > 3 ? std::atol(argv[3]) : 8192;
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
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
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
