// Copy kernel of known size, the calibration point for HBM byte counters:
// a profile of FETCH_SIZE / WRITE_SIZE on this program reads N bytes fetched
// and N bytes written per dispatch, so the ratio of what the counters report
// to N converts counter readings of other kernels to bytes.
//
//   p2z_copy_calibrate [MiB=1024] [launches=5]
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
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "allocation of 2 x %zu MiB failed\n", mib);
    return 1;
  }
  if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  hipEvent_t t0, t1;
  hipEventCreate(&t0);
  hipEventCreate(&t1);
  const unsigned grid = 256 * 8 * 4;
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  hipEventRecord(t0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  hipEventRecord(t1);
  if (hipEventSynchronize(t1) != hipSuccess) return 2;
  float ms = 0;
  hipEventElapsedTime(&ms, t0, t1);
  std::printf("copy %zu MiB x %d: %.4f ms per launch, %.1f GB/s (read + write)\n", mib, reps, ms / reps,
              2.0 * bytes / (ms / reps * 1e-3) / 1e9);
  hipFree(a);
  hipFree(b);
  return 0;
}
