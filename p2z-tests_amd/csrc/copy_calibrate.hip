// Copy kernel of known size, the calibration point for HBM byte counters:
// a profile of FETCH_SIZE / WRITE_SIZE on this program reads N bytes fetched
// and N bytes written per dispatch, so the ratio of what the counters report
// to N converts counter readings of other kernels to bytes.
//
//   p2z_copy_calibrate [MiB=1024] [launches=5] [workgroups=8192; 0 = one element per thread] [reads=1]
//                      [outputs per thread=1] [non-temporal loads=0] [non-temporal stores=0] [buffer sets=1]
//
// Without a profiler it also measures the achievable copy bandwidth.  With
// reads = k > 1 it runs a read-mostly stream instead: each output float4 is
// the sum of k float4 from k slices of MiB / k MiB, so the kernel reads MiB
// and writes MiB / k (the fused pass reads 5.3 bytes per byte it writes);
// a thread computes 1, 2 or 4 outputs, with plain or non-temporal loads and
// stores, over 1 to 8 buffer sets used in turn.
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

// Each thread handles kM outputs (kM * grid threads apart), all kM * k loads
// issued before the sums; kNtLoad / kNtStore: non-temporal loads / stores (as
// the passes' hit loads and track stores).
template <int kM, bool kNtLoad, bool kNtStore>
__global__ void k_stream(const float4* __restrict__ a, float4* __restrict__ b, size_t nout, int k) {
  typedef float native4 __attribute__((ext_vector_type(4)));
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nthr = (size_t)gridDim.x * blockDim.x;
  native4 s[kM];
#pragma unroll
  for (int m = 0; m < kM; ++m) s[m] = native4{0.f, 0.f, 0.f, 0.f};
  for (int j = 0; j < k; ++j) {
#pragma unroll
    for (int m = 0; m < kM; ++m) {
      const size_t i = i0 + m * nthr;
      if (i < nout) {
        const native4* src = reinterpret_cast<const native4*>(a + (size_t)j * nout + i);
        s[m] += kNtLoad ? __builtin_nontemporal_load(src) : *src;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < kM; ++m) {
    const size_t i = i0 + m * nthr;
    if (i < nout) {
      native4* dst = reinterpret_cast<native4*>(b + i);
      if (kNtStore)
        __builtin_nontemporal_store(s[m], dst);
      else
        *dst = s[m];
    }
  }
}

// nsets buffer pairs used in turn, so that with nsets * bytes above the
// 256 MB last-level cache nothing is reused from it across launches.
template <int kM, bool kNtLoad, bool kNtStore>
int stream(size_t bytes, int reps, int k, int nsets) {
  const size_t nout = bytes / sizeof(float4) / (size_t)k;
  float4 *a[8] = {}, *b[8] = {};
  for (int q = 0; q < nsets; ++q) {
    if (hipMalloc(&a[q], nout * k * sizeof(float4)) != hipSuccess ||
        hipMalloc(&b[q], nout * sizeof(float4)) != hipSuccess) {
      std::fprintf(stderr, "allocation failed\n");
      return 1;
    }
    if (hipMemset(a[q], 0, nout * k * sizeof(float4)) != hipSuccess) return 1;
  }
  hipEvent_t t0, t1;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return 1;
  const unsigned grid = (unsigned)((nout + 256 * kM - 1) / (256 * kM));
  auto launch = [&](int r) {
    hipLaunchKernelGGL((k_stream<kM, kNtLoad, kNtStore>), dim3(grid), dim3(256), 0, 0, a[r % nsets], b[r % nsets],
                       nout, k);
  };
  for (int r = 0; r < warmup_launches((double)nout * (k + 1) * sizeof(float4)); ++r) launch(r);
  if (hipEventRecord(t0) != hipSuccess) return 2;
  for (int r = 0; r < reps; ++r) launch(r);
  if (hipEventRecord(t1) != hipSuccess) return 2;
  if (hipEventSynchronize(t1) != hipSuccess) return 2;
  float ms = 0;
  if (hipEventElapsedTime(&ms, t0, t1) != hipSuccess) return 2;
  const double rd = (double)nout * k * sizeof(float4), wr = (double)nout * sizeof(float4);
  std::printf("stream: read %.1f MiB, write %.1f MiB (%d : 1), %d output(s) per thread, %s loads, %s stores, "
              "%d buffer set(s): %.4f ms per launch, %.1f GB/s (read + write)\n",
              rd / 1048576, wr / 1048576, k, kM, kNtLoad ? "non-temporal" : "plain", kNtStore ? "non-temporal" : "plain",
              nsets, ms / reps, (rd + wr) / (ms / reps * 1e-3) / 1e9);
  for (int q = 0; q < nsets; ++q) {
    (void)hipFree(a[q]);
    (void)hipFree(b[q]);
  }
  return 0;
}

template <int kM>
int stream_m(size_t bytes, int reps, int k, bool ntl, bool nts, int nsets) {
  if (ntl) return nts ? stream<kM, true, true>(bytes, reps, k, nsets) : stream<kM, true, false>(bytes, reps, k, nsets);
  return nts ? stream<kM, false, true>(bytes, reps, k, nsets) : stream<kM, false, false>(bytes, reps, k, nsets);
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  const long wg = argc > 3 ? std::atol(argv[3]) : 8192;
  const int k = argc > 4 ? std::atoi(argv[4]) : 1;
  const int per_thread = argc > 5 ? std::atoi(argv[5]) : 1;
  const bool ntl = argc > 6 && std::atoi(argv[6]) != 0;
  const bool nts = argc > 7 && std::atoi(argv[7]) != 0;
  const int nsets = argc > 8 ? std::atoi(argv[8]) : 1;
  if (k < 1 || k > 1024 || reps < 1 || (per_thread != 1 && per_thread != 2 && per_thread != 4) || nsets < 1 ||
      nsets > 8) {
    std::fprintf(stderr,
                 "usage: p2z_copy_calibrate [MiB] [launches >= 1] [workgroups] [reads 1..1024] [outputs per thread "
                 "1|2|4] [non-temporal loads 0|1] [non-temporal stores 0|1] [buffer sets 1..8]\n");
    return 1;
  }
  if (k > 1) {
    const size_t bytes = mib << 20;
    if (per_thread == 1) return stream_m<1>(bytes, reps, k, ntl, nts, nsets);
    if (per_thread == 2) return stream_m<2>(bytes, reps, k, ntl, nts, nsets);
    return stream_m<4>(bytes, reps, k, ntl, nts, nsets);
  }
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
