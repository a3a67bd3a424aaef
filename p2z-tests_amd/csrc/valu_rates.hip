// Issue rates of the VALU instructions the layer step is made of, the
// calibration point for the VALU side of the roofline: what a kernel that
// does nothing but one instruction sustains on this GPU, against the
// 4-cycles-per-wave64-instruction peak at the reported maximum engine clock.
//
//   p2z_valu_rates [workgroups per CU = 8] [iterations = 4096]
//
// Each thread runs `iterations` rounds of 8 independent chains of one
// operation (8 workgroups of 256 threads per CU: 8 waves per SIMD), and
// stores one value so that nothing is optimised away.  Each operation is
// timed over 20 launches (hipEvents) after 3 untimed ones, all after about
// 0.2 s of warm-up.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

enum Op {
  kFmaF32, kMulF32, kRcpF32, kFmaF64, kMulF64, kCvtMulCvt, kRcpF64, kAddF32, kShlXor, kCmpSel, kRotU32,
  kFmaF32FmaF64, kFmaF32RcpF32, kFmaF32ShlXor, kFmaF32RotU32,
  kFmaF32V, kMulF32V, kMulF32Lit, kFmaakF32, kFmaF64V, kMulF64V, kFmaF32VFmaF64V, kNumOps
};
// Operations of one chain step.  The plain FP32 ones take their loop-invariant
// operands from scalar registers (a, b); the "VGPR operands only" ones from
// per-lane values (the compiler keeps the FP64 ones' constants in VGPRs
// either way).  "x | y" interleaves two independent chains of different
// instruction kinds, to show whether they share one issue slot (cost = sum of
// the two) or overlap.
const char* const kNames[kNumOps] = {"v_fma_f32", "v_mul_f32", "v_rcp_f32", "v_fma_f64", "v_mul_f64",
                                     "cvt_f64_f32 + v_mul_f64 + cvt_f32_f64", "v_rcp_f64", "v_add_f32",
                                     "v_lshlrev_b32 + v_xor_b32", "v_add_f32 + v_cmp_gt_f32 + v_cndmask_b32",
                                     "v_alignbit_b32", "v_fma_f32 | v_fma_f64", "v_fma_f32 | v_rcp_f32",
                                     "v_fma_f32 | v_lshlrev_b32 + v_xor_b32", "v_fma_f32 | v_alignbit_b32",
                                     "v_fma_f32, VGPR operands only", "v_mul_f32, VGPR operands only",
                                     "v_mul_f32, literal operand", "v_fmaak_f32 (VGPRs + literal)",
                                     "v_fma_f64, VGPR operands only", "v_mul_f64, VGPR operands only",
                                     "v_fma_f32 | v_fma_f64, VGPR operands only"};

template <int kOp>
__global__ __launch_bounds__(256) void k_rate(float* out, float a, float b, int iters) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  float x[8], y[8];
  double d[8];
  // per-lane values near a and b that the compiler cannot fold into scalars
  const float va = a + (float)(tid & 1) * 1e-7f, vb = b + (float)(tid & 2) * 1e-9f;
  const double vda = va, vdb = vb;
  uint32_t u[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    x[i] = a + tid * 1e-7f + i;
    y[i] = x[i] + 0.5f;
    d[i] = x[i];
    u[i] = (uint32_t)tid * 2654435761u + i;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if constexpr (kOp == kFmaF32) x[i] = __builtin_fmaf(x[i], a, b);
      if constexpr (kOp == kMulF32) x[i] = x[i] * a;
      if constexpr (kOp == kRcpF32) x[i] = __builtin_amdgcn_rcpf(x[i]);
      if constexpr (kOp == kFmaF64) d[i] = __builtin_fma(d[i], (double)a, (double)b);
      if constexpr (kOp == kMulF64) d[i] = d[i] * (double)a;
      if constexpr (kOp == kCvtMulCvt) {
        d[i] = d[i] * (double)x[i];
        x[i] = (float)d[i];
      }
      if constexpr (kOp == kRcpF64) d[i] = __builtin_amdgcn_rcp(d[i]);
      if constexpr (kOp == kAddF32) x[i] = x[i] + a;
      if constexpr (kOp == kShlXor) u[i] ^= u[i] << 1;
      if constexpr (kOp == kCmpSel) x[i] = x[i] > b ? a : x[i] + b;
      if constexpr (kOp == kRotU32) u[i] = (u[i] << 1) | (u[i] >> 31);
      if constexpr (kOp == kFmaF32FmaF64) {
        x[i] = __builtin_fmaf(x[i], a, b);
        d[i] = __builtin_fma(d[i], (double)a, (double)b);
      }
      if constexpr (kOp == kFmaF32RcpF32) {
        x[i] = __builtin_fmaf(x[i], a, b);
        y[i] = __builtin_amdgcn_rcpf(y[i]);
      }
      if constexpr (kOp == kFmaF32ShlXor) {
        x[i] = __builtin_fmaf(x[i], a, b);
        u[i] ^= u[i] << 1;
      }
      if constexpr (kOp == kFmaF32V) x[i] = __builtin_fmaf(x[i], va, vb);
      if constexpr (kOp == kMulF32V) x[i] = x[i] * va;
      if constexpr (kOp == kMulF32Lit) x[i] = x[i] * 0.99951171875f;
      if constexpr (kOp == kFmaakF32) x[i] = __builtin_fmaf(x[i], va, 0.0009765625f);
      if constexpr (kOp == kFmaF64V) d[i] = __builtin_fma(d[i], vda, vdb);
      if constexpr (kOp == kMulF64V) d[i] = d[i] * vda;
      if constexpr (kOp == kFmaF32VFmaF64V) {
        x[i] = __builtin_fmaf(x[i], va, vb);
        d[i] = __builtin_fma(d[i], vda, vdb);
      }
      if constexpr (kOp == kFmaF32RotU32) {
        x[i] = __builtin_fmaf(x[i], a, b);
        u[i] = (u[i] << 1) | (u[i] >> 31);
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += x[i] + y[i] + (float)d[i] + (float)u[i];
  out[tid] = s;
}

template <int kOp>
int time_op(float* out, unsigned grid, int iters, float* ms) {
  hipEvent_t t0, t1;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return 2;
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_rate<kOp>, dim3(grid), dim3(256), 0, 0, out, 0.999f, 1e-3f, iters);
  if (hipEventRecord(t0) != hipSuccess) return 2;
  for (int r = 0; r < 20; ++r) hipLaunchKernelGGL(k_rate<kOp>, dim3(grid), dim3(256), 0, 0, out, 0.999f, 1e-3f, iters);
  if (hipEventRecord(t1) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) return 2;
  if (hipEventElapsedTime(ms, t0, t1) != hipSuccess) return 2;
  *ms /= 20;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // namespace

int main(int argc, char** argv) {
  const int wg_per_cu = argc > 1 ? std::atoi(argv[1]) : 8;
  const int iters = argc > 2 ? std::atoi(argv[2]) : 4096;
  if (wg_per_cu < 1 || wg_per_cu > 64 || iters < 1) {
    std::fprintf(stderr, "usage: p2z_valu_rates [workgroups per CU 1..64] [iterations >= 1]\n");
    return 1;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
  const unsigned grid = (unsigned)(prop.multiProcessorCount * wg_per_cu);
  float* out = nullptr;
  if (hipMalloc(&out, (size_t)grid * 256 * sizeof(float)) != hipSuccess) return 1;
  // one wave64 instruction per SIMD every 4 cycles, 4 SIMDs per CU
  const double peak = prop.multiProcessorCount * 4.0 * (prop.clockRate * 1e3) / 4.0;
  const double wave_instr = (double)grid * 4 * iters * 8;  // of the measured operation, per launch
  float ms[kNumOps];
  // the engine clock ramps up under load: about 0.2 s of FMA work first
  int rc = time_op<kFmaF32>(out, grid, iters, &ms[kFmaF32]);
  for (int r = 0; r < 8 && rc == 0; ++r) rc |= time_op<kFmaF32>(out, grid, iters, &ms[kFmaF32]);
  rc |= time_op<kFmaF32>(out, grid, iters, &ms[kFmaF32]);
  rc |= time_op<kMulF32>(out, grid, iters, &ms[kMulF32]);
  rc |= time_op<kRcpF32>(out, grid, iters, &ms[kRcpF32]);
  rc |= time_op<kFmaF64>(out, grid, iters, &ms[kFmaF64]);
  rc |= time_op<kMulF64>(out, grid, iters, &ms[kMulF64]);
  rc |= time_op<kCvtMulCvt>(out, grid, iters, &ms[kCvtMulCvt]);
  rc |= time_op<kRcpF64>(out, grid, iters, &ms[kRcpF64]);
  rc |= time_op<kAddF32>(out, grid, iters, &ms[kAddF32]);
  rc |= time_op<kShlXor>(out, grid, iters, &ms[kShlXor]);
  rc |= time_op<kCmpSel>(out, grid, iters, &ms[kCmpSel]);
  rc |= time_op<kRotU32>(out, grid, iters, &ms[kRotU32]);
  rc |= time_op<kFmaF32FmaF64>(out, grid, iters, &ms[kFmaF32FmaF64]);
  rc |= time_op<kFmaF32RcpF32>(out, grid, iters, &ms[kFmaF32RcpF32]);
  rc |= time_op<kFmaF32ShlXor>(out, grid, iters, &ms[kFmaF32ShlXor]);
  rc |= time_op<kFmaF32RotU32>(out, grid, iters, &ms[kFmaF32RotU32]);
  rc |= time_op<kFmaF32V>(out, grid, iters, &ms[kFmaF32V]);
  rc |= time_op<kMulF32V>(out, grid, iters, &ms[kMulF32V]);
  rc |= time_op<kMulF32Lit>(out, grid, iters, &ms[kMulF32Lit]);
  rc |= time_op<kFmaakF32>(out, grid, iters, &ms[kFmaakF32]);
  rc |= time_op<kFmaF64V>(out, grid, iters, &ms[kFmaF64V]);
  rc |= time_op<kMulF64V>(out, grid, iters, &ms[kMulF64V]);
  rc |= time_op<kFmaF32VFmaF64V>(out, grid, iters, &ms[kFmaF32VFmaF64V]);
  if (rc) {
    std::fprintf(stderr, "launch failed\n");
    return 2;
  }
  std::printf("%s, %d CUs, max engine clock %d MHz: peak %.3e wave64 instructions/s (4 cycles each per SIMD)\n",
              prop.gcnArchName, prop.multiProcessorCount, prop.clockRate / 1000, peak);
  std::printf("%d workgroups of 256 threads (%d per CU), %d rounds of 8 independent chains per thread\n", grid,
              wg_per_cu, iters);
  for (int i = 0; i < kNumOps; ++i) {
    const double rate = wave_instr / (ms[i] * 1e-3);
    std::printf("  %-40s %8.4f ms  %.3e chain steps/s  %5.1f%% of peak  cost %.2f x v_fma_f32\n", kNames[i], ms[i],
                rate, 100.0 * rate / peak, ms[i] / ms[kFmaF32]);
  }
  (void)hipFree(out);
  return 0;
}
