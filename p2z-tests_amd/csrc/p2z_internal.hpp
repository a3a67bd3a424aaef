// Internal interface between the C-ABI layer (p2z_capi.hip) and the kernel
// translation unit (p2z_kernels.hip).  Not installed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p2z {

// Words per AoSoA block of width W (reference MPTRK / MPHIT, OMP:115-125).
constexpr int kTrkWords = 28;
constexpr int kHitWords = 9;
// Leading words of a hit record the path reads: pos x, y, z and cov xx, xy,
// yy (KalmanUpdate_v2 uses the 2x2 block only, OMP:510-512).
constexpr int kHitUsedWords = 6;

inline bool valid_width(int w) { return w == 16 || w == 32 || w == 64 || w == 128; }
inline bool valid_math(int m) { return m >= 0 && m <= 2; }

struct LaunchCfg {
  int tracks_per_wg = 256;  // threads per workgroup, one track per thread
  int hit_stage = 0;        // layers of hits in flight ahead of the computed one (1 or 2);
                            // 0: per arithmetic mode (default_hit_stage)
};

// Hit-staging depth of the fused pass when none is set: FAST runs 5 waves per
// SIMD at depth 1 (94 VGPRs) against 4 at depth 2 (103) and measures about 1%
// faster; the FP64-gain and reference modes run at 3-4 waves either way and
// keep depth 2.
inline int default_hit_stage(int math) { return math == 0 ? 1 : 2; }

// Fused driver pass: ntracks = nblocks*W tracks, all with nlayer layers.
hipError_t launch_fused(const float* trk, const float* hit, float* out, int64_t nblocks,
                        int nlayer, int width, int math, const LaunchCfg& cfg, hipStream_t s);

hipError_t launch_propagate(const float* inErr, const float* inPar, const int32_t* inChg,
                            const float* msP, float* outErr, float* outPar, int64_t nblocks,
                            int64_t ts, int64_t hs, int width, int math, const LaunchCfg& cfg,
                            hipStream_t s);

hipError_t launch_update(float* trkErr, float* inPar, const float* hitErr, const float* msP,
                         int64_t nblocks, int64_t ts, int64_t hs, int width, int math,
                         const LaunchCfg& cfg, hipStream_t s);

// KalmanUpdate v1 on blocks (fixed: the textbook form).
hipError_t launch_update_v1(float* trkErr, float* inPar, const float* hitErr, const float* msP, int64_t nblocks,
                            int64_t ts, int64_t hs, int width, bool fixed, const LaunchCfg& cfg, hipStream_t s);

// Block descriptor of the ragged format (same layout as p2z_block_desc).
struct BlockDesc {
  int64_t trk_block;
  int64_t hit_block;
  int64_t out_block;
  int32_t nlayer;
  int32_t event;
};
static_assert(sizeof(BlockDesc) == 32, "descriptor layout");

// Ragged pass.  persistent: grid workgroups drain the descriptor table through
// the device counter (reset first on the same stream), claim items per atomic;
// otherwise one wave per 64 track slots.
hipError_t launch_ragged(const float* trk, const float* hit, float* out, const int32_t* lane_nlay,
                         const BlockDesc* blocks, int64_t nblocks, int width, int math, bool persistent,
                         int prefetch, int tpw, int64_t grid, int claim, unsigned* counter, hipStream_t s);
constexpr int kDefaultClaim = 4;

// Report statistics (OMP:758-851) on device.  part: scratch of
// summary_scratch_doubles(nevts, ntrks) doubles.  res: 18 doubles on device
// (9 means, 3 counts, 6 RMS values), written by the final passes.
size_t summary_scratch_doubles(int64_t nevts, int64_t ntrks);
hipError_t launch_summary(const float* out, int64_t nevts, int64_t nb, int width, int64_t ntrks,
                          double hx, double hy, double hz, double* part, double* res,
                          hipStream_t s);

// sin/cos of the reference-typing mode over consecutive float bit patterns.
hipError_t launch_sincosf_range(uint32_t first, int64_t n, float* s, float* c, hipStream_t st);

}  // namespace p2z
