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
// SIMD at depth 1 (89 VGPRs) against 4 at depth 2 (100) and measures 1-3%
// faster; the reference mode (92 / 100 VGPRs: 5 waves against 4) 1-3% faster
// at depth 1 on three boxes; the FP64-gain mode (126 / 138 VGPRs: 4 waves
// against 3) within 1% either way, and keeps depth 2.
inline int default_hit_stage(int math) { return math == 2 ? 2 : 1; }

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

// Work-queue counters of the persistent ragged pass: kQueueStripes counters,
// kQueueStride 32-bit words (128 bytes) apart, kQueueBytes in all.
constexpr int kQueueStripes = 8;
constexpr int kQueueStride = 32;
constexpr size_t kQueueBytes = (size_t)kQueueStripes * kQueueStride * sizeof(unsigned);

// Ragged pass.  persistent: grid workgroups drain the descriptor table through
// the device counters (reset first on the same stream; kQueueBytes), claim
// items per atomic; otherwise one wave per 64 track slots.
hipError_t launch_ragged(const float* trk, const float* hit, float* out, const int32_t* lane_nlay,
                         const BlockDesc* blocks, int64_t nblocks, int width, int math, bool persistent,
                         int prefetch, int tpw, int64_t grid, int claim, unsigned* counter, hipStream_t s);
constexpr int kDefaultClaim = 4;
// Workgroups of tpw threads of the persistent ragged kernel (width, math,
// prefetch) that one CU holds at once (0 if the runtime cannot tell).  A
// persistent grid larger than that many per CU leaves waves that start only
// when resident ones have ended, and their first chunks, handed out without
// a claim, are then the longest items computed last.
int ragged_persistent_groups_per_cu(int width, int math, int prefetch, int tpw);

// Report statistics (OMP:758-851) on device.  part: scratch of
// summary_scratch_doubles(nevts, ntrks) doubles.  res: 18 doubles on device
// (9 means, 3 counts, 6 RMS values), written by the final passes.
size_t summary_scratch_doubles(int64_t nevts, int64_t ntrks);
hipError_t launch_summary(const float* out, int64_t nevts, int64_t nb, int width, int64_t ntrks,
                          double hx, double hy, double hz, double* part, double* res,
                          hipStream_t s);

// ---- device-resident serving loop -------------------------------------------
// Work item of the resident loop: 64 track slots of one request.  W <= 64:
// up to 64/W descriptors (slot s of the item is lane s % W of desc[s / W]);
// W = 128: desc[0] only, lanes half*64 .. half*64+63.  Descriptors past
// ndesc are padding (their slots compute nothing).
struct ServeItem {
  BlockDesc desc[4];
  int32_t ndesc;
  int32_t half;
  int32_t req_slot;   // completion slot of the request
  int32_t req_items;  // items of the request
};
static_assert(sizeof(ServeItem) == 144, "serve item layout");

// Completion record, written by the device into host memory: seq (= its
// index + 1) is written last, after the record, so the host reads a record
// only once seq says it is complete.
struct ServeCompletion {
  int32_t req_slot;
  int32_t pad;
  uint64_t seq;
};

// Control words in host memory (coherent pinned), one per 64-byte line.
struct ServeCtl {
  uint64_t tail;  // items published by the host (item k lives at k % cap)
  uint64_t pad0[7];
  uint64_t heartbeat;  // advanced by the host while it serves
  uint64_t pad1[7];
  uint32_t stop;    // 1: end the loop once the published items are done
  uint32_t epoch;   // launch number, set by the host before each launch
  uint32_t exited;  // the dispatcher copies epoch here when it leaves
  uint32_t pad2[13];
};

// Item claims are striped over kServeStripes counters, one per 128-byte line:
// workgroup g claims the items k with k % kServeStripes == g % kServeStripes,
// so that no single counter takes every claim of the device.
constexpr int kServeStripes = 8;

// Device-side state of the loop (device memory, zeroed before the first launch).
struct ServeDev {
  uint64_t head[kServeStripes][16];  // per stripe: items of the stripe claimed
  uint64_t dev_tail;   // copy of ServeCtl::tail kept by the dispatcher wave
  uint64_t pad1[7];
  uint32_t dev_stop;   // set by the dispatcher: workers end once nothing is left
  uint32_t pad2[15];
  uint64_t comp_tail;  // completion records written
  uint64_t pad3[7];
  // followed by the per-request item counters (uint32_t done[req_slots])
};

struct ServeLaunch {
  const float* trk;
  const float* hit;
  float* out;
  const int32_t* lanes;
  const ServeItem* items;    // host ring (device-visible pointer)
  int64_t item_cap;
  ServeCtl* ctl;             // host control words (device-visible pointer)
  ServeCompletion* comps;    // host completion ring (device-visible pointer)
  int64_t comp_cap;
  ServeDev* dev;             // device state
  uint32_t* done;            // per request slot: items finished (device)
  uint64_t idle_ticks;       // heartbeat timeout, in s_memrealtime ticks (100 MHz)
};

// Prepares a (re)launch on stream s: every stripe's claim counter is set to
// its first item at or past dev_tail (items below it are done), dev_stop to 0.
hipError_t launch_serve_reset(ServeDev* dev, hipStream_t s);

// One persistent launch of the resident loop (grid workgroups of tpw
// threads; wave 0 of workgroup 0 is the dispatcher; at least
// 2 * kServeStripes workgroups, so that every stripe has workers).  Returns
// when launched; the kernel runs until ServeCtl::stop or the heartbeat
// timeout.
hipError_t launch_serve_resident(const ServeLaunch& a, int width, int math, int prefetch, int tpw, int64_t grid,
                                 hipStream_t s);

// sin/cos of the reference-typing mode over consecutive float bit patterns.
hipError_t launch_sincosf_range(uint32_t first, int64_t n, float* s, float* c, hipStream_t st);

}  // namespace p2z
