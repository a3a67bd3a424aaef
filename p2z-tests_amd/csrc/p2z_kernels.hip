// gfx950 kernels of the propagate-to-z + Kalman-update path.
//
// Mapping: one track per lane (wave64), track g = blockIdx.x*blockDim.x +
// threadIdx.x lives in AoSoA block b = g / W at lane l = g % W.  Every load and
// store of a wave touches 64/W contiguous runs of 4W bytes (element-major
// AoSoA, reference OMP:159-161), so track and hit traffic is fully coalesced
// without any transpose.  The track state stays in VGPRs for all layers; the
// hits of layer L+1 are loaded before layer L is computed.
#include "p2z_internal.hpp"
#include "p2z_lane.hpp"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>

namespace p2z {
namespace {

// kStream: non-temporal loads, for the ragged and serving passes, whose track
// input (706 MB for 2000 events) is larger than the 256 MB last-level cache
// and is not read again by the next pass (about 0.8% faster there).  The fused
// pass loads its tracks through the caches: run again on the same buffers it
// takes part of them from the last-level cache (0.115 against 0.121 ms).
template <int W, bool kStream = false>
__device__ inline void load_track(Track& t, const float* __restrict__ tb) {
  auto ld = [&](const float* p) { return kStream ? __builtin_nontemporal_load(p) : *p; };
#pragma unroll
  for (int e = 0; e < 6; ++e) t.p[e] = ld(tb + e * W);
#pragma unroll
  for (int e = 0; e < 21; ++e) t.P[e] = ld(tb + (6 + e) * W);
  const int32_t* q = reinterpret_cast<const int32_t*>(tb) + 27 * W;
  t.q = kStream ? __builtin_nontemporal_load(q) : *q;
}

// The output is written once and not read by the pass: non-temporal stores
// keep it from displacing what the caches hold.  Measured on the 100-event
// pass cycling through three buffer sets (nothing reused across passes from
// the 256 MB last-level cache): 0.131 -> 0.123 ms; on the same buffers every
// pass, where the track input is partly served from that cache, unchanged;
// the 2000-event ragged pass 1-1.5% faster.  Non-temporal track loads as well
// gave the same cold-cache time but lost that reuse (0.121 ms on the same
// buffers).
template <int W>
__device__ inline void store_track(const Track& t, float* __restrict__ ob) {
#pragma unroll
  for (int e = 0; e < 6; ++e) __builtin_nontemporal_store(t.p[e], ob + e * W);
#pragma unroll
  for (int e = 0; e < 21; ++e) __builtin_nontemporal_store(t.P[e], ob + (6 + e) * W);
  __builtin_nontemporal_store(t.q, reinterpret_cast<int32_t*>(ob) + 27 * W);
}

// Hits are read exactly once: non-temporal loads keep them out of the way of
// the track state in the caches.
template <int W>
__device__ inline Hit load_hit(const float* __restrict__ hb) {
  Hit h;
  h.x = __builtin_nontemporal_load(hb + 0 * W);
  h.y = __builtin_nontemporal_load(hb + 1 * W);
  h.z = __builtin_nontemporal_load(hb + 2 * W);
  h.xx = __builtin_nontemporal_load(hb + 3 * W);
  h.xy = __builtin_nontemporal_load(hb + 4 * W);
  h.yy = __builtin_nontemporal_load(hb + 5 * W);
  return h;
}

// Layers [0, n) of one track whose hits start at hb (consecutive layers 9W
// words apart), with the hits of the next kDepth layers (1 or 2) in flight
// while a layer is computed.  The loop is unrolled by kDepth + 1 over
// rotating hit buffers, so a prefetched hit is never copied between
// registers.  The loop runs to nmax, the same for every lane of the wave
// (the fused pass: n == nmax); with kPerLane, lane layer counts n <= nmax
// differ and each lane's loads and steps stop at its own n.  Every load index
// is < n.
template <int W, int kMath, int kDepth, bool kPerLane>
__device__ __forceinline__ void run_layers(Track& t, const float* __restrict__ hb, int n, int nmax) {
  static_assert(kDepth == 1 || kDepth == 2, "prefetch depth 1 or 2");
  if (nmax <= 0) return;
  auto at = [&](int lay) { return load_hit<W>(hb + (int64_t)lay * (kHitWords * W)); };
  auto live = [&](int lay) { return kPerLane ? lay < n : true; };
  Trig g;
  Hit a{}, b{};
  if (live(0)) a = at(0);
  if (1 < nmax && live(1)) b = at(1);
  if constexpr (kDepth == 1) {
    if (live(0)) layer_step<kMath, true>(t, a, g);
    int lay = 1;
    for (; lay + 1 < nmax; lay += 2) {  // b holds layer lay
      if (live(lay + 1)) a = at(lay + 1);
      if (live(lay)) layer_step<kMath, false>(t, b, g);
      if (lay + 2 < nmax && live(lay + 2)) b = at(lay + 2);
      if (live(lay + 1)) layer_step<kMath, false>(t, a, g);
    }
    if (lay < nmax && live(lay)) layer_step<kMath, false>(t, b, g);
  } else {
    Hit c{};
    if (2 < nmax && live(2)) c = at(2);
    if (live(0)) layer_step<kMath, true>(t, a, g);
    int lay = 1;
    for (; lay + 2 < nmax; lay += 3) {  // b holds layer lay, c layer lay + 1
      if (live(lay + 2)) a = at(lay + 2);
      if (live(lay)) layer_step<kMath, false>(t, b, g);
      if (lay + 3 < nmax && live(lay + 3)) b = at(lay + 3);
      if (live(lay + 1)) layer_step<kMath, false>(t, c, g);
      if (lay + 4 < nmax && live(lay + 4)) c = at(lay + 4);
      if (live(lay + 2)) layer_step<kMath, false>(t, a, g);
    }
    if (lay < nmax && live(lay)) layer_step<kMath, false>(t, b, g);
    if (lay + 1 < nmax && live(lay + 1)) layer_step<kMath, false>(t, c, g);
  }
}

// Ragged form of run_layers: lane layer counts n <= nmax differ, the loop
// runs to nmax (the same for every lane of the wave) with one layer step per
// iteration, each lane's loads and steps masked by its own n.  The hits of the
// next kDepth layers are in flight while a layer is computed; a single step
// instance keeps the register count down where the steps are masked.
template <int W, int kMath, int kDepth>
__device__ __forceinline__ void run_layers_lanes(Track& t, const float* __restrict__ hb, int n, int nmax) {
  static_assert(kDepth == 1 || kDepth == 2, "prefetch depth 1 or 2");
  if (nmax <= 0) return;
  auto at = [&](int lay) { return load_hit<W>(hb + (int64_t)lay * (kHitWords * W)); };
  Trig g;
  Hit h0{}, h1{}, h2{};
  if (0 < n) h0 = at(0);
  if (1 < n) h1 = at(1);
  if (kDepth >= 2 && 2 < n) h2 = at(2);
  if (0 < n) layer_step<kMath, true>(t, h0, g);
  for (int lay = 1; lay < nmax; ++lay) {
    h0 = h1;
    if constexpr (kDepth >= 2) {
      h1 = h2;
      if (lay + 2 < n) h2 = at(lay + 2);
    } else if (lay + 1 < n) {
      h1 = at(lay + 1);
    }
    if (lay < n) layer_step<kMath, false>(t, h0, g);
  }
}

// Maximum of v over the 64 lanes of the wave (every lane must be active).
__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}

// ---------------------------------------------------------------------------
// fused driver pass (reference OMP:733-747 for every block at once)
// ---------------------------------------------------------------------------
template <int W, int kMath, int kDepth>
__global__ __launch_bounds__(256) void k_fused(const float* __restrict__ trk,
                                               const float* __restrict__ hit,
                                               float* __restrict__ out, int64_t ntracks,
                                               int nlayer) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntracks) return;
  const int64_t b = g / W;
  const int l = (int)(g % W);
  Track t;
  load_track<W>(t, trk + b * (kTrkWords * W) + l);
  run_layers<W, kMath, kDepth, false>(t, hit + b * (int64_t)nlayer * (kHitWords * W) + l, nlayer, nlayer);
  store_track<W>(t, out + b * (kTrkWords * W) + l);
}

// ---------------------------------------------------------------------------
// ragged events: per-block descriptors, per-lane layer counts
// ---------------------------------------------------------------------------

// Work item = 64 consecutive track slots of the descriptor table (one wave):
// slot s is lane s % W of descriptor s / W.  Each lane runs its own layer
// count; the wave's loop runs to the largest.  All 64 lanes of the wave must
// be active (slots past the table take part with no layers).
template <int W, int kMath, int kPrefetch>
__device__ __forceinline__ void ragged_item(const float* __restrict__ trk, const float* __restrict__ hit,
                                   float* __restrict__ out, const int32_t* __restrict__ lane_nlay,
                                   const BlockDesc* __restrict__ blocks, int64_t nblocks, int64_t item, int lane) {
  const int64_t slot = item * 64 + lane;
  const int64_t di = slot / W;
  const int l = (int)(slot % W);
  const bool valid = di < nblocks;
  Track t{};
  int n = 0;
  const float* hb = hit;
  BlockDesc d{};
  if (valid) {
    d = blocks[di];
    load_track<W, true>(t, trk + d.trk_block * (kTrkWords * W) + l);
    n = lane_nlay[d.trk_block * W + l];
    n = n < d.nlayer ? n : d.nlayer;
    n = n > 0 ? n : 0;
    hb = hit + d.hit_block * (kHitWords * W) + l;
  }
  run_layers_lanes<W, kMath, kPrefetch>(t, hb, n, wave_max(n));
  if (valid) store_track<W>(t, out + d.out_block * (kTrkWords * W) + l);
}

// Persistent form: the items are cut into chunks of `claim` consecutive items.
// Wave w of the grid (nwaves in all) computes chunk w first, with no atomic:
// a counter that every wave hit at start would serialise thousands of atomics
// before the first item is computed.  The chunks after those are striped over
// ns = min(kQueueStripes, workgroups) counters: a wave of workgroup g (stripe
// s = g % ns) takes chunk nwaves + s + ns * atomicAdd(counter[s], 1) until it
// is handed a chunk past the end (every wave reaches that exit).  A single
// counter serialises its atomics where they meet (a few times 1e7 per
// second); striping spreads them, so that one or two items per claim keep the
// load even at the end without the counter becoming the limit.  The atomic is
// issued when the last item of the current chunk starts, so its round trip
// overlaps that item; a wave never holds a whole chunk in reserve while others
// run dry at the end.  Each stripe hands out its chunk indexes in order, once
// each, and every workgroup runs (none waits on another), so every item is
// computed exactly once.  The counters start at 0 (k_reset_counter runs first
// on the same stream).
// FP64 gain: at most 128 VGPRs (4 waves per SIMD instead of 3).
template <int W, int kMath, int kPrefetch>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMath == kFp64Gain ? 4 : 1))) void
k_ragged_persistent(const float* __restrict__ trk, const float* __restrict__ hit, float* __restrict__ out,
                    const int32_t* __restrict__ lane_nlay, const BlockDesc* __restrict__ blocks, int64_t nblocks,
                    unsigned* __restrict__ counter, int claim) {
  const int64_t nitems = (nblocks * W + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
  const int ns = gridDim.x < (unsigned)kQueueStripes ? (int)gridDim.x : kQueueStripes;
  const int stripe = (int)(blockIdx.x % (unsigned)ns);
  unsigned* ctr = counter + stripe * kQueueStride;
  int64_t chunk = (int64_t)blockIdx.x * (blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  while (chunk * claim < nitems) {
    const int64_t first = chunk * claim;
    const int64_t last = first + claim < nitems ? first + claim : nitems;
    unsigned next = 0;
    for (int64_t item = first; item < last; ++item) {
      if (item == last - 1 && lane == 0) next = atomicAdd(ctr, 1u);
      int lane_i = lane;
      // the lane index hidden from loop-invariant code motion, so that
      // lane-dependent values are not hoisted out of the item loop into
      // registers held across all of it (10 VGPRs in every mode: FAST then
      // fits within 96 VGPRs, 5 waves per SIMD, as the one-shot form does)
      asm volatile("" : "+v"(lane_i));
      ragged_item<W, kMath, kPrefetch>(trk, hit, out, lane_nlay, blocks, nblocks, item, lane_i);
    }
    // lane 0's claim read into a scalar: the chunk index, and with it the
    // loop's control, stay wave-uniform scalars
    chunk = nwaves + stripe + (int64_t)ns * (int64_t)__builtin_amdgcn_readlane(next, 0);
  }
}

// One-shot form: one wave per item (the hardware dispatcher hands out the
// items, in descriptor order).
template <int W, int kMath, int kPrefetch>
__global__ __launch_bounds__(256) void k_ragged_oneshot(const float* __restrict__ trk,
                                                        const float* __restrict__ hit,
                                                        float* __restrict__ out,
                                                        const int32_t* __restrict__ lane_nlay,
                                                        const BlockDesc* __restrict__ blocks,
                                                        int64_t nblocks) {
  const int64_t nitems = (nblocks * W + 63) / 64;
  const int64_t item = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (item >= nitems) return;
  ragged_item<W, kMath, kPrefetch>(trk, hit, out, lane_nlay, blocks, nblocks, item, threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// device-resident serving loop
// ---------------------------------------------------------------------------
// One persistent launch serves requests as the host publishes them: the host
// writes work items (ServeItem, 64 track slots of one request) into a ring in
// coherent pinned host memory and advances ServeCtl::tail; wave 0 of
// workgroup 0 (the dispatcher) polls that word over PCIe and copies it into
// device memory (ServeDev::dev_tail), where the worker waves wait for it.  A
// worker claims item indexes in order (ServeDev::head), computes the item as
// the ragged pass does, and counts it against its request; the worker that
// completes a request writes a completion record into host memory.
// Termination: the dispatcher stops when the host sets ServeCtl::stop or when
// the host's heartbeat has not moved for idle_ticks (100 MHz s_memrealtime
// ticks); it then sets dev_stop, and every worker ends once the item it holds
// is past the last published one.  Every wait loop therefore ends.
// Memory: the loop takes no acquire or release at agent or system scope: on
// this GPU those write back or invalidate the L2 of the wave's XCD, and
// thousands of waiting waves would keep every L2 flushing.  Instead
//  - counters and flags are atomics (coherent at the scope they name, so
//    every poll reads the current value);
//  - item fields are read with system-scope loads straight from host memory
//    (the host writes an item before its tail; a worker reads the item only
//    after it has seen the tail);
//  - results are written with system-scope (write-through) stores, and a wave
//    waits for its stores to complete before it counts its item, so a
//    request's results are in memory before its completion is counted, and
//    the completion record's fields before its seq word.

__device__ inline uint32_t lane0_u32(uint32_t v) { return __builtin_amdgcn_readlane(v, 0); }
__device__ inline uint64_t lane0_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, 0);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), 0);
  return ((uint64_t)hi << 32) | lo;
}

template <class T>
__device__ inline T sys_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ inline T agent_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void sys_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// Waits until this wave's vector memory operations have completed
// (s_waitcnt vmcnt(0); gfx9 encoding: vmcnt 0, expcnt 7, lgkmcnt 15): a store
// is then acknowledged at the scope it was issued with.  No cache operation.
__device__ inline void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <int W>
__device__ inline void store_track_through(const Track& t, float* __restrict__ ob) {
  uint32_t* o = reinterpret_cast<uint32_t*>(ob);
#pragma unroll
  for (int e = 0; e < 6; ++e) sys_store(o + e * W, __builtin_bit_cast(uint32_t, t.p[e]));
#pragma unroll
  for (int e = 0; e < 21; ++e) sys_store(o + (6 + e) * W, __builtin_bit_cast(uint32_t, t.P[e]));
  sys_store(o + 27 * W, (uint32_t)t.q);
}

__device__ void serve_dispatch(const ServeLaunch& a) {
  uint64_t seen = ~0ull;
  uint64_t hb = lane0_u64(sys_load(&a.ctl->heartbeat));
  uint64_t t_hb = __builtin_amdgcn_s_memrealtime();
  bool run = true;
  while (run) {
    const uint64_t t = lane0_u64(sys_load(&a.ctl->tail));
    if (t != seen) {
      __hip_atomic_store(&a.dev->dev_tail, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      seen = t;
    }
    const uint32_t stop = lane0_u32(sys_load(&a.ctl->stop));
    const uint64_t h = lane0_u64(sys_load(&a.ctl->heartbeat));
    const uint64_t now = __builtin_amdgcn_s_memrealtime();
    if (h != hb) {
      hb = h;
      t_hb = now;
    }
    run = stop == 0u && now - t_hb <= a.idle_ticks;
    if (run) __builtin_amdgcn_s_sleep(2);
  }
  // the final tail is stored before the stop flag, and the host learns that
  // this launch's dispatcher has left
  wait_stores();
  __hip_atomic_store(&a.dev->dev_stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  sys_store(&a.ctl->exited, lane0_u32(sys_load(&a.ctl->epoch)));
}

template <int W, int kMath, int kPrefetch>
__device__ __forceinline__ void serve_item(const ServeLaunch& a, uint64_t k, int lane) {
  const ServeItem* it = a.items + (int64_t)(k % (uint64_t)a.item_cap);
  const int di = W <= 64 ? lane / W : 0;
  // The whole 144-byte record in one load instruction, one round trip over
  // PCIe: lane j < 18 reads the record's 8-byte word j (18 lanes of
  // consecutive words: a few PCIe requests, where a load per field costs one
  // request each).  Each lane then takes its descriptor's fields from the
  // lanes that hold them: desc[d] is words 4d .. 4d + 3 (trk_block,
  // hit_block, out_block, nlayer | event), the header words 16 (ndesc | half)
  // and 17 (req_slot | req_items).  A padding descriptor is read too: it lies
  // inside the record.
  static_assert(sizeof(ServeItem) == 18 * sizeof(uint64_t), "serve item: 18 words");
  static_assert(offsetof(ServeItem, ndesc) == 16 * 8 && offsetof(ServeItem, req_slot) == 17 * 8, "header words");
  uint64_t word = 0;
  if (lane < 18) word = sys_load(reinterpret_cast<const uint64_t*>(it) + lane);
  const int wlo = (int)(uint32_t)word, whi = (int)(uint32_t)(word >> 32);
  auto pick = [&](int j) -> int64_t {
    return (int64_t)(((uint64_t)(uint32_t)__shfl(whi, j) << 32) | (uint32_t)__shfl(wlo, j));
  };
  const int64_t tb = pick(4 * di);
  const int64_t hblk = pick(4 * di + 1);
  const int64_t ob = pick(4 * di + 2);
  const int nl = __shfl(wlo, 4 * di + 3);
  const int ndesc = (int)__builtin_amdgcn_readlane((uint32_t)wlo, 16);
  const int half = (int)__builtin_amdgcn_readlane((uint32_t)whi, 16);
  const int req_slot = (int)__builtin_amdgcn_readlane((uint32_t)wlo, 17);
  const uint32_t req_items = __builtin_amdgcn_readlane((uint32_t)whi, 17);
  const int l = W <= 64 ? lane % W : half * 64 + lane;
  const bool valid = di < ndesc;
  Track t{};
  int n = 0;
  const float* hb = a.hit;
  if (valid) {
    load_track<W, true>(t, a.trk + tb * (kTrkWords * W) + l);
    n = a.lanes[tb * W + l];
    n = n < nl ? n : nl;
    n = n > 0 ? n : 0;
    hb = a.hit + hblk * (kHitWords * W) + l;
  }
  run_layers_lanes<W, kMath, kPrefetch>(t, hb, n, wave_max(n));
  if (valid) store_track_through<W>(t, a.out + ob * (kTrkWords * W) + l);
  wait_stores();
  if (lane == 0) {
    const uint32_t old =
        __hip_atomic_fetch_add(&a.done[req_slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1u == req_items) {
      // the request's last item: reset its counter for the slot's next
      // request and report it
      __hip_atomic_store(&a.done[req_slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t c =
          __hip_atomic_fetch_add(&a.dev->comp_tail, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ServeCompletion* r = a.comps + (int64_t)(c % (uint64_t)a.comp_cap);
      sys_store(&r->req_slot, req_slot);
      wait_stores();
      sys_store(&r->seq, c + 1);
    }
  }
}

// Claims the item indexes of its workgroup's stripe in order and computes
// them; ends once it holds an index past the last published item after the
// dispatcher has stopped.  Every stripe has workers (the grid has at least two
// workgroups per stripe, dispatched in order), so every published item is
// claimed.  The
// next index is claimed (by lane 0) while the current item starts, as in the
// persistent ragged kernel: the claim's round trip overlaps the item, and the
// lane-0 branch sits inside the loop body, never on the edge back to the
// loop's head (where the compiler may split a loop whose back edges differ by
// lane, letting the other lanes run on without lane 0 and its claim).  Every
// loop condition is a wave-uniform value (lane 0's, read into a scalar).
template <int W, int kMath, int kPrefetch>
__device__ void serve_work(const ServeLaunch& a, int lane) {
  const uint64_t stripe = blockIdx.x % kServeStripes;
  uint64_t* head = &a.dev->head[stripe][0];
  uint64_t claimed = 0;
  if (lane == 0) claimed = __hip_atomic_fetch_add(head, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t k = lane0_u64(claimed) * kServeStripes + stripe;
  bool have = true;
  while (have) {
    bool wait = true;
    while (wait) {
      if (k < lane0_u64(agent_load(&a.dev->dev_tail))) {
        wait = false;
      } else if (lane0_u32(agent_load(&a.dev->dev_stop))) {
        // stop comes after the final tail: an item past it is never published
        have = k < lane0_u64(agent_load(&a.dev->dev_tail));
        wait = false;
      } else {
        __builtin_amdgcn_s_sleep(32);
      }
    }
    if (have) {
      uint64_t next = 0;
      if (lane == 0) next = __hip_atomic_fetch_add(head, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int lane_i = lane;
      asm volatile("" : "+v"(lane_i));  // as in the persistent ragged pass
      serve_item<W, kMath, kPrefetch>(a, k, lane_i);
      k = lane0_u64(next) * kServeStripes + stripe;
    }
  }
}

__global__ void k_serve_reset(ServeDev* dev) {
  const uint64_t t = dev->dev_tail;
  const int j = threadIdx.x;
  if (j < kServeStripes) dev->head[j][0] = t > (uint64_t)j ? (t - j + kServeStripes - 1) / kServeStripes : 0;
  if (j == 0) dev->dev_stop = 0u;
}

template <int W, int kMath, int kPrefetch>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMath == kFp64Gain ? 4 : 1))) void
k_serve_resident(ServeLaunch a) {
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    serve_dispatch(a);
    return;
  }
  serve_work<W, kMath, kPrefetch>(a, threadIdx.x & 63);
}

__global__ void k_reset_counter(unsigned* counter) {
  if (blockIdx.x == 0 && threadIdx.x < kQueueStripes) counter[threadIdx.x * kQueueStride] = 0u;
}

// ---------------------------------------------------------------------------
// per-function kernels (reference propagateToZ / KalmanUpdate_v2 on blocks)
// ---------------------------------------------------------------------------
template <int W, int kMath>
__global__ __launch_bounds__(256) void k_propagate(const float* inErr, const float* inPar,
                                                   const int32_t* inChg, const float* msP,
                                                   float* outErr, float* outPar,
                                                   int64_t ntracks, int64_t ts, int64_t hs) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntracks) return;
  const int64_t b = g / W;
  const int l = (int)(g % W);
  Track t;
  const int64_t o = b * ts + l;
#pragma unroll
  for (int e = 0; e < 21; ++e) t.P[e] = inErr[o + e * W];
#pragma unroll
  for (int e = 0; e < 6; ++e) t.p[e] = inPar[o + e * W];
  t.q = inChg[o];
  const float hz = msP[b * hs + 2 * W + l];
  if constexpr (kMath == kRef)
    propagate_ref(t, hz);
  else
    propagate_fast<false, kMath == kFp64Gain>(t, hz);
#pragma unroll
  for (int e = 0; e < 21; ++e) outErr[o + e * W] = t.P[e];
#pragma unroll
  for (int e = 0; e < 6; ++e) outPar[o + e * W] = t.p[e];
}

template <int W, int kMath>
__global__ __launch_bounds__(256) void k_update(float* trkErr, float* inPar, const float* hitErr,
                                                const float* msP, int64_t ntracks, int64_t ts,
                                                int64_t hs) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntracks) return;
  const int64_t b = g / W;
  const int l = (int)(g % W);
  Track t;
  const int64_t o = b * ts + l;
#pragma unroll
  for (int e = 0; e < 21; ++e) t.P[e] = trkErr[o + e * W];
#pragma unroll
  for (int e = 0; e < 6; ++e) t.p[e] = inPar[o + e * W];
  t.q = 0;
  const int64_t ho = b * hs + l;
  Hit h;
  h.x = msP[ho];
  h.y = msP[ho + W];
  h.z = msP[ho + 2 * W];
  h.xx = hitErr[ho];
  h.xy = hitErr[ho + W];
  h.yy = hitErr[ho + 2 * W];
  if constexpr (kMath == kRef)
    update_ref(t, h);
  else if constexpr (kMath == kFast)
    update_fast<false>(t, h);
  else
    update_fp64gain<false>(t, h);
#pragma unroll
  for (int e = 0; e < 21; ++e) trkErr[o + e * W] = t.P[e];
#pragma unroll
  for (int e = 0; e < 6; ++e) inPar[o + e * W] = t.p[e];
}

// KalmanUpdate v1 (OMP:375-501) on blocks, in place; kFixed selects the
// textbook form.  hitErr: the 6 hit covariance words of each lane.
template <int W, bool kFixed>
__global__ __launch_bounds__(256) void k_update_v1(float* trkErr, float* inPar, const float* hitErr,
                                                   const float* msP, int64_t ntracks, int64_t ts,
                                                   int64_t hs) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntracks) return;
  const int64_t b = g / W;
  const int l = (int)(g % W);
  Track t;
  const int64_t o = b * ts + l;
#pragma unroll
  for (int e = 0; e < 21; ++e) t.P[e] = trkErr[o + e * W];
#pragma unroll
  for (int e = 0; e < 6; ++e) t.p[e] = inPar[o + e * W];
  t.q = 0;
  const int64_t ho = b * hs + l;
  float hc[6];
#pragma unroll
  for (int e = 0; e < 6; ++e) hc[e] = hitErr[ho + e * W];
  if constexpr (kFixed)
    update_v1_fixed(t, hc, msP[ho], msP[ho + W], msP[ho + 2 * W]);
  else
    update_v1_ref(t, hc, msP[ho], msP[ho + W]);
#pragma unroll
  for (int e = 0; e < 21; ++e) trkErr[o + e * W] = t.P[e];
#pragma unroll
  for (int e = 0; e < 6; ++e) inPar[o + e * W] = t.p[e];
}

// ---------------------------------------------------------------------------
// report statistics (OMP:758-851), deterministic fixed-order reduction
// ---------------------------------------------------------------------------
constexpr int kSumThreads = 256;
constexpr int kSumVals = 12;  // pass 1: 9 sums + nnans, nfail, ngood; pass 2: 6 sums

struct TrackView {
  double x, y, z, pt, phi, theta;
  int state;  // 0 good, 1 non-finite, 2 outside the acceptance window
};

__device__ inline TrackView view_track(const float* out, int64_t nb, int W, int64_t ntrks,
                                       int64_t idx, double hx, double hy, double hz) {
  const int64_t ie = idx / ntrks, it = idx % ntrks;
  const float* blk = out + ((it / W) + nb * ie) * (int64_t)(kTrkWords * W);
  const int64_t l = it % W;
  TrackView v;
  v.x = blk[0 * W + l];
  v.y = blk[1 * W + l];
  v.z = blk[2 * W + l];
  v.pt = 1. / (double)blk[3 * W + l];
  v.phi = blk[4 * W + l];
  v.theta = blk[5 * W + l];
  if (!isfinite(v.x) || !isfinite(v.y) || !isfinite(v.z) || !isfinite(v.pt) ||
      !isfinite(v.phi) || !isfinite(v.theta)) {
    v.state = 1;
  } else if (fabs((v.x - hx) / hx) > 1. || fabs((v.y - hy) / hy) > 1. ||
             fabs((v.z - hz) / hz) > 1. || fabs((v.pt - 12.) / 12.) > 1. ||
             fabs((v.phi - 1.3) / 1.3) > 1. || fabs((v.theta - 2.8) / 2.8) > 1.) {
    v.state = 2;
  } else {
    v.state = 0;
  }
  return v;
}

// Tree-reduce kSumVals doubles per thread over the workgroup in a fixed order;
// thread 0 returns the totals in vals.
__device__ inline void block_reduce(double (&vals)[kSumVals], double* lds) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kSumVals; ++k) lds[k * kSumThreads + t] = vals[k];
  __syncthreads();
  for (int stride = kSumThreads / 2; stride > 0; stride >>= 1) {
    if (t < stride) {
#pragma unroll
      for (int k = 0; k < kSumVals; ++k) lds[k * kSumThreads + t] += lds[k * kSumThreads + t + stride];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kSumVals; ++k) vals[k] = lds[k * kSumThreads];
}

template <int W, int kPass>
__global__ __launch_bounds__(kSumThreads) void k_summary(const float* __restrict__ out, int64_t nb,
                                                         int64_t ntrks, int64_t total, double hx,
                                                         double hy, double hz,
                                                         const double* __restrict__ res,
                                                         double* __restrict__ part) {
  __shared__ double lds[kSumVals * kSumThreads];
  double v[kSumVals];
#pragma unroll
  for (int k = 0; k < kSumVals; ++k) v[k] = 0.;
  const int64_t idx = (int64_t)blockIdx.x * kSumThreads + threadIdx.x;
  if (idx < total) {
    const TrackView tv = view_track(out, nb, W, ntrks, idx, hx, hy, hz);
    if (kPass == 1) {
      if (tv.state == 0) {
        v[0] = tv.x; v[1] = tv.y; v[2] = tv.z;
        v[3] = (tv.x - hx) / tv.x; v[4] = (tv.y - hy) / tv.y; v[5] = (tv.z - hz) / tv.z;
        v[6] = tv.pt; v[7] = tv.phi; v[8] = tv.theta;
        v[11] = 1.;
      } else if (tv.state == 1) {
        v[9] = 1.;
      } else {
        v[10] = 1.;
      }
    } else if (tv.state == 0) {
      const double dx = tv.x - res[0], dy = tv.y - res[1], dz = tv.z - res[2];
      const double ddx = (tv.x - hx) / tv.x - res[3], ddy = (tv.y - hy) / tv.y - res[4],
                   ddz = (tv.z - hz) / tv.z - res[5];
      v[0] = dx * dx; v[1] = dy * dy; v[2] = dz * dz;
      v[3] = ddx * ddx; v[4] = ddy * ddy; v[5] = ddz * ddz;
    }
  }
  block_reduce(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kSumVals; ++k) part[(int64_t)blockIdx.x * kSumVals + k] = v[k];
  }
}

// Final reduction of the per-workgroup partials (one workgroup).  Pass 1 turns
// the sums into means (res[0..8]) and counts (res[9..11] = nnans, nfail, ngood);
// pass 2 into the rounded-to-float RMS values (res[12..17]).
template <int kPass>
__global__ __launch_bounds__(kSumThreads) void k_summary_final(const double* __restrict__ part,
                                                               int64_t nparts,
                                                               double* __restrict__ res) {
  __shared__ double lds[kSumVals * kSumThreads];
  double v[kSumVals];
#pragma unroll
  for (int k = 0; k < kSumVals; ++k) v[k] = 0.;
  for (int64_t i = threadIdx.x; i < nparts; i += kSumThreads) {
#pragma unroll
    for (int k = 0; k < kSumVals; ++k) v[k] += part[i * kSumVals + k];
  }
  block_reduce(v, lds);
  if (threadIdx.x == 0) {
    if (kPass == 1) {
      const double n = v[11];
#pragma unroll
      for (int k = 0; k < 9; ++k) res[k] = v[k] / n;
      res[9] = v[9];
      res[10] = v[10];
      res[11] = v[11];
    } else {
      const double n = res[11];
#pragma unroll
      for (int k = 0; k < 6; ++k) res[12 + k] = (double)sqrtf((float)(v[k] / n));
    }
  }
}

// sin and cos of the reference-typing mode (sincosf_glibc) for the float bit
// patterns first, first+1, ... (n of them, n <= 2^32 - first).
__global__ __launch_bounds__(256) void k_sincosf_range(uint32_t first, int64_t n, float* __restrict__ s,
                                                       float* __restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float y = __builtin_bit_cast(float, (uint32_t)(first + (uint64_t)i));
  float sv, cv;
  sincosf_glibc(y, &sv, &cv);
  s[i] = sv;
  c[i] = cv;
}

template <typename F>
hipError_t dispatch_width(int width, F&& f) {
  switch (width) {
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
  }
  return hipErrorInvalidValue;
}

template <typename F>
hipError_t dispatch_math(int math, F&& f) {
  switch (math) {
    case kFast: return f(std::integral_constant<int, kFast>{});
    case kRef: return f(std::integral_constant<int, kRef>{});
    case kFp64Gain: return f(std::integral_constant<int, kFp64Gain>{});
  }
  return hipErrorInvalidValue;
}

inline int64_t grid_for(int64_t n, int tpw) { return (n + tpw - 1) / tpw; }

}  // namespace

hipError_t launch_fused(const float* trk, const float* hit, float* out, int64_t nblocks, int nlayer,
                        int width, int math, const LaunchCfg& cfg, hipStream_t s) {
  const int64_t ntracks = nblocks * width;
  if (ntracks == 0) return hipSuccess;
  const int tpw = cfg.tracks_per_wg;
  return dispatch_width(width, [&](auto wc) {
    return dispatch_math(math, [&](auto mc) {
      constexpr int W = decltype(wc)::value, M = decltype(mc)::value;
      const int depth = cfg.hit_stage ? cfg.hit_stage : default_hit_stage(math);
      if (depth == 1)
        hipLaunchKernelGGL((k_fused<W, M, 1>), dim3((unsigned)grid_for(ntracks, tpw)), dim3(tpw), 0, s,
                           trk, hit, out, ntracks, nlayer);
      else
        hipLaunchKernelGGL((k_fused<W, M, 2>), dim3((unsigned)grid_for(ntracks, tpw)), dim3(tpw), 0, s,
                           trk, hit, out, ntracks, nlayer);
      return hipGetLastError();
    });
  });
}

hipError_t launch_propagate(const float* inErr, const float* inPar, const int32_t* inChg,
                            const float* msP, float* outErr, float* outPar, int64_t nblocks,
                            int64_t ts, int64_t hs, int width, int math, const LaunchCfg& cfg,
                            hipStream_t s) {
  const int64_t ntracks = nblocks * width;
  if (ntracks == 0) return hipSuccess;
  const int tpw = cfg.tracks_per_wg;
  return dispatch_width(width, [&](auto wc) {
    return dispatch_math(math, [&](auto mc) {
      constexpr int W = decltype(wc)::value, M = decltype(mc)::value;
      hipLaunchKernelGGL((k_propagate<W, M>), dim3((unsigned)grid_for(ntracks, tpw)), dim3(tpw), 0,
                         s, inErr, inPar, inChg, msP, outErr, outPar, ntracks, ts, hs);
      return hipGetLastError();
    });
  });
}

hipError_t launch_update(float* trkErr, float* inPar, const float* hitErr, const float* msP,
                         int64_t nblocks, int64_t ts, int64_t hs, int width, int math,
                         const LaunchCfg& cfg, hipStream_t s) {
  const int64_t ntracks = nblocks * width;
  if (ntracks == 0) return hipSuccess;
  const int tpw = cfg.tracks_per_wg;
  return dispatch_width(width, [&](auto wc) {
    return dispatch_math(math, [&](auto mc) {
      constexpr int W = decltype(wc)::value, M = decltype(mc)::value;
      hipLaunchKernelGGL((k_update<W, M>), dim3((unsigned)grid_for(ntracks, tpw)), dim3(tpw), 0, s,
                         trkErr, inPar, hitErr, msP, ntracks, ts, hs);
      return hipGetLastError();
    });
  });
}

hipError_t launch_update_v1(float* trkErr, float* inPar, const float* hitErr, const float* msP, int64_t nblocks,
                            int64_t ts, int64_t hs, int width, bool fixed, const LaunchCfg& cfg, hipStream_t s) {
  const int64_t ntracks = nblocks * width;
  if (ntracks == 0) return hipSuccess;
  const int tpw = cfg.tracks_per_wg;
  return dispatch_width(width, [&](auto wc) {
    constexpr int W = decltype(wc)::value;
    if (fixed)
      hipLaunchKernelGGL((k_update_v1<W, true>), dim3((unsigned)grid_for(ntracks, tpw)), dim3(tpw), 0, s, trkErr,
                         inPar, hitErr, msP, ntracks, ts, hs);
    else
      hipLaunchKernelGGL((k_update_v1<W, false>), dim3((unsigned)grid_for(ntracks, tpw)), dim3(tpw), 0, s, trkErr,
                         inPar, hitErr, msP, ntracks, ts, hs);
    return hipGetLastError();
  });
}

int ragged_persistent_groups_per_cu(int width, int math, int prefetch, int tpw) {
  static std::mutex mu;
  static std::map<std::array<int, 4>, int> cache;
  const std::array<int, 4> key{width, math, prefetch >= 2 ? 2 : 1, tpw};
  std::lock_guard<std::mutex> lock(mu);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int n = 0;
  const hipError_t e = dispatch_width(width, [&](auto wc) {
    return dispatch_math(math, [&](auto mc) {
      constexpr int W = decltype(wc)::value, M = decltype(mc)::value;
      const void* k = prefetch >= 2 ? reinterpret_cast<const void*>(&k_ragged_persistent<W, M, 2>)
                                    : reinterpret_cast<const void*>(&k_ragged_persistent<W, M, 1>);
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, tpw, 0);
    });
  });
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  cache.emplace(key, n);
  return n;
}

hipError_t launch_ragged(const float* trk, const float* hit, float* out, const int32_t* lane_nlay,
                         const BlockDesc* blocks, int64_t nblocks, int width, int math, bool persistent,
                         int prefetch, int tpw, int64_t grid, int claim, unsigned* counter, hipStream_t s) {
  const int64_t nitems = (nblocks * width + 63) / 64;
  if (nitems == 0) return hipSuccess;
  return dispatch_width(width, [&](auto wc) {
    return dispatch_math(math, [&](auto mc) {
      constexpr int W = decltype(wc)::value, M = decltype(mc)::value;
      if (!persistent) {
        const dim3 grid1((unsigned)grid_for(nitems * 64, tpw));
        if (prefetch >= 2)
          hipLaunchKernelGGL((k_ragged_oneshot<W, M, 2>), grid1, dim3(tpw), 0, s, trk, hit, out, lane_nlay, blocks,
                             nblocks);
        else
          hipLaunchKernelGGL((k_ragged_oneshot<W, M, 1>), grid1, dim3(tpw), 0, s, trk, hit, out, lane_nlay, blocks,
                             nblocks);
        return hipGetLastError();
      }
      hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, counter);
      const int64_t items_per_wg = (int64_t)(tpw / 64) * claim;
      int64_t g = grid > 0 ? grid : 1;
      g = std::min<int64_t>(g, (nitems + items_per_wg - 1) / items_per_wg);
      if (prefetch >= 2)
        hipLaunchKernelGGL((k_ragged_persistent<W, M, 2>), dim3((unsigned)g), dim3(tpw), 0, s, trk, hit, out,
                           lane_nlay, blocks, nblocks, counter, claim);
      else
        hipLaunchKernelGGL((k_ragged_persistent<W, M, 1>), dim3((unsigned)g), dim3(tpw), 0, s, trk, hit, out,
                           lane_nlay, blocks, nblocks, counter, claim);
      return hipGetLastError();
    });
  });
}

hipError_t launch_serve_reset(ServeDev* dev, hipStream_t s) {
  hipLaunchKernelGGL(k_serve_reset, dim3(1), dim3(64), 0, s, dev);
  return hipGetLastError();
}

hipError_t launch_serve_resident(const ServeLaunch& a, int width, int math, int prefetch, int tpw, int64_t grid,
                                 hipStream_t s) {
  if (grid < 2 * kServeStripes) return hipErrorInvalidValue;  // every stripe needs workers
  return dispatch_width(width, [&](auto wc) {
    return dispatch_math(math, [&](auto mc) {
      constexpr int W = decltype(wc)::value, M = decltype(mc)::value;
      if (prefetch >= 2)
        hipLaunchKernelGGL((k_serve_resident<W, M, 2>), dim3((unsigned)grid), dim3(tpw), 0, s, a);
      else
        hipLaunchKernelGGL((k_serve_resident<W, M, 1>), dim3((unsigned)grid), dim3(tpw), 0, s, a);
      return hipGetLastError();
    });
  });
}

hipError_t launch_sincosf_range(uint32_t first, int64_t n, float* s, float* c, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sincosf_range, dim3((unsigned)grid_for(n, 256)), dim3(256), 0, st, first, n, s, c);
  return hipGetLastError();
}

size_t summary_scratch_doubles(int64_t nevts, int64_t ntrks) {
  return (size_t)grid_for(nevts * ntrks, kSumThreads) * kSumVals;
}

hipError_t launch_summary(const float* out, int64_t nevts, int64_t nb, int width, int64_t ntrks,
                          double hx, double hy, double hz, double* part, double* res,
                          hipStream_t s) {
  const int64_t total = nevts * ntrks;
  const int64_t nparts = grid_for(total, kSumThreads);
  if (nparts == 0) return hipErrorInvalidValue;
  return dispatch_width(width, [&](auto wc) {
    constexpr int W = decltype(wc)::value;
    hipLaunchKernelGGL((k_summary<W, 1>), dim3((unsigned)nparts), dim3(kSumThreads), 0, s, out, nb,
                       ntrks, total, hx, hy, hz, res, part);
    hipLaunchKernelGGL((k_summary_final<1>), dim3(1), dim3(kSumThreads), 0, s, part, nparts, res);
    hipLaunchKernelGGL((k_summary<W, 2>), dim3((unsigned)nparts), dim3(kSumThreads), 0, s, out, nb,
                       ntrks, total, hx, hy, hz, res, part);
    hipLaunchKernelGGL((k_summary_final<2>), dim3(1), dim3(kSumThreads), 0, s, part, nparts, res);
    return hipGetLastError();
  });
}

}  // namespace p2z
