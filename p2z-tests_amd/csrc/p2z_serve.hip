// Serving loop: admission by an HBM output arena, continuous batching of
// events into persistent launches on several streams, per-event timestamps
// (declared in include/p2z.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <random>
#include <unordered_map>
#include <vector>

#include "p2z.h"
#include "p2z_capi_util.hpp"
#include "p2z_ctx.hpp"
#include "p2z_internal.hpp"

namespace {

// First-fit allocator of arena blocks (ranges start -> length).
class Arena {
 public:
  explicit Arena(int64_t n) : total_(n) {
    if (n > 0) free_[0] = n;
  }
  int64_t alloc(int64_t n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t at = it->first, rest = it->second - n;
        free_.erase(it);
        if (rest > 0) free_[at + n] = rest;
        used_ += n;
        return at;
      }
    }
    return -1;
  }
  void release(int64_t at, int64_t n) {
    auto it = free_.emplace(at, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && at + n == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    used_ -= n;
  }
  int64_t total() const { return total_; }
  int64_t used() const { return used_; }

 private:
  std::map<int64_t, int64_t> free_;
  int64_t total_ = 0, used_ = 0;
};

struct Request {
  std::vector<p2z_block_desc> blocks;  // batches form: out_block rewritten to arena positions
  int64_t out_block = 0, nblocks = 0;
  double t_arrival = 0, t_admit = 0, t_launch = 0, t_done = 0;
  bool done = false;
  int32_t slot = -1;        // resident loop: completion slot
  uint64_t first_item = 0;  // resident loop: first ring item
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  unsigned* counter = nullptr;
  p2z::BlockDesc* d_desc = nullptr;
  p2z_block_desc* h_desc = nullptr;  // pinned staging
  int64_t cap = 0;
  std::vector<int64_t> requests;
  bool busy = false;
};

// The device-resident loop (P2Z_SERVE_RESIDENT): host rings, device state and
// the bookkeeping of what may be overwritten.
constexpr int64_t kResItems = 65536;  // work-item ring (9 MiB of pinned memory)
constexpr int32_t kResSlots = 4096;   // requests in flight; also the completion ring
constexpr double kResIdleSeconds = 2.0;
constexpr double kBeatSeconds = 0.05;

struct Resident {
  hipStream_t stream = nullptr;
  p2z::ServeItem* items_h = nullptr;  // pinned, coherent
  p2z::ServeItem* items_d = nullptr;  // its device view
  p2z::ServeCtl* ctl_h = nullptr;
  p2z::ServeCtl* ctl_d = nullptr;
  p2z::ServeCompletion* comps_h = nullptr;
  p2z::ServeCompletion* comps_d = nullptr;
  p2z::ServeDev* dev = nullptr;       // device memory: ServeDev + done[kResSlots]
  uint64_t published = 0;             // items published (ServeCtl::tail)
  uint64_t comp_read = 0;             // completion records consumed
  uint64_t heartbeat = 0;
  double t_beat = 0;                  // host time of the last heartbeat store
  std::vector<int64_t> slot_req;      // request of each slot (-1: free)
  std::vector<int32_t> free_slots;
  std::map<uint64_t, int64_t> unfinished;  // first item -> request, in publication order
  std::vector<p2z_block_desc> sorted;      // a request's blocks put longest first
  bool running = false;
  uint32_t epoch = 0;  // launches so far (ServeCtl::epoch of the current one)
  int64_t launches = 0;
  int tpw = 256, prefetch = 2;
  int64_t grid = 0;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct p2z_server {
  p2z_ctx* ctx = nullptr;
  p2z_server_opts opts{};
  const float* trk = nullptr;
  const float* hit = nullptr;
  const int32_t* lanes = nullptr;
  float* arena_mem = nullptr;
  Arena* arena = nullptr;
  std::unordered_map<int64_t, Request> reqs;
  std::deque<int64_t> pending;
  std::vector<Slot> slots;
  std::vector<p2z_completion> ready;
  int64_t batches = 0, batch_events = 0;
  std::vector<int64_t> sort_start;  // counting-sort offsets of launch_batches
  Resident* res = nullptr;  // non-null: the device-resident loop serves
};

namespace {

// Descriptor room of a slot.  Reserved up front for batches of up to
// kSlotDescriptors blocks (hundreds of typical events) and grown geometrically
// past that: a reallocation frees device memory, which waits for the device,
// and allocates pinned memory, which is slow -- neither belongs in the loop.
constexpr int64_t kSlotDescriptors = 65536;

int slot_reserve(Slot& s, int64_t n) {
  if (s.cap >= n) return P2Z_OK;
  n = std::max<int64_t>(n, 2 * s.cap);
  if (s.d_desc) P2Z_HIP_CHECK(hipFree(s.d_desc));
  if (s.h_desc) P2Z_HIP_CHECK(hipHostFree(s.h_desc));
  s.d_desc = nullptr;
  s.h_desc = nullptr;
  s.cap = 0;
  const int64_t cap = std::max<int64_t>(n, kSlotDescriptors);
  P2Z_HIP_CHECK(hipMalloc(&s.d_desc, cap * sizeof(p2z::BlockDesc)));
  P2Z_HIP_CHECK(hipHostMalloc(&s.h_desc, cap * sizeof(p2z_block_desc), hipHostMallocDefault));
  s.cap = cap;
  return P2Z_OK;
}

int harvest(p2z_server* sv) {
  for (auto& s : sv->slots) {
    if (!s.busy) continue;
    const hipError_t q = hipEventQuery(s.done);
    if (q == hipErrorNotReady) continue;
    P2Z_HIP_CHECK(q);
    const double t = now_s();
    for (int64_t id : s.requests) {
      Request& r = sv->reqs.at(id);
      r.done = true;
      r.t_done = t;
      sv->ready.push_back(p2z_completion{id, r.out_block, r.nblocks, r.t_arrival, r.t_admit, r.t_launch, r.t_done});
    }
    s.requests.clear();
    s.busy = false;
  }
  return P2Z_OK;
}

// The blocks of the requests `ids`, longest first (stable: requests in order,
// each request's blocks in its order), written to dst.  Layer counts span a
// few dozen values, so a counting sort over them does it in two passes; a
// comparison sort of a large batch (tens of thousands of blocks) took
// milliseconds of host time while the device waited for the batch.
void gather_longest_first(p2z_server* sv, const std::vector<int64_t>& ids, p2z_block_desc* dst, int64_t nblk) {
  constexpr int32_t kMaxCounted = 4096;
  int32_t top = 0;
  for (int64_t id : ids)
    for (const auto& d : sv->reqs.at(id).blocks) top = std::max(top, d.nlayer);
  if (top >= kMaxCounted) {
    int64_t k = 0;
    for (int64_t id : ids)
      for (const auto& d : sv->reqs.at(id).blocks) dst[k++] = d;
    std::stable_sort(dst, dst + nblk,
                     [](const p2z_block_desc& a, const p2z_block_desc& b) { return a.nlayer > b.nlayer; });
    return;
  }
  // start[c]: first position of the blocks with c layers (descending c)
  std::vector<int64_t>& start = sv->sort_start;
  start.assign((size_t)top + 2, 0);
  for (int64_t id : ids)
    for (const auto& d : sv->reqs.at(id).blocks) ++start[(size_t)d.nlayer];
  int64_t at = 0;
  for (int32_t c = top; c >= 0; --c) {
    const int64_t n = start[(size_t)c];
    start[(size_t)c] = at;
    at += n;
  }
  for (int64_t id : ids)
    for (const auto& d : sv->reqs.at(id).blocks) dst[start[(size_t)d.nlayer]++] = d;
}

int launch_batches(p2z_server* sv) {
  const p2z_server_opts& o = sv->opts;
  for (auto& s : sv->slots) {
    if (sv->pending.empty()) break;
    if (s.busy) continue;
    // every admitted request waiting, up to the block limit (at least one)
    std::vector<int64_t> ids;
    int64_t nblk = 0;
    while (!sv->pending.empty()) {
      const Request& r = sv->reqs.at(sv->pending.front());
      if (!ids.empty() && o.max_batch_blocks > 0 && nblk + r.nblocks > o.max_batch_blocks) break;
      ids.push_back(sv->pending.front());
      nblk += r.nblocks;
      sv->pending.pop_front();
    }
    const int rc = [&]() -> int {
      P2Z_TRY(slot_reserve(s, nblk));
      gather_longest_first(sv, ids, s.h_desc, nblk);
      const int tpw = o.tracks_per_wg ? o.tracks_per_wg : sv->ctx->cfg.tracks_per_wg;
      const int wpc = o.waves_per_cu ? o.waves_per_cu : 8;
      const int pf = o.prefetch ? o.prefetch : 2;
      int64_t grid = ((int64_t)sv->ctx->cus * wpc * 64 + tpw - 1) / tpw;
      const int per_cu = p2z::ragged_persistent_groups_per_cu(o.width, o.math, pf, tpw);
      if (per_cu > 0) grid = std::min<int64_t>(grid, (int64_t)per_cu * sv->ctx->cus);
      P2Z_HIP_CHECK(
          hipMemcpyAsync(s.d_desc, s.h_desc, nblk * sizeof(p2z_block_desc), hipMemcpyHostToDevice, s.stream));
      P2Z_HIP_CHECK(p2z::launch_ragged(sv->trk, sv->hit, sv->arena_mem, sv->lanes, s.d_desc, nblk, o.width, o.math,
                                       true, pf, tpw, grid,
                                       o.items_per_claim ? o.items_per_claim : 1, s.counter, s.stream));
      P2Z_HIP_CHECK(hipEventRecord(s.done, s.stream));
      return P2Z_OK;
    }();
    if (rc != P2Z_OK) {
      // the batch did not start: its requests wait again, in their order
      for (auto it = ids.rbegin(); it != ids.rend(); ++it) sv->pending.push_front(*it);
      return rc;
    }
    const double t = now_s();
    for (int64_t id : ids) sv->reqs.at(id).t_launch = t;
    s.requests = std::move(ids);
    s.busy = true;
    sv->batches += 1;
    sv->batch_events += (int64_t)s.requests.size();
  }
  return P2Z_OK;
}

// ---- device-resident loop ----------------------------------------------------

template <class T>
void host_store(T* p, T v) {
  __atomic_store_n(p, v, __ATOMIC_RELEASE);
}
template <class T>
T host_load(const T* p) {
  return __atomic_load_n(p, __ATOMIC_ACQUIRE);
}

// Launches the loop on the server's stream.  A relaunch (the previous loop
// ended on the idle timeout) resumes at the first item the previous loop did
// not compute: it computed every item below its final tail and none at or
// above it (the reset kernel, queued behind it, sets the claim counters).
int resident_launch(p2z_server* sv) {
  Resident& R = *sv->res;
  const p2z_server_opts& o = sv->opts;
  P2Z_HIP_CHECK(p2z::launch_serve_reset(R.dev, R.stream));
  host_store(&R.ctl_h->stop, 0u);
  host_store(&R.ctl_h->epoch, ++R.epoch);
  host_store(&R.ctl_h->heartbeat, ++R.heartbeat);
  R.t_beat = now_s();
  p2z::ServeLaunch a{};
  a.trk = sv->trk;
  a.hit = sv->hit;
  a.out = sv->arena_mem;
  a.lanes = sv->lanes;
  a.items = R.items_d;
  a.item_cap = kResItems;
  a.ctl = R.ctl_d;
  a.comps = R.comps_d;
  a.comp_cap = kResSlots;
  a.dev = R.dev;
  a.done = reinterpret_cast<uint32_t*>(R.dev + 1);
  a.idle_ticks = (uint64_t)(kResIdleSeconds * 1e8);
  P2Z_HIP_CHECK(p2z::launch_serve_resident(a, o.width, o.math, R.prefetch, R.tpw, R.grid, R.stream));
  R.running = true;
  R.launches += 1;
  return P2Z_OK;
}

// Keeps the loop alive: heartbeat, and a relaunch when its dispatcher has left
// (idle timeout) while requests wait.  The relaunch queues on the loop's
// stream behind the kernel that is ending, whose workers still finish every
// item below its final tail.
int resident_keepalive(p2z_server* sv) {
  Resident& R = *sv->res;
  // the dispatcher polls this word over PCIe: one store per kBeatSeconds is
  // enough against an idle timeout of kResIdleSeconds
  const double now = now_s();
  if (now - R.t_beat >= kBeatSeconds) {
    host_store(&R.ctl_h->heartbeat, ++R.heartbeat);
    R.t_beat = now;
  }
  if (R.running && host_load(&R.ctl_h->exited) == R.epoch) R.running = false;
  if (!R.running && !R.unfinished.empty()) P2Z_TRY(resident_launch(sv));
  return P2Z_OK;
}

// Collects the completion records the loop has written.
int resident_harvest(p2z_server* sv) {
  Resident& R = *sv->res;
  for (;;) {
    p2z::ServeCompletion& c = R.comps_h[R.comp_read % kResSlots];
    if (host_load(&c.seq) != R.comp_read + 1) break;
    const int32_t slot = host_load(&c.req_slot);
    if (slot < 0 || slot >= kResSlots || R.slot_req[slot] < 0)
      return p2z::fail(P2Z_ESTATE, "completion record %llu names slot %d, which has no request",
                       (unsigned long long)R.comp_read, slot);
    const int64_t id = R.slot_req[slot];
    Request& r = sv->reqs.at(id);
    r.done = true;
    r.t_done = now_s();
    sv->ready.push_back(p2z_completion{id, r.out_block, r.nblocks, r.t_arrival, r.t_admit, r.t_launch, r.t_done});
    R.unfinished.erase(r.first_item);
    R.slot_req[slot] = -1;
    R.free_slots.push_back(slot);
    R.comp_read += 1;
  }
  return P2Z_OK;
}

// Items a request of nblocks descriptors takes at width W.
int64_t resident_items(int64_t nblocks, int W) {
  return W <= 64 ? (nblocks * W + 63) / 64 : 2 * nblocks;
}

// Publishes an admitted request: its descriptors longest-first, packed into
// work items in the ring, then one release store of the new tail.
void resident_publish(p2z_server* sv, int64_t id, const p2z_block_desc* blocks, Request& r, int32_t slot) {
  Resident& R = *sv->res;
  const int W = sv->opts.width;
  const int64_t nb = r.nblocks;
  // longest first; an event's blocks usually arrive that way already
  const p2z_block_desc* d = blocks;
  auto longer = [](const p2z_block_desc& a, const p2z_block_desc& b) { return a.nlayer > b.nlayer; };
  if (!std::is_sorted(blocks, blocks + nb, longer)) {
    R.sorted.assign(blocks, blocks + nb);
    std::stable_sort(R.sorted.begin(), R.sorted.end(), longer);
    d = R.sorted.data();
  }
  const int64_t n = resident_items(nb, W);
  const int per = W <= 64 ? 64 / W : 1;
  const int64_t at = r.out_block;
  for (int64_t i = 0; i < n; ++i) {
    p2z::ServeItem& it = R.items_h[(R.published + i) % kResItems];
    const int64_t first = W <= 64 ? i * per : i / 2;
    const int nd = (int)std::min<int64_t>(per, nb - first);
    for (int k = 0; k < nd; ++k) {
      const p2z_block_desc& s = d[first + k];
      it.desc[k] = p2z::BlockDesc{s.trk_block, s.hit_block, s.out_block + at, s.nlayer, s.event};
    }
    it.ndesc = nd;
    it.half = W <= 64 ? 0 : (int32_t)(i % 2);
    it.req_slot = slot;
    it.req_items = (int32_t)n;
  }
  r.slot = slot;
  r.first_item = R.published;
  R.slot_req[slot] = id;
  R.unfinished.emplace(R.published, id);
  R.published += (uint64_t)n;
  host_store(&R.ctl_h->tail, R.published);
  r.t_launch = now_s();
}

int resident_create(p2z_server* sv) {
  auto* R = new Resident;
  sv->res = R;
  const p2z_server_opts& o = sv->opts;
  R->tpw = o.tracks_per_wg ? o.tracks_per_wg : sv->ctx->cfg.tracks_per_wg;
  R->prefetch = o.prefetch ? o.prefetch : 2;
  const int wpc = o.waves_per_cu ? o.waves_per_cu : 8;
  R->grid = std::max<int64_t>(2 * p2z::kServeStripes, ((int64_t)sv->ctx->cus * wpc * 64 + R->tpw - 1) / R->tpw);
  R->slot_req.assign(kResSlots, -1);
  for (int32_t i = kResSlots - 1; i >= 0; --i) R->free_slots.push_back(i);
  P2Z_HIP_CHECK(hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking));
  const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
  P2Z_HIP_CHECK(hipHostMalloc((void**)&R->items_h, kResItems * sizeof(p2z::ServeItem), fl));
  P2Z_HIP_CHECK(hipHostMalloc((void**)&R->ctl_h, sizeof(p2z::ServeCtl), fl));
  P2Z_HIP_CHECK(hipHostMalloc((void**)&R->comps_h, kResSlots * sizeof(p2z::ServeCompletion), fl));
  std::memset(R->items_h, 0, kResItems * sizeof(p2z::ServeItem));
  std::memset(R->ctl_h, 0, sizeof(p2z::ServeCtl));
  std::memset(R->comps_h, 0, kResSlots * sizeof(p2z::ServeCompletion));
  P2Z_HIP_CHECK(hipHostGetDevicePointer((void**)&R->items_d, R->items_h, 0));
  P2Z_HIP_CHECK(hipHostGetDevicePointer((void**)&R->ctl_d, R->ctl_h, 0));
  P2Z_HIP_CHECK(hipHostGetDevicePointer((void**)&R->comps_d, R->comps_h, 0));
  const size_t dev_bytes = sizeof(p2z::ServeDev) + kResSlots * sizeof(uint32_t);
  P2Z_HIP_CHECK(hipMalloc((void**)&R->dev, dev_bytes));
  P2Z_HIP_CHECK(hipMemsetAsync(R->dev, 0, dev_bytes, R->stream));
  P2Z_TRY(resident_launch(sv));
  return P2Z_OK;
}

// Ends the loop (after the published items) and frees its memory.
void resident_destroy(p2z_server* sv) {
  Resident* R = sv->res;
  if (!R) return;
  if (R->ctl_h) host_store(&R->ctl_h->stop, 1u);
  if (R->stream) (void)hipStreamSynchronize(R->stream);
  if (R->dev) (void)hipFree(R->dev);
  if (R->items_h) (void)hipHostFree(R->items_h);
  if (R->ctl_h) (void)hipHostFree(R->ctl_h);
  if (R->comps_h) (void)hipHostFree(R->comps_h);
  if (R->stream) (void)hipStreamDestroy(R->stream);
  delete R;
  sv->res = nullptr;
}

}  // namespace

extern "C" {

double p2z_now(void) { return now_s(); }

static int server_create_impl(p2z_ctx* ctx, const void* trk, const void* hit, const int32_t* lane_nlayer,
                              const p2z_server_opts* opts, p2z_server** out) {
  if (!ctx || !trk || !hit || !lane_nlayer || !opts || !out) return p2z::fail(P2Z_EINVAL, "null argument");
  *out = nullptr;
  p2z_server_opts o = *opts;
  if (!p2z::valid_width(o.width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", o.width);
  if (!p2z::valid_math(o.math)) return p2z::fail(P2Z_EINVAL, "unknown math mode %d", o.math);
  if (o.nstreams == 0) o.nstreams = 4;
  if (o.nstreams < 1 || o.nstreams > 16) return p2z::fail(P2Z_EINVAL, "nstreams must be in [1,16]");
  if (o.max_batch_blocks < 0) return p2z::fail(P2Z_EINVAL, "max_batch_blocks must be >= 0");
  // by default a batch fits the descriptor room each stream reserves here: a
  // larger one would reallocate it inside the loop, and freeing device memory
  // waits for every batch in flight (a request larger than the limit still
  // runs, as a batch of its own)
  if (o.max_batch_blocks == 0) o.max_batch_blocks = (int32_t)kSlotDescriptors;
  if (o.tracks_per_wg && (o.tracks_per_wg < 64 || o.tracks_per_wg > 256 || o.tracks_per_wg % 64))
    return p2z::fail(P2Z_EINVAL, "tracks_per_wg must be 64, 128, 192 or 256");
  if (o.prefetch < 0 || o.prefetch > 2) return p2z::fail(P2Z_EINVAL, "prefetch must be 0, 1 or 2");
  if (o.waves_per_cu < 0 || o.waves_per_cu > 32) return p2z::fail(P2Z_EINVAL, "waves_per_cu in [0,32]");
  if (o.items_per_claim < 0 || o.items_per_claim > 64) return p2z::fail(P2Z_EINVAL, "items_per_claim in [0,64]");
  if (o.pool_blocks <= 0 || o.pool_hit_blocks <= 0)
    return p2z::fail(P2Z_EINVAL, "pool_blocks and pool_hit_blocks (the pool's sizes) are required");
  if (o.flags & ~P2Z_SERVE_RESIDENT) return p2z::fail(P2Z_EINVAL, "unknown server flags 0x%x", o.flags);
  const int64_t blk_bytes = (int64_t)p2z::kTrkWords * o.width * sizeof(float);
  const int64_t nblk = o.budget_bytes / blk_bytes;
  if (nblk <= 0) return p2z::fail(P2Z_EINVAL, "budget of %lld bytes holds no result block", (long long)o.budget_bytes);
  P2Z_TRY(p2z::use_device(ctx));
  std::unique_ptr<p2z_server, int (*)(p2z_server*)> owner(new p2z_server, &p2z_server_destroy);
  p2z_server* sv = owner.get();
  sv->ctx = ctx;
  sv->opts = o;
  sv->trk = static_cast<const float*>(trk);
  sv->hit = static_cast<const float*>(hit);
  sv->lanes = lane_nlayer;
  int rc = P2Z_OK;
  const hipError_t e = hipMalloc(&sv->arena_mem, nblk * blk_bytes);
  if (e == hipErrorOutOfMemory) rc = p2z::fail(P2Z_ENOMEM, "arena of %lld bytes does not fit", (long long)(nblk * blk_bytes));
  else rc = p2z::hip_status(e, "hipMalloc");
  if (rc == P2Z_OK) sv->arena = new Arena(nblk);
  if (rc == P2Z_OK && (o.flags & P2Z_SERVE_RESIDENT)) {
    rc = resident_create(sv);
  } else if (rc == P2Z_OK) {
    sv->slots.resize(o.nstreams);
    for (auto& s : sv->slots) {
      if (rc == P2Z_OK) rc = p2z::hip_status(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
      if (rc == P2Z_OK) rc = p2z::hip_status(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate");
      if (rc == P2Z_OK) rc = p2z::hip_status(hipMalloc(&s.counter, p2z::kQueueBytes), "hipMalloc");
      if (rc == P2Z_OK) rc = slot_reserve(s, kSlotDescriptors);
    }
  }
  if (rc != P2Z_OK) return rc;
  *out = owner.release();
  return P2Z_OK;
}

int p2z_server_create(p2z_ctx* ctx, const void* trk, const void* hit, const int32_t* lane_nlayer,
                      const p2z_server_opts* opts, p2z_server** out) {
  return p2z::guarded("p2z_server_create",
                      [&] { return server_create_impl(ctx, trk, hit, lane_nlayer, opts, out); });
}

int p2z_server_destroy(p2z_server* sv) {
  if (!sv) return P2Z_OK;
  if (sv->ctx) (void)p2z::use_device(sv->ctx);
  resident_destroy(sv);
  for (auto& s : sv->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.d_desc) (void)hipFree(s.d_desc);
    if (s.h_desc) (void)hipHostFree(s.h_desc);
    if (s.counter) (void)hipFree(s.counter);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
  }
  if (sv->arena_mem) (void)hipFree(sv->arena_mem);
  delete sv->arena;
  delete sv;
  return P2Z_OK;
}

static int server_submit_impl(p2z_server* sv, int64_t request, const p2z_block_desc* blocks, int64_t nblocks,
                              double t_arrival) {
  if (!sv || (!blocks && nblocks)) return p2z::fail(P2Z_EINVAL, "null argument");
  if (nblocks <= 0) return p2z::fail(P2Z_EINVAL, "a request needs at least one block");
  if (sv->reqs.count(request)) return p2z::fail(P2Z_EINVAL, "request %lld already submitted", (long long)request);
  if (nblocks > sv->arena->total())
    return p2z::fail(P2Z_EINVAL, "request %lld needs %lld result blocks, the arena holds %lld", (long long)request,
                     (long long)nblocks, (long long)sv->arena->total());
  if (p2z_check_blocks(blocks, nblocks, sv->opts.pool_blocks, sv->opts.pool_hit_blocks, nblocks) != P2Z_OK)
    return p2z::fail(P2Z_EINVAL, "request %lld: %s", (long long)request, p2z_last_error());
  if (sv->res) {
    Resident& R = *sv->res;
    const int64_t n = resident_items(nblocks, sv->opts.width);
    if (n > kResItems)
      return p2z::fail(P2Z_EINVAL, "request %lld needs %lld work items, the ring holds %lld", (long long)request,
                       (long long)n, (long long)kResItems);
    P2Z_TRY(p2z::use_device(sv->ctx));
    P2Z_TRY(resident_harvest(sv));
    const uint64_t oldest = R.unfinished.empty() ? R.published : R.unfinished.begin()->first;
    if (R.free_slots.empty() || R.published + (uint64_t)n - oldest > (uint64_t)kResItems)
      return p2z::fail(P2Z_EBUSY, "resident loop full (%lld requests, %llu items in flight)",
                       (long long)(kResSlots - (int64_t)R.free_slots.size()),
                       (unsigned long long)(R.published - oldest));
  }
  const int64_t at = sv->arena->alloc(nblocks);
  if (at < 0) return p2z::fail(P2Z_EBUSY, "arena full (%lld of %lld blocks in use)", (long long)sv->arena->used(),
                               (long long)sv->arena->total());
  Request r;
  r.out_block = at;
  r.nblocks = nblocks;
  r.t_arrival = t_arrival;
  r.t_admit = now_s();
  if (sv->res) {
    // the loop reads its descriptors from the ring: the request keeps none
    Resident& R = *sv->res;
    const int32_t slot = R.free_slots.back();
    R.free_slots.pop_back();
    Request& rr = sv->reqs.emplace(request, std::move(r)).first->second;
    resident_publish(sv, request, blocks, rr, slot);
    return resident_keepalive(sv);
  }
  r.blocks.assign(blocks, blocks + nblocks);
  for (auto& d : r.blocks) d.out_block += at;
  sv->reqs.emplace(request, std::move(r));
  sv->pending.push_back(request);
  return P2Z_OK;
}

int p2z_server_submit(p2z_server* sv, int64_t request, const p2z_block_desc* blocks, int64_t nblocks,
                      double t_arrival) {
  return p2z::guarded("p2z_server_submit",
                      [&] { return server_submit_impl(sv, request, blocks, nblocks, t_arrival); });
}

static int server_pump_impl(p2z_server* sv) {
  if (!sv) return p2z::fail(P2Z_EINVAL, "null server");
  P2Z_TRY(p2z::use_device(sv->ctx));
  if (sv->res) {
    P2Z_TRY(resident_harvest(sv));
    P2Z_TRY(resident_keepalive(sv));
    sv->batches = sv->res->launches;
    return (int)sv->ready.size();
  }
  P2Z_TRY(harvest(sv));
  P2Z_TRY(launch_batches(sv));
  return (int)sv->ready.size();
}

int p2z_server_pump(p2z_server* sv) {
  return p2z::guarded("p2z_server_pump", [&] { return server_pump_impl(sv); });
}

static int server_poll_impl(p2z_server* sv, p2z_completion* out, int32_t cap) {
  if (!sv || (!out && cap > 0)) return p2z::fail(P2Z_EINVAL, "null argument");
  if (cap < 0) return p2z::fail(P2Z_EINVAL, "negative capacity %d", cap);
  const int n = std::min<int>(cap, (int)sv->ready.size());
  for (int i = 0; i < n; ++i) out[i] = sv->ready[i];
  sv->ready.erase(sv->ready.begin(), sv->ready.begin() + n);
  return n;
}

int p2z_server_poll(p2z_server* sv, p2z_completion* out, int32_t cap) {
  return p2z::guarded("p2z_server_poll", [&] { return server_poll_impl(sv, out, cap); });
}

int p2z_server_release(p2z_server* sv, int64_t request) {
  if (!sv) return p2z::fail(P2Z_EINVAL, "null server");
  auto it = sv->reqs.find(request);
  if (it == sv->reqs.end()) return p2z::fail(P2Z_EINVAL, "unknown request %lld", (long long)request);
  if (!it->second.done) return p2z::fail(P2Z_ESTATE, "request %lld is not finished", (long long)request);
  sv->arena->release(it->second.out_block, it->second.nblocks);
  sv->reqs.erase(it);
  return P2Z_OK;
}

const void* p2z_server_arena(const p2z_server* sv) { return sv ? sv->arena_mem : nullptr; }

static int serve_poisson_impl(p2z_server* sv, const p2z_block_desc* blocks, const int64_t* event_blocks,
                              int64_t nevents, double rate, uint64_t seed, int32_t release, p2z_serve_report* rep,
                              double* latency, int64_t* out_block) {
  if (!sv || !blocks || !event_blocks || !rep) return p2z::fail(P2Z_EINVAL, "null argument");
  if (nevents <= 0 || !(rate > 0)) return p2z::fail(P2Z_EINVAL, "need nevents > 0 and rate > 0");
  if (!sv->reqs.empty()) return p2z::fail(P2Z_ESTATE, "the server must be idle");
  // per-event descriptors with out_block counted from 0
  std::vector<std::vector<p2z_block_desc>> ev(nevents);
  for (int64_t e = 0; e < nevents; ++e) {
    const int64_t b0 = event_blocks[e], b1 = event_blocks[e + 1];
    if (b1 <= b0) return p2z::fail(P2Z_EINVAL, "event %lld has no blocks", (long long)e);
    ev[e].assign(blocks + b0, blocks + b1);
    for (int64_t i = 0; i < b1 - b0; ++i) ev[e][i].out_block = i;
  }
  std::mt19937_64 rng(seed);
  std::exponential_distribution<double> gap(rate);
  std::vector<double> arrive(nevents), lat(nevents, 0.0);
  const double t0 = now_s() + 1e-3;
  double t = t0;
  for (int64_t e = 0; e < nevents; ++e) {
    t += gap(rng);
    arrive[e] = t;
  }
  std::deque<int64_t> backlog;
  int64_t next = 0, ndone = 0, inflight = 0, max_inflight = 0, rejections = 0;
  double t_last = t0, t_submit = 0, t_pump = 0;
  std::vector<p2z_completion> comp(256);
  sv->batches = sv->batch_events = 0;
  int64_t launches0 = 0;
  if (sv->res) {
    // the loop that serves the run counts as one launch, relaunches as more
    P2Z_TRY(p2z::use_device(sv->ctx));
    P2Z_TRY(resident_keepalive(sv));
    if (!sv->res->running) P2Z_TRY(resident_launch(sv));
    launches0 = sv->res->launches - 1;
  }
  while (ndone < nevents) {
    const double now = now_s();
    while (next < nevents && arrive[next] <= now) backlog.push_back(next++);
    while (!backlog.empty()) {
      const int64_t e = backlog.front();
      const double ts = now_s();
      const int rc = p2z_server_submit(sv, e, ev[e].data(), (int64_t)ev[e].size(), arrive[e]);
      t_submit += now_s() - ts;
      if (rc == P2Z_EBUSY) {
        // nothing in flight and still no room: the kept results fill the arena
        if (inflight == 0 && !release)
          return p2z::fail(P2Z_ENOMEM, "the arena cannot hold the results kept so far (event %lld)", (long long)e);
        ++rejections;
        break;
      }
      P2Z_TRY(rc);
      backlog.pop_front();
      max_inflight = std::max(max_inflight, ++inflight);
    }
    const double tp = now_s();
    const int r = p2z_server_pump(sv);
    t_pump += now_s() - tp;
    if (r < 0) return r;
    int n;
    while ((n = p2z_server_poll(sv, comp.data(), (int32_t)comp.size())) > 0) {
      for (int i = 0; i < n; ++i) {
        const p2z_completion& c = comp[i];
        lat[c.request] = c.t_done - c.t_arrival;
        if (out_block) out_block[c.request] = c.out_block;
        t_last = std::max(t_last, c.t_done);
        if (release) P2Z_TRY(p2z_server_release(sv, c.request));
        --inflight;
        ++ndone;
      }
    }
  }
  std::vector<double> s = lat;
  std::sort(s.begin(), s.end());
  auto q = [&](double p) { return s[(size_t)std::min<double>((double)(s.size() - 1), std::floor(p * (s.size() - 1) + 0.5))]; };
  double mean = 0;
  for (double v : s) mean += v;
  mean /= (double)s.size();
  *rep = p2z_serve_report{};
  rep->nevents = nevents;
  rep->track_layers = 0;  // the lane layer counts live on the device: the caller fills these two
  rep->tl_per_s = 0;
  rep->batches = sv->res ? sv->res->launches - launches0 : sv->batches;
  rep->max_inflight_events = max_inflight;
  rep->busy_rejections = rejections;
  rep->rate_offered = rate;
  rep->t_span = t_last - arrive[0];
  rep->rate_achieved = nevents / (arrive[nevents - 1] - arrive[0] > 0 ? arrive[nevents - 1] - arrive[0] : 1.0);
  rep->lat_p50 = q(0.5);
  rep->lat_p99 = q(0.99);
  rep->lat_p999 = q(0.999);
  rep->lat_mean = mean;
  rep->lat_max = s.back();
  rep->mean_batch_events = sv->res ? (double)nevents / (double)rep->batches
                                   : (sv->batches ? (double)sv->batch_events / sv->batches : 0.0);
  rep->host_submit_s = t_submit;
  rep->host_pump_s = t_pump;
  if (latency)
    for (int64_t e = 0; e < nevents; ++e) latency[e] = lat[e];
  return P2Z_OK;
}

int p2z_serve_poisson(p2z_server* sv, const p2z_block_desc* blocks, const int64_t* event_blocks, int64_t nevents,
                      double rate, uint64_t seed, int32_t release, p2z_serve_report* rep, double* latency,
                      int64_t* out_block) {
  return p2z::guarded("p2z_serve_poisson", [&] {
    return serve_poisson_impl(sv, blocks, event_blocks, nevents, rate, seed, release, rep, latency, out_block);
  });
}

}  // extern "C"
