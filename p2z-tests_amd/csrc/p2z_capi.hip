// C ABI of the p2z library: contexts, memory, launches (declared in include/p2z.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "p2z.h"
#include "p2z_capi_util.hpp"
#include "p2z_internal.hpp"
#include "p2z_ctx.hpp"

using p2z::LaunchCfg;



namespace p2z {
int use_device(p2z_ctx* c) {
  P2Z_HIP_CHECK(hipSetDevice(c->device));
  return P2Z_OK;
}
}  // namespace p2z

namespace {

// Device scratch of p2z_summary_device_noalloc: results in the first
// kNoallocResDoubles doubles, the report kernels' per-workgroup partial sums
// after them.  A file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, until the process ends, and every call on
// that GPU shares it.  Host code never passes the array's name to a kernel or
// a copy (on the host it is only a stand-in for the symbol); it uses the
// device address hipGetSymbolAddress returns, fetched once per device when
// the first context on that device is created.
constexpr size_t kNoallocDoubles = (size_t)P2Z_SUMMARY_NOALLOC_BYTES / sizeof(double);
constexpr size_t kNoallocResDoubles = 32;
static_assert(P2Z_SUMMARY_NOALLOC_MAX_TRACKS ==
                  (P2Z_SUMMARY_NOALLOC_BYTES / 8 - 32) / 12 * 256,
              "maximum tracks of the header = partial slots of the scratch x 256 tracks");
__device__ __align__(256) double scratch[kNoallocDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // by device ordinal; null until fetched

// Fetches (once per device) the device address of `scratch` on the current
// device, which must be `device`.
int noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return P2Z_OK;
  void* ptr = nullptr;
  P2Z_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return p2z::fail(P2Z_EHIP, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return P2Z_OK;
}

double* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

namespace {
using p2z::use_device;

// Releases the compute streams, events and staging slots of the host-buffer
// pipeline (the caller has waited for them).
void free_pipeline_slots(p2z_ctx* c) {
  for (auto s : c->pipe) (void)hipStreamDestroy(s);
  for (auto* ev : {&c->ev_in, &c->ev_done, &c->ev_out}) {
    for (auto e : *ev) (void)hipEventDestroy(e);
    ev->clear();
  }
  for (auto p : c->slot_trk) (void)hipFree(p);
  for (auto p : c->slot_hit) (void)hipFree(p);
  for (auto p : c->slot_out) (void)hipFree(p);
  c->pipe.clear();
  c->slot_trk.clear(); c->slot_hit.clear(); c->slot_out.clear();
  c->slot_trk_bytes = c->slot_hit_bytes = 0;
}

int apply_opts(p2z_ctx* c, const p2z_opts* o) {
  LaunchCfg cfg;
  int ns = 4;
  if (o) {
    if (o->tracks_per_wg != 0) {
      if (o->tracks_per_wg < 64 || o->tracks_per_wg > 256 || o->tracks_per_wg % 64 != 0)
        return p2z::fail(P2Z_EINVAL, "tracks_per_wg must be 64, 128, 192 or 256 (got %d)",
                         o->tracks_per_wg);
      cfg.tracks_per_wg = o->tracks_per_wg;
    }
    if (o->nstreams != 0) {
      if (o->nstreams < 1 || o->nstreams > 16)
        return p2z::fail(P2Z_EINVAL, "nstreams must be in [1,16] (got %d)", o->nstreams);
      ns = o->nstreams;
    }
    if (o->hit_stage < 0 || o->hit_stage > 2)
      return p2z::fail(P2Z_EINVAL, "hit_stage must be 1 or 2, or 0 for the default (got %d)", o->hit_stage);
    if (o->hit_stage) cfg.hit_stage = o->hit_stage;
  }
  c->cfg = cfg;
  if (!c->pipe_in) P2Z_HIP_CHECK(hipStreamCreateWithFlags(&c->pipe_in, hipStreamNonBlocking));
  if (!c->pipe_out) P2Z_HIP_CHECK(hipStreamCreateWithFlags(&c->pipe_out, hipStreamNonBlocking));
  if ((int)c->pipe.size() != ns) {
    free_pipeline_slots(c);
    c->pipe.assign(ns, nullptr);
    for (auto& s : c->pipe) P2Z_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto* ev : {&c->ev_in, &c->ev_done, &c->ev_out}) {
      ev->assign(ns, nullptr);
      for (auto& e : *ev) P2Z_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
  }
  return P2Z_OK;
}

int check_blocks(const void* a, const void* b, const void* c, int64_t nevts, int64_t nb,
                 int32_t nlayer, int32_t width, int32_t math) {
  if (!a || !b || !c) return p2z::fail(P2Z_EINVAL, "null buffer");
  if (nevts < 0 || nb < 0 || nlayer < 0)
    return p2z::fail(P2Z_EINVAL, "negative size (nevts=%lld nb=%lld nlayer=%d)", (long long)nevts,
                     (long long)nb, nlayer);
  if (!p2z::valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (!p2z::valid_math(math)) return p2z::fail(P2Z_EINVAL, "unknown math mode %d", math);
  if (nevts * nb * (int64_t)width > (int64_t)1 << 31)
    return p2z::fail(P2Z_EINVAL, "more than 2^31 tracks in one call");
  return P2Z_OK;
}

}  // namespace

extern "C" {

int p2z_version(void) { return P2Z_ABI_VERSION; }

int p2z_device_count(int* n) {
  if (!n) return p2z::fail(P2Z_EINVAL, "null argument");
  *n = 0;
  const hipError_t e = hipGetDeviceCount(n);
  if (e == hipErrorNoDevice) { *n = 0; return P2Z_OK; }
  P2Z_HIP_CHECK(e);
  return P2Z_OK;
}

static int create_impl(p2z_ctx** out, int device, const p2z_opts* opts) {
  if (!out) return p2z::fail(P2Z_EINVAL, "null ctx pointer");
  *out = nullptr;
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || device < 0 || device >= n)
    return p2z::fail(P2Z_ENODEV, "no GPU %d (%d visible)", device, e == hipSuccess ? n : 0);
  std::unique_ptr<p2z_ctx, int (*)(p2z_ctx*)> owner(new p2z_ctx, &p2z_destroy);
  p2z_ctx* c = owner.get();
  c->device = device;
  int rc = use_device(c);
  if (rc == P2Z_OK) rc = p2z::hip_status(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  if (rc == P2Z_OK) rc = p2z::hip_status(hipEventCreate(&c->t0), "hipEventCreate");
  if (rc == P2Z_OK) rc = p2z::hip_status(hipEventCreate(&c->t1), "hipEventCreate");
  if (rc == P2Z_OK) rc = noalloc_scratch_init(device);
  if (rc == P2Z_OK) rc = apply_opts(c, opts);
  if (rc == P2Z_OK) rc = p2z::hip_status(hipMalloc((void**)&c->sum_res, 32 * sizeof(double)), "hipMalloc");
  if (rc == P2Z_OK) rc = p2z::hip_status(hipMalloc((void**)&c->counter, p2z::kQueueBytes), "hipMalloc");
  if (rc == P2Z_OK)
    rc = p2z::hip_status(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device),
                         "hipDeviceGetAttribute");
  if (rc != P2Z_OK) return rc;
  *out = owner.release();
  return P2Z_OK;
}

int p2z_create(p2z_ctx** out, int device, const p2z_opts* opts) {
  return p2z::guarded("p2z_create", [&] { return create_impl(out, device, opts); });
}

int p2z_destroy(p2z_ctx* c) {
  if (!c) return P2Z_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)sync_pipeline(c);
  free_pipeline_slots(c);
  if (c->pipe_in) (void)hipStreamDestroy(c->pipe_in);
  if (c->pipe_out) (void)hipStreamDestroy(c->pipe_out);
  if (c->sum_part) (void)hipFree(c->sum_part);
  if (c->sum_res) (void)hipFree(c->sum_res);
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  if (c->counter) (void)hipFree(c->counter);
  if (c->t0) (void)hipEventDestroy(c->t0);
  if (c->t1) (void)hipEventDestroy(c->t1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return P2Z_OK;
}

static int set_opts_impl(p2z_ctx* c, const p2z_opts* o) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(hipStreamSynchronize(c->stream));
  P2Z_HIP_CHECK(sync_pipeline(c));
  return apply_opts(c, o);
}

int p2z_set_opts(p2z_ctx* c, const p2z_opts* o) {
  return p2z::guarded("p2z_set_opts", [&] { return set_opts_impl(c, o); });
}

int p2z_sync(p2z_ctx* c) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(hipStreamSynchronize(c->stream));
  P2Z_HIP_CHECK(sync_pipeline(c));
  return P2Z_OK;
}

int p2z_get_stream(p2z_ctx* c, void** s) {
  if (!c || !s) return p2z::fail(P2Z_EINVAL, "null argument");
  *s = (void*)c->stream;
  return P2Z_OK;
}

int p2z_dev_alloc(p2z_ctx* c, void** p, size_t bytes) {
  if (!c || !p) return p2z::fail(P2Z_EINVAL, "null argument");
  *p = nullptr;
  P2Z_TRY(use_device(c));
  if (bytes == 0) return P2Z_OK;
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipErrorOutOfMemory) return p2z::fail(P2Z_ENOMEM, "hipMalloc(%zu) out of memory", bytes);
  P2Z_HIP_CHECK(e);
  return P2Z_OK;
}

int p2z_dev_free(p2z_ctx* c, void* p) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (!p) return P2Z_OK;
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(hipFree(p));
  return P2Z_OK;
}

int p2z_host_alloc(p2z_ctx* c, void** p, size_t bytes) {
  if (!c || !p) return p2z::fail(P2Z_EINVAL, "null argument");
  *p = nullptr;
  P2Z_TRY(use_device(c));
  if (bytes == 0) return P2Z_OK;
  const hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e == hipErrorOutOfMemory) return p2z::fail(P2Z_ENOMEM, "hipHostMalloc(%zu) out of memory", bytes);
  P2Z_HIP_CHECK(e);
  return P2Z_OK;
}

int p2z_host_free(p2z_ctx* c, void* p) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (!p) return P2Z_OK;
  P2Z_HIP_CHECK(hipHostFree(p));
  return P2Z_OK;
}

int p2z_memcpy_h2d(p2z_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (bytes && (!dst || !src))) return p2z::fail(P2Z_EINVAL, "null argument");
  P2Z_TRY(use_device(c));
  if (bytes) P2Z_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return P2Z_OK;
}

int p2z_memcpy_d2h(p2z_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (bytes && (!dst || !src))) return p2z::fail(P2Z_EINVAL, "null argument");
  P2Z_TRY(use_device(c));
  if (bytes) P2Z_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return P2Z_OK;
}

int p2z_timer_start(p2z_ctx* c) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(hipEventRecord(c->t0, c->stream));
  return P2Z_OK;
}

int p2z_timer_stop(p2z_ctx* c, float* ms) {
  if (!c || !ms) return p2z::fail(P2Z_EINVAL, "null argument");
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(hipEventRecord(c->t1, c->stream));
  P2Z_HIP_CHECK(hipEventSynchronize(c->t1));
  P2Z_HIP_CHECK(hipEventElapsedTime(ms, c->t0, c->t1));
  return P2Z_OK;
}

int p2z_propagate_to_z(p2z_ctx* c, const float* inErr, const float* inPar, const int32_t* inChg,
                       const float* msP, float* outErr, float* outPar, int64_t nblocks,
                       int64_t trk_stride, int64_t hit_stride, int32_t width, int32_t math) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (!inErr || !inPar || !inChg || !msP || !outErr || !outPar)
    return p2z::fail(P2Z_EINVAL, "null buffer");
  if (!p2z::valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (!p2z::valid_math(math)) return p2z::fail(P2Z_EINVAL, "unknown math mode %d", math);
  if (nblocks < 0 || nblocks * (int64_t)width > (int64_t)1 << 31)
    return p2z::fail(P2Z_EINVAL, "bad block count %lld", (long long)nblocks);
  if (trk_stride < 0 || hit_stride < 3 * width)
    return p2z::fail(P2Z_EINVAL, "bad strides (trk %lld, hit %lld)", (long long)trk_stride, (long long)hit_stride);
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(p2z::launch_propagate(inErr, inPar, inChg, msP, outErr, outPar, nblocks, trk_stride,
                                      hit_stride, width, math, c->cfg, c->stream));
  return P2Z_OK;
}

int p2z_kalman_update(p2z_ctx* c, float* trkErr, float* inPar, const float* hitErr, const float* msP,
                      int64_t nblocks, int64_t trk_stride, int64_t hit_stride, int32_t width,
                      int32_t math) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (!trkErr || !inPar || !hitErr || !msP) return p2z::fail(P2Z_EINVAL, "null buffer");
  if (!p2z::valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (!p2z::valid_math(math)) return p2z::fail(P2Z_EINVAL, "unknown math mode %d", math);
  if (nblocks < 0 || nblocks * (int64_t)width > (int64_t)1 << 31)
    return p2z::fail(P2Z_EINVAL, "bad block count %lld", (long long)nblocks);
  if (trk_stride < 0 || hit_stride < 0)
    return p2z::fail(P2Z_EINVAL, "bad strides (trk %lld, hit %lld)", (long long)trk_stride, (long long)hit_stride);
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(p2z::launch_update(trkErr, inPar, hitErr, msP, nblocks, trk_stride, hit_stride, width,
                                   math, c->cfg, c->stream));
  return P2Z_OK;
}

int p2z_kalman_update_v1(p2z_ctx* c, float* trkErr, float* inPar, const float* hitErr, const float* msP,
                         int64_t nblocks, int64_t trk_stride, int64_t hit_stride, int32_t width, uint32_t flags) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (!trkErr || !inPar || !hitErr || !msP) return p2z::fail(P2Z_EINVAL, "null buffer");
  if (!p2z::valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (flags & ~P2Z_V1_FIXED) return p2z::fail(P2Z_EINVAL, "unknown v1 flags 0x%x", flags);
  if (nblocks < 0 || nblocks * (int64_t)width > (int64_t)1 << 31)
    return p2z::fail(P2Z_EINVAL, "bad block count %lld", (long long)nblocks);
  if (trk_stride < 0 || hit_stride < 0)
    return p2z::fail(P2Z_EINVAL, "bad strides (trk %lld, hit %lld)", (long long)trk_stride, (long long)hit_stride);
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(p2z::launch_update_v1(trkErr, inPar, hitErr, msP, nblocks, trk_stride, hit_stride, width,
                                      (flags & P2Z_V1_FIXED) != 0, c->cfg, c->stream));
  return P2Z_OK;
}

int p2z_run(p2z_ctx* c, const void* trk, const void* hit, void* out, int64_t nevts, int64_t nb,
            int32_t nlayer, int32_t width, int32_t math) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  P2Z_TRY(check_blocks(trk, hit, out, nevts, nb, nlayer, width, math));
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(p2z::launch_fused((const float*)trk, (const float*)hit, (float*)out, nevts * nb, nlayer,
                                  width, math, c->cfg, c->stream));
  return P2Z_OK;
}

int p2z_sincosf_ref(p2z_ctx* c, uint32_t first_bits, int64_t n, float* s, float* co) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (n < 0 || (uint64_t)first_bits + (uint64_t)n > (1ull << 32))
    return p2z::fail(P2Z_EINVAL, "bit range %u + %lld leaves the 2^32 float patterns", first_bits, (long long)n);
  if (n > 0 && (!s || !co)) return p2z::fail(P2Z_EINVAL, "null buffer");
  P2Z_TRY(use_device(c));
  P2Z_HIP_CHECK(p2z::launch_sincosf_range(first_bits, n, s, co, c->stream));
  return P2Z_OK;
}

namespace {

// True for page-locked host memory (hipHostMalloc); pageable memory makes
// hipPointerGetAttributes fail or report unregistered memory.
bool host_pinned(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

}  // namespace

static int run_host_impl(p2z_ctx* c, const void* trk, const void* hit, void* out, int64_t nevts, int64_t nb,
                         int32_t nlayer, int32_t width, int32_t math) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  P2Z_TRY(check_blocks(trk, hit, out, nevts, nb, nlayer, width, math));
  P2Z_TRY(use_device(c));
  if (nevts == 0 || nb == 0) return P2Z_OK;
  const int ns = (int)c->pipe.size();
  const bool hit_pinned = nlayer > 0 && host_pinned(hit);
  // Chunks of whole events through ns staging slots, in three stages: every
  // copy in runs on pipe_in, one after another at the full PCIe rate (copies
  // in on several streams would share the link and reach their kernels, then
  // their copies out, at the same time, leaving the link idle while they
  // drain); chunk k's kernel runs on slot k % ns's stream; every copy out runs
  // on pipe_out, beside the copies in.  Four chunks per slot keep the last
  // kernel and copy out, which nothing overlaps, short.  A slot is reused once
  // its previous kernel has read the inputs (ev_done) and its previous output
  // has been copied out (ev_out).
  const int64_t nchunks = std::min<int64_t>(nevts, 4 * (int64_t)ns);
  const int64_t per = (nevts + nchunks - 1) / nchunks;
  const size_t trk_ev = (size_t)nb * p2z::kTrkWords * width * sizeof(float);
  const size_t hit_ev = (size_t)nb * nlayer * p2z::kHitWords * width * sizeof(float);
  const size_t need_trk = trk_ev * per, need_hit = std::max<size_t>(hit_ev * per, 1);
  if (c->slot_trk_bytes < need_trk || c->slot_hit_bytes < need_hit || (int)c->slot_trk.size() != ns) {
    P2Z_HIP_CHECK(sync_pipeline(c));
    for (auto p : c->slot_trk) (void)hipFree(p);
    for (auto p : c->slot_hit) (void)hipFree(p);
    for (auto p : c->slot_out) (void)hipFree(p);
    c->slot_trk.assign(ns, nullptr); c->slot_hit.assign(ns, nullptr); c->slot_out.assign(ns, nullptr);
    c->slot_trk_bytes = c->slot_hit_bytes = 0;
    for (int i = 0; i < ns; ++i) {
      hipError_t e = hipMalloc(&c->slot_trk[i], need_trk);
      if (e == hipSuccess) e = hipMalloc(&c->slot_hit[i], need_hit);
      if (e == hipSuccess) e = hipMalloc(&c->slot_out[i], need_trk);
      if (e == hipErrorOutOfMemory) return p2z::fail(P2Z_ENOMEM, "staging buffers do not fit");
      P2Z_HIP_CHECK(e);
    }
    c->slot_trk_bytes = need_trk;
    c->slot_hit_bytes = need_hit;
  }
  // every path out waits for the pipeline streams: no copy from or to the
  // caller's host buffers is still in flight when this returns
  const int rc = [&]() -> int {
    for (int64_t k = 0; k < nchunks; ++k) {
      const int64_t e0 = k * per, e1 = std::min(nevts, e0 + per);
      if (e0 >= e1) break;
      const int i = (int)(k % ns);
      const bool reuse = k >= ns;  // slot i held chunk k - ns
      const int64_t ne = e1 - e0;
      if (reuse) P2Z_HIP_CHECK(hipStreamWaitEvent(c->pipe_in, c->ev_done[i], 0));
      P2Z_HIP_CHECK(hipMemcpyAsync(c->slot_trk[i], (const char*)trk + e0 * trk_ev, ne * trk_ev,
                                   hipMemcpyHostToDevice, c->pipe_in));
      if (hit_pinned) {
        // only the words the path reads: pos x, y, z and cov xx, xy, yy of each
        // hit record (the record's last 3W words, cov xz, yz, zz, are never
        // read, OMP:510-512); one row per record, same pitch on both sides
        const size_t pitch = (size_t)p2z::kHitWords * width * sizeof(float);
        P2Z_HIP_CHECK(hipMemcpy2DAsync(c->slot_hit[i], pitch, (const char*)hit + e0 * hit_ev, pitch,
                                       (size_t)p2z::kHitUsedWords * width * sizeof(float),
                                       (size_t)(ne * nb * nlayer), hipMemcpyHostToDevice, c->pipe_in));
      } else {
        // pageable memory: one contiguous copy (a strided copy from pageable
        // memory goes row by row through the runtime's staging)
        P2Z_HIP_CHECK(hipMemcpyAsync(c->slot_hit[i], (const char*)hit + e0 * hit_ev, ne * hit_ev,
                                     hipMemcpyHostToDevice, c->pipe_in));
      }
      P2Z_HIP_CHECK(hipEventRecord(c->ev_in[i], c->pipe_in));
      hipStream_t s = c->pipe[i];
      P2Z_HIP_CHECK(hipStreamWaitEvent(s, c->ev_in[i], 0));
      if (reuse) P2Z_HIP_CHECK(hipStreamWaitEvent(s, c->ev_out[i], 0));
      P2Z_HIP_CHECK(p2z::launch_fused((const float*)c->slot_trk[i], (const float*)c->slot_hit[i],
                                      (float*)c->slot_out[i], ne * nb, nlayer, width, math, c->cfg, s));
      P2Z_HIP_CHECK(hipEventRecord(c->ev_done[i], s));
      P2Z_HIP_CHECK(hipStreamWaitEvent(c->pipe_out, c->ev_done[i], 0));
      P2Z_HIP_CHECK(hipMemcpyAsync((char*)out + e0 * trk_ev, c->slot_out[i], ne * trk_ev,
                                   hipMemcpyDeviceToHost, c->pipe_out));
      P2Z_HIP_CHECK(hipEventRecord(c->ev_out[i], c->pipe_out));
    }
    return P2Z_OK;
  }();
  const hipError_t e = sync_pipeline(c);
  if (rc == P2Z_OK) P2Z_HIP_CHECK(e);
  return rc;
}

int p2z_run_host(p2z_ctx* c, const void* trk, const void* hit, void* out, int64_t nevts, int64_t nb,
                 int32_t nlayer, int32_t width, int32_t math) {
  return p2z::guarded("p2z_run_host",
                      [&] { return run_host_impl(c, trk, hit, out, nevts, nb, nlayer, width, math); });
}

static int run_ragged_impl(p2z_ctx* c, const void* trk, const void* hit, void* out, const int32_t* lane_nlayer,
                           const p2z_block_desc* blocks, int64_t nblocks, int32_t width, int32_t math, uint32_t flags,
                           const p2z_ragged_opts* opts) {
  if (!c) return p2z::fail(P2Z_EINVAL, "null ctx");
  if (!trk || !hit || !out || !lane_nlayer || !blocks) return p2z::fail(P2Z_EINVAL, "null buffer");
  if (!p2z::valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (!p2z::valid_math(math)) return p2z::fail(P2Z_EINVAL, "unknown math mode %d", math);
  if (nblocks < 0 || nblocks * (int64_t)width > ((int64_t)1 << 31) - 64)
    return p2z::fail(P2Z_EINVAL, "bad block count %lld", (long long)nblocks);
  if (flags & ~(P2Z_RAGGED_PERSISTENT | P2Z_RAGGED_GRAPH)) return p2z::fail(P2Z_EINVAL, "unknown flags 0x%x", flags);
  int tpw = c->cfg.tracks_per_wg, prefetch = 2, wpc = 8, claim = p2z::kDefaultClaim;
  if (opts) {
    if (opts->tracks_per_wg) {
      if (opts->tracks_per_wg < 64 || opts->tracks_per_wg > 256 || opts->tracks_per_wg % 64)
        return p2z::fail(P2Z_EINVAL, "tracks_per_wg must be 64, 128, 192 or 256");
      tpw = opts->tracks_per_wg;
    }
    if (opts->prefetch) {
      if (opts->prefetch < 1 || opts->prefetch > 2) return p2z::fail(P2Z_EINVAL, "prefetch must be 1 or 2");
      prefetch = opts->prefetch;
    }
    if (opts->waves_per_cu) {
      if (opts->waves_per_cu < 1 || opts->waves_per_cu > 32) return p2z::fail(P2Z_EINVAL, "waves_per_cu in [1,32]");
      wpc = opts->waves_per_cu;
    }
    if (opts->items_per_claim) {
      if (opts->items_per_claim < 1 || opts->items_per_claim > 64)
        return p2z::fail(P2Z_EINVAL, "items_per_claim in [1,64]");
      claim = opts->items_per_claim;
    }
  }
  P2Z_TRY(use_device(c));
  if (nblocks == 0) return P2Z_OK;
  const bool persistent = flags & P2Z_RAGGED_PERSISTENT;
  int64_t grid = ((int64_t)c->cus * wpc * 64 + tpw - 1) / tpw;
  if (persistent) {
    // no more workgroups than the CUs hold at once
    const int per_cu = p2z::ragged_persistent_groups_per_cu(width, math, prefetch, tpw);
    if (per_cu > 0) grid = std::min<int64_t>(grid, (int64_t)per_cu * c->cus);
  }
  const auto* bd = reinterpret_cast<const p2z::BlockDesc*>(blocks);
  auto launch = [&](hipStream_t s) {
    return p2z::launch_ragged((const float*)trk, (const float*)hit, (float*)out, lane_nlayer, bd, nblocks, width,
                              math, persistent, prefetch, tpw, grid, claim, c->counter, s);
  };
  if (!(flags & P2Z_RAGGED_GRAPH)) {
    P2Z_HIP_CHECK(launch(c->stream));
    return P2Z_OK;
  }
  // the launch (counter reset + persistent kernel, or the one-shot kernel),
  // captured once per argument set
  const p2z_ctx::GraphKey key{trk, hit, out, lane_nlayer, blocks, nblocks, width, math, prefetch, tpw,
                              persistent ? grid : -1, claim};
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    if (c->graphs.size() >= p2z_ctx::kMaxGraphs) {
      // a full cache is emptied once the graphs launched from it have finished
      P2Z_HIP_CHECK(hipStreamSynchronize(c->stream));
      for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
      c->graphs.clear();
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    P2Z_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = launch(c->stream);
    const hipError_t ce = hipStreamEndCapture(c->stream, &g);
    P2Z_HIP_CHECK(le);
    P2Z_HIP_CHECK(ce);
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    P2Z_HIP_CHECK(ie);
    it = c->graphs.emplace(key, ge).first;
  }
  P2Z_HIP_CHECK(hipGraphLaunch(it->second, c->stream));
  return P2Z_OK;
}

int p2z_run_ragged(p2z_ctx* c, const void* trk, const void* hit, void* out, const int32_t* lane_nlayer,
                   const p2z_block_desc* blocks, int64_t nblocks, int32_t width, int32_t math, uint32_t flags,
                   const p2z_ragged_opts* opts) {
  return p2z::guarded("p2z_run_ragged", [&] {
    return run_ragged_impl(c, trk, hit, out, lane_nlayer, blocks, nblocks, width, math, flags, opts);
  });
}

namespace {

int summary_args(p2z_ctx* c, const void* out, int64_t nevts, int64_t nb, int32_t width, int64_t ntrks,
                 const float* lasthit, p2z_summary* s) {
  if (!c || !out || !lasthit || !s) return p2z::fail(P2Z_EINVAL, "null argument");
  if (!p2z::valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (nevts <= 0 || ntrks <= 0 || ntrks > nb * width)
    return p2z::fail(P2Z_EINVAL, "need nevts > 0 and 0 < ntrks <= nb*width");
  return P2Z_OK;
}

// Runs the report kernels with the given scratch, then copies the 18 results
// to the host and unpacks them.
int summary_run(p2z_ctx* c, const void* out, int64_t nevts, int64_t nb, int32_t width, int64_t ntrks,
                const float* lasthit, double* part, double* res, p2z_summary* s) {
  P2Z_HIP_CHECK(p2z::launch_summary((const float*)out, nevts, nb, width, ntrks, lasthit[0], lasthit[1],
                                    lasthit[2], part, res, c->stream));
  double r[18];
  P2Z_HIP_CHECK(hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, c->stream));
  P2Z_HIP_CHECK(hipStreamSynchronize(c->stream));
  std::memset(s, 0, sizeof *s);
  s->avgx = r[0]; s->avgy = r[1]; s->avgz = r[2];
  s->avgdx = r[3]; s->avgdy = r[4]; s->avgdz = r[5];
  s->avgpt = r[6]; s->avgphi = r[7]; s->avgtheta = r[8];
  s->nnans = (int32_t)r[9]; s->nfail = (int32_t)r[10]; s->ngood = (int32_t)r[11];
  s->stdx = r[12]; s->stdy = r[13]; s->stdz = r[14];
  s->stddx = r[15]; s->stddy = r[16]; s->stddz = r[17];
  return P2Z_OK;
}

}  // namespace

int p2z_summary_device(p2z_ctx* c, const void* out, int64_t nevts, int64_t nb, int32_t width,
                       int64_t ntrks, const float lasthit[3], p2z_summary* s) {
  P2Z_TRY(summary_args(c, out, nevts, nb, width, ntrks, lasthit, s));
  P2Z_TRY(use_device(c));
  const size_t need = p2z::summary_scratch_doubles(nevts, ntrks);
  if (c->sum_part_len < need) {
    if (c->sum_part) P2Z_HIP_CHECK(hipFree(c->sum_part));
    c->sum_part = nullptr;
    c->sum_part_len = 0;
    P2Z_HIP_CHECK(hipMalloc((void**)&c->sum_part, need * sizeof(double)));
    c->sum_part_len = need;
  }
  return summary_run(c, out, nevts, nb, width, ntrks, lasthit, c->sum_part, c->sum_res, s);
}

int p2z_summary_device_noalloc(p2z_ctx* c, const void* out, int64_t nevts, int64_t nb, int32_t width,
                               int64_t ntrks, const float lasthit[3], p2z_summary* s) {
  P2Z_TRY(summary_args(c, out, nevts, nb, width, ntrks, lasthit, s));
  if (nevts > P2Z_SUMMARY_NOALLOC_MAX_TRACKS / ntrks)
    return p2z::fail(P2Z_EINVAL, "%lld events x %lld tracks exceed the %lld tracks of the static report scratch",
                     (long long)nevts, (long long)ntrks, (long long)P2Z_SUMMARY_NOALLOC_MAX_TRACKS);
  P2Z_TRY(use_device(c));
  double* base = noalloc_scratch(c->device);
  if (!base) return p2z::fail(P2Z_ESTATE, "static report scratch of GPU %d not set up", c->device);
  // every partial slot the kernels read is written first in the same call,
  // and the results are written before they are read: nothing to clear
  return summary_run(c, out, nevts, nb, width, ntrks, lasthit, base + kNoallocResDoubles, base, s);
}

}  // extern "C"
