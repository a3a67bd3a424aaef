// The p2z context (opaque in the C ABI); shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <vector>

#include "p2z_internal.hpp"

struct p2z_ctx {
  int device = 0;
  hipStream_t stream = nullptr;            // main stream: device-pointer entry points
  // host-buffer pipeline: every copy in on pipe_in, every copy out on
  // pipe_out, each staging slot's compute on its own stream in pipe; per slot,
  // events mark its inputs copied, its kernel done and its output copied
  std::vector<hipStream_t> pipe;
  hipStream_t pipe_in = nullptr, pipe_out = nullptr;
  std::vector<hipEvent_t> ev_in, ev_done, ev_out;
  hipEvent_t t0 = nullptr, t1 = nullptr;   // timer
  p2z::LaunchCfg cfg;
  // host-buffer pipeline staging, one slot per compute stream
  std::vector<void*> slot_trk, slot_hit, slot_out;
  size_t slot_trk_bytes = 0, slot_hit_bytes = 0;
  // device report scratch
  double* sum_part = nullptr;
  size_t sum_part_len = 0;
  double* sum_res = nullptr;
  // ragged path: work-queue counter, compute units, captured graphs
  unsigned* counter = nullptr;
  int cus = 0;
  using GraphKey = std::tuple<const void*, const void*, void*, const void*, const void*, int64_t, int, int, int,
                              int, int64_t, int>;
  std::map<GraphKey, hipGraphExec_t> graphs;  // at most kMaxGraphs argument sets
  static constexpr size_t kMaxGraphs = 64;
};

namespace p2z {
// Makes the context's device current.
int use_device(p2z_ctx* c);
}  // namespace p2z

// Waits for every stream of the host-buffer pipeline; the first error, if any.
inline hipError_t sync_pipeline(p2z_ctx* c) {
  hipError_t first = hipSuccess;
  auto wait = [&](hipStream_t s) {
    if (!s) return;
    const hipError_t e = hipStreamSynchronize(s);
    if (first == hipSuccess) first = e;
  };
  wait(c->pipe_in);
  for (auto s : c->pipe) wait(s);
  wait(c->pipe_out);
  return first;
}
