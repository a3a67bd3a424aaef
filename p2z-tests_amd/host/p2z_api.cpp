// The reference host API (p2z.hpp) implemented over the C ABI.
#include "p2z.hpp"

#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

P2zConfig& p2zConfig() {
  static P2zConfig cfg;
  return cfg;
}

void p2zCheck(int rc, const char* what) {
  if (rc != P2Z_OK)
    throw std::runtime_error(std::string(what) + ": " + p2z_last_error() + " (rc=" + std::to_string(rc) + ")");
}

namespace {
struct CtxHolder {
  p2z_ctx* ctx = nullptr;
  ~CtxHolder() { p2z_destroy(ctx); }
};
CtxHolder g_ctx;

// Device buffer owned by the shared context.
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { p2zCheck(p2z_dev_alloc(p2zContext(), &p, bytes), "p2z_dev_alloc"); }
  ~DevBuf() { p2z_dev_free(p2zContext(), p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};
}  // namespace

p2z_ctx* p2zContext() {
  if (!g_ctx.ctx) p2zCheck(p2z_create(&g_ctx.ctx, p2zConfig().device, nullptr), "p2z_create");
  return g_ctx.ctx;
}

float randn(float mu, float sigma) { return p2z_randn(mu, sigma); }

template <int W>
MPTRK_<W>* prepareTracks(ATRK inputtrk) {
  const P2zConfig& c = p2zConfig();
  const int64_t nb = nbOf<W>();
  auto* res = static_cast<MPTRK_<W>*>(std::malloc(sizeof(MPTRK_<W>) * c.nevts * nb));
  p2zCheck(p2z_prepare_tracks(reinterpret_cast<const p2z_atrk*>(&inputtrk), res, c.nevts, nb, W, c.smear),
           "prepareTracks");
  return res;
}

template <int W>
MPHIT_<W>* prepareHits(std::vector<AHIT>& inputhits) {
  const P2zConfig& c = p2zConfig();
  const int64_t nb = nbOf<W>();
  auto* res = static_cast<MPHIT_<W>*>(std::malloc(sizeof(MPHIT_<W>) * c.nlayer * c.nevts * nb));
  p2zCheck(p2z_prepare_hits(reinterpret_cast<const p2z_ahit*>(inputhits.data()),
                            (int32_t)inputhits.size(), res, c.nevts, nb, W, c.nlayer, c.smear),
           "prepareHits");
  return res;
}

// One block: stage the pieces into a device MPTRK/MPHIT pair, run, copy back.
template <int W>
void propagateToZ(const MP6x6SF_<W>* inErr, const MP6F_<W>* inPar, const MP1I_<W>* inChg,
                  const MP3F_<W>* msP, MP6x6SF_<W>* outErr, MP6F_<W>* outPar) {
  p2z_ctx* ctx = p2zContext();
  MPTRK_<W> t;
  MPHIT_<W> h;
  std::memcpy(&t.cov, inErr, sizeof t.cov);
  std::memcpy(&t.par, inPar, sizeof t.par);
  std::memcpy(&t.q, inChg, sizeof t.q);
  std::memcpy(&h.pos, msP, sizeof h.pos);
  std::memset(&h.cov, 0, sizeof h.cov);
  DevBuf dt(sizeof t), dh(sizeof h);
  p2zCheck(p2z_memcpy_h2d(ctx, dt.p, &t, sizeof t), "h2d");
  p2zCheck(p2z_memcpy_h2d(ctx, dh.p, &h, sizeof h), "h2d");
  auto* d = static_cast<MPTRK_<W>*>(dt.p);
  auto* e = static_cast<MPHIT_<W>*>(dh.p);
  p2zCheck(p2z_propagate_to_z(ctx, d->cov.data, d->par.data, d->q.data, e->pos.data, d->cov.data,
                              d->par.data, 1, 28 * W, 9 * W, W, p2zConfig().math),
           "propagateToZ");
  p2zCheck(p2z_memcpy_d2h(ctx, &t, dt.p, sizeof t), "d2h");
  p2zCheck(p2z_sync(ctx), "sync");
  std::memcpy(outErr, &t.cov, sizeof t.cov);
  std::memcpy(outPar, &t.par, sizeof t.par);
}

template <int W>
void KalmanUpdate_v2(MP6x6SF_<W>* trkErr, MP6F_<W>* inPar, const MP3x3SF_<W>* hitErr,
                     const MP3F_<W>* msP) {
  p2z_ctx* ctx = p2zContext();
  MPTRK_<W> t;
  MPHIT_<W> h;
  std::memcpy(&t.cov, trkErr, sizeof t.cov);
  std::memcpy(&t.par, inPar, sizeof t.par);
  std::memset(&t.q, 0, sizeof t.q);
  std::memcpy(&h.pos, msP, sizeof h.pos);
  std::memcpy(&h.cov, hitErr, sizeof h.cov);
  DevBuf dt(sizeof t), dh(sizeof h);
  p2zCheck(p2z_memcpy_h2d(ctx, dt.p, &t, sizeof t), "h2d");
  p2zCheck(p2z_memcpy_h2d(ctx, dh.p, &h, sizeof h), "h2d");
  auto* d = static_cast<MPTRK_<W>*>(dt.p);
  auto* e = static_cast<MPHIT_<W>*>(dh.p);
  p2zCheck(p2z_kalman_update(ctx, d->cov.data, d->par.data, e->cov.data, e->pos.data, 1, 28 * W, 9 * W, W,
                             p2zConfig().math),
           "KalmanUpdate_v2");
  p2zCheck(p2z_memcpy_d2h(ctx, &t, dt.p, sizeof t), "d2h");
  p2zCheck(p2z_sync(ctx), "sync");
  std::memcpy(trkErr, &t.cov, sizeof t.cov);
  std::memcpy(inPar, &t.par, sizeof t.par);
}

template <int W>
void KalmanUpdate(MP6x6SF_<W>* trkErr, MP6F_<W>* inPar, const MP3x3SF_<W>* hitErr, const MP3F_<W>* msP,
                  bool fixed) {
  p2z_ctx* ctx = p2zContext();
  MPTRK_<W> t;
  MPHIT_<W> h;
  std::memcpy(&t.cov, trkErr, sizeof t.cov);
  std::memcpy(&t.par, inPar, sizeof t.par);
  std::memset(&t.q, 0, sizeof t.q);
  std::memcpy(&h.pos, msP, sizeof h.pos);
  std::memcpy(&h.cov, hitErr, sizeof h.cov);
  DevBuf dt(sizeof t), dh(sizeof h);
  p2zCheck(p2z_memcpy_h2d(ctx, dt.p, &t, sizeof t), "h2d");
  p2zCheck(p2z_memcpy_h2d(ctx, dh.p, &h, sizeof h), "h2d");
  auto* d = static_cast<MPTRK_<W>*>(dt.p);
  auto* e = static_cast<MPHIT_<W>*>(dh.p);
  p2zCheck(p2z_kalman_update_v1(ctx, d->cov.data, d->par.data, e->cov.data, e->pos.data, 1, 28 * W, 9 * W, W,
                                fixed ? P2Z_V1_FIXED : 0u),
           "KalmanUpdate");
  p2zCheck(p2z_memcpy_d2h(ctx, &t, dt.p, sizeof t), "d2h");
  p2zCheck(p2z_sync(ctx), "sync");
  std::memcpy(trkErr, &t.cov, sizeof t.cov);
  std::memcpy(inPar, &t.par, sizeof t.par);
}

template <int W>
double p2zRunDriver(const MPTRK_<W>* trk, const MPHIT_<W>* hit, MPTRK_<W>* outtrk, int niter,
                    bool include_data) {
  const P2zConfig& c = p2zConfig();
  p2z_ctx* ctx = p2zContext();
  const int64_t nb = nbOf<W>();
  const size_t trk_bytes = sizeof(MPTRK_<W>) * c.nevts * nb;
  const size_t hit_bytes = sizeof(MPHIT_<W>) * c.nlayer * c.nevts * nb;
  float ms = 0.f;
  if (include_data) {
    // transfers inside the timed region (CU4's include-data mode)
    p2zCheck(p2z_timer_start(ctx), "timer");
    for (int it = 0; it < niter; ++it)
      p2zCheck(p2z_run_host(ctx, trk, hit, outtrk, c.nevts, nb, c.nlayer, W, c.math), "p2z_run_host");
    p2zCheck(p2z_timer_stop(ctx, &ms), "timer");
    return ms * 1e-3;
  }
  DevBuf dt(trk_bytes), dh(hit_bytes), dout(trk_bytes);
  p2zCheck(p2z_memcpy_h2d(ctx, dt.p, trk, trk_bytes), "h2d");
  p2zCheck(p2z_memcpy_h2d(ctx, dh.p, hit, hit_bytes), "h2d");
  p2zCheck(p2z_sync(ctx), "sync");
  p2zCheck(p2z_timer_start(ctx), "timer");
  for (int it = 0; it < niter; ++it)
    p2zCheck(p2z_run(ctx, dt.p, dh.p, dout.p, c.nevts, nb, c.nlayer, W, c.math), "p2z_run");
  p2zCheck(p2z_timer_stop(ctx, &ms), "timer");
  p2zCheck(p2z_memcpy_d2h(ctx, outtrk, dout.p, trk_bytes), "d2h");
  p2zCheck(p2z_sync(ctx), "sync");
  return ms * 1e-3;
}

#define P2Z_INSTANTIATE(W)                                                                        \
  template MPTRK_<W>* prepareTracks<W>(ATRK);                                                     \
  template MPHIT_<W>* prepareHits<W>(std::vector<AHIT>&);                                         \
  template void propagateToZ<W>(const MP6x6SF_<W>*, const MP6F_<W>*, const MP1I_<W>*,             \
                                const MP3F_<W>*, MP6x6SF_<W>*, MP6F_<W>*);                        \
  template void KalmanUpdate_v2<W>(MP6x6SF_<W>*, MP6F_<W>*, const MP3x3SF_<W>*, const MP3F_<W>*); \
  template void KalmanUpdate<W>(MP6x6SF_<W>*, MP6F_<W>*, const MP3x3SF_<W>*, const MP3F_<W>*, bool); \
  template double p2zRunDriver<W>(const MPTRK_<W>*, const MPHIT_<W>*, MPTRK_<W>*, int, bool);
P2Z_INSTANTIATE(16)
P2Z_INSTANTIATE(32)
P2Z_INSTANTIATE(64)
P2Z_INSTANTIATE(128)
#undef P2Z_INSTANTIATE
