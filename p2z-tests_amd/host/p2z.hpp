// The reference's host API (src/propagate-toz-test_OMP.cpp of cerati/p2z-tests)
// on top of the p2z C ABI (include/p2z.h).
//
// Same type names, layouts and function names as the reference, so code
// written against it keeps compiling:
//   ATRK, AHIT                                   (OMP:57-67)
//   MP1I MP3F MP6F MP3x3SF MP6x6SF, MPTRK MPHIT  (OMP:69-125)
//   bTk, bHit and the par/pos accessors          (OMP:147-241)
//   prepareTracks, prepareHits, randn            (OMP:127-145, 243-293)
//   propagateToZ, KalmanUpdate_v2                (OMP:623, 503)
//   KalmanUpdate (v1, off the reference's path)  (OMP:440)
// Differences, all forced by moving from compile-time to runtime sizes:
//   * the AoSoA width is the template argument W of the MP* types; the
//     unqualified names use P2Z_BSIZE (default 32), the reference's bsize;
//   * nevts / ntrks / nlayer are fields of p2zConfig() instead of macros;
//   * propagateToZ / KalmanUpdate_v2 run on the GPU (one call per block, as
//     the reference's signature implies); p2zRunDriver runs the whole driver
//     loop (OMP:731-748) in one fused launch, which is what a program that
//     wants speed calls.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "p2z.h"

#ifndef P2Z_BSIZE
#define P2Z_BSIZE 32
#endif

struct ATRK {
  float par[6];
  float cov[21];
  int q;
};
struct AHIT {
  float pos[3];
  float cov[6];
};
static_assert(sizeof(ATRK) == sizeof(p2z_atrk), "ATRK layout");
static_assert(sizeof(AHIT) == sizeof(p2z_ahit), "AHIT layout");

template <int W> struct MP1I_ { int data[1 * W]; };
template <int W> struct MP3F_ { float data[3 * W]; };
template <int W> struct MP6F_ { float data[6 * W]; };
template <int W> struct MP3x3SF_ { float data[6 * W]; };
template <int W> struct MP6x6SF_ { float data[21 * W]; };
template <int W> struct MPTRK_ {
  MP6F_<W> par;
  MP6x6SF_<W> cov;
  MP1I_<W> q;
};
template <int W> struct MPHIT_ {
  MP3F_<W> pos;
  MP3x3SF_<W> cov;
};

using MP1I = MP1I_<P2Z_BSIZE>;
using MP3F = MP3F_<P2Z_BSIZE>;
using MP6F = MP6F_<P2Z_BSIZE>;
using MP3x3SF = MP3x3SF_<P2Z_BSIZE>;
using MP6x6SF = MP6x6SF_<P2Z_BSIZE>;
using MPTRK = MPTRK_<P2Z_BSIZE>;
using MPHIT = MPHIT_<P2Z_BSIZE>;

// Runtime replacements of the reference's size macros (OMP:18-41).
struct P2zConfig {
  int64_t nevts = 100;
  int64_t ntrks = 9600;
  int32_t nlayer = 20;
  int32_t niter = 10;
  double smear = 1e-7;
  int32_t math = P2Z_MATH_REF;  // the reference's own arithmetic; --math fast to opt out
  int device = 0;
};
P2zConfig& p2zConfig();
template <int W> inline int64_t nbOf() { return p2zConfig().ntrks / W; }  // nb (OMP:29)

// ---- accessors (OMP:147-241) ----------------------------------------------
template <int W> inline MPTRK_<W>* bTk(MPTRK_<W>* t, size_t ev, size_t ib) { return &t[ib + nbOf<W>() * ev]; }
template <int W> inline const MPTRK_<W>* bTk(const MPTRK_<W>* t, size_t ev, size_t ib) { return &t[ib + nbOf<W>() * ev]; }
template <int W> inline const MPHIT_<W>* bHit(const MPHIT_<W>* h, size_t ev, size_t ib, size_t lay) {
  return &h[lay + ib * p2zConfig().nlayer + ev * p2zConfig().nlayer * nbOf<W>()];
}
template <int W> inline int q(const MP1I_<W>* b, size_t it) { return b->data[it]; }
template <int W> inline float par(const MP6F_<W>* b, size_t it, size_t ip) { return b->data[it + ip * W]; }
template <int W> inline float par(const MPTRK_<W>* b, size_t it, size_t ip) { return par(&b->par, it, ip); }
template <int W> inline float par(const MPTRK_<W>* t, size_t ev, size_t tk, size_t ip) {
  return par(bTk(t, ev, tk / W), tk % W, ip);
}
template <int W> inline void setpar(MP6F_<W>* b, size_t it, size_t ip, float v) { b->data[it + ip * W] = v; }
template <int W> inline float pos(const MP3F_<W>* b, size_t it, size_t ip) { return b->data[it + ip * W]; }
template <int W> inline float pos(const MPHIT_<W>* b, size_t it, size_t ip) { return pos(&b->pos, it, ip); }
#define P2Z_PAR_ACCESSOR(name, ip)                                                                  \
  template <int W> inline float name(const MP6F_<W>* b, size_t it) { return par(b, it, ip); }        \
  template <int W> inline float name(const MPTRK_<W>* b, size_t it) { return par(b, it, ip); }       \
  template <int W> inline float name(const MPTRK_<W>* t, size_t ev, size_t tk) { return par(t, ev, tk, ip); }
P2Z_PAR_ACCESSOR(x, 0)
P2Z_PAR_ACCESSOR(y, 1)
P2Z_PAR_ACCESSOR(z, 2)
P2Z_PAR_ACCESSOR(ipt, 3)
P2Z_PAR_ACCESSOR(phi, 4)
P2Z_PAR_ACCESSOR(theta, 5)
#undef P2Z_PAR_ACCESSOR

// ---- generator (OMP:127-145, 243-293) -----------------------------------------
float randn(float mu, float sigma);
// Return malloc'd arrays; the caller frees them, as in the reference.
template <int W> MPTRK_<W>* prepareTracks(ATRK inputtrk);
template <int W> MPHIT_<W>* prepareHits(std::vector<AHIT>& inputhits);
inline MPTRK* prepareTracks(ATRK inputtrk) { return prepareTracks<P2Z_BSIZE>(inputtrk); }
inline MPHIT* prepareHits(std::vector<AHIT>& inputhits) { return prepareHits<P2Z_BSIZE>(inputhits); }

// ---- the path, on the GPU ---------------------------------------------------
// One block each, host memory in and out (the reference signatures).
template <int W>
void propagateToZ(const MP6x6SF_<W>* inErr, const MP6F_<W>* inPar, const MP1I_<W>* inChg,
                  const MP3F_<W>* msP, MP6x6SF_<W>* outErr, MP6F_<W>* outPar);
template <int W>
void KalmanUpdate_v2(MP6x6SF_<W>* trkErr, MP6F_<W>* inPar, const MP3x3SF_<W>* hitErr,
                     const MP3F_<W>* msP);
// KalmanUpdate (v1, OMP:375-501; not called by the reference's driver): the
// reference's code as written, or with fixed = true the textbook 3-D update
// in the path's packing (p2z_kalman_update_v1).
template <int W>
void KalmanUpdate(MP6x6SF_<W>* trkErr, MP6F_<W>* inPar, const MP3x3SF_<W>* hitErr, const MP3F_<W>* msP,
                  bool fixed = false);

// The driver loop (OMP:731-748) for all events: outtrk = trk, then nlayer x
// (propagateToZ, KalmanUpdate_v2) per block, on device-resident copies.
// Returns the GPU time of the niter passes in seconds (hipEvents).
template <int W>
double p2zRunDriver(const MPTRK_<W>* trk, const MPHIT_<W>* hit, MPTRK_<W>* outtrk, int niter,
                    bool include_data);

// The shared context used by the functions above (created on first use).
p2z_ctx* p2zContext();
// Throws std::runtime_error with p2z_last_error() when rc != P2Z_OK.
void p2zCheck(int rc, const char* what);
