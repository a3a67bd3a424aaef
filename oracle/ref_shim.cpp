// Reference build wrapper -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference translation unit src/propagate-toz-test_OMP.cpp of
// cerati/p2z-tests from where it lies (the path comes from oracle/Makefile as
// P2Z_REF_SOURCE) and exposes its generator and per-block path functions with C
// linkage, so the oracle restatement can be compared with the real reference
// code per track.  The reference's compile-time sizes (nevts, ntrks, bsize,
// nlayer, NITER, nthreads) are set by the Makefile per build.  Outputs go to
// oracle/_ref/ only.
#define main p2z_reference_main
#include P2Z_REF_SOURCE
#undef main
// The input track and hits (the reference includes this inside main; here at
// file scope so the generator can be driven without running main).
#include "input_track.h"

#include <cstring>

namespace {
// Layer order the reference builds in main (OMP:693-695): hit 25 first.
std::vector<AHIT> ReferenceLayerHits() {
  return {inputhit25, inputhit24, inputhit23, inputhit22, inputhit21, inputhit20, inputhit19,
          inputhit18, inputhit17, inputhit16, inputhit15, inputhit14, inputhit13, inputhit12,
          inputhit11, inputhit10, inputhit09, inputhit08, inputhit07, inputhit06, inputhit05,
          inputhit04, inputhit03, inputhit02, inputhit01, inputhit00};
}
}  // namespace

extern "C" {

int ref_bsize() { return bsize; }
int ref_nevts() { return nevts; }
int ref_ntrks() { return ntrks; }
int ref_nb() { return nb; }
int ref_nlayer() { return nlayer; }
int ref_niter() { return NITER; }

// srand(1) + prepareTracks + prepareHits exactly as main does (OMP:715-720).
void ref_generate(float* trk_out, float* hit_out) {
  std::vector<AHIT> hits = ReferenceLayerHits();
  srand(1);
  MPTRK* trk = prepareTracks(inputtrk);
  MPHIT* hit = prepareHits(hits);
  std::memcpy(trk_out, trk, sizeof(MPTRK) * nevts * nb);
  std::memcpy(hit_out, hit, sizeof(MPHIT) * nlayer * nevts * nb);
  free(trk);
  free(hit);
}

// The input records in layer order: trk28 = par[6], cov[21], q (as float);
// hits = nhit x (pos[3], cov[6]).  Returns the number of hits.
int ref_inputs(float* trk28, float* hits) {
  for (int i = 0; i < 6; ++i) trk28[i] = inputtrk.par[i];
  for (int i = 0; i < 21; ++i) trk28[6 + i] = inputtrk.cov[i];
  trk28[27] = static_cast<float>(inputtrk.q);
  std::vector<AHIT> h = ReferenceLayerHits();
  for (size_t l = 0; l < h.size(); ++l) {
    for (int i = 0; i < 3; ++i) hits[l * 9 + i] = h[l].pos[i];
    for (int i = 0; i < 6; ++i) hits[l * 9 + 3 + i] = h[l].cov[i];
  }
  return static_cast<int>(h.size());
}

void ref_srand(unsigned s) { srand(s); }
float ref_randn(float mu, float sigma) { return randn(mu, sigma); }

// One pass of the reference driver body (OMP:733-747), single-threaded.
void ref_run(const float* trk_in, const float* hit_in, float* out) {
  const MPTRK* trk = reinterpret_cast<const MPTRK*>(trk_in);
  const MPHIT* hit = reinterpret_cast<const MPHIT*>(hit_in);
  MPTRK* outtrk = reinterpret_cast<MPTRK*>(out);
  for (size_t ie = 0; ie < nevts; ++ie)
    for (size_t ib = 0; ib < nb; ++ib) {
      const MPTRK* btracks = bTk(trk, ie, ib);
      MPTRK* obtracks = bTk(outtrk, ie, ib);
      (*obtracks) = (*btracks);
      for (size_t layer = 0; layer < nlayer; ++layer) {
        const MPHIT* bhits = bHit(hit, ie, ib, layer);
        propagateToZ(&(*obtracks).cov, &(*obtracks).par, &(*obtracks).q, &(*bhits).pos,
                     &(*obtracks).cov, &(*obtracks).par);
        KalmanUpdate_v2(&(*obtracks).cov, &(*obtracks).par, &(*bhits).cov, &(*bhits).pos);
      }
    }
}

// The two path functions on one block each (OMP:623, OMP:503).
void ref_propagate_block(const float* trk_in, const float* hit_in, float* trk_out) {
  const MPTRK* t = reinterpret_cast<const MPTRK*>(trk_in);
  const MPHIT* h = reinterpret_cast<const MPHIT*>(hit_in);
  MPTRK* o = reinterpret_cast<MPTRK*>(trk_out);
  o->q = t->q;
  propagateToZ(&t->cov, &t->par, &t->q, &h->pos, &o->cov, &o->par);
}
void ref_update_block(float* trk_inout, const float* hit_in) {
  MPTRK* t = reinterpret_cast<MPTRK*>(trk_inout);
  const MPHIT* h = reinterpret_cast<const MPHIT*>(hit_in);
  KalmanUpdate_v2(&t->cov, &t->par, &h->cov, &h->pos);
}

// KalmanUpdate v1 (OMP:375-501, not called by the reference's driver).
void ref_update_v1_block(float* trk_inout, const float* hit_in) {
  MPTRK* t = reinterpret_cast<MPTRK*>(trk_inout);
  const MPHIT* h = reinterpret_cast<const MPHIT*>(hit_in);
  KalmanUpdate(&t->cov, &t->par, &h->cov, &h->pos);
}

// The reference program itself (prints the full report to stdout).
int ref_main() { return p2z_reference_main(0, nullptr); }

}  // extern "C"
