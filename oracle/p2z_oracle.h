/*
 * p2z oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU algorithm for the propagate-to-z +
 * Kalman-update path (src/propagate-toz-test_OMP.cpp of cerati/p2z-tests).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as a checker or as the timed CPU baseline.  The
 * product (p2z-tests_amd/) never links or calls it.
 *
 * Layout (reference OMP:57-125, accessors OMP:155-241): one AoSoA block of
 * width W stores element e of lane l at word e*W + l.
 *   MPTRK block: par[6W] | cov[21W] | q[W] (int32)   = 28W words
 *   MPHIT block: pos[3W] | cov[6W]                   =  9W words
 * Track block index  b = ib + nb*ie              (bTk, OMP:147-153)
 * Hit block index    lay + nlayer*b              (bHit, OMP:212-217, 282)
 *
 * Pinning of each function (what checks it) is listed with the function.
 */
#ifndef P2Z_ORACLE_H
#define P2Z_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_TRK_WORDS 28
#define ORACLE_HIT_WORDS 9

/* Reference input record types (OMP:57-67). */
typedef struct { float par[6]; float cov[21]; int q; } oracle_atrk;
typedef struct { float pos[3]; float cov[6]; } oracle_ahit;

/* ---- generator (OMP:127-145, 243-293) -------------------------------- */
void  oracle_srand(unsigned seed);
float oracle_randn(float mu, float sigma);
/* Fills nevts*nb MPTRK blocks exactly as prepareTracks (OMP:243-263). */
void  oracle_prepare_tracks(const oracle_atrk* in, float* out, int64_t nevts,
                            int64_t nb, int W, double smear);
/* Fills nlayer*nevts*nb MPHIT blocks exactly as prepareHits (OMP:265-293);
 * layer lay uses hits[lay].  Returns -1 (the reference exits) if nlayer > nhits. */
int   oracle_prepare_hits(const oracle_ahit* hits, int nhits, float* out,
                          int64_t nevts, int64_t nb, int W, int nlayer, double smear);

/* ---- the path, one block of width W -------------------------------------- */
/* propagateToZ (OMP:622-682 with MultHelixPropEndcap OMP:296-341 and
 * MultHelixPropTranspEndcap OMP:343-373).  in and out may alias. */
void oracle_propagate_to_z(const float* inErr, const float* inPar, const int32_t* inChg,
                           const float* msP, float* outErr, float* outPar, int W);
/* KalmanUpdate_v2 (OMP:503-620), in place. */
void oracle_kalman_update_v2(float* trkErr, float* inPar, const float* hitErr,
                             const float* msP, int W);

/* KalmanUpdate v1 (OMP:375-501), in place; hitErr has the 6 hit covariance
 * words per lane.  fixed == 0: as written (upper-triangle addressing of both
 * covariances, the (0,2) cofactor of OMP:395, pt written to the ipt slot);
 * fixed != 0: the textbook 3-D update in the path's lower-triangle packing. */
void oracle_kalman_update_v1(float* trkErr, float* inPar, const float* hitErr,
                             const float* msP, int W, int fixed);

/* Driver (OMP:731-748): out = trk, then nlayer x (propagateToZ, KalmanUpdate_v2)
 * per block.  Parallel over (event x block) with nthreads OpenMP threads
 * (<=0: runtime default).  Hits may be per-track ragged via nlay_lane (NULL =
 * every lane runs nlayer layers). */
void oracle_run(const float* trk, const float* hit, float* out, int64_t nevts,
                int64_t nb, int W, int nlayer, int nthreads);

/* Same driver with the reference's OpenMP structure: parallel over events only
 * (OMP:733), used to time the reference threading beside the all-core form. */
void oracle_run_events_only(const float* trk, const float* hit, float* out,
                            int64_t nevts, int64_t nb, int W, int nlayer, int nthreads);

/* ---- full-FP64 oracle ---------------------------------------------------- */
/* propagateToZ / KalmanUpdate_v2 on one block computed entirely in double,
 * results rounded to float (same block layout and aliasing rules). */
void oracle_propagate_to_z_fp64(const float* inErr, const float* inPar, const int32_t* inChg,
                                const float* msP, float* outErr, float* outPar, int W);
void oracle_kalman_update_v2_fp64(float* trkErr, float* inPar, const float* hitErr,
                                  const float* msP, int W);

/* Same algorithm with every quantity in double (libm sin/cos in double); the
 * state is kept in double between layers.  Inputs are the FP32 blocks; the
 * outputs are 28 doubles per track in track-major order:
 *   out64[t*28 + e], t = b*W + l, e = 0..5 par, 6..26 cov, 27 q. */
void oracle_run_fp64(const float* trk, const float* hit, double* out64, int64_t nevts,
                     int64_t nb, int W, int nlayer, int nthreads);

/* Ragged driver: blocks described by block tables.
 *   blk_trk[b]   : MPTRK block index of block b (input and output)
 *   blk_hit[b]   : index of the first MPHIT block of b (its layers follow)
 *   blk_nlay[b]  : number of layer slots of block b
 *   lane_nlay    : per-track layer count, laid out like the MPTRK blocks
 *                  (nblocks*W int32, lane l of block b at b*W+l); a lane stops
 *                  updating after its own count (slots past it are padding).
 * fp64 != 0 selects the double algorithm, writing out64 (track-major, 28 doubles);
 * otherwise the float algorithm writes out (MPTRK blocks). */
void oracle_run_ragged(const float* trk, const float* hit, float* out, double* out64,
                       int64_t nblocks, int W, const int64_t* blk_trk, const int64_t* blk_hit,
                       const int32_t* blk_nlay, const int32_t* lane_nlay, int fp64,
                       int nthreads);

/* ---- the C library's sin / cos ------------------------------------------ */
/* s[i] = sinf(x), c[i] = cosf(x) of the C library (separate calls, as the
 * reference's source writes them at OMP:633-636, 643-644, 652-653) for the
 * float bit patterns x = first_bits + i, i < n; OpenMP threads (<=0: all). */
void oracle_libm_sincosf_range(uint32_t first_bits, int64_t n, float* s, float* c, int nthreads);

/* ---- report (OMP:758-876) -------------------------------------------------- */
typedef struct {
  double avgx, avgy, avgz, avgdx, avgdy, avgdz, avgpt, avgphi, avgtheta;
  double stdx, stdy, stdz, stddx, stddy, stddz;  /* already sqrtf'd, as OMP:846-851 */
  int nnans, nfail, ngood;
} oracle_summary_t;

/* ntrks is the reference's per-event track count (the stats loop runs it<ntrks
 * even when nb*W < ntrks); lasthit is the unsmeared hit of layer nlayer-1. */
void oracle_summary(const float* out, int64_t nevts, int64_t nb, int W, int64_t ntrks,
                    const float lasthit[3], oracle_summary_t* s);
/* Prints the 11 summary lines (OMP:866-876) into buf; returns chars written. */
int  oracle_format_summary(const oracle_summary_t* s, char* buf, int cap);

#ifdef __cplusplus
}
#endif
#endif
