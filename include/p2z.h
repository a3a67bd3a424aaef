/*
 * p2z -- MI355X (gfx950) propagate-to-z + Kalman-update path, C ABI.
 *
 * Drop-in for the hot path of cerati/p2z-tests: the per-block functions
 * propagateToZ / KalmanUpdate_v2 and the driver loop of
 * src/propagate-toz-test_OMP.cpp (cited as OMP:<line> below), with sizes as
 * runtime arguments instead of the reference's compile-time macros.
 *
 * Data layout is the reference's, unchanged (zero-copy):
 *   MPTRK block (OMP:115-120): par[6W] | cov[21W] | q[W] (int32), 28W words
 *   MPHIT block (OMP:122-125): pos[3W] | cov[6W],                  9W words
 *   element e of lane l of a block is word e*W + l (OMP:159-161)
 *   track block of (event ie, block ib): ib + nb*ie          (bTk, OMP:147)
 *   hit block of (ie, ib, layer lay):   lay + nlayer*(ib+nb*ie) (bHit, OMP:215)
 * W (the reference's bsize) is one of 16, 32, 64, 128.
 *
 * Conventions: every function returns P2Z_OK (0) or a negative P2Z_E* code and
 * never exits or throws; p2z_last_error() describes the last failure of the
 * calling thread.  Functions taking "device" pointers enqueue work on the
 * context's stream and return before it finishes (p2z_sync waits).  A context
 * is used by one host thread at a time; distinct contexts are independent.
 */
#ifndef P2Z_H
#define P2Z_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P2Z_ABI_VERSION 2

/* status codes */
#define P2Z_OK 0
#define P2Z_EINVAL (-1)   /* bad argument (size, width, pointer, mode)      */
#define P2Z_ENOMEM (-2)   /* device or host allocation failed, or budget    */
#define P2Z_EHIP (-3)     /* HIP runtime error (message in p2z_last_error)  */
#define P2Z_ENODEV (-4)   /* no such GPU                                    */
#define P2Z_EBUSY (-5)    /* serving: no free slot / budget right now       */
#define P2Z_ESTATE (-6)   /* call not valid in the object's current state   */

/* arithmetic modes */
#define P2Z_MATH_FAST 0     /* FP32, FMA contraction, sparse forms (default)  */
#define P2Z_MATH_REF 1      /* the reference CPU typing: FP64 sub-expressions,
                               no contraction, reference summation order      */
#define P2Z_MATH_FP64GAIN 2 /* FP32 propagation; FP64 2x2 inverse, gain and
                               Joseph-form covariance update                  */

/* Input records (reference ATRK / AHIT, OMP:57-67). */
typedef struct { float par[6]; float cov[21]; int32_t q; } p2z_atrk;
typedef struct { float pos[3]; float cov[6]; } p2z_ahit;

typedef struct p2z_ctx p2z_ctx;

typedef struct {
  int32_t tracks_per_wg; /* workgroup size in tracks: 64, 128, 192 or 256; 0 = 256 */
  int32_t nstreams;      /* staging slots of the host-buffer pipeline; 0 = 4      */
  int32_t hit_stage;     /* hit-staging depth: layers of hits in flight ahead of
                            the one being computed, 1 or 2; 0 = per mode
                            (2 for P2Z_MATH_FP64GAIN, 1 otherwise)          */
  int32_t reserved;
} p2z_opts;

/* ---- library, context, memory ------------------------------------------ */
int p2z_version(void);
const char* p2z_last_error(void);
int p2z_device_count(int* n);
int p2z_create(p2z_ctx** ctx, int device, const p2z_opts* opts /* NULL = defaults */);
int p2z_destroy(p2z_ctx* ctx);
int p2z_sync(p2z_ctx* ctx);
int p2z_set_opts(p2z_ctx* ctx, const p2z_opts* opts);
int p2z_get_stream(p2z_ctx* ctx, void** hip_stream);
int p2z_dev_alloc(p2z_ctx* ctx, void** dptr, size_t bytes);
int p2z_dev_free(p2z_ctx* ctx, void* dptr);
int p2z_host_alloc(p2z_ctx* ctx, void** hptr, size_t bytes); /* pinned */
int p2z_host_free(p2z_ctx* ctx, void* hptr);
int p2z_memcpy_h2d(p2z_ctx* ctx, void* dst, const void* src, size_t bytes); /* async */
int p2z_memcpy_d2h(p2z_ctx* ctx, void* dst, const void* src, size_t bytes); /* async */
/* hipEvent timer on the context's stream: start, then stop returns elapsed ms */
int p2z_timer_start(p2z_ctx* ctx);
int p2z_timer_stop(p2z_ctx* ctx, float* ms);

/* ---- generator (host) ----------------------------------------------------
 * Replaces randn / prepareTracks / prepareHits (OMP:127-145, 243-263, 265-293):
 * same libc rand() stream, same draw order and typing.  p2z_srand resets the
 * stream (srand) and the cached second variate. */
void p2z_srand(unsigned seed);
float p2z_randn(float mu, float sigma);
/* out: nevts*nb MPTRK blocks */
int p2z_prepare_tracks(const p2z_atrk* in, void* out, int64_t nevts, int64_t nb, int32_t width,
                       double smear);
/* out: nlayer*nevts*nb MPHIT blocks; layer lay uses hits[lay]; P2Z_EINVAL when
 * nlayer > nhits (the reference calls exit(1), OMP:271-274). */
int p2z_prepare_hits(const p2z_ahit* hits, int32_t nhits, void* out, int64_t nevts, int64_t nb,
                     int32_t width, int32_t nlayer, double smear);

/* Input presets of the drop-in program.  P2Z_PRESET_REF is the reference's
 * own input (src/input_track.h:1-113 in main's layer order, OMP:693-695;
 * smear 1e-7, OMP:30); P2Z_PRESET_CONFIG0 the nominal track and hit of
 * propagate-toz-pseudocode.txt:2-15 (charge -1, same hit on every layer,
 * smear 1e-5).  Fills trk and hits[0..25] (cap >= 26) and *smear; returns the
 * number of hits (26) or P2Z_EINVAL. */
#define P2Z_PRESET_REF 0
#define P2Z_PRESET_CONFIG0 1
int p2z_preset_input(int32_t preset, p2z_atrk* trk, p2z_ahit* hits, int32_t cap, double* smear);

/* ---- the path (device pointers) -----------------------------------------
 * propagateToZ (OMP:623-682) on nblocks blocks.  inErr/inPar/inChg/outErr/
 * outPar point at block 0 of arrays whose consecutive blocks are trk_stride
 * 32-bit words apart (28*W for an MPTRK array); msP likewise with hit_stride
 * (9*W for an MPHIT array).  in and out may alias (the reference calls it in
 * place, OMP:742). */
int p2z_propagate_to_z(p2z_ctx* ctx, const float* inErr, const float* inPar, const int32_t* inChg,
                       const float* msP, float* outErr, float* outPar, int64_t nblocks,
                       int64_t trk_stride, int64_t hit_stride, int32_t width, int32_t math);
/* KalmanUpdate_v2 (OMP:503-620), in place, same stride convention. */
int p2z_kalman_update(p2z_ctx* ctx, float* trkErr, float* inPar, const float* hitErr,
                      const float* msP, int64_t nblocks, int64_t trk_stride, int64_t hit_stride,
                      int32_t width, int32_t math);
/* KalmanUpdate v1 (OMP:375-501: 3-D hit, 3x3 inverse), in place, same stride
 * convention; hitErr points at the 6 hit covariance words of each lane.  The
 * reference does not call it (commented out at OMP:743; the SYCL variant does,
 * src/propagate-toz-test_sycl.cpp:718).
 *   flags 0: as written -- it addresses both covariances as upper triangles
 *     stored row by row (the rest of the path stores lower triangles), its
 *     inverse entry (0,2) is m01*m12 - m02*m12 (OMP:395), and it writes
 *     pt + K*res into the ipt slot (OMP:447, 494); reference typing.
 *   P2Z_V1_FIXED: the textbook 3-D update in the path's packing (lower
 *     triangles, correct inverse in double, residual in x, y, z, ipt updated
 *     as ipt). */
#define P2Z_V1_FIXED 1u
int p2z_kalman_update_v1(p2z_ctx* ctx, float* trkErr, float* inPar, const float* hitErr,
                         const float* msP, int64_t nblocks, int64_t trk_stride, int64_t hit_stride,
                         int32_t width, uint32_t flags);
/* The driver loop (OMP:731-748) for one pass: out = trk, then nlayer x
 * (propagateToZ, KalmanUpdate_v2) per block, fused in one launch.  trk, hit,
 * out are device pointers (MPTRK[nevts*nb], MPHIT[nlayer*nevts*nb],
 * MPTRK[nevts*nb]).  Asynchronous. */
int p2z_run(p2z_ctx* ctx, const void* trk, const void* hit, void* out, int64_t nevts, int64_t nb,
            int32_t nlayer, int32_t width, int32_t math);
/* Same with host pointers (the reference's include-data mode, CU4:830-875):
 * chunks of events are copied in, computed and copied out on the context's
 * streams so that transfers overlap compute.  Returns when out is complete.
 * Pinned buffers (p2z_host_alloc) give full PCIe rate; from a pinned hit
 * buffer only the 6 words of each record the path reads (pos, cov xx xy yy)
 * cross PCIe, not the unused cov xz yz zz (OMP:510-512). */
int p2z_run_host(p2z_ctx* ctx, const void* trk, const void* hit, void* out, int64_t nevts,
                 int64_t nb, int32_t nlayer, int32_t width, int32_t math);

/* The sin and cos P2Z_MATH_REF computes (the bits of glibc's sinf / cosf /
 * sincosf, which the reference calls at OMP:633-636, 643-644, 652-653) for
 * the n consecutive float bit patterns first_bits, first_bits+1, ...:
 * s[i], c[i] (device pointers, n floats each).  first_bits + n <= 2^32.
 * Asynchronous.  A check of the device arithmetic against the C library. */
int p2z_sincosf_ref(p2z_ctx* ctx, uint32_t first_bits, int64_t n, float* s, float* c);

/* ---- ragged events ----------------------------------------------------------
 * No counterpart in the reference, whose sizes are uniform compile-time
 * macros (OMP:18-41).  Events of any track count, tracks of any layer count:
 * the tracks of an event are sorted by layer count and packed into width-W
 * MPTRK blocks (blocks never mix events; the last one is padded with lanes of
 * layer count 0, copied through unchanged).  A block with n layer slots owns n
 * consecutive MPHIT blocks; slot L holds layer L of the lanes that have it.
 * lane_nlayer holds each track's layer count in MP1I layout (block b lane l at
 * b*W + l); a lane stops after its own count. */
typedef struct {
  int64_t trk_block; /* input MPTRK block (and lane_nlayer block)         */
  int64_t hit_block; /* first of the block's MPHIT blocks                 */
  int64_t out_block; /* output MPTRK block                                */
  int32_t nlayer;    /* layer slots of the block (max of its lanes)       */
  int32_t event;     /* owning event                                      */
} p2z_block_desc;

#define P2Z_RAGGED_PERSISTENT 1u /* one persistent launch draining a work queue
                                    (descriptor order = processing order);
                                    otherwise one wave per 64 track slots     */
#define P2Z_RAGGED_GRAPH 2u      /* launch through a captured hipGraph, cached
                                    per argument set (either form); a context
                                    keeps up to 64 sets and empties its cache
                                    (after its stream drains) when it is full */

typedef struct {
  int32_t tracks_per_wg; /* 0: the context's setting                          */
  int32_t prefetch;      /* hit layers loaded ahead: 1 or 2; 0 = 2             */
  int32_t waves_per_cu;  /* persistent grid = CUs * waves_per_cu waves; 0 = 8  */
  int32_t items_per_claim; /* persistent form: work items (64 track slots
                              each) per chunk, 1..64; 0 = 4.  Wave w of the
                              grid computes chunk w first, then takes one
                              chunk per atomic on the queue counter       */
} p2z_ragged_opts;

/* Runs every block of the descriptor table (device memory) through its layers.
 * trk, hit, out, lane_nlayer, blocks: device pointers.  Asynchronous.
 * The descriptors are not range-checked on the device: every trk_block,
 * hit_block .. hit_block + nlayer - 1 and out_block must lie inside the
 * buffers (p2z_check_blocks checks a host copy of a table). */
/* Host-side check of a descriptor table (host memory) against the buffers it
 * addresses: 0 <= trk_block < pool_blocks, nlayer >= 0, 0 <= hit_block and
 * hit_block + nlayer <= pool_hit_blocks, 0 <= out_block < out_blocks.
 * P2Z_OK, or P2Z_EINVAL naming the first descriptor that is outside. */
int p2z_check_blocks(const p2z_block_desc* blocks, int64_t nblocks, int64_t pool_blocks,
                     int64_t pool_hit_blocks, int64_t out_blocks);
int p2z_run_ragged(p2z_ctx* ctx, const void* trk, const void* hit, void* out,
                   const int32_t* lane_nlayer, const p2z_block_desc* blocks, int64_t nblocks,
                   int32_t width, int32_t math, uint32_t flags, const p2z_ragged_opts* opts);

/* ---- synthetic workloads (host) ----------------------------------------------
 * Helix-consistent events: each track's true helix (the path's own helix
 * model, in double) is intersected with its layer planes and the hits are
 * the true points smeared by their covariance; the starting state is the
 * truth smeared by its covariance.  Deterministic for a seed, independent of
 * thread count (one random stream per event).
 *   P2Z_WORKLOAD_PRECISION: nevents x ntracks x nlayer (uniform), 1/pT
 *     log-uniform [0.01,2] GeV^-1, theta U[0.1,3.04], phi U[-pi,pi], hit
 *     covariance diagonals log-uniform [1e-8,1e-3] cm^2.
 *   P2Z_WORKLOAD_RAGGED: tracks per event log-normal (median 2000, sigma 1)
 *     clipped to [50,20000], layers per track U{8..26}, layer spacing
 *     U[5,25] cm, |cos theta| >= 0.35, hit covariance diagonals
 *     log-uniform [1e-6,1e-3] cm^2. */
#define P2Z_WORKLOAD_PRECISION 2
#define P2Z_WORKLOAD_RAGGED 3

typedef struct {
  int32_t kind;    /* P2Z_WORKLOAD_*                                 */
  int32_t width;   /* AoSoA width                                    */
  int64_t nevents;
  uint64_t seed;
  int32_t ntracks; /* PRECISION: tracks per event (RAGGED: ignored)   */
  int32_t nlayer;  /* PRECISION: layers per track (RAGGED: ignored)   */
  int32_t nthreads;/* host threads for generation; 0 = all            */
  int32_t reserved;
} p2z_workload_spec;

typedef struct {
  int64_t nevents, ntracks, nblocks, nhit_blocks, track_layers;
  int32_t width, kind;
} p2z_workload_info;

typedef struct p2z_workload p2z_workload;
int p2z_workload_create(const p2z_workload_spec* spec, p2z_workload** out);
int p2z_workload_destroy(p2z_workload* w);
int p2z_workload_get_info(const p2z_workload* w, p2z_workload_info* info);
/* Host arrays owned by the workload (valid until destroy):
 *   tracks  : nblocks MPTRK blocks           hits   : nhit_blocks MPHIT blocks
 *   lane_nlayer : nblocks*W int32            blocks : nblocks descriptors in
 *   event order; blocks_lpt: the same descriptors sorted by layer count,
 *   longest first (the persistent kernel's processing order);
 *   event_blocks: nevents+1 offsets of each event's blocks in event order;
 *   event_tracks: nevents track counts. */
const void* p2z_workload_tracks(const p2z_workload* w);
const void* p2z_workload_hits(const p2z_workload* w);
const int32_t* p2z_workload_lane_nlayer(const p2z_workload* w);
const p2z_block_desc* p2z_workload_blocks(const p2z_workload* w);
const p2z_block_desc* p2z_workload_blocks_lpt(const p2z_workload* w);
const int64_t* p2z_workload_event_blocks(const p2z_workload* w);
const int32_t* p2z_workload_event_tracks(const p2z_workload* w);
/* Binary fixture: a JSON header line (sizes, spec, checksums) followed by the
 * arrays above; load verifies the checksums. */
int p2z_workload_save(const p2z_workload* w, const char* path);
int p2z_workload_load(const char* path, p2z_workload** out);

/* ---- serving: continuous batching of events ---------------------------------
 * No counterpart in the reference (a batch program).  The inputs of all events
 * are resident on the device (one ragged pool: trk, hit, lane_nlayer); a
 * request names one event by its block descriptors (host memory; out_block
 * counted from 0 within the event).  Admission is bounded by an output arena
 * of budget_bytes in HBM: a request is admitted when its result blocks fit,
 * otherwise p2z_server_submit returns P2Z_EBUSY (retry after releasing
 * results) or P2Z_EINVAL (it can never fit).  p2z_server_pump launches every
 * admitted request not yet running as one batch on a free stream (one
 * persistent launch, descriptors longest-first), so requests arriving while
 * others run are admitted as slots and memory free up and their work
 * overlaps on the device; it also collects finished batches.  Results stay
 * in the arena until p2z_server_release.  Per-lane arithmetic does not depend
 * on the batch: a request's result is the same alone and in any mix.
 * With P2Z_SERVE_RESIDENT the requests are served by a device-resident loop
 * instead: one persistent kernel, launched when the server is created, stays
 * on the GPU and takes work as it is published.  p2z_server_submit writes an
 * admitted request's work items (64 track slots each, descriptors
 * longest-first) into a ring in pinned host memory and publishes them with
 * one store; the kernel's dispatcher wave sees them over PCIe and its worker
 * waves compute them as they come, so no launch or copy stands between a
 * request and the GPU; the worker that finishes a request's last item writes
 * a completion record into host memory, which p2z_server_pump collects.  The
 * loop ends when the server is destroyed, or after about two seconds in which
 * neither submit nor pump is called (the next call launches it again).  At
 * most 4096 requests and 65536 work items are in flight at once (P2Z_EBUSY
 * beyond); nstreams and max_batch_blocks do not apply.
 * A server is used by one host thread at a time. */
#define P2Z_SERVE_RESIDENT 1u
typedef struct {
  int64_t budget_bytes;     /* output arena size                               */
  int32_t nstreams;         /* batches in flight at once, 1..16; 0 = 4         */
  int32_t max_batch_blocks; /* block limit of one batch (a request above it
                               runs as a batch of its own); 0 = 65536, the
                               descriptor room reserved per stream at creation */
  int32_t width, math;      /* AoSoA width of the pool, arithmetic mode        */
  int32_t tracks_per_wg, prefetch, waves_per_cu, items_per_claim; /* as p2z_ragged_opts, except
                               items_per_claim 0 = 1: a batch of one event
                               then runs one work item per wave (latency)  */
  int64_t pool_blocks;      /* MPTRK blocks in trk (W-lane rows of lane_nlayer) */
  int64_t pool_hit_blocks;  /* MPHIT blocks in hit; requests are checked
                               against both (p2z_check_blocks) on submit     */
  uint32_t flags;           /* P2Z_SERVE_RESIDENT or 0 (batches of launches) */
  int32_t reserved;         /* 0                                             */
} p2z_server_opts;

typedef struct {
  int64_t request;    /* the caller's id                                    */
  int64_t out_block;  /* first arena block of the result                    */
  int64_t nblocks;    /* result blocks                                      */
  double t_arrival, t_admit, t_launch, t_done; /* seconds, p2z_now() clock  */
} p2z_completion;

typedef struct p2z_server p2z_server;
double p2z_now(void); /* monotonic host clock, seconds */
int p2z_server_create(p2z_ctx* ctx, const void* trk, const void* hit, const int32_t* lane_nlayer,
                      const p2z_server_opts* opts, p2z_server** out);
int p2z_server_destroy(p2z_server* s); /* waits for running batches */
int p2z_server_submit(p2z_server* s, int64_t request, const p2z_block_desc* blocks, int64_t nblocks,
                      double t_arrival);
int p2z_server_pump(p2z_server* s);                      /* returns completions ready  */
int p2z_server_poll(p2z_server* s, p2z_completion* out, int32_t cap); /* returns count */
int p2z_server_release(p2z_server* s, int64_t request);  /* frees its arena blocks     */
const void* p2z_server_arena(const p2z_server* s);       /* device base of the arena   */

typedef struct {
  int64_t nevents, track_layers, batches, max_inflight_events, busy_rejections;
  double rate_offered, rate_achieved;        /* events/s                        */
  double t_span;                             /* first arrival to last done, s   */
  double tl_per_s;                           /* track-layers / t_span           */
  double lat_p50, lat_p99, lat_p999, lat_mean, lat_max; /* seconds            */
  double mean_batch_events;
  double host_submit_s, host_pump_s;         /* host time inside submit / pump  */
} p2z_serve_report;

/* Poisson arrivals at `rate` events/s of the pool's events (event e = blocks
 * [event_blocks[e], event_blocks[e+1]) of the event-ordered descriptor table),
 * submitted as they arrive, admitted as the arena allows, pumped until all
 * are done.  release != 0 frees each result when done (steady state);
 * otherwise results stay (the arena must hold them all).  latency: optional
 * per-event done - arrival (nevents doubles); out_block: optional per-event
 * arena position of the result. */
int p2z_serve_poisson(p2z_server* s, const p2z_block_desc* blocks, const int64_t* event_blocks,
                      int64_t nevents, double rate, uint64_t seed, int32_t release, p2z_serve_report* rep,
                      double* latency, int64_t* out_block);

/* ---- report (device) ------------------------------------------------------
 * The statistics of OMP:758-876 computed on the GPU from device-resident
 * output: double accumulation in a fixed order (deterministic), same
 * acceptance window and denominators.  lasthit = unsmeared hit of the last
 * layer.  ntrks <= nb*W.  Synchronous. */
typedef struct {
  double avgx, avgy, avgz, avgdx, avgdy, avgdz, avgpt, avgphi, avgtheta;
  double stdx, stdy, stdz, stddx, stddy, stddz;
  int32_t nnans, nfail, ngood, reserved;
} p2z_summary;
int p2z_summary_device(p2z_ctx* ctx, const void* out, int64_t nevts, int64_t nb, int32_t width,
                       int64_t ntrks, const float lasthit[3], p2z_summary* s);
/* p2z_summary_device without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory at call time
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and result (bit for bit) as
 * p2z_summary_device, with their scratch in a static device array of
 * P2Z_SUMMARY_NOALLOC_BYTES per GPU (set up by p2z_create, held until the
 * process ends) instead of context memory grown by hipMalloc.  At most
 * P2Z_SUMMARY_NOALLOC_MAX_TRACKS tracks (nevts * ntrks), else P2Z_EINVAL.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from two threads or two contexts) race and must be serialised by the
 * caller.  Synchronous. */
#define P2Z_SUMMARY_NOALLOC_BYTES (4 << 20)
#define P2Z_SUMMARY_NOALLOC_MAX_TRACKS 11184128LL /* (4 MiB / 8 - 32) / 12 * 256 */
int p2z_summary_device_noalloc(p2z_ctx* ctx, const void* out, int64_t nevts, int64_t nb, int32_t width,
                               int64_t ntrks, const float lasthit[3], p2z_summary* s);
/* The same statistics on host memory (used by the drop-in report). */
int p2z_summary_host(const void* out, int64_t nevts, int64_t nb, int32_t width, int64_t ntrks,
                     const float lasthit[3], p2z_summary* s);
/* The 11 report lines of OMP:866-876 into buf; returns length or P2Z_EINVAL. */
int p2z_format_summary(const p2z_summary* s, char* buf, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* P2Z_H */
