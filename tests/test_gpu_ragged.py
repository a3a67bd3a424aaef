"""GPU tests of the ragged-event path (p2z_run_ragged): persistent work-queue
launch, hipGraph replay, one-shot launch, all widths, against the oracle's
ragged driver and against the plain fused path."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import p2z
import pyoracle as O
from parity import compare, same_words, words_equal

pytestmark = pytest.mark.gpu


def log(msg: str) -> None:
    print(msg, flush=True)


def by_track(w: "p2z.Workload", out: np.ndarray) -> np.ndarray:
    """(ntracks, 28) float64 in event / sorted-track order, padding lanes dropped."""
    W = w.width
    t = out.reshape(w.nblocks, 28, W).transpose(0, 2, 1).reshape(-1, 28)
    keep = np.zeros(w.nblocks * W, bool)
    for e in range(w.nevents):
        b0 = int(w.event_blocks[e])
        keep[b0 * W: b0 * W + int(w.event_tracks[e])] = True
    r = t[keep].astype(np.float64)
    r[:, 27] = t[keep][:, 27].copy().view(np.int32)
    return r


def oracle_ragged(w, fp64=False):
    b = w.blocks
    res = O.run_ragged(w.tracks, w.hits, w.nblocks, w.width, b["trk_block"], b["hit_block"], b["nlayer"],
                       w.lane_nlayer, fp64=fp64)
    if fp64:
        keep = np.zeros(w.nblocks * w.width, bool)
        for e in range(w.nevents):
            b0 = int(w.event_blocks[e])
            keep[b0 * w.width: b0 * w.width + int(w.event_tracks[e])] = True
        return res[keep]
    return res


class Dev:
    """A workload's arrays on the device."""

    def __init__(self, ctx, w, lpt=True):
        self.trk = ctx.upload(w.tracks)
        self.hit = ctx.upload(w.hits)
        self.lanes = ctx.upload(np.ascontiguousarray(w.lane_nlayer))
        self.blk = ctx.upload(np.ascontiguousarray(w.blocks_lpt if lpt else w.blocks))
        self.out = ctx.alloc(w.tracks.nbytes)

    def free(self):
        for b in (self.trk, self.hit, self.lanes, self.blk, self.out):
            b.free()


@pytest.fixture(scope="module")
def small_ragged():
    return p2z.Workload("ragged", nevents=12, width=32, seed=5)


@pytest.mark.parametrize("form", ["persistent", "graph", "oneshot"])
def test_ref_mode_bitwise_against_oracle(ctx, small_ragged, form):
    w = small_ragged
    log(f"case ragged-ref/{form} tracks={w.ntracks}")
    d = Dev(ctx, w)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, w.width, "ref",
                   persistent=form != "oneshot", graph=form == "graph")
    out = ctx.download(d.out)
    d.free()
    exp = oracle_ragged(w)
    log(f"  words equal {words_equal(out, exp):.6f}")
    assert same_words(out, exp)


@pytest.mark.parametrize("persistent", [True, False])
def test_descriptor_edge_cases_against_oracle(ctx, small_ragged, persistent):
    """Descriptors cut to 0 layers (the track is copied through) and to 1
    layer, beside full ones, so that waves mix lanes with no layers and lanes
    with many: reference typing equals the oracle's ragged driver with the same
    per-block layer counts.  An empty table launches nothing."""
    w = small_ragged
    blocks = np.array(w.blocks, copy=True)
    blocks["nlayer"][::7] = 0
    blocks["nlayer"][3::5] = np.minimum(blocks["nlayer"][3::5], 1)
    assert (blocks["nlayer"] == 0).any() and (blocks["nlayer"] == 1).any()
    d = Dev(ctx, w, lpt=False)
    dblk = ctx.upload(np.ascontiguousarray(blocks))
    ctx.h2d(d.out, np.zeros(w.tracks.size, np.float32))
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, dblk, w.nblocks, w.width, "ref", persistent=persistent)
    out = ctx.download(d.out)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, dblk, 0, w.width, "ref", persistent=persistent)   # no-op
    again = ctx.download(d.out)
    dblk.free()
    d.free()
    exp = O.run_ragged(w.tracks, w.hits, w.nblocks, w.width, blocks["trk_block"], blocks["hit_block"],
                       blocks["nlayer"], w.lane_nlayer)
    log(f"case ragged-edges persistent={persistent}: words equal {words_equal(out, exp):.6f}")
    assert same_words(out, exp)
    W = w.width
    zero = np.flatnonzero(blocks["nlayer"] == 0)
    trk_b, out_b = w.tracks.reshape(w.nblocks, -1), out.reshape(w.nblocks, -1)
    assert np.array_equal(out_b[zero].view(np.uint32), trk_b[zero].view(np.uint32))
    assert np.array_equal(again.view(np.uint32), out.view(np.uint32))


@pytest.mark.parametrize("math", ["fast", "fp64gain"])
def test_widths_give_the_same_tracks(ctx, math):
    """The packing is width independent (tracks sorted, then drawn), so every
    width must produce the same per-track results."""
    res = {}
    for W in (16, 32, 64, 128):
        w = p2z.Workload("ragged", nevents=6, width=W, seed=9)
        d = Dev(ctx, w)
        ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, W, math)
        res[W] = by_track(w, ctx.download(d.out))
        d.free()
    for W in (16, 64, 128):
        assert np.array_equal(res[W], res[32]), W


def test_ragged_equals_plain_path_on_uniform_events(ctx):
    """A uniform (precision) workload through the ragged kernel equals the
    plain fused kernel bit for bit."""
    w = p2z.Workload("precision", nevents=4, width=32, seed=3, ntracks=960, nlayer=20)
    d = Dev(ctx, w, lpt=False)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, 32, "fast")
    a = ctx.download(d.out)
    ctx.run(d.trk, d.hit, d.out, 4, 960 // 32, 20, 32, "fast")
    b = ctx.download(d.out)
    d.free()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# A persistent grid smaller than the work: beyond each wave's first chunk the
# items come from the work-queue counter, so the counter reset matters.
SMALL_GRID = dict(tracks_per_wg=64, waves_per_cu=1, items_per_claim=1)


@pytest.mark.parametrize("persistent,opts", [(True, {}), (True, SMALL_GRID), (False, {})])
def test_graph_replays_equal_eager(ctx, small_ragged, persistent, opts):
    """Captured graph (counter reset kernel + persistent kernel, or the
    one-shot kernel) replayed several times with other work in between: every
    replay equals the eager launch."""
    w = small_ragged
    assert (w.nblocks * w.width + 63) // 64 > 256  # more items than the small grid has waves (256 CUs)
    d = Dev(ctx, w)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, w.width, "fast")
    eager = ctx.download(d.out)
    scratch = ctx.alloc(w.tracks.nbytes)
    for rep in range(4):
        ctx.run_ragged(d.trk, d.hit, scratch, d.lanes, d.blk, w.nblocks, w.width, "ref")  # other work
        ctx.h2d(d.out, np.zeros_like(eager))
        ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, w.width, "fast", graph=True,
                       persistent=persistent, **opts)
        got = ctx.download(d.out)
        assert np.array_equal(got.view(np.uint32), eager.view(np.uint32)), rep
    scratch.free()
    d.free()


def test_graph_cache_is_bounded(ctx, small_ragged):
    """More distinct argument sets than a context caches graphs for (64): the
    first k descriptors of the table, for 70 values of k, each captured and
    launched.  The cache is emptied when full; every launch writes exactly the
    output blocks of its k descriptors, equal to the eager launch's."""
    w = small_ragged
    W = w.width
    d = Dev(ctx, w)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, W, "fast")
    eager = ctx.download(d.out).reshape(w.nblocks, -1)
    out_of = np.asarray(w.blocks_lpt)["out_block"]
    assert w.nblocks > 70
    for k in range(1, 71):
        ctx.h2d(d.out, np.zeros(w.tracks.size, np.float32))
        ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, k, W, "fast", graph=True)
        got = ctx.download(d.out).reshape(w.nblocks, -1)
        done = np.zeros(w.nblocks, bool)
        done[out_of[:k]] = True
        assert np.array_equal(got[done].view(np.uint32), eager[done].view(np.uint32)), k
        assert not got[~done].any(), k
    d.free()


@pytest.mark.parametrize("opts", [dict(prefetch=1), dict(tracks_per_wg=64), dict(waves_per_cu=2),
                                  dict(tracks_per_wg=128, prefetch=2, waves_per_cu=16),
                                  dict(items_per_claim=1), dict(items_per_claim=16, waves_per_cu=1),
                                  SMALL_GRID, dict(SMALL_GRID, items_per_claim=3),
                                  dict(SMALL_GRID, items_per_claim=7, prefetch=1),
                                  dict(persistent=False, prefetch=1), dict(persistent=False, tracks_per_wg=256)])
@pytest.mark.parametrize("math", ["fast", "fp64gain"])
def test_launch_shapes_are_bitwise_equal(ctx, small_ragged, opts, math):
    w = small_ragged
    d = Dev(ctx, w)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, w.width, math)
    a = ctx.download(d.out)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, w.width, math, **opts)
    b = ctx.download(d.out)
    d.free()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("kind", ["ragged", "precision"])
def test_accuracy_against_fp64(ctx, kind):
    """Error against the FP64 oracle: the FAST path is within 2x the reference
    FP32 algorithm's own error (the float oracle against FP64); the FP64-gain
    path is no worse than the FP32 reference and improves the covariance."""
    if kind == "ragged":
        w = p2z.Workload("ragged", nevents=6, width=32, seed=21)
    else:
        w = p2z.Workload("precision", nevents=2, width=32, seed=21, ntracks=2048, nlayer=20)
    o64 = oracle_ragged(w, fp64=True)
    o32 = by_track(w, oracle_ragged(w))
    base = compare(o32, o64)
    d = Dev(ctx, w)
    res = {}
    for math in ("fast", "fp64gain"):
        ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, w.width, math)
        res[math] = compare(by_track(w, ctx.download(d.out)), o64)
    d.free()
    log(f"  {kind}: reference fp32 vs fp64 {base}")
    for m, st in res.items():
        log(f"  {kind}: {m} vs fp64 {st}")
    f, g = res["fast"], res["fp64gain"]
    assert f["nonfinite_mismatch"] <= max(2, base["nonfinite_mismatch"] * 2)
    assert f["par_median"] <= 2 * base["par_median"] + 1e-9
    assert f["cov_median"] <= 2 * base["cov_median"] + 1e-9
    assert g["par_median"] <= base["par_median"] * 1.01 + 1e-9
    assert g["cov_median"] <= base["cov_median"]


def test_ref_mode_full_size_against_oracle(ctx):
    """The 2000-event ragged workload at full size (6.3e6 tracks, 1.1e8
    track-layers) in reference typing, through the persistent graph-captured
    launch: every word equals the oracle's ragged driver (the reference
    algorithm with glibc's sin/cos, on all host cores)."""
    W = 32
    w = p2z.Workload("ragged", nevents=2000, width=W, seed=2024)
    log(f"case ragged-ref/full-size tracks={w.ntracks} track-layers={w.track_layers}")
    d = Dev(ctx, w)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, W, "ref", graph=True)
    out = ctx.download(d.out)
    d.free()
    exp = oracle_ragged(w)
    log(f"  words equal {words_equal(out, exp):.8f}")
    assert same_words(out, exp)


def test_full_size_forms_agree(ctx):
    """The 2000-event ragged workload at full size (about 1.1e8 track-layers):
    the persistent launch (default shape and a grid smaller than the work),
    the one-shot launch and one launch per event give the same bits, since
    each lane's arithmetic does not depend on the launch; every live track
    ends on its last hit's z plane."""
    W = 32
    w = p2z.Workload("ragged", nevents=2000, width=W, seed=2024)
    log(f"case ragged/full-size tracks={w.ntracks} track-layers={w.track_layers}")
    d = Dev(ctx, w)
    ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, W, "fast")
    ref = ctx.download(d.out)
    for opts in (dict(graph=True), dict(SMALL_GRID, items_per_claim=8), dict(persistent=False, graph=True)):
        ctx.h2d(d.out, np.zeros_like(ref))
        ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, d.blk, w.nblocks, W, "fast", **opts)
        got = ctx.download(d.out)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), opts
    # one launch per event, over the event-ordered descriptors
    evt = ctx.upload(np.ascontiguousarray(w.blocks))
    ctx.h2d(d.out, np.zeros_like(ref))
    eb = w.event_blocks
    for e in range(w.nevents):
        b0, b1 = int(eb[e]), int(eb[e + 1])
        ctx.run_ragged(d.trk, d.hit, d.out, d.lanes, evt.ptr + b0 * 32, b1 - b0, W, "fast", persistent=False)
    got = ctx.download(d.out)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    evt.free()
    # z on the last hit plane, finite parameters, for every live lane
    t = ref.reshape(w.nblocks, 28, W)
    hits = np.asarray(w.hits).reshape(-1, 9, W)
    n = np.asarray(w.lane_nlayer).reshape(w.nblocks, W)
    hb = np.asarray(w.blocks)["hit_block"]
    live = n > 0
    bi, li = np.nonzero(live)
    zlast = hits[hb[bi] + n[bi, li] - 1, 2, li]
    assert np.array_equal(t[bi, 2, li], zlast)
    assert np.all(np.isfinite(t[bi, :6, li]))
    d.free()


def test_autotuner_refines_its_closest_candidates(ctx):
    """The tuner's sweep and its second, longer timing of the closest
    candidates of each form, on a small sample: every returned best is one of
    the re-timed candidates, and each form's best is the fastest of them."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "p2z-tests_amd"))
    import autotune

    res = autotune.tune(ctx, "ragged", nevents=24, math="fast", seed=3, reps=2, widths=(32, 64), tpws=(64,),
                        prefetch=(1,), waves=(8, 16), claims=(1, 2), refine=2)
    log(f"case autotune small sample: {res['candidates']} candidates, {res['refined']} refined")
    assert res["candidates"] == 2 * (4 + 1)
    assert res["refined"] == 2 + 2
    for form in ("persistent", "oneshot"):
        rs = [r for r in res["refined_results"] if r["form"] == form]
        assert len(rs) == 2
        assert res["best_" + form] == min(rs, key=lambda r: r["ms"])
    assert res["best"] in res["refined_results"]
