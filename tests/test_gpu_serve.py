"""GPU tests of the serving loop (p2z_server_*), in both of its forms: batches
of persistent launches, and the device-resident loop (one persistent kernel
fed through a ring in host memory).  Results are the same alone and in the
mix, admission by the HBM arena, request bookkeeping, and the Poisson
driver's report.  Failure conditions are forced from the host (budgets, ids);
every kernel launched runs to completion normally (the resident loop ends on
the server's stop flag or on its idle timeout)."""
from __future__ import annotations

import numpy as np
import pytest

import p2z

pytestmark = pytest.mark.gpu
FORMS = pytest.mark.parametrize("resident", [False, True], ids=["batches", "resident"])


def log(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def pool(ctx):
    """A ragged workload resident on the device."""
    w = p2z.Workload("ragged", nevents=40, width=32, seed=11)
    d = dict(w=w, trk=ctx.upload(w.tracks), hit=ctx.upload(w.hits), lanes=ctx.upload(np.ascontiguousarray(w.lane_nlayer)))
    yield d
    for k in ("trk", "hit", "lanes"):
        d[k].free()


def event_blocks(w, e):
    b = np.array(w.blocks[w.event_blocks[e]:w.event_blocks[e + 1]])
    b["out_block"] -= b["out_block"][0]
    return b


def blk_bytes(W):
    return 28 * W * 4


@FORMS
def test_result_is_the_same_alone_and_in_the_mix(ctx, pool, resident):
    w = pool["w"]
    W = w.width
    budget = (w.nblocks + 16) * blk_bytes(W)
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], budget, W, "fast", nstreams=4, resident=resident) as sv:
        # mixed: Poisson arrivals fast enough that several events share batches and overlap
        rep, lat, ob = sv.poisson(w.blocks, w.event_blocks, rate=50000.0, seed=3, release=False)
        log(f"  mix: {rep.batches} batches, {rep.mean_batch_events:.2f} events/batch, "
            f"max in flight {rep.max_inflight_events}")
        nb = np.diff(w.event_blocks)
        mixed = {}
        for e in range(w.nevents):
            c = p2z.Completion(e, int(ob[e]), int(nb[e]), 0, 0, 0, 0)
            mixed[e] = sv.result(c)
        for e in range(w.nevents):
            sv.release(e)
        # alone: one request at a time
        for e in range(w.nevents):
            rid = 1000 + e
            assert sv.submit(rid, event_blocks(w, e)) == p2z.OK
            c = sv.wait([rid])[rid]
            alone = sv.result(c)
            sv.release(rid)
            assert np.array_equal(alone.view(np.uint32), mixed[e].view(np.uint32)), e
    assert rep.batches < w.nevents   # the mix really batched events together (resident: one loop)


@FORMS
def test_served_in_reference_typing_equals_the_oracle(ctx, pool, resident):
    """Reference typing through the server (Poisson arrivals, events batched
    and overlapping): every served event equals the oracle's ragged driver on
    that event's blocks, word for word."""
    import pyoracle as O
    w = pool["w"]
    W = w.width
    b = w.blocks
    exp = O.run_ragged(w.tracks, w.hits, w.nblocks, W, b["trk_block"], b["hit_block"], b["nlayer"],
                       w.lane_nlayer)
    nb = np.diff(w.event_blocks)
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], 64 * 1024 * 1024, W, "ref", resident=resident) as sv:
        rep, lat, ob = sv.poisson(w.blocks, w.event_blocks, rate=20000.0, seed=5, release=False)
        log(f"  ref served: {rep.batches} batches for {w.nevents} events")
        differ = []
        for e in range(w.nevents):
            got = sv.result(p2z.Completion(e, int(ob[e]), int(nb[e]), 0, 0, 0, 0))
            b0, b1 = int(w.event_blocks[e]), int(w.event_blocks[e + 1])
            if not np.array_equal(got.view(np.uint32), exp[b0 * 28 * W:b1 * 28 * W].view(np.uint32)):
                differ.append(e)
            sv.release(e)
        log(f"  events differing from the oracle: {differ}")
        assert not differ


@FORMS
def test_matches_the_batch_path(ctx, pool, resident):
    """A served event equals the same event in one ragged batch launch."""
    w = pool["w"]
    W = w.width
    blk = ctx.upload(np.ascontiguousarray(w.blocks))
    out = ctx.alloc(w.tracks.nbytes)
    ctx.run_ragged(pool["trk"], pool["hit"], out, pool["lanes"], blk, w.nblocks, W, "fast")
    ref = ctx.download(out)
    blk.free()
    out.free()
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], 64 * 1024 * 1024, W, "fast", resident=resident) as sv:
        for e in (0, 7, w.nevents - 1):
            assert sv.submit(e, event_blocks(w, e)) == p2z.OK
            c = sv.wait([e])[e]
            b0, b1 = int(w.event_blocks[e]), int(w.event_blocks[e + 1])
            exp = ref[b0 * 28 * W: b1 * 28 * W]
            assert np.array_equal(sv.result(c).view(np.uint32), exp.view(np.uint32)), e
            sv.release(e)


@FORMS
def test_admission_is_bounded_by_the_arena(ctx, pool, resident):
    w = pool["w"]
    W = w.width
    nb = np.diff(w.event_blocks)
    e0, e1 = 0, 1
    budget = int(max(nb[e0], nb[e1])) * blk_bytes(W)          # room for one of the two at a time
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], budget, W, "fast", nstreams=2, resident=resident) as sv:
        assert sv.submit(e0, event_blocks(w, e0)) == p2z.OK
        assert sv.submit(e1, event_blocks(w, e1)) == p2z.EBUSY      # no room until e0 is released
        sv.wait([e0])
        assert sv.submit(e1, event_blocks(w, e1)) == p2z.EBUSY      # done but not released: still held
        sv.release(e0)
        assert sv.submit(e1, event_blocks(w, e1)) == p2z.OK
        sv.wait([e1])
        sv.release(e1)
    # an event larger than the whole arena is rejected outright
    small = int(nb.max() - 1) * blk_bytes(W)
    big = int(np.argmax(nb))
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], small, W, "fast", resident=resident) as sv:
        with pytest.raises(p2z.P2zError) as err:
            sv.submit(big, event_blocks(w, big))
        assert err.value.rc == p2z.EINVAL


@FORMS
def test_request_bookkeeping_errors(ctx, pool, resident):
    w = pool["w"]
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], 64 * 1024 * 1024, w.width, "fast", resident=resident) as sv:
        assert sv.submit(5, event_blocks(w, 5)) == p2z.OK
        with pytest.raises(p2z.P2zError) as err:
            sv.submit(5, event_blocks(w, 5))                  # duplicate id
        assert err.value.rc == p2z.EINVAL
        with pytest.raises(p2z.P2zError) as err:
            sv.release(99)                                    # unknown id
        assert err.value.rc == p2z.EINVAL
        bad = event_blocks(w, 6)
        bad["out_block"][0] = 10 ** 6                         # result block outside the request
        with pytest.raises(p2z.P2zError):
            sv.submit(6, bad)
        for field, value in (("trk_block", w.nblocks), ("hit_block", w.nhit_blocks)):
            bad = event_blocks(w, 6)
            bad[field][-1] = value                            # a block outside the resident pool
            with pytest.raises(p2z.P2zError) as err:
                sv.submit(6, bad)
            assert err.value.rc == p2z.EINVAL and "outside" in str(err.value)
        sv.wait([5])
        sv.release(5)
    with pytest.raises(p2z.P2zError):                         # budget below one block
        p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], 100, w.width, "fast", resident=resident)


@FORMS
def test_kept_results_overflowing_the_arena_fail_cleanly(ctx, pool, resident):
    """Poisson run keeping every result in an arena too small for all of them:
    an error once nothing is in flight, not a wait forever."""
    w = pool["w"]
    budget = int(np.diff(w.event_blocks)[:8].sum()) * blk_bytes(w.width)
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], budget, w.width, "fast", resident=resident) as sv:
        with pytest.raises(p2z.P2zError) as err:
            sv.poisson(w.blocks, w.event_blocks, rate=20000.0, release=False)
        assert err.value.rc == p2z.ENOMEM


@FORMS
@pytest.mark.parametrize("rate", [2000.0, 20000.0])
def test_poisson_report(ctx, pool, rate, resident):
    w = pool["w"]
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], 256 * 1024 * 1024, w.width, "fp64gain", resident=resident) as sv:
        rep, lat, ob = sv.poisson(w.blocks, w.event_blocks, rate=rate, seed=1, release=True)
    d = rep.as_dict()
    log(f"  rate {rate}: {d}")
    assert rep.nevents == w.nevents and rep.batches >= 1
    assert 0 < rep.lat_p50 <= rep.lat_p99 <= rep.lat_p999 <= rep.lat_max
    assert np.all(lat > 0)
    assert abs(np.median(lat) - rep.lat_p50) <= 1e-3 + 0.5 * rep.lat_p50


@FORMS
def test_full_size_served_equals_batch(resident):
    """The serving workload at full size: 2000 ragged events with the FP64
    gain steps, arriving at 200k events/s so that many share batches and
    overlap on the streams.  Every served event equals the same event in one
    ragged batch launch, bit for bit."""
    with p2z.Context(0) as c:
        w = p2z.Workload("ragged", nevents=2000, width=32, seed=2024)
        W = w.width
        log(f"case serve/full-size events={w.nevents} tracks={w.ntracks}")
        trk, hit = c.upload(w.tracks), c.upload(w.hits)
        lanes = c.upload(np.ascontiguousarray(w.lane_nlayer))
        blk = c.upload(np.ascontiguousarray(w.blocks))
        out = c.alloc(w.tracks.nbytes)
        c.run_ragged(trk, hit, out, lanes, blk, w.nblocks, W, "fp64gain")
        ref = c.download(out)
        out.free()
        blk.free()
        with p2z.Server(c, trk, hit, lanes, 64 * 2**30, W, "fp64gain", nstreams=4, resident=resident) as sv:
            rep, lat, ob = sv.poisson(w.blocks, w.event_blocks, rate=200000.0, seed=5, release=False)
            log(f"  {rep.batches} batches, {rep.mean_batch_events:.2f} events/batch, "
                f"max in flight {rep.max_inflight_events}")
            assert rep.nevents == w.nevents and rep.batches < w.nevents
            nb = np.diff(w.event_blocks)
            for e in range(w.nevents):
                got = sv.result(p2z.Completion(e, int(ob[e]), int(nb[e]), 0, 0, 0, 0))
                b0, b1 = int(w.event_blocks[e]), int(w.event_blocks[e + 1])
                assert np.array_equal(got.view(np.uint32), ref[b0 * 28 * W: b1 * 28 * W].view(np.uint32)), e
            for e in range(w.nevents):
                sv.release(e)
        for b in (trk, hit, lanes):
            b.free()


def test_resident_loop_restarts_after_idle(ctx, pool):
    """The resident loop ends by itself when the host neither submits nor
    pumps for about two seconds (its bounded wait); the next call launches it
    again, and requests are served as before, with the same bits."""
    import time
    w = pool["w"]
    W = w.width
    with p2z.Server(ctx, pool["trk"], pool["hit"], pool["lanes"], 64 * 1024 * 1024, W, "fast",
                    resident=True) as sv:
        assert sv.submit(1, event_blocks(w, 1)) == p2z.OK
        first = sv.result(sv.wait([1])[1])
        sv.release(1)
        time.sleep(3.0)                              # the loop's idle timeout passes
        assert sv.submit(2, event_blocks(w, 1)) == p2z.OK
        again = sv.result(sv.wait([2])[2])
        sv.release(2)
    assert np.array_equal(first.view(np.uint32), again.view(np.uint32))
