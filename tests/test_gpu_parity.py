"""GPU parity: the HIP path through the C ABI against the pinned oracle and
the reference build's golden vectors, at the reference's tolerances
(parameters 1e-5, covariance 1e-4 relative; tests/parity.py)."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

import p2z
import pyoracle as O
from parity import compare, ok, track_major

pytestmark = pytest.mark.gpu


def log(msg: str) -> None:
    print(msg, flush=True)


def gpu_run(ctx, trk, hit, nevts, nb, nlayer, W, math):
    dt, dh = ctx.upload(trk), ctx.upload(hit)
    do = ctx.alloc(trk.nbytes)
    ctx.run(dt, dh, do, nevts, nb, nlayer, W, math)
    ctx.sync()
    out = ctx.download(do)
    for b in (dt, dh, do):
        b.free()
    return out


@pytest.mark.parametrize("math", ["ref", "fast"])
@pytest.mark.parametrize("name", ["small", "w16"])
def test_fused_matches_reference_build(ctx, golden, name, math):
    """The reference compiled from its own source wrote these outputs."""
    g = golden[name]
    nevts, nb, W, nl = (int(g[k]) for k in ("nevts", "nb", "width", "nlayer"))
    log(f"case fused/{name}/{math}")
    out = gpu_run(ctx, g["trk"], g["hit"], nevts, nb, nl, W, math)
    st = compare(track_major(out, W), track_major(g["out"], W))
    log(f"  {st}")
    assert ok(st), st


@pytest.mark.parametrize("W", [16, 32, 64, 128])
@pytest.mark.parametrize("math", ["ref", "fast"])
def test_widths_against_oracle(ctx, W, math):
    nevts, ntrks, nl = 2, 512, 20
    log(f"case widths/W={W}/{math}")
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    out = gpu_run(ctx, trk, hit, nevts, nb, nl, W, math)
    ref = O.run(trk, hit, nevts, nb, W, nl)
    st = compare(track_major(out, W), track_major(ref, W))
    log(f"  {st}")
    assert ok(st), st


def _random_blocks(rng, nblk, W):
    from test_oracle import _random_blocks as rb
    trk = np.concatenate([rb(rng, W, zrow=bool(i % 2))[0] for i in range(nblk)])
    hit = np.concatenate([rb(rng, W)[1] for _ in range(nblk)])
    return trk, hit


@pytest.mark.parametrize("math", ["ref", "fast"])
def test_per_function_entry_points(ctx, math):
    """propagateToZ and KalmanUpdate_v2 entry points on random blocks (with a
    non-zero z row in half of them) against the oracle's block functions."""
    rng = np.random.default_rng(11)
    W, nblk = 32, 8
    trk, hit = _random_blocks(rng, nblk, W)
    log(f"case per-function/{math}")
    dt, dh, do = ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)
    ctx.propagate_to_z(dt, dh, do, nblk, W, math)
    prop = ctx.download(do)
    ctx.kalman_update(dt, dh, nblk, W, math)
    upd = ctx.download(dt)
    for b in (dt, dh, do):
        b.free()
    eprop = np.concatenate([O.propagate_block(trk[b * 28 * W:(b + 1) * 28 * W], hit[b * 9 * W:(b + 1) * 9 * W], W)
                            for b in range(nblk)])
    eupd = np.concatenate([O.update_block(trk[b * 28 * W:(b + 1) * 28 * W], hit[b * 9 * W:(b + 1) * 9 * W], W)
                           for b in range(nblk)])
    for got, exp, what in ((prop, eprop, "propagate"), (upd, eupd, "update")):
        st = compare(track_major(got, W), track_major(exp, W))
        log(f"  {what}: {st}")
        assert ok(st), (what, st)


def test_fp64gain_update_is_double_textbook(ctx):
    """The FP64-gain update of float inputs equals the textbook update done in
    double, to float rounding of the outputs."""
    from test_oracle import _textbook_update, _unpack
    rng = np.random.default_rng(12)
    W, nblk = 32, 4
    trk, hit = _random_blocks(rng, nblk, W)
    dt, dh = ctx.upload(trk), ctx.upload(hit)
    ctx.kalman_update(dt, dh, nblk, W, "fp64gain")
    upd = ctx.download(dt)
    dt.free()
    dh.free()
    for b in range(nblk):
        tb, hb, ub = (a[b * n * W:(b + 1) * n * W] for a, n in ((trk, 28), (hit, 9), (upd, 28)))
        for l in range(W):
            p0, P0, _ = _unpack(tb, W, l)
            p1, P1, _ = _unpack(ub, W, l)
            R = np.array([[hb[3 * W + l], hb[4 * W + l]], [hb[4 * W + l], hb[5 * W + l]]], np.float64)
            pe, Pe = _textbook_update(p0, P0, np.array([hb[l], hb[W + l]], np.float64), R)
            assert np.allclose(p1, pe, rtol=2e-7, atol=1e-30), (b, l)
            scale = np.sqrt(np.outer(np.abs(np.diag(Pe)), np.abs(np.diag(Pe)))) + 1e-30
            assert np.max(np.abs(P1 - Pe) / scale) < 1e-6, (b, l)


def test_fp64gain_error_against_fp64_oracle(ctx, golden):
    """Precision report on the reference inputs: error of the plain FP32 path
    and of the FP64-gain path against the full-FP64 oracle."""
    g = golden["small"]
    nevts, nb, W, nl = (int(g[k]) for k in ("nevts", "nb", "width", "nlayer"))
    o64 = O.run_fp64(g["trk"], g["hit"], nevts, nb, W, nl)
    res = {}
    for math in ("fast", "fp64gain"):
        out = gpu_run(ctx, g["trk"], g["hit"], nevts, nb, nl, W, math)
        res[math] = compare(track_major(out, W), o64)
        log(f"  {math} vs fp64: {res[math]}")
    assert res["fp64gain"]["nonfinite_mismatch"] == 0
    assert res["fp64gain"]["par_max"] <= res["fast"]["par_max"] * 1.01


def test_ref_typing_is_bitwise_close(ctx, golden):
    """Reference typing, no contraction: results equal the reference build's
    except where a device sin/cos differs from glibc's by an ulp."""
    g = golden["small"]
    nevts, nb, W, nl = (int(g[k]) for k in ("nevts", "nb", "width", "nlayer"))
    out = gpu_run(ctx, g["trk"], g["hit"], nevts, nb, nl, W, "ref")
    same = np.mean(out.view(np.uint32) == g["out"].view(np.uint32))
    log(f"  bitwise-equal words: {same:.4f}")
    assert same > 0.5


@pytest.mark.parametrize("math", ["ref", "fast"])
def test_known_answer_full_size(ctx, golden, math):
    """The reference configuration end to end on the GPU: the printed report
    equals the reference's known answer to the last printed digit."""
    nevts, ntrks, W, nl = 100, 9600, 32, 20
    log(f"case known-answer/{math}")
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    out = gpu_run(ctx, trk, hit, nevts, nb, nl, W, math)
    s = p2z.summary_host(out, nevts, nb, W, ntrks, last)
    text = p2z.format_summary(s).splitlines()
    log("\n".join(text))
    from test_oracle import _parse
    for a, b in zip(_parse(text), _parse(golden["ref_summary_full"])):
        assert abs(a - b) <= 1.01e-6, (a, b)
    assert s.nnans == 0 and s.nfail == 0 and s.ngood == nevts * ntrks


def test_device_summary_matches_host(ctx):
    nevts, ntrks, W, nl = 10, 9600, 32, 20
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    dt, dh, do = ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)
    ctx.run(dt, dh, do, nevts, nb, nl, W, "fast")
    sd = ctx.summary_device(do, nevts, nb, W, ntrks, last)
    out = ctx.download(do)
    for b in (dt, dh, do):
        b.free()
    sh = p2z.summary_host(out, nevts, nb, W, ntrks, last)
    a, b = sd.as_dict(), sh.as_dict()
    for k in a:
        assert abs(a[k] - b[k]) <= 1e-12 * max(1.0, abs(b[k])), (k, a[k], b[k])
    assert p2z.format_summary(sd) == p2z.format_summary(sh)


def test_host_pipeline_equals_device_run(ctx):
    """The include-data pipeline (chunks over streams) computes the same bits."""
    nevts, ntrks, W, nl = 13, 1024, 32, 20
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    a = gpu_run(ctx, trk, hit, nevts, nb, nl, W, "fast")
    b = np.empty_like(trk)
    ctx.run_host(trk, hit, b, nevts, nb, nl, W, "fast")
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_deterministic_across_runs_and_launch_shapes(ctx):
    nevts, ntrks, W, nl = 3, 4096, 32, 20
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    a = gpu_run(ctx, trk, hit, nevts, nb, nl, W, "fast")
    b = gpu_run(ctx, trk, hit, nevts, nb, nl, W, "fast")
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for tpw in (64, 128, 192):
        c = p2z.Context(0, tracks_per_wg=tpw)
        try:
            d = gpu_run(c, trk, hit, nevts, nb, nl, W, "fast")
        finally:
            c.close()
        assert np.array_equal(a.view(np.uint32), d.view(np.uint32)), tpw


def test_edge_sizes(ctx):
    """Empty inputs are no-ops; one layer; a width larger than the track count."""
    W = 32
    trk = np.zeros(p2z.trk_words(1, W), np.float32)
    dt = ctx.upload(trk)
    ctx.run(dt, dt, dt, 0, 1, 20, W)        # no events
    ctx.run(dt, dt, dt, 5, 0, 20, W)        # no blocks
    ctx.sync()
    dt.free()
    for nl in (1, 2):
        trk, hit, last, nb = p2z.generate("ref", 1, 64, W, nl)
        out = gpu_run(ctx, trk, hit, 1, nb, nl, W, "ref")
        st = compare(track_major(out, W), track_major(O.run(trk, hit, 1, nb, W, nl), W))
        assert ok(st), (nl, st)
    # zero layers: the output is the input
    trk, hit, last, nb = p2z.generate("ref", 1, 64, W, 1)
    out = gpu_run(ctx, trk, hit, 1, nb, 0, W, "fast")
    assert np.array_equal(out.view(np.uint32), trk.view(np.uint32))


def test_invalid_arguments_do_not_launch(ctx):
    buf = ctx.alloc(4096)
    with pytest.raises(p2z.P2zError) as e:
        ctx.run(buf, buf, buf, 1, 1, 20, 48)
    assert e.value.rc == p2z.EINVAL
    with pytest.raises(p2z.P2zError):
        ctx.run(buf, buf, buf, 1, 1, 20, 32, math=7)
    with pytest.raises(p2z.P2zError):
        p2z.Context(0, tracks_per_wg=100)
    buf.free()


def test_dropin_program_prints_known_answer(golden):
    """bin/p2z_main, the reference program's drop-in, at a reduced size: its
    report equals the oracle's on the same inputs."""
    nevts, ntrks = 4, 960
    res = subprocess.run([p2z.MAIN_PATH, "--nevts", str(nevts), "--ntrks", str(ntrks), "--niter", "2",
                          "--math", "ref"], capture_output=True, text=True, timeout=120)
    log(res.stdout[-2000:])
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    assert any(ln.startswith("formatted 2 4 960 32 30 ") for ln in lines)
    rep = [ln for ln in lines if (ln.startswith("track ") and not ln.startswith("track in")) or ln.startswith("number of")]
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, 32, 20)
    s = O.summary(O.run(trk, hit, nevts, nb, 32, 20), nevts, nb, 32, ntrks, last)
    from test_oracle import _parse
    for a, b in zip(_parse(rep), _parse(O.format_summary(s).splitlines())):
        assert abs(a - b) <= 1.01e-6, (a, b)
