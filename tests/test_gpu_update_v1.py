"""GPU parity of KalmanUpdate v1 (OMP:375-501, the optional 3-D update)
through p2z_kalman_update_v1: both forms bitwise equal to the oracle (no
contraction, the same typing on both sides, no libm in the update), and the
form as written bitwise equal to the reference build's fixture."""
from __future__ import annotations

import numpy as np
import pytest

import p2z
import pyoracle as O
from parity import v1_blocks

pytestmark = pytest.mark.gpu


def log(msg: str) -> None:
    print(msg, flush=True)


def gpu_v1(ctx, trk, hit, W, fixed):
    nblk = trk.size // (28 * W)
    dt, dh = ctx.upload(trk), ctx.upload(hit)
    ctx.kalman_update_v1(dt, dh, nblk, W, fixed=fixed)
    ctx.sync()
    out = ctx.download(dt)
    dt.free()
    dh.free()
    return out


def oracle_v1(trk, hit, W, fixed):
    nblk = trk.size // (28 * W)
    return np.concatenate([O.update_v1_block(np.ascontiguousarray(trk[b * 28 * W:(b + 1) * 28 * W]),
                                             np.ascontiguousarray(hit[b * 9 * W:(b + 1) * 9 * W]), W, fixed)
                           for b in range(nblk)])


def same_words(a, b) -> bool:
    """Bitwise equal, NaN matching NaN (the device and the host may write
    different NaN payloads)."""
    ua, ub = a.view(np.uint32), b.view(np.uint32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((ua == ub) | nan))


def test_as_written_equals_reference_fixture(ctx, golden):
    g = golden["update_v1"]
    W = int(g["width"])
    log("case v1/as-written/fixture")
    out = gpu_v1(ctx, g["trk"], g["hit"], W, False)
    assert same_words(out, g["out"])


@pytest.mark.parametrize("fixed", [False, True])
@pytest.mark.parametrize("W", [16, 32, 64, 128])
def test_bitwise_against_oracle(ctx, W, fixed):
    rng = np.random.default_rng(100 + W)
    blocks = [v1_blocks(rng, W) for _ in range(6)]
    trk = np.concatenate([b[0] for b in blocks])
    hit = np.concatenate([b[1] for b in blocks])
    log(f"case v1/{'fixed' if fixed else 'as-written'}/W={W}")
    out = gpu_v1(ctx, trk, hit, W, fixed)
    exp = oracle_v1(trk, hit, W, fixed)
    eq = float(np.mean(out.view(np.uint32) == exp.view(np.uint32)))
    log(f"  words equal {eq:.6f}")
    assert same_words(out, exp)


def test_invalid_flags(ctx):
    buf = ctx.alloc(28 * 32 * 4 + 9 * 32 * 4)
    with pytest.raises(p2z.P2zError) as e:
        p2z.check(p2z.lib().p2z_kalman_update_v1(ctx.handle, buf.ptr, buf.ptr, buf.ptr, buf.ptr, 1, 28 * 32,
                                                 9 * 32, 32, 4), "p2z_kalman_update_v1")
    assert e.value.rc == p2z.EINVAL
    buf.free()
