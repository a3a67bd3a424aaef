"""GPU test of p2z_summary_device_noalloc, the report entry point that
allocates nothing (its scratch is the library's static device array): through
the C ABI its result equals p2z_summary_device's, bit for bit, on two calls in
a row, the second on a smaller input (the static array keeps the first call's
contents); sizes beyond the array are rejected before any launch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import p2z

pytestmark = pytest.mark.gpu


def _summaries(ctx, nevts, ntrks, W=32, nl=20):
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    dt, dh, do = ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)
    try:
        ctx.run(dt, dh, do, nevts, nb, nl, W, "ref")
        a = ctx.summary_device(do, nevts, nb, W, ntrks, last).as_dict()
        b = ctx.summary_device_noalloc(do, nevts, nb, W, ntrks, last).as_dict()
        host = p2z.summary_host(ctx.download(do), nevts, nb, W, ntrks, last).as_dict()
    finally:
        for buf in (dt, dh, do):
            buf.free()
    return a, b, host


def test_noalloc_summary_equals_summary_device():
    with p2z.Context(0) as ctx:
        for nevts, ntrks in ((12, 9600), (3, 960)):
            print(f"case noalloc-summary {nevts}x{ntrks}", flush=True)
            a, b, host = _summaries(ctx, nevts, ntrks)
            assert a == b, (a, b)                                  # same kernels, same order: same bits
            assert b["ngood"] + b["nnans"] + b["nfail"] == nevts * ntrks
            for k in a:
                assert abs(b[k] - host[k]) <= 1e-12 * max(1.0, abs(host[k])), (k, b[k], host[k])


def test_noalloc_summary_rejects_oversize():
    with p2z.Context(0) as ctx:
        buf = ctx.alloc(4096)
        s = p2z.Summary()
        last = (C.c_float * 3)(0.0, 0.0, 0.0)
        ntrks = 9600
        nevts = p2z.SUMMARY_NOALLOC_MAX_TRACKS // ntrks + 1
        rc = p2z.lib().p2z_summary_device_noalloc(ctx.handle, buf.ptr, nevts, ntrks // 32, 32, ntrks, last,
                                                  C.byref(s))
        assert rc == p2z.EINVAL and "static report scratch" in p2z.last_error()
        buf.free()
