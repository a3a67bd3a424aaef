"""GPU test of the reference-typing mode's sin/cos (p2z_sincosf_ref, the
device build of p2z-tests_amd/csrc/p2z_libm.hpp) against the C library the
reference calls (glibc sinf / cosf, OMP:633-636, 643-644, 652-653) on every
one of the 2^32 float inputs, bit for bit (NaN results compare equal)."""
from __future__ import annotations

import time

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _same(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


def test_device_sincosf_equals_c_library_on_every_float(ctx):
    chunk = 1 << 27
    ds, dc = ctx.alloc(4 * chunk), ctx.alloc(4 * chunk)
    hs, hc = np.empty(chunk, np.float32), np.empty(chunk, np.float32)
    bad = 0
    t0 = time.time()
    try:
        for k in range(1 << 32 >> 27):
            first = k * chunk
            ctx.sincosf_ref(first, chunk, ds, dc)
            ctx.d2h(hs, ds)
            ctx.d2h(hc, dc)
            es, ec = O.libm_sincosf_range(first, chunk)
            ok = _same(hs, es) & _same(hc, ec)
            if not ok.all():
                i = int(np.flatnonzero(~ok)[0])
                print(f"  first difference at bits {first + i:08x}: device {hs[i]!r} {hc[i]!r}, "
                      f"library {es[i]!r} {ec[i]!r}", flush=True)
                bad += int((~ok).sum())
            if k % 8 == 7:
                print(f"  {(k + 1) * chunk:#x} inputs checked, {bad} differ ({time.time() - t0:.1f} s)", flush=True)
    finally:
        ds.free()
        dc.free()
    assert bad == 0
