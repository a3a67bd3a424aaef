"""CPU tests of the product library: it loads, exports every entry point the
header declares, and its host parts (generator, presets, report, argument
checks) agree with the pinned oracle.  No GPU compute is called here."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import p2z
import pyoracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions() -> list[str]:
    names = []
    for h in os.listdir(os.path.join(REPO, "include")):
        if h.endswith(".h"):
            with open(os.path.join(REPO, "include", h)) as f:
                text = f.read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(p2z_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(p2z.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert p2z.lib().p2z_version() == 2


def test_binding_covers_header():
    assert set(declared_functions()) <= set(p2z.exported_symbols()) | EXTRA_BOUND


EXTRA_BOUND = set()  # entry points bound outside p2z._SIGS (none)


def test_presets_match_reference_input(golden):
    t, hits, smear = p2z.preset_input("ref")
    g = golden["ref_input"]
    assert smear == g["smear"]
    assert list(np.float32(t.par)) == list(np.float32(g["track"][:6]))
    assert list(np.float32(t.cov)) == list(np.float32(g["track"][6:27]))
    assert t.q == int(g["track"][27])
    for h, row in zip(hits, g["hits"]):
        assert list(np.float32(h.pos)) == list(np.float32(row[:3]))
        assert list(np.float32(h.cov)) == list(np.float32(row[3:]))
    t0, h0, s0 = p2z.preset_input("config0")
    assert s0 == 1e-5 and t0.q == -1 and len(h0) == 26
    assert all(list(h.pos) == list(h0[0].pos) for h in h0)


@pytest.mark.parametrize("name", ["small", "w16"])
def test_generator_matches_reference(golden, name):
    g = golden[name]
    trk, hit, last, nb = p2z.generate("ref", int(g["nevts"]), int(g["ntrks"]), int(g["width"]), int(g["nlayer"]))
    assert nb == int(g["nb"])
    assert np.array_equal(trk.view(np.uint32), g["trk"].view(np.uint32))
    assert np.array_equal(hit.view(np.uint32), g["hit"].view(np.uint32))


def test_randn_stream_matches_oracle():
    # both draw from libc's rand(): reseed before each stream
    p2z.srand(123)
    a = [p2z.lib().p2z_randn(0.5, 2.0) for _ in range(1001)]
    O.lib().oracle_srand(123)
    b = [O.lib().oracle_randn(0.5, 2.0) for _ in range(1001)]
    assert np.array_equal(np.float32(a).view(np.uint32), np.float32(b).view(np.uint32))


def test_host_summary_matches_oracle(golden):
    g = golden["small"]
    nevts, nb, W, ntrks = (int(g[k]) for k in ("nevts", "nb", "width", "ntrks"))
    last = np.array(golden["ref_input"]["hits"][19][:3], np.float32)
    mine = p2z.summary_host(g["out"], nevts, nb, W, ntrks, last)
    theirs = O.summary(g["out"], nevts, nb, W, ntrks, last)
    assert mine.as_dict() == theirs.as_dict()
    assert p2z.format_summary(mine) == O.format_summary(theirs)


def test_summary_counts_nans_and_failures(golden):
    g = golden["small"]
    nevts, nb, W, ntrks = (int(g[k]) for k in ("nevts", "nb", "width", "ntrks"))
    out = g["out"].copy()
    out[0 * W + 3] = np.nan          # block 0 lane 3: x is NaN
    out[3 * W + 5] = 1e-3            # block 0 lane 5: ipt -> pt = 1000, outside the window
    last = np.array(golden["ref_input"]["hits"][19][:3], np.float32)
    mine = p2z.summary_host(out, nevts, nb, W, ntrks, last)
    theirs = O.summary(out, nevts, nb, W, ntrks, last)
    assert mine.nnans == 1 and mine.nfail == 1 and mine.ngood == nevts * ntrks - 2
    assert mine.as_dict() == theirs.as_dict()


def test_argument_errors_are_reported_not_fatal():
    t, hits, smear = p2z.preset_input("ref")
    out = np.empty(p2z.trk_words(1, 32), np.float32)
    rc = p2z.lib().p2z_prepare_tracks(C.byref(t), out.ctypes.data, 1, 1, 33, smear)
    assert rc == p2z.EINVAL and "width" in p2z.last_error()
    arr = (p2z.AHIT * 26)(*hits)
    hout = np.empty(p2z.hit_words(1, 27, 32), np.float32)
    rc = p2z.lib().p2z_prepare_hits(arr, 26, hout.ctypes.data, 1, 1, 32, 27, smear)
    assert rc == p2z.EINVAL and "nlayer" in p2z.last_error()
    assert p2z.lib().p2z_run(None, None, None, None, 1, 1, 1, 32, 0) == p2z.EINVAL
    assert p2z.lib().p2z_summary_host(None, 1, 1, 32, 1, None, None) == p2z.EINVAL
    s = p2z.Summary()
    last = (C.c_float * 3)(1, 2, 3)
    assert p2z.lib().p2z_summary_host(out.ctypes.data, 1, 1, 32, 33, last, C.byref(s)) == p2z.EINVAL


def test_create_without_gpu_fails_cleanly():
    if p2z.device_count() > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert p2z.lib().p2z_create(C.byref(h), 0, None) == p2z.ENODEV
    assert "no GPU" in p2z.last_error()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message before
    any GPU work (no context is created)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def test_bench_reads_the_committed_profile_figures():
    """The roofline fields bench.py adds from the committed rocprofv3 and
    calibration summaries parse, and stay consistent with each other: HBM
    traffic within 1% of the algorithmic bytes of the main workload, every
    mode's VALU count per track-layer present, and the achievable streaming
    bandwidth between the copy kernel's range and the 8 TB/s peak."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    alg = 100 * 9600 * 20 * bench.bytes_per_track_layer(20)
    for math in ("fast", "ref"):
        traffic = bench.read_traffic(math)
        assert traffic is not None and abs(traffic / alg - 1) < 0.01, (math, traffic, alg)
    for math in p2z.MATH:
        v = bench.valu_roofline(math, 1e11)
        assert v is not None and 150 < v["lane_instr_per_tl"] < 1000, (math, v)
    a = bench.achievable_hbm(5000.0)
    assert a is not None and 4000 < a["value"] < bench.HBM_PEAK_GBS
    assert abs(a["frac"] - 5000.0 / a["value"]) < 1e-3
