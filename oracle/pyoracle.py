"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle_p2z.so (the C restatement of the reference path) and,
where it was built, the reference itself from oracle/_ref/.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, and
only as a checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle_p2z.so")
# the same oracle with correctly rounded sin/cos in the path
ORACLE_LIB_CR = os.path.join(HERE, "liboracle_p2z_crlibm.so")
REF_DIR = os.path.join(HERE, "_ref")

TRK_WORDS, HIT_WORDS = 28, 9
vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double


class OATRK(C.Structure):
    _fields_ = [("par", C.c_float * 6), ("cov", C.c_float * 21), ("q", C.c_int32)]


class OAHIT(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("cov", C.c_float * 6)]


FIELDS = ("avgx avgy avgz avgdx avgdy avgdz avgpt avgphi avgtheta "
          "stdx stdy stdz stddx stddy stddz").split()


class OSummary(C.Structure):
    _fields_ = [(n, C.c_double) for n in FIELDS] + [("nnans", C.c_int), ("nfail", C.c_int), ("ngood", C.c_int)]

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n in FIELDS}
        d.update(nnans=self.nnans, nfail=self.nfail, ngood=self.ngood)
        return d


_SIGS = {
    "oracle_srand": ([C.c_uint], None),
    "oracle_randn": ([C.c_float, C.c_float], C.c_float),
    "oracle_prepare_tracks": ([C.POINTER(OATRK), vp, i64, i64, C.c_int, dbl], None),
    "oracle_prepare_hits": ([C.POINTER(OAHIT), C.c_int, vp, i64, i64, C.c_int, C.c_int, dbl], C.c_int),
    "oracle_propagate_to_z": ([vp, vp, vp, vp, vp, vp, C.c_int], None),
    "oracle_kalman_update_v2": ([vp, vp, vp, vp, C.c_int], None),
    "oracle_kalman_update_v1": ([vp, vp, vp, vp, C.c_int, C.c_int], None),
    "oracle_propagate_to_z_fp64": ([vp, vp, vp, vp, vp, vp, C.c_int], None),
    "oracle_kalman_update_v2_fp64": ([vp, vp, vp, vp, C.c_int], None),
    "oracle_run": ([vp, vp, vp, i64, i64, C.c_int, C.c_int, C.c_int], None),
    "oracle_run_events_only": ([vp, vp, vp, i64, i64, C.c_int, C.c_int, C.c_int], None),
    "oracle_run_fp64": ([vp, vp, vp, i64, i64, C.c_int, C.c_int, C.c_int], None),
    "oracle_run_ragged": ([vp, vp, vp, vp, i64, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int], None),
    "oracle_summary": ([vp, i64, i64, C.c_int, i64, C.POINTER(C.c_float), C.POINTER(OSummary)], None),
    "oracle_format_summary": ([C.POINTER(OSummary), C.c_char_p, C.c_int], C.c_int),
    "oracle_libm_sincosf_range": ([C.c_uint32, i64, vp, vp, C.c_int], None),
}

_libs: dict = {}
_variant = ["glibc"]


def lib(variant: str | None = None):
    """The oracle library: variant 'glibc' (sinf/cosf of the C library, as the
    reference) or 'cr' (correctly rounded sin/cos).  Default: the current
    variant (see use_variant)."""
    v = variant or _variant[0]
    if v not in _libs:
        path = ORACLE_LIB if v == "glibc" else ORACLE_LIB_CR
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        for n, (a, r) in _SIGS.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r
        _libs[v] = L
    return _libs[v]


class use_variant:
    """with use_variant('cr'): ... runs the oracle functions on that variant."""

    def __init__(self, v: str):
        self.v = v

    def __enter__(self):
        self.prev = _variant[0]
        _variant[0] = self.v

    def __exit__(self, *exc):
        _variant[0] = self.prev


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def atrk(a) -> OATRK:
    t = OATRK()
    for i in range(6):
        t.par[i] = float(a[i])
    for i in range(21):
        t.cov[i] = float(a[6 + i])
    t.q = int(a[27])
    return t


def ahits(a):
    a = np.asarray(a, np.float64).reshape(-1, 9)
    arr = (OAHIT * len(a))()
    for k, row in enumerate(a):
        for i in range(3):
            arr[k].pos[i] = float(row[i])
        for i in range(6):
            arr[k].cov[i] = float(row[3 + i])
    return arr


def generate(trk28, hits9, nevts: int, nb: int, width: int, nlayer: int, smear: float, seed: int = 1):
    """srand(seed) + prepareTracks + prepareHits of the reference (OMP:715-720)."""
    L = lib()
    t = atrk(trk28)
    h = ahits(hits9)
    trk = np.empty(nevts * nb * TRK_WORDS * width, np.float32)
    hit = np.empty(nevts * nb * nlayer * HIT_WORDS * width, np.float32)
    L.oracle_srand(seed)
    L.oracle_prepare_tracks(C.byref(t), _p(trk), nevts, nb, width, smear)
    rc = L.oracle_prepare_hits(h, len(h), _p(hit), nevts, nb, width, nlayer, smear)
    if rc != 0:
        raise ValueError("nlayer exceeds the number of input hits")
    return trk, hit


def run(trk, hit, nevts, nb, width, nlayer, nthreads=0) -> np.ndarray:
    out = np.empty_like(trk)
    lib().oracle_run(_p(trk), _p(hit), _p(out), nevts, nb, width, nlayer, nthreads)
    return out


def run_events_only(trk, hit, nevts, nb, width, nlayer, nthreads=0) -> np.ndarray:
    out = np.empty_like(trk)
    lib().oracle_run_events_only(_p(trk), _p(hit), _p(out), nevts, nb, width, nlayer, nthreads)
    return out


def run_fp64(trk, hit, nevts, nb, width, nlayer, nthreads=0) -> np.ndarray:
    """Full-FP64 oracle; returns (ntracks, 28) float64, track-major."""
    out = np.empty((nevts * nb * width, 28), np.float64)
    lib().oracle_run_fp64(_p(trk), _p(hit), _p(out), nevts, nb, width, nlayer, nthreads)
    return out


def fp64_blocks(o64: np.ndarray, trk: np.ndarray, width: int) -> np.ndarray:
    """Track-major FP64 results (run_fp64) as AoSoA float32 blocks with the
    input's charges, the layout the report (summary) reads."""
    nblk = o64.shape[0] // width
    out = np.ascontiguousarray(o64.reshape(nblk, width, 28).transpose(0, 2, 1)).astype(np.float32)
    out[:, 27, :] = trk.reshape(nblk, 28, width)[:, 27, :]
    return out.reshape(-1)


def run_ragged(trk, hit, nblocks, width, blk_trk, blk_hit, blk_nlay, lane_nlay, fp64=False, nthreads=0):
    blk_trk = np.ascontiguousarray(blk_trk, np.int64)
    blk_hit = np.ascontiguousarray(blk_hit, np.int64)
    blk_nlay = np.ascontiguousarray(blk_nlay, np.int32)
    lane_nlay = np.ascontiguousarray(lane_nlay, np.int32)
    out = np.zeros_like(trk)
    out64 = np.zeros((trk.size // TRK_WORDS, 28), np.float64) if fp64 else np.zeros(1, np.float64)
    lib().oracle_run_ragged(_p(trk), _p(hit), _p(out), _p(out64), nblocks, width, _p(blk_trk), _p(blk_hit),
                            _p(blk_nlay), _p(lane_nlay), 1 if fp64 else 0, nthreads)
    return out64 if fp64 else out


def propagate_block(trk_blk, hit_blk, width) -> np.ndarray:
    """oracle propagateToZ on one MPTRK block (returns a new block)."""
    W = width
    out = trk_blk.copy()
    lib().oracle_propagate_to_z(_p(trk_blk) + 6 * W * 4, _p(trk_blk), _p(trk_blk) + 27 * W * 4, _p(hit_blk),
                                _p(out) + 6 * W * 4, _p(out), W)
    return out


def update_block(trk_blk, hit_blk, width) -> np.ndarray:
    """oracle KalmanUpdate_v2 on one MPTRK block (returns a new block)."""
    W = width
    out = trk_blk.copy()
    lib().oracle_kalman_update_v2(_p(out) + 6 * W * 4, _p(out), _p(hit_blk) + 3 * W * 4, _p(hit_blk), W)
    return out


def update_v1_block(trk_blk, hit_blk, width, fixed: bool) -> np.ndarray:
    """oracle KalmanUpdate v1 (OMP:375-501) on one MPTRK block; fixed selects
    the textbook form in the path's packing (returns a new block)."""
    W = width
    out = trk_blk.copy()
    lib().oracle_kalman_update_v1(_p(out) + 6 * W * 4, _p(out), _p(hit_blk) + 3 * W * 4, _p(hit_blk), W,
                                  int(bool(fixed)))
    return out


def propagate_block_fp64(trk_blk, hit_blk, width) -> np.ndarray:
    """propagateToZ on one block in double, rounded to float."""
    W = width
    out = trk_blk.copy()
    lib().oracle_propagate_to_z_fp64(_p(trk_blk) + 6 * W * 4, _p(trk_blk), _p(trk_blk) + 27 * W * 4, _p(hit_blk),
                                     _p(out) + 6 * W * 4, _p(out), W)
    return out


def update_block_fp64(trk_blk, hit_blk, width) -> np.ndarray:
    """KalmanUpdate_v2 on one block in double, rounded to float."""
    W = width
    out = trk_blk.copy()
    lib().oracle_kalman_update_v2_fp64(_p(out) + 6 * W * 4, _p(out), _p(hit_blk) + 3 * W * 4, _p(hit_blk), W)
    return out


def summary(out, nevts, nb, width, ntrks, lasthit) -> dict:
    s = OSummary()
    last = (C.c_float * 3)(*[float(v) for v in lasthit])
    lib().oracle_summary(_p(out), nevts, nb, width, ntrks, last, C.byref(s))
    return s


def libm_sincosf_range(first_bits: int, n: int, nthreads: int = 0):
    """The C library's sinf and cosf for the float bit patterns first_bits ..
    first_bits + n - 1; returns (s, c) float32 arrays."""
    s = np.empty(n, np.float32)
    c = np.empty(n, np.float32)
    lib("glibc").oracle_libm_sincosf_range(first_bits, n, _p(s), _p(c), nthreads)
    return s, c


def format_summary(s: OSummary) -> str:
    buf = C.create_string_buffer(4096)
    lib().oracle_format_summary(C.byref(s), buf, 4096)
    return buf.value.decode()


# ---------------------------------------------------------------------------
# the reference compiled from its own source (oracle/_ref, built where the
# reference tree exists)
# ---------------------------------------------------------------------------

def ref_lib(name: str):
    """Loads oracle/_ref/libref_<name>.so or returns None when absent."""
    path = os.path.join(REF_DIR, f"libref_{name}.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    for n in ("ref_bsize", "ref_nevts", "ref_ntrks", "ref_nb", "ref_nlayer", "ref_niter"):
        getattr(L, n).restype = C.c_int
    L.ref_generate.argtypes = [vp, vp]
    L.ref_run.argtypes = [vp, vp, vp]
    L.ref_inputs.argtypes = [vp, vp]
    L.ref_inputs.restype = C.c_int
    L.ref_propagate_block.argtypes = [vp, vp, vp]
    L.ref_update_block.argtypes = [vp, vp]
    L.ref_update_v1_block.argtypes = [vp, vp]
    L.ref_randn.argtypes = [C.c_float, C.c_float]
    L.ref_randn.restype = C.c_float
    L.ref_srand.argtypes = [C.c_uint]
    return L


def ref_config(L) -> dict:
    return dict(nevts=L.ref_nevts(), ntrks=L.ref_ntrks(), width=L.ref_bsize(), nb=L.ref_nb(),
                nlayer=L.ref_nlayer(), niter=L.ref_niter())


def ref_generate(L):
    c = ref_config(L)
    trk = np.empty(c["nevts"] * c["nb"] * TRK_WORDS * c["width"], np.float32)
    hit = np.empty(c["nevts"] * c["nb"] * c["nlayer"] * HIT_WORDS * c["width"], np.float32)
    L.ref_generate(_p(trk), _p(hit))
    return trk, hit


def ref_run(L, trk, hit) -> np.ndarray:
    out = np.empty_like(trk)
    L.ref_run(_p(trk), _p(hit), _p(out))
    return out


def ref_inputs(L):
    t = np.empty(28, np.float32)
    h = np.empty(26 * 9, np.float32)
    n = L.ref_inputs(_p(t), _p(h))
    return t, h[: n * 9].reshape(n, 9)
