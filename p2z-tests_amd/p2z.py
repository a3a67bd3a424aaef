"""ctypes binding of libp2z.so, the C ABI declared in include/p2z.h.

Host arrays are numpy arrays in the reference's AoSoA layout (float32 words;
the charge lanes hold int32 bit patterns):
    MPTRK block: par[6W] | cov[21W] | q[W]   (28W words)
    MPHIT block: pos[3W] | cov[6W]           ( 9W words)
Device memory is owned by DeviceBuffer objects allocated through the library.

The library must be built (``python -c "import __graft_entry__ as g; g.build()"``
at the repository root); loading fails loudly otherwise.  There is no CPU
fallback: every compute call runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# P2Z_LIBRARY selects another build of the library (same ABI), e.g. a variant.
LIB_PATH = os.environ.get("P2Z_LIBRARY") or os.path.join(HERE, "lib", "libp2z.so")
MAIN_PATH = os.path.join(HERE, "bin", "p2z_main")

TRK_WORDS = 28
HIT_WORDS = 9
WIDTHS = (16, 32, 64, 128)
MATH = {"fast": 0, "ref": 1, "fp64gain": 2}
PRESET = {"ref": 0, "config0": 1}

OK, EINVAL, ENOMEM, EHIP, ENODEV, EBUSY, ESTATE = 0, -1, -2, -3, -4, -5, -6


class P2zError(RuntimeError):
    def __init__(self, rc: int, what: str, msg: str):
        super().__init__(f"{what} failed (rc={rc}): {msg}")
        self.rc = rc


class ATRK(C.Structure):
    _fields_ = [("par", C.c_float * 6), ("cov", C.c_float * 21), ("q", C.c_int32)]


class AHIT(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("cov", C.c_float * 6)]


class Opts(C.Structure):
    _fields_ = [("tracks_per_wg", C.c_int32), ("nstreams", C.c_int32),
                ("hit_stage", C.c_int32), ("reserved", C.c_int32)]


SUMMARY_FIELDS = ("avgx avgy avgz avgdx avgdy avgdz avgpt avgphi avgtheta "
                  "stdx stdy stdz stddx stddy stddz").split()


class Summary(C.Structure):
    _fields_ = [(n, C.c_double) for n in SUMMARY_FIELDS] + [
        ("nnans", C.c_int32), ("nfail", C.c_int32), ("ngood", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n in SUMMARY_FIELDS}
        d.update(nnans=self.nnans, nfail=self.nfail, ngood=self.ngood)
        return d


class BlockDesc(C.Structure):
    _fields_ = [("trk_block", C.c_int64), ("hit_block", C.c_int64), ("out_block", C.c_int64),
                ("nlayer", C.c_int32), ("event", C.c_int32)]


BLOCK_DESC_DTYPE = np.dtype([("trk_block", "<i8"), ("hit_block", "<i8"), ("out_block", "<i8"), ("nlayer", "<i4"),
                             ("event", "<i4")])


class RaggedOpts(C.Structure):
    _fields_ = [("tracks_per_wg", C.c_int32), ("prefetch", C.c_int32), ("waves_per_cu", C.c_int32),
                ("items_per_claim", C.c_int32)]


class WorkloadSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("width", C.c_int32), ("nevents", C.c_int64), ("seed", C.c_uint64),
                ("ntracks", C.c_int32), ("nlayer", C.c_int32), ("nthreads", C.c_int32), ("reserved", C.c_int32)]


class WorkloadInfo(C.Structure):
    _fields_ = [("nevents", C.c_int64), ("ntracks", C.c_int64), ("nblocks", C.c_int64), ("nhit_blocks", C.c_int64),
                ("track_layers", C.c_int64), ("width", C.c_int32), ("kind", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("budget_bytes", C.c_int64), ("nstreams", C.c_int32), ("max_batch_blocks", C.c_int32),
                ("width", C.c_int32), ("math", C.c_int32), ("tracks_per_wg", C.c_int32), ("prefetch", C.c_int32),
                ("waves_per_cu", C.c_int32), ("items_per_claim", C.c_int32), ("pool_blocks", C.c_int64),
                ("pool_hit_blocks", C.c_int64), ("flags", C.c_uint32), ("reserved", C.c_int32)]


class Completion(C.Structure):
    _fields_ = [("request", C.c_int64), ("out_block", C.c_int64), ("nblocks", C.c_int64), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_launch", C.c_double), ("t_done", C.c_double)]


class ServeReport(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("nevents", "track_layers", "batches", "max_inflight_events",
                                          "busy_rejections")] + \
               [(n, C.c_double) for n in ("rate_offered", "rate_achieved", "t_span", "tl_per_s", "lat_p50", "lat_p99",
                                          "lat_p999", "lat_mean", "lat_max", "mean_batch_events",
                                          "host_submit_s", "host_pump_s")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


RAGGED_PERSISTENT, RAGGED_GRAPH = 1, 2
V1_FIXED = 1
SUMMARY_NOALLOC_MAX_TRACKS = 11184128
SERVE_RESIDENT = 1
WORKLOAD = {"precision": 2, "ragged": 3}

_lib = None
vp, i32, i64, u32, dbl, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_double, C.c_size_t

_SIGS = {
    "p2z_version": ([], i32),
    "p2z_last_error": ([], C.c_char_p),
    "p2z_device_count": ([C.POINTER(C.c_int)], i32),
    "p2z_create": ([C.POINTER(vp), C.c_int, C.POINTER(Opts)], i32),
    "p2z_destroy": ([vp], i32),
    "p2z_sync": ([vp], i32),
    "p2z_set_opts": ([vp, C.POINTER(Opts)], i32),
    "p2z_get_stream": ([vp, C.POINTER(vp)], i32),
    "p2z_dev_alloc": ([vp, C.POINTER(vp), sz], i32),
    "p2z_dev_free": ([vp, vp], i32),
    "p2z_host_alloc": ([vp, C.POINTER(vp), sz], i32),
    "p2z_host_free": ([vp, vp], i32),
    "p2z_memcpy_h2d": ([vp, vp, vp, sz], i32),
    "p2z_memcpy_d2h": ([vp, vp, vp, sz], i32),
    "p2z_timer_start": ([vp], i32),
    "p2z_timer_stop": ([vp, C.POINTER(C.c_float)], i32),
    "p2z_srand": ([C.c_uint], None),
    "p2z_randn": ([C.c_float, C.c_float], C.c_float),
    "p2z_prepare_tracks": ([C.POINTER(ATRK), vp, i64, i64, i32, dbl], i32),
    "p2z_prepare_hits": ([C.POINTER(AHIT), i32, vp, i64, i64, i32, i32, dbl], i32),
    "p2z_preset_input": ([i32, C.POINTER(ATRK), C.POINTER(AHIT), i32, C.POINTER(dbl)], i32),
    "p2z_propagate_to_z": ([vp, vp, vp, vp, vp, vp, vp, i64, i64, i64, i32, i32], i32),
    "p2z_kalman_update": ([vp, vp, vp, vp, vp, i64, i64, i64, i32, i32], i32),
    "p2z_kalman_update_v1": ([vp, vp, vp, vp, vp, i64, i64, i64, i32, u32], i32),
    "p2z_run": ([vp, vp, vp, vp, i64, i64, i32, i32, i32], i32),
    "p2z_run_host": ([vp, vp, vp, vp, i64, i64, i32, i32, i32], i32),
    "p2z_run_ragged": ([vp, vp, vp, vp, vp, vp, i64, i32, i32, u32, C.POINTER(RaggedOpts)], i32),
    "p2z_check_blocks": ([vp, i64, i64, i64, i64], i32),
    "p2z_workload_create": ([C.POINTER(WorkloadSpec), C.POINTER(vp)], i32),
    "p2z_workload_destroy": ([vp], i32),
    "p2z_workload_get_info": ([vp, C.POINTER(WorkloadInfo)], i32),
    "p2z_workload_tracks": ([vp], vp),
    "p2z_workload_hits": ([vp], vp),
    "p2z_workload_lane_nlayer": ([vp], vp),
    "p2z_workload_blocks": ([vp], vp),
    "p2z_workload_blocks_lpt": ([vp], vp),
    "p2z_workload_event_blocks": ([vp], vp),
    "p2z_workload_event_tracks": ([vp], vp),
    "p2z_workload_save": ([vp, C.c_char_p], i32),
    "p2z_workload_load": ([C.c_char_p, C.POINTER(vp)], i32),
    "p2z_now": ([], dbl),
    "p2z_server_create": ([vp, vp, vp, vp, C.POINTER(ServerOpts), C.POINTER(vp)], i32),
    "p2z_server_destroy": ([vp], i32),
    "p2z_server_submit": ([vp, i64, vp, i64, dbl], i32),
    "p2z_server_pump": ([vp], i32),
    "p2z_server_poll": ([vp, C.POINTER(Completion), i32], i32),
    "p2z_server_release": ([vp, i64], i32),
    "p2z_server_arena": ([vp], vp),
    "p2z_serve_poisson": ([vp, vp, vp, i64, dbl, C.c_uint64, i32, C.POINTER(ServeReport), vp, vp], i32),
    "p2z_summary_device": ([vp, vp, i64, i64, i32, i64, C.POINTER(C.c_float), C.POINTER(Summary)], i32),
    "p2z_summary_host": ([vp, i64, i64, i32, i64, C.POINTER(C.c_float), C.POINTER(Summary)], i32),
    "p2z_format_summary": ([C.POINTER(Summary), C.c_char_p, i32], i32),
    "p2z_sincosf_ref": ([vp, u32, i64, vp, vp], i32),
    "p2z_summary_device_noalloc": ([vp, vp, i64, i64, i32, i64, C.POINTER(C.c_float), C.POINTER(Summary)], i32),
}


def lib():
    """Loads libp2z.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() at the repository root")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def last_error() -> str:
    return lib().p2z_last_error().decode()


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise P2zError(rc, what, last_error())
    return rc


def _ptr(a) -> int:
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if isinstance(a, int):
        return a
    raise TypeError(f"cannot pass {type(a)} as a buffer")


def _need(a, nbytes: int, what: str) -> None:
    """Raises ValueError when a device buffer or array holds fewer than nbytes
    (a raw int pointer is the caller's responsibility)."""
    have = a.nbytes if isinstance(a, (DeviceBuffer, np.ndarray)) else None
    if have is not None and have < nbytes:
        raise ValueError(f"{what}: {nbytes} bytes needed, buffer holds {have}")


# ---------------------------------------------------------------------------
# host-side inputs
# ---------------------------------------------------------------------------

def trk_words(nblocks: int, width: int) -> int:
    return nblocks * TRK_WORDS * width


def hit_words(nblocks: int, nlayer: int, width: int) -> int:
    return nblocks * nlayer * HIT_WORDS * width


def preset_input(name: str):
    """Returns (ATRK, [AHIT]*26, smear) of a preset ('ref' or 'config0')."""
    t = ATRK()
    hits = (AHIT * 26)()
    smear = C.c_double()
    n = check(lib().p2z_preset_input(PRESET[name], C.byref(t), hits, 26, C.byref(smear)), "p2z_preset_input")
    return t, list(hits)[:n], smear.value


def atrk_from_array(a) -> ATRK:
    """ATRK from 28 numbers: par[6], cov[21], q."""
    t = ATRK()
    for i in range(6):
        t.par[i] = float(a[i])
    for i in range(21):
        t.cov[i] = float(a[6 + i])
    t.q = int(a[27])
    return t


def ahits_from_array(a) -> list:
    a = np.asarray(a, dtype=np.float64).reshape(-1, 9)
    hits = []
    for row in a:
        h = AHIT()
        for i in range(3):
            h.pos[i] = float(row[i])
        for i in range(6):
            h.cov[i] = float(row[3 + i])
        hits.append(h)
    return hits


def srand(seed: int) -> None:
    lib().p2z_srand(seed)


def prepare_tracks(trk: ATRK, nevts: int, nb: int, width: int, smear: float) -> np.ndarray:
    """prepareTracks (OMP:243-263) on the current rand() stream."""
    out = np.empty(trk_words(nevts * nb, width), np.float32)
    check(lib().p2z_prepare_tracks(C.byref(trk), out.ctypes.data, nevts, nb, width, smear), "p2z_prepare_tracks")
    return out


def prepare_hits(hits: list, nevts: int, nb: int, width: int, nlayer: int, smear: float) -> np.ndarray:
    """prepareHits (OMP:265-293) on the current rand() stream."""
    arr = (AHIT * len(hits))(*hits)
    out = np.empty(hit_words(nevts * nb, nlayer, width), np.float32)
    check(lib().p2z_prepare_hits(arr, len(hits), out.ctypes.data, nevts, nb, width, nlayer, smear),
          "p2z_prepare_hits")
    return out


def generate(preset: str, nevts: int, ntrks: int, width: int, nlayer: int, seed: int = 1):
    """srand(seed) + prepareTracks + prepareHits as the reference's main does
    (OMP:715-720).  Returns (trk, hit, lasthit[3], nb)."""
    t, hits, smear = preset_input(preset)
    nb = ntrks // width
    srand(seed)
    trk = prepare_tracks(t, nevts, nb, width, smear)
    hit = prepare_hits(hits, nevts, nb, width, nlayer, smear)
    last = np.array(list(hits[nlayer - 1].pos), np.float32)
    return trk, hit, last, nb


def summary_host(out: np.ndarray, nevts: int, nb: int, width: int, ntrks: int, lasthit) -> Summary:
    s = Summary()
    last = (C.c_float * 3)(*[float(v) for v in lasthit])
    check(lib().p2z_summary_host(out.ctypes.data, nevts, nb, width, ntrks, last, C.byref(s)), "p2z_summary_host")
    return s


def format_summary(s: Summary) -> str:
    buf = C.create_string_buffer(4096)
    check(lib().p2z_format_summary(C.byref(s), buf, 4096), "p2z_format_summary")
    return buf.value.decode()


def device_count() -> int:
    n = C.c_int(0)
    check(lib().p2z_device_count(C.byref(n)), "p2z_device_count")
    return n.value


# ---------------------------------------------------------------------------
# synthetic workloads (ragged-event format)
# ---------------------------------------------------------------------------

def _view(ptr: int, dtype, count: int) -> np.ndarray:
    if count == 0:
        return np.empty(0, dtype)
    buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype, count=count)


class Workload:
    """A generated (or loaded) workload in the ragged format.  The numpy
    arrays are views of library memory, valid while the object lives."""

    def __init__(self, kind: str = "ragged", nevents: int = 2000, width: int = 32, seed: int = 1,
                 ntracks: int = 9600, nlayer: int = 20, nthreads: int = 0, _handle=None):
        if _handle is None:
            spec = WorkloadSpec(WORKLOAD[kind], width, nevents, seed, ntracks, nlayer, nthreads, 0)
            h = vp()
            check(lib().p2z_workload_create(C.byref(spec), C.byref(h)), "p2z_workload_create")
            self.handle = h
        else:
            self.handle = _handle
        info = WorkloadInfo()
        check(lib().p2z_workload_get_info(self.handle, C.byref(info)), "p2z_workload_get_info")
        self.info = info
        W = info.width
        L = lib()
        self.width = W
        self.nevents, self.ntracks, self.nblocks = info.nevents, info.ntracks, info.nblocks
        self.nhit_blocks, self.track_layers = info.nhit_blocks, info.track_layers
        self.tracks = _view(L.p2z_workload_tracks(self.handle), np.float32, info.nblocks * TRK_WORDS * W)
        self.hits = _view(L.p2z_workload_hits(self.handle), np.float32, info.nhit_blocks * HIT_WORDS * W)
        self.lane_nlayer = _view(L.p2z_workload_lane_nlayer(self.handle), np.int32, info.nblocks * W)
        self.blocks = _view(L.p2z_workload_blocks(self.handle), BLOCK_DESC_DTYPE, info.nblocks)
        self.blocks_lpt = _view(L.p2z_workload_blocks_lpt(self.handle), BLOCK_DESC_DTYPE, info.nblocks)
        self.event_blocks = _view(L.p2z_workload_event_blocks(self.handle), np.int64, info.nevents + 1)
        self.event_tracks = _view(L.p2z_workload_event_tracks(self.handle), np.int32, info.nevents)

    @classmethod
    def load(cls, path: str) -> "Workload":
        h = vp()
        check(lib().p2z_workload_load(path.encode(), C.byref(h)), "p2z_workload_load")
        return cls(_handle=h)

    def save(self, path: str) -> None:
        check(lib().p2z_workload_save(self.handle, path.encode()), "p2z_workload_save")

    def close(self) -> None:
        if self.handle:
            lib().p2z_workload_destroy(self.handle)
            self.handle = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# device side
# ---------------------------------------------------------------------------

def check_blocks(blocks: np.ndarray, pool_blocks: int, pool_hit_blocks: int, out_blocks: int) -> None:
    """p2z_check_blocks on a descriptor table (BLOCK_DESC_DTYPE array):
    raises P2zError (EINVAL) naming the first descriptor outside the buffers."""
    b = np.ascontiguousarray(blocks, dtype=BLOCK_DESC_DTYPE)
    check(lib().p2z_check_blocks(b.ctypes.data if b.size else None, b.size, pool_blocks, pool_hit_blocks, out_blocks),
          "p2z_check_blocks")


@dataclass
class DeviceBuffer:
    ctx: "Context"
    ptr: int
    nbytes: int

    def free(self) -> None:
        if self.ptr:
            check(lib().p2z_dev_free(self.ctx.handle, self.ptr), "p2z_dev_free")
            self.ptr = 0
            self.ctx._bufs = [b for b in self.ctx._bufs if b is not self]


class Server:
    """Continuous-batching server over a device-resident ragged pool
    (p2z_server_*): requests are events given by their block descriptors.
    resident=True serves them with the device-resident loop (one persistent
    kernel fed through a ring in pinned host memory) instead of batches of
    launches."""

    def __init__(self, ctx: "Context", trk, hit, lane_nlayer, budget_bytes: int, width: int, math="fast",
                 nstreams: int = 0, max_batch_blocks: int = 0, tracks_per_wg: int = 0, prefetch: int = 0,
                 waves_per_cu: int = 0, items_per_claim: int = 0, resident: bool = False):
        self.ctx = ctx
        self.width = width
        o = ServerOpts(budget_bytes, nstreams, max_batch_blocks, width, MATH.get(math, math), tracks_per_wg,
                       prefetch, waves_per_cu, items_per_claim, trk.nbytes // (TRK_WORDS * width * 4),
                       hit.nbytes // (HIT_WORDS * width * 4), SERVE_RESIDENT if resident else 0, 0)
        self.handle = vp()
        check(lib().p2z_server_create(ctx.handle, _ptr(trk), _ptr(hit), _ptr(lane_nlayer), C.byref(o),
                                      C.byref(self.handle)), "p2z_server_create")

    def close(self) -> None:
        if self.handle:
            lib().p2z_server_destroy(self.handle)
            self.handle = vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, request: int, blocks: np.ndarray, t_arrival: float | None = None) -> int:
        """Returns OK or EBUSY; raises on other errors."""
        blocks = np.ascontiguousarray(blocks, BLOCK_DESC_DTYPE)
        t = lib().p2z_now() if t_arrival is None else t_arrival
        rc = lib().p2z_server_submit(self.handle, request, blocks.ctypes.data, len(blocks), t)
        if rc == EBUSY:
            return rc
        return check(rc, "p2z_server_submit")

    def pump(self) -> int:
        return check(lib().p2z_server_pump(self.handle), "p2z_server_pump")

    def poll(self, cap: int = 256) -> list:
        arr = (Completion * cap)()
        n = check(lib().p2z_server_poll(self.handle, arr, cap), "p2z_server_poll")
        return [arr[i] for i in range(n)]

    def release(self, request: int) -> None:
        check(lib().p2z_server_release(self.handle, request), "p2z_server_release")

    def arena_ptr(self) -> int:
        return lib().p2z_server_arena(self.handle)

    def wait(self, requests, timeout: float = 60.0) -> dict:
        """Pumps until every request in `requests` is done; returns id -> completion."""
        want, got = set(requests), {}
        t0 = lib().p2z_now()
        while want - set(got):
            self.pump()
            for c in self.poll():
                got[c.request] = c
            if lib().p2z_now() - t0 > timeout:
                raise TimeoutError(f"requests not done after {timeout} s: {sorted(want - set(got))[:5]}")
        return got

    def result(self, completion, dtype=np.float32) -> np.ndarray:
        """Copies one finished request's result blocks from the arena."""
        W = self.width
        nbytes = completion.nblocks * TRK_WORDS * W * 4
        out = np.empty(nbytes // 4, dtype)
        src = self.arena_ptr() + completion.out_block * TRK_WORDS * W * 4
        check(lib().p2z_memcpy_d2h(self.ctx.handle, out.ctypes.data, src, nbytes), "p2z_memcpy_d2h")
        self.ctx.sync()
        return out

    def poisson(self, blocks: np.ndarray, event_blocks: np.ndarray, rate: float, seed: int = 1,
                release: bool = True, nevents: int | None = None):
        """Poisson arrivals through the native serving loop; returns
        (ServeReport, latencies[nevents], out_block[nevents])."""
        blocks = np.ascontiguousarray(blocks, BLOCK_DESC_DTYPE)
        event_blocks = np.ascontiguousarray(event_blocks, np.int64)
        n = len(event_blocks) - 1 if nevents is None else nevents
        rep = ServeReport()
        lat = np.zeros(n, np.float64)
        ob = np.zeros(n, np.int64)
        check(lib().p2z_serve_poisson(self.handle, blocks.ctypes.data, event_blocks.ctypes.data, n, rate, seed,
                                      1 if release else 0, C.byref(rep), lat.ctypes.data, ob.ctypes.data),
              "p2z_serve_poisson")
        return rep, lat, ob


class Context:
    """A p2z context: one GPU, its streams and buffers."""

    def __init__(self, device: int = 0, tracks_per_wg: int = 0, nstreams: int = 0, hit_stage: int = 0):
        self.handle = vp()
        self._opts = Opts(tracks_per_wg, nstreams, hit_stage, 0)
        check(lib().p2z_create(C.byref(self.handle), device, C.byref(self._opts)), "p2z_create")
        self._bufs: list[DeviceBuffer] = []
        self._pinned: list[int] = []

    # lifetime
    def close(self) -> None:
        if self.handle:
            for b in self._bufs:
                if b.ptr:
                    lib().p2z_dev_free(self.handle, b.ptr)
                    b.ptr = 0
            for p in self._pinned:
                lib().p2z_host_free(self.handle, p)
            self._pinned = []
            check(lib().p2z_destroy(self.handle), "p2z_destroy")
            self.handle = vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_opts(self, tracks_per_wg: int = 0, nstreams: int = 0, hit_stage: int = 0) -> None:
        self._opts = Opts(tracks_per_wg, nstreams, hit_stage, 0)
        check(lib().p2z_set_opts(self.handle, C.byref(self._opts)), "p2z_set_opts")

    def sync(self) -> None:
        check(lib().p2z_sync(self.handle), "p2z_sync")

    # memory
    def alloc(self, nbytes: int) -> DeviceBuffer:
        p = vp()
        check(lib().p2z_dev_alloc(self.handle, C.byref(p), nbytes), "p2z_dev_alloc")
        b = DeviceBuffer(self, p.value or 0, nbytes)
        self._bufs.append(b)
        return b

    def pinned_empty(self, n: int, dtype=np.float32) -> np.ndarray:
        """A numpy array of n elements over pinned host memory
        (p2z_host_alloc), valid until the context is closed."""
        nbytes = max(int(n) * np.dtype(dtype).itemsize, 1)
        p = vp()
        check(lib().p2z_host_alloc(self.handle, C.byref(p), nbytes), "p2z_host_alloc")
        self._pinned.append(p.value)
        return np.frombuffer((C.c_char * nbytes).from_address(p.value), dtype=dtype, count=int(n))

    def pinned_copy(self, a: np.ndarray) -> np.ndarray:
        b = self.pinned_empty(a.size, a.dtype)
        b[...] = a.reshape(-1)
        return b

    def upload(self, a: np.ndarray) -> DeviceBuffer:
        b = self.alloc(a.nbytes)
        self.h2d(b, a)
        return b

    def h2d(self, dst: DeviceBuffer, a: np.ndarray) -> None:
        if a.nbytes > dst.nbytes:
            raise ValueError(f"{a.nbytes} bytes do not fit a device buffer of {dst.nbytes}")
        check(lib().p2z_memcpy_h2d(self.handle, dst.ptr, _ptr(a), a.nbytes), "p2z_memcpy_h2d")
        self.sync()

    def d2h(self, a: np.ndarray, src: DeviceBuffer) -> np.ndarray:
        if a.nbytes > src.nbytes:
            raise ValueError(f"{a.nbytes} bytes requested from a device buffer of {src.nbytes}")
        check(lib().p2z_memcpy_d2h(self.handle, _ptr(a), src.ptr, a.nbytes), "p2z_memcpy_d2h")
        self.sync()
        return a

    def download(self, src: DeviceBuffer, dtype=np.float32) -> np.ndarray:
        a = np.empty(src.nbytes // np.dtype(dtype).itemsize, dtype)
        return self.d2h(a, src)

    # timing (hipEvents on the context's stream)
    def timer_start(self) -> None:
        check(lib().p2z_timer_start(self.handle), "p2z_timer_start")

    def timer_stop(self) -> float:
        ms = C.c_float()
        check(lib().p2z_timer_stop(self.handle, C.byref(ms)), "p2z_timer_stop")
        return ms.value

    # the path
    def _check_run(self, trk, hit, out, nevts, nb, nlayer, width) -> None:
        _need(trk, 4 * trk_words(nevts * nb, width), "tracks")
        _need(out, 4 * trk_words(nevts * nb, width), "out")
        _need(hit, 4 * hit_words(nevts * nb, nlayer, width), "hits")

    def run(self, trk, hit, out, nevts: int, nb: int, nlayer: int, width: int, math="fast") -> None:
        self._check_run(trk, hit, out, nevts, nb, nlayer, width)
        check(lib().p2z_run(self.handle, _ptr(trk), _ptr(hit), _ptr(out), nevts, nb, nlayer, width,
                            MATH.get(math, math)), "p2z_run")

    def run_host(self, trk: np.ndarray, hit: np.ndarray, out: np.ndarray, nevts: int, nb: int, nlayer: int,
                 width: int, math="fast") -> None:
        self._check_run(trk, hit, out, nevts, nb, nlayer, width)
        check(lib().p2z_run_host(self.handle, _ptr(trk), _ptr(hit), _ptr(out), nevts, nb, nlayer, width,
                                 MATH.get(math, math)), "p2z_run_host")

    def propagate_to_z(self, trk: DeviceBuffer, hit: DeviceBuffer, out: DeviceBuffer, nblocks: int,
                       width: int, math="fast", hit_stride: int | None = None) -> None:
        """propagateToZ on MPTRK blocks trk -> out with one MPHIT block per
        track block (hit block b at b*hit_stride words)."""
        W = width
        hs = HIT_WORDS * W if hit_stride is None else hit_stride
        self._check_blocks(trk, hit, nblocks, W, hs, out)
        f4 = 4
        check(lib().p2z_propagate_to_z(self.handle, trk.ptr + 6 * W * f4, trk.ptr, trk.ptr + 27 * W * f4, hit.ptr,
                                       out.ptr + 6 * W * f4, out.ptr, nblocks, TRK_WORDS * W, hs, W,
                                       MATH.get(math, math)), "p2z_propagate_to_z")

    @staticmethod
    def _check_blocks(trk, hit, nblocks: int, W: int, hs: int, out=None) -> None:
        if nblocks <= 0:
            return
        _need(trk, 4 * trk_words(nblocks, W), "tracks")
        if out is not None:
            _need(out, 4 * trk_words(nblocks, W), "out")
        _need(hit, 4 * ((nblocks - 1) * hs + HIT_WORDS * W), "hits")

    def kalman_update(self, trk: DeviceBuffer, hit: DeviceBuffer, nblocks: int, width: int, math="fast",
                      hit_stride: int | None = None) -> None:
        """KalmanUpdate_v2 in place on MPTRK blocks with one MPHIT block each."""
        W = width
        hs = HIT_WORDS * W if hit_stride is None else hit_stride
        self._check_blocks(trk, hit, nblocks, W, hs)
        f4 = 4
        check(lib().p2z_kalman_update(self.handle, trk.ptr + 6 * W * f4, trk.ptr, hit.ptr + 3 * W * f4, hit.ptr,
                                      nblocks, TRK_WORDS * W, hs, W, MATH.get(math, math)), "p2z_kalman_update")

    def kalman_update_v1(self, trk: DeviceBuffer, hit: DeviceBuffer, nblocks: int, width: int,
                         fixed: bool = False, hit_stride: int | None = None) -> None:
        """KalmanUpdate v1 (3-D hit) in place on MPTRK blocks with one MPHIT
        block each; fixed selects the textbook form (P2Z_V1_FIXED)."""
        W = width
        hs = HIT_WORDS * W if hit_stride is None else hit_stride
        self._check_blocks(trk, hit, nblocks, W, hs)
        f4 = 4
        check(lib().p2z_kalman_update_v1(self.handle, trk.ptr + 6 * W * f4, trk.ptr, hit.ptr + 3 * W * f4, hit.ptr,
                                         nblocks, TRK_WORDS * W, hs, W, V1_FIXED if fixed else 0),
              "p2z_kalman_update_v1")

    def run_ragged(self, trk, hit, out, lane_nlayer, blocks, nblocks: int, width: int, math="fast",
                   persistent: bool = True, graph: bool = False, tracks_per_wg: int = 0, prefetch: int = 0,
                   waves_per_cu: int = 0, items_per_claim: int = 0) -> None:
        """All blocks of a descriptor table (device buffer of BlockDesc) through
        their layers: one persistent work-queue launch (optionally replayed from
        a captured hipGraph), or one wave per 64 track slots."""
        _need(blocks, nblocks * C.sizeof(BlockDesc), "block descriptors")
        flags = (RAGGED_PERSISTENT if persistent else 0) | (RAGGED_GRAPH if graph else 0)
        o = RaggedOpts(tracks_per_wg, prefetch, waves_per_cu, items_per_claim)
        check(lib().p2z_run_ragged(self.handle, _ptr(trk), _ptr(hit), _ptr(out), _ptr(lane_nlayer), _ptr(blocks),
                                   nblocks, width, MATH.get(math, math), flags, C.byref(o)), "p2z_run_ragged")

    def summary_device_noalloc(self, out: DeviceBuffer, nevts: int, nb: int, width: int, ntrks: int,
                               lasthit) -> Summary:
        """summary_device with its scratch in the library's static device
        array (no allocation; at most SUMMARY_NOALLOC_MAX_TRACKS tracks)."""
        _need(out, 4 * trk_words(nevts * nb, width), "out")
        s = Summary()
        last = (C.c_float * 3)(*[float(v) for v in lasthit])
        check(lib().p2z_summary_device_noalloc(self.handle, _ptr(out), nevts, nb, width, ntrks, last, C.byref(s)),
              "p2z_summary_device_noalloc")
        return s

    def sincosf_ref(self, first_bits: int, n: int, s: DeviceBuffer, c: DeviceBuffer) -> None:
        """sin and cos of the reference-typing mode for the float bit patterns
        first_bits .. first_bits + n - 1, into device buffers s and c."""
        _need(s, 4 * n, "sin out")
        _need(c, 4 * n, "cos out")
        check(lib().p2z_sincosf_ref(self.handle, first_bits, n, _ptr(s), _ptr(c)), "p2z_sincosf_ref")

    def summary_device(self, out: DeviceBuffer, nevts: int, nb: int, width: int, ntrks: int, lasthit) -> Summary:
        _need(out, 4 * trk_words(nevts * nb, width), "out")
        s = Summary()
        last = (C.c_float * 3)(*[float(v) for v in lasthit])
        check(lib().p2z_summary_device(self.handle, out.ptr, nevts, nb, width, ntrks, last, C.byref(s)),
              "p2z_summary_device")
        return s
