"""GPU test of the library's threading contract (include/p2z.h): distinct
contexts used from distinct host threads at the same time are independent.
Each thread's passes give the same bits as one context alone, and each thread
reads its own last error (the message is kept per thread).  ctypes releases
the interpreter lock during the library calls, so the threads' calls overlap."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import p2z

pytestmark = pytest.mark.gpu

NTHREADS = 4


def test_contexts_on_threads_are_independent():
    nevts, ntrks, W, nl = 4, 640, 32, 20
    trk, hit, _, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    w = p2z.Workload("ragged", nevents=6, width=W, seed=11)
    lanes = np.ascontiguousarray(w.lane_nlayer)
    lpt = np.ascontiguousarray(w.blocks_lpt)

    def passes(c, reps):
        dt, dh, do = c.upload(trk), c.upload(hit), c.alloc(trk.nbytes)
        rt, rh, rl, rb = c.upload(w.tracks), c.upload(w.hits), c.upload(lanes), c.upload(lpt)
        ro = c.alloc(w.tracks.nbytes)
        for _ in range(reps):
            c.run(dt, dh, do, nevts, nb, nl, W, "fast")
            c.run_ragged(rt, rh, ro, rl, rb, w.nblocks, W, "fp64gain", graph=True)
        c.sync()
        return c.download(do), c.download(ro)

    with p2z.Context(0) as c:
        want_fused, want_ragged = passes(c, 1)

    results, errors, failures = {}, {}, {}
    start = threading.Barrier(NTHREADS)

    def work(k):
        try:
            with p2z.Context(0) as c:
                start.wait(timeout=60)
                results[k] = passes(c, 10)
                # an error of this thread's own, read back in this thread
                bad_math = 10 + k
                try:
                    c.run(c.alloc(trk.nbytes), c.alloc(hit.nbytes), c.alloc(trk.nbytes), nevts, nb, nl, W, bad_math)
                except p2z.P2zError as e:
                    errors[k] = str(e)
        except Exception as e:  # reported below, in the main thread
            failures[k] = repr(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(NTHREADS)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not failures, failures
    for k in range(NTHREADS):
        fused, ragged = results[k]
        assert np.array_equal(fused.view(np.uint32), want_fused.view(np.uint32)), k
        assert np.array_equal(ragged.view(np.uint32), want_ragged.view(np.uint32)), k
        assert f"unknown math mode {10 + k}" in errors[k], errors.get(k)
    w.close()
