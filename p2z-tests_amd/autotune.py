"""Autotuner of the ragged launch (persistent work queue or one-shot).

Sweeps the launch form (persistent work queue or one-shot), the AoSoA width
W in {16, 32, 64, 128}, tracks per workgroup, hit prefetch depth and, for the
persistent form, the grid size (waves per CU) and the items per claim on a
sample of the workload, timing each candidate with hipEvents (sustained time
per launch over back-to-back launches, median of three rounds).  The closest
candidates of each form are then timed again with rounds four times as long
(short rounds rank near-ties by noise), and the fastest of those is
returned.  The sample is generated at each width by the same
generator (the tracks do not depend on W, so every candidate does the same
work).  Results are cached per device and workload kind in
~/.cache/p2z/autotune.json unless cache=False.

    python autotune.py [--kind ragged] [--events 200] [--math fp64gain] [--seed 17]
    python autotune.py --run FORM,W,TPW,PREFETCH[,WAVES_PER_CU,CLAIM] [--reps N]
        (one configuration, N graph-captured launches: for profiling)
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import p2z  # noqa: E402

WIDTHS = (16, 32, 64, 128)
TPWS = (64, 256)
PREFETCH = (1, 2)
WAVES_PER_CU = (8, 12, 16, 20, 32)
CLAIMS = (1, 2, 4, 8)
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "p2z", "autotune.json")


def _settle(ctx, fn, seconds: float = 0.25) -> None:
    """Runs fn() back to back, untimed, for `seconds` of wall time: after the
    GPU idles (the host generating a sample), its engine clock takes tens of
    milliseconds of sustained work to come back up."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(4):
            fn()
        ctx.sync()


def _time(ctx, fn, reps: int, rounds: int = 3) -> float:
    """Sustained time per launch: the median over `rounds` of the hipEvent
    time of `reps` back-to-back launches, divided by reps.  Back-to-back
    launches keep the GPU under load, as in a production pass, and rank the
    candidates as that pass would.  reps untimed launches come first (the
    sweep's candidates follow each other, so the clock stays up; a single
    configuration is settled first, _settle)."""
    for _ in range(reps):
        fn()
    ctx.sync()
    ts = []
    for _ in range(rounds):
        ctx.timer_start()
        for _ in range(reps):
            fn()
        ts.append(ctx.timer_stop() / reps)
    return float(np.median(ts))


def tune(ctx, kind: str = "ragged", nevents: int = 200, math: str = "fp64gain", seed: int = 17,
         reps: int = 5, widths=WIDTHS, tpws=TPWS, prefetch=PREFETCH, waves=WAVES_PER_CU, claims=CLAIMS,
         forms=("persistent", "oneshot"), refine: int = 6, log=lambda s: None) -> dict:
    """Times every candidate launch on the sample and returns the fastest
    overall (best) and the fastest of each launch form (best_<form>).  The
    persistent form sweeps tracks per workgroup x prefetch depth x waves per
    CU x items per claim; the one-shot form (one wave per 64 track slots, dispatched by the
    hardware) sweeps tracks per workgroup x prefetch depth.  The `refine`
    fastest candidates of each form are timed again with 4 * reps launches
    per round, and the best_* entries come from that second timing."""

    def timed(W, cands, r):
        w = p2z.Workload(kind, nevents=nevents, width=W, seed=seed)
        trk, hit = ctx.upload(w.tracks), ctx.upload(w.hits)
        lanes = ctx.upload(np.ascontiguousarray(w.lane_nlayer))
        blk = ctx.upload(np.ascontiguousarray(w.blocks_lpt))
        out = ctx.alloc(w.tracks.nbytes)
        # the GPU idled while the host generated the sample: its clock comes
        # back up under the first candidate before anything is timed
        f0, tpw0, pf0, wpc0, cl0 = cands[0]
        _settle(ctx, lambda: ctx.run_ragged(trk, hit, out, lanes, blk, w.nblocks, W, math,
                                            persistent=f0 == "persistent", tracks_per_wg=tpw0, prefetch=pf0,
                                            waves_per_cu=wpc0, items_per_claim=cl0))
        recs = []
        for form, tpw, pf, wpc, cl in cands:
            pers = form == "persistent"
            ms = _time(ctx, lambda: ctx.run_ragged(trk, hit, out, lanes, blk, w.nblocks, W, math, persistent=pers,
                                                   tracks_per_wg=tpw, prefetch=pf, waves_per_cu=wpc,
                                                   items_per_claim=cl), r)
            rec = dict(form=form, width=W, tracks_per_wg=tpw, prefetch=pf, waves_per_cu=wpc, items_per_claim=cl,
                       ms=ms, tl_per_s=w.track_layers / (ms * 1e-3))
            recs.append(rec)
            log(f"  {form:10s} W={W:3d} tpw={tpw:3d} prefetch={pf} waves/CU={wpc:2d} claim={cl}: {ms:.4f} ms  "
                f"{rec['tl_per_s']:.3e} TL/s" + ("  (refined)" if r != reps else ""))
        for b in (trk, hit, lanes, blk, out):
            b.free()
        w.close()
        return recs

    results = []
    for W in widths:
        cands = []
        if "persistent" in forms:
            cands += [("persistent", tpw, pf, wpc, cl)
                      for tpw, pf, wpc, cl in itertools.product(tpws, prefetch, waves, claims)]
        if "oneshot" in forms:
            cands += [("oneshot", tpw, pf, 0, 0) for tpw, pf in itertools.product(tpws, prefetch)]
        results += timed(W, cands, reps)
    key = lambda r: (r["form"], r["tracks_per_wg"], r["prefetch"], r["waves_per_cu"], r["items_per_claim"])
    refined = []
    if refine > 0:
        top = [r for form in forms
               for r in sorted((r for r in results if r["form"] == form), key=lambda r: r["ms"])[:refine]]
        for W in sorted({r["width"] for r in top}):
            refined += timed(W, [key(r) for r in top if r["width"] == W], 4 * reps)
    ranked = refined or results
    res = dict(best=min(ranked, key=lambda r: r["ms"]), kind=kind, math=math, sample_events=nevents,
               candidates=len(results), refined=len(refined), results=results, refined_results=refined)
    for form in forms:
        rs = [r for r in ranked if r["form"] == form]
        if rs:
            res["best_" + form] = min(rs, key=lambda r: r["ms"])
    return res


def cached_or_tune(ctx, key: str, cache: bool = True, **kw) -> dict:
    db = {}
    if cache and os.path.exists(CACHE):
        with open(CACHE) as f:
            db = json.load(f)
    if key in db:
        return db[key]
    res = tune(ctx, **kw)
    if cache:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        db[key] = res
        with open(CACHE, "w") as f:
            json.dump(db, f, indent=1)
    return res


def run_one(ctx, spec: str, kind: str, nevents: int, math: str, reps: int, seed: int = 2024) -> dict:
    """Times one launch configuration (graph-captured; sustained time per
    launch over reps back-to-back launches, median of three rounds)."""
    f = spec.split(",")
    form, W, tpw, pf = f[0], int(f[1]), int(f[2]), int(f[3])
    wpc = int(f[4]) if len(f) > 4 else 0
    cl = int(f[5]) if len(f) > 5 else 0
    w = p2z.Workload(kind, nevents=nevents, width=W, seed=seed)
    trk, hit = ctx.upload(w.tracks), ctx.upload(w.hits)
    lanes = ctx.upload(np.ascontiguousarray(w.lane_nlayer))
    blk = ctx.upload(np.ascontiguousarray(w.blocks_lpt))
    out = ctx.alloc(w.tracks.nbytes)
    fn = lambda: ctx.run_ragged(trk, hit, out, lanes, blk, w.nblocks, W, math, persistent=form == "persistent",
                                graph=True, tracks_per_wg=tpw, prefetch=pf, waves_per_cu=wpc, items_per_claim=cl)
    _settle(ctx, fn)
    ms = _time(ctx, fn, reps)
    return dict(form=form, width=W, tracks_per_wg=tpw, prefetch=pf, waves_per_cu=wpc, items_per_claim=cl, ms=ms,
                tl_per_s=w.track_layers / (ms * 1e-3), track_layers=w.track_layers, tracks=w.ntracks)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--kind", default="ragged", choices=list(p2z.WORKLOAD))
    ap.add_argument("--events", type=int, default=200)
    ap.add_argument("--math", default="fp64gain", choices=list(p2z.MATH))
    ap.add_argument("--run", default="", help="FORM,W,TPW,PREFETCH[,WAVES_PER_CU,CLAIM]: time one configuration")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=17, help="workload seed of the sweep")
    a = ap.parse_args()
    with p2z.Context(0) as ctx:
        if a.run:
            print(json.dumps(run_one(ctx, a.run, a.kind, a.events, a.math, a.reps)))
            return
        res = tune(ctx, a.kind, a.events, a.math, seed=a.seed, log=print)
    print(json.dumps(res["best"]))


if __name__ == "__main__":
    main()
