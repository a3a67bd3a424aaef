"""Throughput of the p2z propagate-to-z + Kalman-update path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--math fast|ref|fp64gain]
                    [--sections main,ref,config0,ragged,serving,ingest,precision]

Headline workload (value): the reference's full configuration, 100 events x
9600 tracks x 20 layers, AoSoA bsize 32, FP32, inputs from the reference
generator (preset 'ref': src/input_track.h, srand(1), smear 1e-7).  One step
is one pass of the reference driver loop (src/propagate-toz-test_OMP.cpp:
733-747) over all events: a single fused launch with tracks, hits and output
resident in HBM.

One JSON line: value = track-layer updates/s, a roofline object for the fused
kernel (algorithmic HBM bytes per launch / hipEvent launch time; its `valu`
member is the VALU issue fraction from the lane-instructions per track-layer
that rocprofv3 counted on this workload), the CPU
baseline (the reference program compiled from its own source, oracle/_ref,
OpenMP over events, on a bounded sample of the same workload), and extra
sections:
  ref_mode  : the parity-checked arithmetic (P2Z_MATH_REF, bit for bit the
              reference build) on the headline workload
  config0   : 10 events x 9600 tracks x 20 layers of the nominal inputs of
              propagate-toz-pseudocode.txt, in each arithmetic mode, beside the
              reference program at 10 events on the CPU
  ragged    : 2000 uneven events (log-normal track counts, 8-26 layers per
              track), FP32 with the FP64 gain steps and FP32 alone; the launch
              is autotuned on a held-out event set, then the persistent and
              one-shot launches, each replayed from a hipGraph, are timed
              against one plain launch per event on the evaluation set
  serving   : the same events arriving as a Poisson stream at 1k..3M events/s
              through the continuous-batching server (HBM budget 64 GB), in
              the device-resident loop and in batches of launches: p50 / p99 /
              p999 event latency, sustained throughput, and the first rate
              the server no longer keeps up with
  ingest    : the headline workload from host buffers (p2z_run_host: chunked
              copies in and out overlapping the compute on several streams),
              pinned and pageable, beside the parts alone
  precision : 100 events x 9600 tracks x 20 layers with wide angle / momentum /
              hit-error ranges: FP32 path against the FP64-gain path
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "p2z-tests_amd"))

import numpy as np  # noqa: E402

import p2z  # noqa: E402

HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (non-packed lane-instructions/s)
VALU_PEAK = 256 * 4 * 16 * 2.4e9
PROFILE_TRAFFIC = os.path.join(HERE, "profiles", "traffic_fused_w32_100ev.txt")
PROFILE_VALU = os.path.join(HERE, "profiles", "valu_counters_fused_w32_100ev.md")
PROFILE_STREAM = os.path.join(HERE, "profiles", "hbm_copy_bandwidth.txt")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def bytes_per_track_layer(nlayer: int) -> float:
    """Algorithmic HBM bytes per track-layer: the 6 hit words the path reads
    (x, y, z, cov xx, xy, yy) plus the 28-word track read once and written
    once, spread over the layers."""
    return 24.0 + 224.0 / nlayer


def _profile_value(path: str, key: str, math: str):
    """The figure `key[math] = value` of a committed profile summary, if any."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            m = re.match(rf"{key}\[{math}\]\s*=\s*([0-9.eE+]+)", ln.strip())
            if m:
                return float(m.group(1))
    return None


def read_traffic(math: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if the committed
    profile summary for this workload and math mode exists."""
    return _profile_value(PROFILE_TRAFFIC, "traffic_bytes_per_launch", math)


def achievable_hbm(achieved_gbs: float, which: str = "main"):
    """The bandwidth the best plain streaming kernel reaches on this GPU with
    the main workload's read:write mix (5:1) and size (committed
    calibration): on one buffer set ("main", as the timed passes run) or
    cycling through three ("cold", no reuse from the last-level cache); and
    the fraction of it the pass reaches.  The practical HBM ceiling beside the
    datasheet peak."""
    gbs = _profile_value(PROFILE_STREAM, "achievable_gbs", which)
    if gbs is None:
        return None
    sets = "one buffer set" if which == "main" else "three buffer sets in turn"
    return {"value": gbs, "unit": "GB/s", "frac": round(achieved_gbs / gbs, 4),
            "kernel": f"plain stream, 540 MiB read + 108 MiB written per launch, {sets}"}


def valu_roofline(math: str, tl_per_s: float):
    """VALU issue fraction: lane-instructions per track-layer counted by
    rocprofv3 --pmc SQ_INSTS_VALU on this workload (committed profile), times
    the measured rate, against the VALU issue peak."""
    per_tl = _profile_value(PROFILE_VALU, "valu_lane_instr_per_tl", math)
    if per_tl is None:
        return None
    rate = per_tl * tl_per_s
    return {"lane_instr_per_tl": per_tl, "achieved": rate, "peak": VALU_PEAK, "unit": "lane-instr/s",
            "frac": round(rate / VALU_PEAK, 4)}


def host_cpu() -> dict:
    """The host the CPU baseline runs on: CPU model, sockets, cores and
    hardware threads (from /proc/cpuinfo), and the CPUs this process may use."""
    info = {"model": None, "sockets": None, "cores": None, "threads": None,
            "usable_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            text = f.read()
    except OSError:
        return info
    procs = [p for p in text.split("\n\n") if p.strip()]
    field = lambda p, k: next((ln.split(":", 1)[1].strip() for ln in p.splitlines() if ln.startswith(k)), None)
    info["model"] = field(procs[0], "model name") if procs else None
    phys = {(field(p, "physical id"), field(p, "core id")) for p in procs}
    info["sockets"] = len({s for s, _ in phys})
    info["cores"] = len(phys)
    info["threads"] = len(procs)
    info["smt"] = info["threads"] > info["cores"]
    return info


def cpu_baseline(nevts: int, ntrks: int, nlayer: int, name: str = "ref_bench") -> dict | None:
    """The reference program (src/propagate-toz-test_OMP.cpp, built by
    oracle/Makefile from the reference's own source with its flags) on the
    same workload with NITER=10 (the reference Makefile's value) and 16
    OpenMP threads, of which its loop over events (OMP:733) keeps at most
    nevts busy; its own timer brackets the NITER region only (OMP:729-753).
    name: ref_bench (100 events) or ref_bench_cfg0 (10 events)."""
    cpuflags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    has512 = all(f in cpuflags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))
    exe = os.path.join(HERE, "oracle", "_ref", name + ("_v4" if has512 else "_v3"))
    if not os.path.exists(exe):
        return cpu_baseline_port(nevts, ntrks, nlayer)
    log(f"cpu baseline: {os.path.basename(exe)}")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        log(res.stderr)
        return None
    m = re.search(r"done ntracks=(\d+) tot time=([0-9.]+) \(s\)", res.stdout)
    f = re.search(r"formatted (\d+) (\d+) (\d+) (\d+) (\d+) [0-9.]+ 0 [0-9.]+ (\d+)", res.stdout)
    ntr, secs = int(m.group(1)), float(m.group(2))
    niter, ne, nt, threads = int(f.group(1)), int(f.group(2)), int(f.group(3)), int(f.group(6))
    assert (ne, nt) == (nevts, ntrks), (ne, nt)
    return {"value": ntr * nlayer / secs, "unit": "track-layer updates/s",
            "cores": min(threads, ne), "threads_used": min(threads, ne), "kind": "reference",
            "sample": f"reference OpenMP program built from its source ({os.path.basename(exe)}), "
                      f"{ne} events x {nt} tracks x {nlayer} layers, NITER={niter}, "
                      f"{threads} OpenMP threads over events, {secs:.3f} s timed",
            "host": host_cpu()}


def cpu_baseline_port(nevts: int, ntrks: int, nlayer: int) -> dict | None:
    """Fallback when the reference build is absent: the oracle restatement
    with the reference's threading (parallel over events), 2 passes."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    try:
        import pyoracle
        pyoracle.lib()
    except (ImportError, FileNotFoundError):
        return None
    W, threads = 32, 16
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nlayer)
    t0 = time.perf_counter()
    for _ in range(2):
        pyoracle.run_events_only(trk, hit, nevts, nb, W, nlayer, threads)
    secs = time.perf_counter() - t0
    return {"value": 2 * nevts * nb * W * nlayer / secs, "unit": "track-layer updates/s",
            "cores": min(threads, nevts), "kind": "port",
            "sample": f"oracle restatement, {nevts} events x {ntrks} tracks x {nlayer} layers, 2 passes, "
                      f"{threads} threads over events, {secs:.3f} s",
            "host": host_cpu()}


# The GPU's clocks rise over the first tens of milliseconds of sustained work
# after an idle spell (measured on the headline pass: 0.33 ms for the first
# launches, 0.12 ms after 30 ms, a steady 0.116 ms from about 40 ms on). Every
# timed figure follows at least this much untimed work of the same kind, so
# it measures the sustained rate, not the clock ramp.
SETTLE_S = 0.25


def settle(ctx, fn, seconds: float = SETTLE_S) -> int:
    """Runs fn() back to back, untimed, for `seconds` of wall time; returns
    the number of calls."""
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            fn()
        n += 10
        ctx.sync()
    return n


def time_ms(ctx, fn, reps: int) -> float:
    """Average hipEvent time of fn() over reps back-to-back calls, after the
    settling run and reps untimed calls: the GPU idles while the host
    generates a workload, and its clocks take tens of milliseconds of work to
    come back up."""
    settle(ctx, fn)
    for _ in range(reps):
        fn()
    ctx.sync()
    ctx.timer_start()
    for _ in range(reps):
        fn()
    return ctx.timer_stop() / reps


def ragged_bytes(w) -> float:
    """Algorithmic HBM bytes of one ragged pass: per track 24 B of hit per
    layer plus its 112-byte state read and written once."""
    return 24.0 * w.track_layers + 224.0 * w.ntracks


# Seeds of the ragged workload: the autotuner sweeps its candidates on a
# held-out event set of the same distribution and size (a persistent grid
# needs several work chunks per wave before its timing carries over, which a
# small sample does not give); the chosen launch is then timed on the
# evaluation set, which the tuner never saw.
TUNE_SEED, EVAL_SEED = 7, 2024


def cold_cache_pass(ctx, trk, hit, first, run, tl_per_step: int, b_tl: float) -> dict:
    """The same pass cycling through three copies of its buffers: together
    they exceed the GPU's 256 MB last-level cache, so no pass reads what the
    previous one left there (the timed passes on one buffer set can take part
    of the 107.5 MB track input from it).  A side figure beside `value`."""
    sets = [first] + [(ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)) for _ in range(2)]
    it = [0]

    def step():
        run(sets[it[0] % 3])
        it[0] += 1

    ms = time_ms(ctx, step, 60)
    for bufs in sets[1:]:
        for b in bufs:
            b.free()
    tlps = tl_per_step / (ms * 1e-3)
    gbs = tlps * b_tl / 1e9
    return {"buffer_sets": 3, "ms_per_step": round(ms, 5), "value": tlps,
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "achievable": achievable_hbm(gbs, "cold")}


def section_ragged(ctx, nevents: int, sample_events: int) -> tuple[dict, dict]:
    """Ragged events (2000 events, log-normal track counts, 8-26 layers):
    autotune per arithmetic mode on a held-out event set (TUNE_SEED), then
    the evaluation set (EVAL_SEED) through the persistent and the one-shot
    launch, each replayed from a captured hipGraph, and through one plain
    launch per event."""
    sys.path.insert(0, os.path.join(HERE, "p2z-tests_amd"))
    import autotune

    res = {"math_modes": {}}
    pools = {}

    def pool(W):
        if W not in pools:
            log(f"ragged: generating {nevents} events at width {W}")
            w = p2z.Workload("ragged", nevents=nevents, width=W, seed=EVAL_SEED)
            pools[W] = dict(w=w, d=dict(trk=ctx.upload(w.tracks), hit=ctx.upload(w.hits),
                                        lanes=ctx.upload(np.ascontiguousarray(w.lane_nlayer)),
                                        lpt=ctx.upload(np.ascontiguousarray(w.blocks_lpt)),
                                        evt=ctx.upload(np.ascontiguousarray(w.blocks))))
        return pools[W]["w"], pools[W]["d"]

    for math in ("fp64gain", "fast"):
        log(f"ragged: autotuning {math} on {sample_events} held-out events (seed {TUNE_SEED})")
        t_tune = time.perf_counter()
        tuned = autotune.tune(ctx, "ragged", nevents=sample_events, math=math, reps=3, seed=TUNE_SEED, log=log)
        t_tune = time.perf_counter() - t_tune
        best = tuned["best"]
        # each launch form at its own best width
        bp, bo = tuned["best_persistent"], tuned["best_oneshot"]
        opts = dict(tracks_per_wg=bp["tracks_per_wg"], prefetch=bp["prefetch"], waves_per_cu=bp["waves_per_cu"],
                    items_per_claim=bp["items_per_claim"])
        oopts = dict(tracks_per_wg=bo["tracks_per_wg"], prefetch=bo["prefetch"])
        log(f"ragged: {math} best {best}")
        row = {"tuned": {"form": best["form"], "persistent": dict(width=bp["width"], **opts),
                         "oneshot": dict(width=bo["width"], **oopts),
                         "candidates": tuned["candidates"], "refined": tuned["refined"],
                         "sample_events": sample_events,
                         "tuning_set": f"seed {TUNE_SEED} (held out; evaluation set seed {EVAL_SEED})",
                         "tuning_s": round(t_tune, 2)}}

        def rate(w, ms):
            return {"ms": round(ms, 4), "tl_per_s": w.track_layers / (ms * 1e-3),
                    "hbm_frac": round(ragged_bytes(w) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}

        # the two forms timed in alternation, three rounds each (median):
        # the clock history of a long run shifts a single round by a few %
        Wp, Wo = bp["width"], bo["width"]
        wp, dp = pool(Wp)
        wo, do = pool(Wo)
        res["workload"] = (f"{nevents} events, {wp.ntracks} tracks, {wp.track_layers} track-layers "
                           f"(tracks/event log-normal median 2000 sigma 1 in [50,20000], layers U{{8..26}})")
        outp, outo = ctx.alloc(wp.tracks.nbytes), ctx.alloc(wo.tracks.nbytes)
        tp, to = [], []
        for _ in range(3):
            tp.append(time_ms(ctx, lambda: ctx.run_ragged(dp["trk"], dp["hit"], outp, dp["lanes"], dp["lpt"],
                                                          wp.nblocks, Wp, math, graph=True, **opts), 20))
            to.append(time_ms(ctx, lambda: ctx.run_ragged(do["trk"], do["hit"], outo, do["lanes"], do["lpt"],
                                                          wo.nblocks, Wo, math, persistent=False, graph=True,
                                                          **oopts), 20))
        row["persistent_graph"] = dict(width=Wp, **rate(wp, float(np.median(tp))),
                                       rounds_ms=[round(t, 4) for t in tp])
        row["oneshot_graph"] = dict(width=Wo, **rate(wo, float(np.median(to))), rounds_ms=[round(t, 4) for t in to])
        outp.free()
        W, w, d, out = Wo, wo, do, outo
        eb = w.event_blocks

        def per_event():
            for e in range(nevents):
                b0, b1 = int(eb[e]), int(eb[e + 1])
                ctx.run_ragged(d["trk"], d["hit"], out, d["lanes"], d["evt"].ptr + b0 * 32, b1 - b0, W, math,
                               persistent=False, **oopts)

        row["one_launch_per_event"] = {**rate(w, time_ms(ctx, per_event, 1)), "width": W, "launches": nevents}
        best_row = max((row["persistent_graph"], row["oneshot_graph"]), key=lambda r: r["tl_per_s"])
        row["speedup_vs_per_event"] = round(best_row["tl_per_s"] / row["one_launch_per_event"]["tl_per_s"], 2)
        out.free()
        res["math_modes"][math] = row
    res["bytes_per_pass"] = "24 B x track-layers + 224 B x tracks (algorithmic)"
    opts0 = dict(res["math_modes"]["fp64gain"]["tuned"]["persistent"])
    W0 = opts0.pop("width")
    for W, pl in pools.items():
        if W != W0:
            for b in pl["d"].values():
                b.free()
            pl["w"].close()
    return res, dict(w=pools[W0]["w"], d=pools[W0]["d"], W=W0, opts=opts0)


def section_serving(ctx, pool: dict, rates, budget_gb: float, resident: bool) -> list:
    """Serving: Poisson arrivals of the resident ragged events through the
    continuous-batching server, in the device-resident loop (one persistent
    kernel fed through a ring in host memory) or in batches of launches."""
    w, d, W, opts = pool["w"], pool["d"], pool["W"], dict(pool["opts"])
    # latency first: one work item per wave and two layers of hits in flight
    opts.update(items_per_claim=1, prefetch=2)
    rows = []
    with p2z.Server(ctx, d["trk"], d["hit"], d["lanes"], int(budget_gb * 2**30), W, "fp64gain", nstreams=4,
                    resident=resident, **opts) as sv:
        for rate in rates:
            log(f"serving ({'resident loop' if resident else 'batches'}): {rate:.0f} events/s")
            rep, lat, ob = sv.poisson(w.blocks, w.event_blocks, rate=rate, seed=int(rate), release=True)
            served = rep.nevents / rep.t_span
            rows.append({"rate_offered": rate, "events": int(rep.nevents),
                         "served_events_per_s": round(served, 1), "served_over_offered": round(served / rate, 3),
                         "tl_per_s": w.track_layers / rep.t_span,
                         "p50_ms": rep.lat_p50 * 1e3, "p99_ms": rep.lat_p99 * 1e3, "p999_ms": rep.lat_p999 * 1e3,
                         "mean_ms": rep.lat_mean * 1e3, "launches" if resident else "batches": int(rep.batches),
                         "busy_rejections": int(rep.busy_rejections),
                         "max_inflight_events": int(rep.max_inflight_events),
                         "span_ms": rep.t_span * 1e3,
                         "host_submit_us_per_event": rep.host_submit_s / rep.nevents * 1e6,
                         "host_pump_us_per_event": rep.host_pump_s / rep.nevents * 1e6})
    return rows


def serving_knee(rows: list) -> dict | None:
    """The first offered rate the server no longer keeps up with (served
    below 90% of offered), with its sustained throughput and latencies."""
    for r in rows:
        if r["served_over_offered"] < 0.9:
            return {k: r[k] for k in ("rate_offered", "served_events_per_s", "tl_per_s", "p50_ms", "p99_ms")}
    return None


def section_ingest(ctx, trk, hit, nevts: int, nb: int, nl: int, W: int, math: str) -> dict:
    """The include-data mode (CU4:830-875): host buffers in, host buffer out,
    through p2z_run_host, which copies chunks of events in, computes and copies
    out on separate streams so that the transfers overlap each other and the
    compute.  Timed with pinned and with pageable host memory, beside the
    parts alone (one H2D of the whole inputs, one D2H of the output, the
    compute).  From pinned memory the pipeline copies only the 6 of 9 words
    of each hit record that the path reads."""
    tl = nevts * nb * W * nl
    out = np.empty_like(trk)
    hit_used = hit.nbytes * 6 // 9
    res = {"workload": f"{nevts} events x {nb * W} tracks x {nl} layers per pass: tracks in {trk.nbytes / 1e6:.1f} MB, "
                       f"hits {hit.nbytes / 1e6:.1f} MB ({hit_used / 1e6:.1f} MB of them read by the path), "
                       f"tracks out {trk.nbytes / 1e6:.1f} MB"}

    def wall(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    ms = wall(lambda: ctx.run_host(trk, hit, out, nevts, nb, nl, W, math))
    res["pageable"] = {"ms": round(ms, 3), "tl_per_s": tl / (ms * 1e-3), "h2d_bytes": trk.nbytes + hit.nbytes}
    pt, ph, po = ctx.pinned_copy(trk), ctx.pinned_copy(hit), ctx.pinned_empty(trk.size)
    ms = wall(lambda: ctx.run_host(pt, ph, po, nevts, nb, nl, W, math))
    res["pinned"] = {"ms": round(ms, 3), "tl_per_s": tl / (ms * 1e-3), "h2d_bytes": trk.nbytes + hit_used,
                     "h2d_gb_per_s": round((trk.nbytes + hit_used) / (ms * 1e-3) / 1e9, 1)}
    dt, dh, do = ctx.alloc(trk.nbytes), ctx.alloc(hit.nbytes), ctx.alloc(trk.nbytes)
    h2d = wall(lambda: (ctx.h2d(dt, pt), ctx.h2d(dh, ph)))
    d2h = wall(lambda: ctx.d2h(po, do))
    comp = time_ms(ctx, lambda: ctx.run(dt, dh, do, nevts, nb, nl, W, math), 10)
    for b in (dt, dh, do):
        b.free()
    res["parts_ms"] = {"h2d_pinned_whole_inputs": round(h2d, 3), "d2h_pinned": round(d2h, 3), "compute": round(comp, 4)}
    res["h2d_gb_per_s"] = round((trk.nbytes + hit.nbytes) / (h2d * 1e-3) / 1e9, 1)
    return res


def section_ref_mode(ctx, trk, hit, nevts: int, nb: int, nl: int, W: int) -> dict:
    """The parity-checked arithmetic (P2Z_MATH_REF: bit for bit the reference
    build) on the headline workload: same timing as the headline value."""
    dt, dh, do = ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)
    ms = time_ms(ctx, lambda: ctx.run(dt, dh, do, nevts, nb, nl, W, "ref"), 50)
    for b in (dt, dh, do):
        b.free()
    tl = nevts * nb * W * nl
    tlps = tl / (ms * 1e-3)
    gbs = tlps * bytes_per_track_layer(nl) / 1e9
    return {"kernel": f"k_fused<{W},REF,1>", "ms": round(ms, 5), "tl_per_s": tlps,
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "achieved_gb_s": round(gbs, 1),
            "traffic": read_traffic("ref"), "valu": valu_roofline("ref", tlps)}


def section_config0(ctx, no_cpu: bool) -> dict:
    """Config 0: the nominal track and hit of propagate-toz-pseudocode.txt
    (smear 1e-5), 10 events x 9600 tracks x 20 layers, on the GPU in each
    arithmetic mode, beside the reference program at 10 events on the CPU."""
    nevts, ntrks, nl, W = 10, 9600, 20, 32
    trk, hit, last, nb = p2z.generate("config0", nevts, ntrks, W, nl)
    dt, dh, do = ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)
    tl = nevts * nb * W * nl
    res = {"workload": f"{nevts} events x {ntrks} tracks x {nl} layers, config0 preset, W={W}"}
    for math in ("fast", "ref", "fp64gain"):
        ms = time_ms(ctx, lambda: ctx.run(dt, dh, do, nevts, nb, nl, W, math), 50)
        res[math] = {"ms": round(ms, 5), "tl_per_s": tl / (ms * 1e-3)}
    for b in (dt, dh, do):
        b.free()
    if not no_cpu:
        log("cpu baseline at config 0 (10 events)")
        res["cpu_baseline"] = cpu_baseline(nevts, ntrks, nl, "ref_bench_cfg0")
        if res["cpu_baseline"]:
            res["gpu_fast_over_cpu"] = round(res["fast"]["tl_per_s"] / res["cpu_baseline"]["value"], 1)
    return res


def section_precision(ctx, nevents: int) -> dict:
    """Precision workload throughput: the FP32 path and the FP64-gain path on the wide
    (ill-conditioned) workload; the error report lives in the tests."""
    w = p2z.Workload("precision", nevents=nevents, width=32, seed=42, ntracks=9600, nlayer=20)
    trk, hit, out = ctx.upload(w.tracks), ctx.upload(w.hits), ctx.alloc(w.tracks.nbytes)
    res = {"workload": f"{nevents} events x 9600 tracks x 20 layers, 1/pT log-U[0.01,2], theta U[0.1,3.04], "
                       f"hit variances log-U[1e-8,1e-3]"}
    for math in ("fast", "fp64gain"):
        ms = time_ms(ctx, lambda: ctx.run(trk, hit, out, nevents, 9600 // 32, 20, 32, math), 20)
        res[math] = {"ms": round(ms, 4), "tl_per_s": w.track_layers / (ms * 1e-3)}
    for b in (trk, hit, out):
        b.free()
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--math", default="fast", choices=list(p2z.MATH))
    ap.add_argument("--nevts", type=int, default=100)
    ap.add_argument("--ntrks", type=int, default=9600)
    ap.add_argument("--nlayer", type=int, default=20)
    ap.add_argument("--bsize", type=int, default=32)
    ap.add_argument("--tracks-per-wg", type=int, default=0)
    ap.add_argument("--hit-stage", type=int, default=0, choices=(0, 1, 2),
                    help="layers of hits in flight ahead of the one computed (0: the library's default per mode)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sections", default="main,ref,config0,ragged,serving,ingest,precision",
                    help="comma list of main, ref, config0, ragged, serving, ingest, precision")
    ap.add_argument("--ragged-events", type=int, default=2000)
    ap.add_argument("--serving-rates", default="1000,2000,5000,10000,20000,50000,200000,1000000,3000000")
    ap.add_argument("--budget-gb", type=float, default=64.0)
    a = ap.parse_args()
    sections = set(a.sections.split(","))
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this path shards by event with no exchange step, and "
              "this build runs on one GPU only (--gpus 1).", file=sys.stderr)
        return 2

    nevts, ntrks, nl, W = a.nevts, a.ntrks, a.nlayer, a.bsize
    log(f"generating {nevts} events x {ntrks} tracks x {nl} layers (reference generator)")
    trk, hit, last, nb = p2z.generate("ref", nevts, ntrks, W, nl)
    ctx = p2z.Context(0, tracks_per_wg=a.tracks_per_wg, hit_stage=a.hit_stage)
    dt, dh, do = ctx.upload(trk), ctx.upload(hit), ctx.alloc(trk.nbytes)
    tl_per_step = nevts * nb * W * nl

    log(f"settling the device ({SETTLE_S} s of untimed passes), then warmup {a.warmup} steps")
    settle_passes = settle(ctx, lambda: ctx.run(dt, dh, do, nevts, nb, nl, W, a.math))
    for _ in range(a.warmup):
        ctx.run(dt, dh, do, nevts, nb, nl, W, a.math)
    ctx.sync()
    log(f"timing {a.steps} steps")
    wall0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(a.steps):
        ctx.run(dt, dh, do, nevts, nb, nl, W, a.math)
    ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - wall0
    ms_per_step = ms / a.steps
    value = tl_per_step / (ms_per_step * 1e-3)

    cold = cold_cache_pass(ctx, trk, hit, (dt, dh, do), lambda b: ctx.run(*b, nevts, nb, nl, W, a.math),
                           tl_per_step, bytes_per_track_layer(nl))

    # sanity on the result (product only: no oracle here): finite, z on the last hit plane
    out = ctx.download(do)
    z_ok = bool(np.all(out.reshape(nevts * nb, 28, W)[:, 2, :] == hit.reshape(nevts * nb, nl, 9, W)[:, nl - 1, 2, :]))
    ctx.close()

    b_tl = bytes_per_track_layer(nl)
    achieved = tl_per_step * b_tl / (ms_per_step * 1e-3) / 1e9
    default_workload = (nevts, ntrks, nl, W) == (100, 9600, 20, 32)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(a.math) if default_workload else None,
                "bytes_per_track_layer": b_tl, "launch_ms": round(ms_per_step, 5),
                "valu": valu_roofline(a.math, value) if default_workload else None,
                "achievable": achievable_hbm(achieved) if default_workload else None}

    extra = {}
    if "ref" in sections:
        ctx = p2z.Context(0)
        log("reference arithmetic (P2Z_MATH_REF) on the headline workload")
        extra["ref_mode"] = section_ref_mode(ctx, trk, hit, nevts, nb, nl, W)
        ctx.close()
    if "config0" in sections:
        ctx = p2z.Context(0)
        log("config 0")
        extra["config0"] = section_config0(ctx, a.no_cpu_baseline)
        ctx.close()
    if "ragged" in sections or "serving" in sections:
        ctx = p2z.Context(0)
        rag, pool = section_ragged(ctx, a.ragged_events, a.ragged_events)
        if "ragged" in sections:
            extra["ragged"] = rag
        if "serving" in sections:
            rates = [float(r) for r in a.serving_rates.split(",") if r]
            res_rows = section_serving(ctx, pool, rates, a.budget_gb, resident=True)
            batch_rates = [r for r in rates if r >= 20000]
            bat_rows = section_serving(ctx, pool, batch_rates, a.budget_gb, resident=False)
            best = max(rag["math_modes"]["fp64gain"]["persistent_graph"]["tl_per_s"],
                       rag["math_modes"]["fp64gain"]["oneshot_graph"]["tl_per_s"]) if "math_modes" in rag else None
            extra["serving"] = {"budget_gb": a.budget_gb, "events_per_rate": a.ragged_events, "math": "fp64gain",
                                "resident_loop": {"rates": res_rows, "knee": serving_knee(res_rows)},
                                "batches_of_launches": {"rates": bat_rows, "knee": serving_knee(bat_rows)},
                                "batch_path_tl_per_s": best}
        for b in pool["d"].values():
            b.free()
        pool["w"].close()
        ctx.close()
    if "ingest" in sections:
        ctx = p2z.Context(0)
        log("ingest (host buffers through p2z_run_host)")
        extra["ingest"] = section_ingest(ctx, trk, hit, nevts, nb, nl, W, a.math)
        ctx.close()
    if "precision" in sections:
        ctx = p2z.Context(0)
        log("precision workload")
        extra["precision"] = section_precision(ctx, 100)
        ctx.close()

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(nevts, ntrks, nl)

    rec = {
        "metric": "track-layer updates/s on 1 MI355X",
        "value": value,
        "unit": "track-layer updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.math != "fp64gain" else "f32+f64gain",
        "data": "synthetic: reference generator (src/input_track.h, srand(1), smear 1e-7)",
        "config": {"workload": f"propagateToZ+KalmanUpdate_v2 driver pass, {nevts} events x {ntrks} tracks x "
                               f"{nl} layers, AoSoA bsize {W}, fused single launch, inputs resident in HBM",
                   "nevts": nevts, "ntrks": ntrks, "nlayer": nl, "bsize": W, "math": a.math,
                   "hit_stage": a.hit_stage or "default", "parallelism": "one GPU, one track per lane"},
        "roofline": roofline,
        "cold_cache": cold,
        "cpu_baseline": cpu,
        "wall_s_timed": round(wall, 4),
        "device_settle": {"seconds": SETTLE_S, "passes": settle_passes,
                          "why": "untimed passes before the warmup steps: the GPU's clocks take tens of "
                                 "milliseconds of sustained work to rise after the idle spell of the "
                                 "host-side input generation"},
        "z_on_last_hit_plane": z_ok,
    }
    if cpu:
        rec["gpu_over_cpu"] = round(value / cpu["value"], 1)
    rec.update(extra)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
