"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference (src/propagate-toz-test_OMP.cpp of cerati/p2z-tests) is compiled
from its own source by `make -C oracle ref` into oracle/_ref/, once per size
configuration, with -ffp-contract=off.  This script runs those builds and
stores their inputs and outputs as small data files, so the parity tests can
check the oracle and the GPU path where the reference tree is absent.

    python tests/golden/make_golden.py          (needs oracle/_ref built)

Files:
  ref_input.json        the reference's input track and 26 hits in main's
                        layer order (src/input_track.h, OMP:693-695)
  known_answer.txt      the 11 summary lines of logs/propagate_gcc_omp.log:30-40
  ref_small.npz         nevts=3 ntrks=96 bsize=32 nlayer=20: inputs and one
                        driver pass of outputs (every track)
  ref_w16.npz           nevts=2 ntrks=64 bsize=16 nlayer=20: same
  ref_track_layers.npz  nevts=1 ntrks=1 bsize=1: the state after each of the
                        20 layers, after propagateToZ and after KalmanUpdate_v2
  ref_summary_full.txt  the reference's own report at the full configuration
                        (nevts=100 ntrks=9600 nlayer=20), contraction off
  fp64_summary_full.txt the same report from the FP64 oracle (every
                        operation of the algorithm in double) on the same
                        inputs: the answer without FP32 rounding
  ref_steps.npz         8 random blocks (bsize 32, seeded): one propagateToZ and
                        one KalmanUpdate_v2 each, from both reference builds
  ref_update_v1.npz     8 random blocks (bsize 32, seeded, all 6 hit covariance
                        words set): one KalmanUpdate (v1, OMP:375-501) each
  ref_spread.json       the reference's own per-track spread on the driver
                        outputs (small, w16 and mid = 10 events x 960 tracks)
                        and on the single steps, under
                          <name>      : its Makefile's FP model (FMA build)
                                        against the contraction-off build
                          <name>_libm : the same algorithm with correctly
                                        rounded sin/cos (the oracle's 'cr'
                                        variant) against glibc's sinf/cosf

Every *.npz holding driver outputs has `out` (contraction off) and `out_fma`
(the reference's default FP model, FMA contraction).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests"))
import pyoracle  # noqa: E402
from parity import compare, random_blocks, track_major  # noqa: E402

REF_LOG = "/root/reference/logs/propagate_gcc_omp.log"


def need(name):
    L = pyoracle.ref_lib(name)
    if L is None:
        sys.exit(f"oracle/_ref/libref_{name}.so missing: run `make -C oracle ref` first")
    return L


def main() -> None:
    L = need("small")
    t, h = pyoracle.ref_inputs(L)
    with open(os.path.join(HERE, "ref_input.json"), "w") as f:
        json.dump({"source": "src/input_track.h:1-113 in main's layer order (OMP:693-695)",
                   "track": [float(v) for v in t], "hits": [[float(v) for v in row] for row in h],
                   "smear": 1e-7, "seed": 1}, f, indent=1)

    with open(REF_LOG) as f:
        lines = f.read().splitlines()
    ka = [ln for ln in lines if (ln.startswith("track ") and not ln.startswith("track in ")) or ln.startswith("number of tracks")]
    assert len(ka) == 11, ka
    with open(os.path.join(HERE, "known_answer.txt"), "w") as f:
        f.write("\n".join(ka) + "\n")

    spread = {}
    for name in ("small", "w16", "mid"):
        L, Lf = need(name), need(name + "_fma")
        c = pyoracle.ref_config(L)
        assert c == pyoracle.ref_config(Lf)
        trk, hit = pyoracle.ref_generate(L)
        trk_f, hit_f = pyoracle.ref_generate(Lf)
        out = pyoracle.ref_run(L, trk, hit)
        # the FMA build driven on the contraction-off build's inputs
        out_fma = pyoracle.ref_run(Lf, trk, hit)
        W = c["width"]
        spread[name] = compare(track_major(out_fma, W), track_major(out, W))
        with pyoracle.use_variant("cr"):
            out_cr = pyoracle.run(trk, hit, c["nevts"], c["nb"], W, c["nlayer"])
        spread[name + "_libm"] = compare(track_major(out_cr, W), track_major(out, W))
        spread[name]["inputs_bitwise_equal_across_builds"] = bool(
            np.array_equal(trk.view(np.uint32), trk_f.view(np.uint32))
            and np.array_equal(hit.view(np.uint32), hit_f.view(np.uint32)))
        if name != "mid":
            np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), trk=trk, hit=hit, out=out, out_fma=out_fma,
                                **{k: np.int64(v) for k, v in c.items()})

    # single steps on random blocks through both builds
    L, Lf = need("small"), need("small_fma")
    W = L.ref_bsize()
    rng = np.random.default_rng(2024)
    blocks = [random_blocks(rng, W, zrow=bool(i % 2)) for i in range(8)]
    trk = np.concatenate([b[0] for b in blocks])
    hit = np.concatenate([b[1] for b in blocks])
    res = {}
    for tag, lib in (("", L), ("_fma", Lf)):
        prop, upd = np.empty_like(trk), trk.copy()
        for b in range(len(blocks)):
            t = slice(b * 28 * W, (b + 1) * 28 * W)
            h = slice(b * 9 * W, (b + 1) * 9 * W)
            tb, hb = np.ascontiguousarray(trk[t]), np.ascontiguousarray(hit[h])
            pb = np.empty_like(tb)
            lib.ref_propagate_block(tb.ctypes.data, hb.ctypes.data, pb.ctypes.data)
            prop[t] = pb
            ub = tb.copy()
            lib.ref_update_block(ub.ctypes.data, hb.ctypes.data)
            upd[t] = ub
        res["prop" + tag], res["upd" + tag] = prop, upd
    np.savez_compressed(os.path.join(HERE, "ref_steps.npz"), trk=trk, hit=hit, width=np.int64(W), **res)
    spread["step_propagate"] = compare(track_major(res["prop_fma"], W), track_major(res["prop"], W))
    spread["step_update"] = compare(track_major(res["upd_fma"], W), track_major(res["upd"], W))
    with pyoracle.use_variant("cr"):
        prop_cr = np.concatenate([pyoracle.propagate_block(np.ascontiguousarray(trk[b * 28 * W:(b + 1) * 28 * W]),
                                                           np.ascontiguousarray(hit[b * 9 * W:(b + 1) * 9 * W]), W)
                                  for b in range(len(blocks))])
    spread["step_propagate_libm"] = compare(track_major(prop_cr, W), track_major(res["prop"], W))
    spread["step_update_libm"] = compare(track_major(res["upd"], W), track_major(res["upd"], W))
    with open(os.path.join(HERE, "ref_spread.json"), "w") as f:
        json.dump(spread, f, indent=1)

    # one track, layer by layer, through the reference's own per-block functions
    L = need("w1")
    c = pyoracle.ref_config(L)
    assert c["width"] == 1 and c["nevts"] == 1 and c["nb"] == 1
    trk, hit = pyoracle.ref_generate(L)
    state = trk.copy()
    after_prop, after_upd = [], []
    for lay in range(c["nlayer"]):
        hb = np.ascontiguousarray(hit[lay * 9:(lay + 1) * 9])
        nxt = np.empty_like(state)
        L.ref_propagate_block(state.ctypes.data, hb.ctypes.data, nxt.ctypes.data)
        after_prop.append(nxt.copy())
        L.ref_update_block(nxt.ctypes.data, hb.ctypes.data)
        after_upd.append(nxt.copy())
        state = nxt
    np.savez_compressed(os.path.join(HERE, "ref_track_layers.npz"), trk=trk, hit=hit,
                        after_prop=np.array(after_prop), after_update=np.array(after_upd))

    # the reference's own report at full size (runs its main)
    full = os.path.join(REPO, "oracle", "_ref", "libref_full.so")
    code = f"import ctypes; ctypes.CDLL({full!r}).ref_main()"
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         env=dict(os.environ, OMP_NUM_THREADS="8"))
    rep = [ln for ln in res.stdout.splitlines() if ln.startswith("track ") or ln.startswith("number of")]
    rep = [ln for ln in rep if not ln.startswith("track in ")]
    assert len(rep) == 11, res.stdout
    with open(os.path.join(HERE, "ref_summary_full.txt"), "w") as f:
        f.write("\n".join(rep) + "\n")
    write_fp64_summary(t, h)
    write_update_v1(need("small"))
    print("golden fixtures written to", HERE)


def write_update_v1(lib) -> None:
    from parity import v1_blocks
    rng = np.random.default_rng(2027)
    W = 32
    blocks = [v1_blocks(rng, W) for _ in range(8)]
    trk = np.concatenate([b[0] for b in blocks])
    hit = np.concatenate([b[1] for b in blocks])
    out = trk.copy()
    for t in range(len(blocks)):
        tb = np.ascontiguousarray(out[t * 28 * W:(t + 1) * 28 * W])
        hb = np.ascontiguousarray(hit[t * 9 * W:(t + 1) * 9 * W])
        lib.ref_update_v1_block(tb.ctypes.data, hb.ctypes.data)
        out[t * 28 * W:(t + 1) * 28 * W] = tb
    np.savez_compressed(os.path.join(HERE, "ref_update_v1.npz"), trk=trk, hit=hit, out=out, width=np.int64(W))


def write_fp64_summary(t, h) -> None:
    nevts, ntrks, W, nl = 100, 9600, 32, 20
    nb = ntrks // W
    trk, hit = pyoracle.generate(t, h, nevts, nb, W, nl, 1e-7)
    o64 = pyoracle.run_fp64(trk, hit, nevts, nb, W, nl)
    s = pyoracle.summary(pyoracle.fp64_blocks(o64, trk, W), nevts, nb, W, ntrks,
                         np.asarray(h, np.float32)[nl - 1, :3])
    with open(os.path.join(HERE, "fp64_summary_full.txt"), "w") as f:
        f.write(pyoracle.format_summary(s))


if __name__ == "__main__":
    main()
