"""CPU tests of the synthetic workloads and the ragged format
(p2z_workload_*, p2z_check_blocks): the distributions the configurations ask
for, determinism, the packing invariants the kernels rely on, and the save /
load round trip with its checksums.  Host code only: no GPU is needed."""
from __future__ import annotations

import os

import numpy as np
import pytest

import p2z


@pytest.fixture(scope="module")
def ragged():
    w = p2z.Workload("ragged", nevents=120, width=32, seed=7)
    yield w
    w.close()


def lanes_of_event(w, e):
    b0, b1 = int(w.event_blocks[e]), int(w.event_blocks[e + 1])
    W = w.width
    tb = w.blocks["trk_block"][b0:b1]
    return np.concatenate([w.lane_nlayer[t * W:(t + 1) * W] for t in tb])


def test_ragged_event_sizes_and_layer_counts(ragged):
    """Tracks per event log-normal around 2000 clipped to [50, 20000]; layers
    per track in 8..26; padding lanes carry no layers."""
    w = ragged
    nt = np.asarray(w.event_tracks)
    assert nt.min() >= 50 and nt.max() <= 20000
    assert 1000 < np.median(nt) < 4000
    assert w.ntracks == nt.sum()
    n = np.asarray(w.lane_nlayer)
    live = n[n > 0]
    assert live.size == w.ntracks and live.min() >= 8 and live.max() <= 26
    assert set(np.unique(n)) <= {0} | set(range(8, 27))
    assert w.track_layers == int(live.sum())
    assert 12 < live.mean() < 22                      # U{8..26}: mean 17


def test_ragged_packing_invariants(ragged):
    """Within an event the lanes are packed longest first, each descriptor's
    nlayer is its block's longest lane, descriptors address their own blocks,
    and blocks_lpt holds the same descriptors sorted longest first."""
    w = ragged
    W = w.width
    for e in range(w.nevents):
        n = lanes_of_event(w, e)
        k = int(w.event_tracks[e])
        assert np.all(np.diff(n[:k]) <= 0) and np.all(n[k:] == 0), e
    b = np.asarray(w.blocks)
    for i, d in enumerate(b):
        assert d["out_block"] == i and d["trk_block"] == i
        assert d["nlayer"] == w.lane_nlayer[i * W:(i + 1) * W].max()
    p2z.check_blocks(b, w.nblocks, w.nhit_blocks, w.nblocks)
    lpt = np.asarray(w.blocks_lpt)
    assert np.all(np.diff(lpt["nlayer"]) <= 0)
    assert sorted(lpt["trk_block"].tolist()) == list(range(w.nblocks))
    assert np.array_equal(np.sort(lpt, order="trk_block"), b)


def test_ragged_hits_follow_the_layers(ragged):
    """Consecutive layers of a track step 5..25 cm in z in one direction; hit
    variances in [1e-6, 1e-3] cm^2 with |correlation| <= 0.5."""
    w = ragged
    W = w.width
    hits = np.asarray(w.hits).reshape(-1, 9, W)
    rng = np.random.default_rng(0)
    for i in rng.choice(w.nblocks, 40, replace=False):
        d = w.blocks[i]
        for l in range(W):
            n = int(w.lane_nlayer[i * W + l])
            if n == 0:
                continue
            h = hits[d["hit_block"]:d["hit_block"] + n, :, l].astype(np.float64)
            dz = np.diff(h[:, 2])
            assert np.all(np.abs(dz) >= 5 - 1e-3) and np.all(np.abs(dz) <= 25 + 1e-3)
            assert np.all(np.sign(dz) == np.sign(dz[0]))
            xx, xy, yy = h[:, 3], h[:, 4], h[:, 5]
            assert np.all((xx >= 1e-6 * 0.999) & (xx <= 1e-3 * 1.001) & (yy >= 1e-6 * 0.999) & (yy <= 1e-3 * 1.001))
            assert np.all(np.abs(xy) <= 0.5 * np.sqrt(xx * yy) * 1.001)


def test_precision_workload_ranges():
    """The wide-range workload: uniform events of ntracks x nlayer, 1/pT in
    [0.01, 2] (starting states smeared by their covariance), theta in
    [0.1, 3.04], hit variances in [1e-8, 1e-3]."""
    w = p2z.Workload("precision", nevents=2, width=32, seed=3, ntracks=640, nlayer=20)
    assert w.ntracks == 1280 and w.track_layers == 1280 * 20
    assert np.all(np.asarray(w.lane_nlayer) == 20)
    t = np.asarray(w.tracks).reshape(-1, 28, 32)
    ipt = np.abs(t[:, 3, :])
    assert np.quantile(ipt, 0.01) > 0.005 and np.quantile(ipt, 0.99) < 2.5
    theta = t[:, 5, :]
    assert np.quantile(theta, 0.01) > 0.05 and np.quantile(theta, 0.99) < 3.1
    h = np.asarray(w.hits).reshape(-1, 9, 32)
    assert h[:, 3].min() >= 1e-8 * 0.999 and h[:, 3].max() <= 1e-3 * 1.001
    w.close()


def test_determinism_across_threads():
    """One random stream per event: the same seed gives the same workload
    whatever the number of generator threads."""
    a = p2z.Workload("ragged", nevents=24, width=64, seed=99, nthreads=1)
    b = p2z.Workload("ragged", nevents=24, width=64, seed=99, nthreads=0)
    c = p2z.Workload("ragged", nevents=24, width=64, seed=100, nthreads=0)
    for k in ("tracks", "hits", "lane_nlayer", "event_blocks", "event_tracks"):
        assert np.array_equal(np.asarray(getattr(a, k)).view(np.uint8), np.asarray(getattr(b, k)).view(np.uint8)), k
    assert np.array_equal(np.asarray(a.blocks), np.asarray(b.blocks))
    assert not np.array_equal(np.asarray(a.tracks), np.asarray(c.tracks))
    for w in (a, b, c):
        w.close()


def test_widths_hold_the_same_tracks():
    """The generator gives the same tracks at every AoSoA width (the autotuner
    compares widths on equal work)."""
    ws = {W: p2z.Workload("ragged", nevents=5, width=W, seed=4) for W in (16, 128)}

    def by_track(w):
        W = w.width
        t = np.asarray(w.tracks).reshape(-1, 28, W).transpose(0, 2, 1).reshape(-1, 28)
        return t[np.asarray(w.lane_nlayer) > 0]

    a, b = by_track(ws[16]), by_track(ws[128])
    assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for w in ws.values():
        w.close()


def test_save_load_round_trip(ragged, tmp_path):
    path = str(tmp_path / "events.p2zw")
    ragged.save(path)
    w = p2z.Workload.load(path)
    for k in ("tracks", "hits", "lane_nlayer", "event_blocks", "event_tracks"):
        assert np.array_equal(np.asarray(getattr(w, k)).view(np.uint8),
                              np.asarray(getattr(ragged, k)).view(np.uint8)), k
    assert np.array_equal(np.asarray(w.blocks), np.asarray(ragged.blocks))
    assert np.array_equal(np.asarray(w.blocks_lpt), np.asarray(ragged.blocks_lpt))
    assert (w.nevents, w.ntracks, w.track_layers, w.width) == (ragged.nevents, ragged.ntracks, ragged.track_layers,
                                                              ragged.width)
    w.close()
    # a flipped byte in the arrays fails the checksum
    with open(path, "r+b") as f:
        f.seek(os.path.getsize(path) - 100)
        v = f.read(1)
        f.seek(-1, 1)
        f.write(bytes([v[0] ^ 0x5A]))
    with pytest.raises(p2z.P2zError) as e:
        p2z.Workload.load(path)
    assert e.value.rc == p2z.EINVAL
    with pytest.raises(p2z.P2zError):
        p2z.Workload.load(str(tmp_path / "missing.p2zw"))


def test_load_of_an_impossible_header_fails_cleanly(ragged, tmp_path):
    """A header that claims more blocks than host memory holds: the loader's
    failed allocation comes back as a status (no exception crosses the C ABI,
    no crash), and the library keeps working."""
    path = str(tmp_path / "events.p2zw")
    ragged.save(path)
    with open(path, "rb") as f:
        hdr, rest = f.read().split(b"\n", 1)
    big = hdr.replace(b'"nblocks":%d' % ragged.nblocks, b'"nblocks":%d' % (1 << 40))
    assert big != hdr
    bad = str(tmp_path / "huge.p2zw")
    with open(bad, "wb") as f:
        f.write(big + b"\n" + rest)
    with pytest.raises(p2z.P2zError) as e:
        p2z.Workload.load(bad)
    assert e.value.rc in (p2z.ENOMEM, p2z.EINVAL)
    assert "p2z_workload_load" in str(e.value) or "allocation" in str(e.value)
    w = p2z.Workload.load(path)
    assert w.nblocks == ragged.nblocks
    w.close()


def test_check_blocks_rejects_outside_descriptors(ragged):
    w = ragged
    b = np.array(w.blocks[:8])
    p2z.check_blocks(b, w.nblocks, w.nhit_blocks, w.nblocks)
    p2z.check_blocks(b[:0], w.nblocks, w.nhit_blocks, w.nblocks)
    for field, value in (("trk_block", w.nblocks), ("trk_block", -1), ("hit_block", w.nhit_blocks - 1),
                         ("out_block", w.nblocks), ("nlayer", -1)):
        bad = b.copy()
        bad[3][field] = value
        with pytest.raises(p2z.P2zError) as e:
            p2z.check_blocks(bad, w.nblocks, w.nhit_blocks, w.nblocks)
        assert e.value.rc == p2z.EINVAL and "descriptor 3" in str(e.value), field
