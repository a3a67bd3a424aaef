"""Pins the oracle (oracle/p2z_oracle.c) before anything is checked against it.

What pins what:
  * generator + path, bitwise: the golden vectors written by the reference
    compiled from its own source (tests/golden/make_golden.py), and, where the
    reference build is present, random inputs through both;
  * the report: the reference's printed known answer (logs/propagate_gcc_omp.log)
    and the report of the reference build at full size;
  * propagateToZ parameters: helix geometry (circle of radius |k pt| through
    the start point, tangent phi, transverse arc dz tan(theta)) in double;
  * the sparse covariance products: a dense Jt P Jt^T in double;
  * KalmanUpdate_v2: the textbook gain form in double and its limits
    (R -> infinity: no change; R -> 0: snap to the hit);
  * the FP64 variant: agreement with the float oracle at float precision;
  * the ragged driver: equivalence with the plain driver.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O

KFACT = np.float32(100.0 / (-0.299792458 * 3.8112))   # the reference's float kfact
KFACT64 = 100.0 / (-0.299792458 * 3.8112)              # the FP64 oracle's kfact


def ref_inputs(golden):
    g = golden["ref_input"]
    return np.array(g["track"], np.float64), np.array(g["hits"], np.float64)


# ---------------------------------------------------------------------------
# bitwise against the reference build's golden vectors
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["small", "w16"])
def test_generator_matches_reference(golden, name):
    g = golden[name]
    t, h = ref_inputs(golden)
    trk, hit = O.generate(t, h, int(g["nevts"]), int(g["nb"]), int(g["width"]), int(g["nlayer"]), 1e-7)
    assert np.array_equal(trk.view(np.uint32), g["trk"].view(np.uint32))
    assert np.array_equal(hit.view(np.uint32), g["hit"].view(np.uint32))


@pytest.mark.parametrize("name", ["small", "w16"])
def test_path_matches_reference(golden, name):
    g = golden[name]
    out = O.run(g["trk"], g["hit"], int(g["nevts"]), int(g["nb"]), int(g["width"]), int(g["nlayer"]))
    assert np.array_equal(out.view(np.uint32), g["out"].view(np.uint32))


def test_layer_by_layer_matches_reference(golden):
    g = golden["layers"]
    state = g["trk"].copy()
    for lay in range(g["after_prop"].shape[0]):
        hb = np.ascontiguousarray(g["hit"][lay * 9:(lay + 1) * 9])
        state = O.propagate_block(state, hb, 1)
        assert np.array_equal(state.view(np.uint32), g["after_prop"][lay].view(np.uint32)), lay
        state = O.update_block(state, hb, 1)
        assert np.array_equal(state.view(np.uint32), g["after_update"][lay].view(np.uint32)), lay


def _random_blocks(rng, W, zrow=True):
    from parity import random_blocks
    return random_blocks(rng, W, zrow)


def test_random_blocks_match_reference_build():
    L = O.ref_lib("small")
    if L is None:
        pytest.skip("reference build (oracle/_ref) not present")
    W = L.ref_bsize()
    rng = np.random.default_rng(7)
    for trial in range(20):
        blk, hit = _random_blocks(rng, W, zrow=bool(trial % 2))
        mine = O.propagate_block(blk, hit, W)
        theirs = np.empty_like(blk)
        L.ref_propagate_block(blk.ctypes.data, hit.ctypes.data, theirs.ctypes.data)
        assert np.array_equal(mine.view(np.uint32), theirs.view(np.uint32)), trial
        mine_u = O.update_block(blk, hit, W)
        theirs_u = blk.copy()
        L.ref_update_block(theirs_u.ctypes.data, hit.ctypes.data)
        assert np.array_equal(mine_u.view(np.uint32), theirs_u.view(np.uint32)), trial


# ---------------------------------------------------------------------------
# the report
# ---------------------------------------------------------------------------

def _parse(lines):
    vals = []
    for ln in lines:
        for tok in ln.replace("=", " ").split():
            try:
                vals.append(float(tok))
            except ValueError:
                pass
    return vals


def test_known_answer_full_size(golden):
    """Full reference configuration: the oracle's report is the reference
    build's report line for line, and the published log's to the last printed
    digit (the log's own host shows +-1 flips in that digit: FMA contraction
    and libm differ there)."""
    t, h = ref_inputs(golden)
    nevts, ntrks, W, nl = 100, 9600, 32, 20
    nb = ntrks // W
    trk, hit = O.generate(t, h, nevts, nb, W, nl, 1e-7)
    out = O.run(trk, hit, nevts, nb, W, nl)
    s = O.summary(out, nevts, nb, W, ntrks, h[nl - 1, :3].astype(np.float32))
    text = O.format_summary(s).splitlines()
    assert text == golden["ref_summary_full"]
    mine, ka = _parse(text), _parse(golden["known_answer"])
    assert len(mine) == len(ka)
    for a, b in zip(mine, ka):
        assert abs(a - b) <= 1.01e-6, (a, b)
    assert s.nnans == 0 and s.nfail == 0 and s.ngood == nevts * ntrks


def test_fp64_known_answer_full_size(golden):
    """The FP64 oracle's report at the full reference configuration is the
    committed fixture (the answer without FP32 rounding).  The reference's
    FP32 report departs from it by its own rounding bias: the tracks are
    copies of one track smeared by 1e-7, so their rounding errors do not
    average out (pt: 12.328097 against 12.327897)."""
    t, h = ref_inputs(golden)
    nevts, ntrks, W, nl = 100, 9600, 32, 20
    nb = ntrks // W
    trk, hit = O.generate(t, h, nevts, nb, W, nl, 1e-7)
    o64 = O.run_fp64(trk, hit, nevts, nb, W, nl)
    s = O.summary(O.fp64_blocks(o64, trk, W), nevts, nb, W, ntrks, h[nl - 1, :3].astype(np.float32))
    assert O.format_summary(s).splitlines() == golden["fp64_summary_full"]
    bias = [abs(a - b) for a, b in zip(_parse(golden["ref_summary_full"]), _parse(golden["fp64_summary_full"]))]
    assert max(bias) < 1e-3


# ---------------------------------------------------------------------------
# mathematics
# ---------------------------------------------------------------------------

def _unpack(blk, W, l):
    par = np.array([blk[e * W + l] for e in range(6)], np.float64)
    c = np.array([blk[(6 + e) * W + l] for e in range(21)], np.float64)
    P = np.zeros((6, 6))
    k = 0
    for i in range(6):
        for j in range(i + 1):
            P[i, j] = P[j, i] = c[k]
            k += 1
    q = blk[27 * W:28 * W].view(np.int32)[l]
    return par, P, q


def test_propagation_is_a_helix():
    rng = np.random.default_rng(3)
    W = 16
    blk, hit = _random_blocks(rng, W)
    out = O.propagate_block(blk, hit, W)
    for l in range(W):
        p0, _, q = _unpack(blk, W, l)
        p1, P1, _ = _unpack(out, W, l)
        x, y, z, ipt, phi, th = p0
        rho = float(q * KFACT) / ipt            # signed radius k * pt
        cx, cy = x - rho * np.sin(phi), y + rho * np.cos(phi)
        dz = float(hit[2 * W + l]) - z
        s_arc = dz * np.tan(th)                 # transverse arc length
        alpha = s_arc / rho
        assert p1[2] == np.float32(hit[2 * W + l])                       # z lands on the hit plane
        assert p1[3] == np.float32(ipt) and p1[5] == np.float32(th)      # ipt, theta unchanged
        assert abs(np.hypot(p1[0] - cx, p1[1] - cy) - abs(rho)) <= 2e-5 * (abs(rho) + 1)
        assert abs((p1[4] - phi) - alpha) <= 2e-5 * (1 + abs(alpha))
        # the new point is the start point turned by alpha about the centre
        ex = cx + rho * np.sin(phi + alpha)
        ey = cy - rho * np.cos(phi + alpha)
        assert abs(p1[0] - ex) <= 2e-5 * (1 + abs(ex)) and abs(p1[1] - ey) <= 2e-5 * (1 + abs(ey))
        assert np.all(P1[2, :] == 0) and np.all(P1[:, 2] == 0)           # z row zeroed


def test_zero_step_leaves_track_unchanged():
    rng = np.random.default_rng(4)
    W = 16
    blk, hit = _random_blocks(rng, W, zrow=False)
    hit[2 * W:3 * W] = blk[2 * W:3 * W]                                   # hit z = track z
    out = O.propagate_block(blk, hit, W)
    assert np.array_equal(out[:6 * W], blk[:6 * W])
    assert np.array_equal(out[6 * W:27 * W], blk[6 * W:27 * W])


def _jac_dense(p, q, hz, kfact=KFACT):
    """The 6x6 Jacobian of the reference with row 2 zeroed, in double."""
    x, y, z, ipt, phi, th = p
    k = float(q * kfact)
    dZ = hz - z
    pt = 1 / ipt
    cP, sP, cT, sT = np.cos(phi), np.sin(phi), np.cos(th), np.sin(th)
    iT = 1 / cT
    iTk = iT / k
    a = dZ * sT * ipt * iTk
    sa, ca = np.sin(a), np.cos(a)
    sC, cC = np.sin(cP * sa), np.cos(cP * sa)
    J = np.eye(6)
    J[2, 2] = 0
    J[0, 2] = cP * sT * (sP * ca * sC - ca) * iT
    J[0, 3] = cP * sT * dZ * ca * (1 - sP * sC) * (iT * pt) - k * (cP * sa - sP * (1 - cC)) * pt * pt
    J[0, 4] = (k * pt) * (-sP * sa + sP * sP * sa * sC - cP * (1 - cC))
    J[0, 5] = cP * dZ * ca * (1 - sP * sC) * iT * iT
    J[1, 2] = ca * sT * (cP * cP * sC - sP) * iT
    J[1, 3] = sT * dZ * ca * (cP * cP * sC + sP) * (iT * pt) - k * (sP * sa + cP * (1 - cC)) * pt * pt
    J[1, 4] = (k * pt) * (-sP * (1 - cC) - sP * cP * sa * sC + cP * sa)
    J[1, 5] = dZ * ca * (cP * cP * sC + sP) * iT * iT
    J[4, 2] = -ipt * sT * iTk
    J[4, 3] = sT * dZ * iTk
    J[4, 5] = ipt * dZ * iT * iTk
    return J


def test_covariance_products_are_dense_JPJt():
    """The sparse MultHelixPropEndcap / MultHelixPropTranspEndcap pair equals a
    dense Jt P Jt^T (catches a transposed operand or a wrong index)."""
    rng = np.random.default_rng(5)
    W = 16
    blk, hit = _random_blocks(rng, W)
    out = O.propagate_block(blk, hit, W)
    for l in range(W):
        p0, P0, q = _unpack(blk, W, l)
        _, P1, _ = _unpack(out, W, l)
        J = _jac_dense(p0, q, float(hit[2 * W + l]))
        ref = J @ P0 @ J.T
        scale = np.sqrt(np.outer(np.diag(ref), np.diag(ref))) + 1e-30
        assert np.max(np.abs(P1 - ref) / scale) < 1e-4, l
        assert np.allclose(P1, P1.T)


def _textbook_update(p, P, m, R):
    H = np.zeros((2, 6))
    H[0, 0] = H[1, 1] = 1
    S = H @ P @ H.T + R
    K = P @ H.T @ np.linalg.inv(S)
    return p + K @ (m - H @ p), P - K @ H @ P


def test_update_is_textbook_kalman():
    rng = np.random.default_rng(6)
    W = 16
    blk, hit = _random_blocks(rng, W, zrow=False)
    out = O.update_block(blk, hit, W)
    for l in range(W):
        p0, P0, _ = _unpack(blk, W, l)
        p1, P1, _ = _unpack(out, W, l)
        m = np.array([hit[l], hit[W + l]], np.float64)
        R = np.array([[hit[3 * W + l], hit[4 * W + l]], [hit[4 * W + l], hit[5 * W + l]]], np.float64)
        pe, Pe = _textbook_update(p0, P0, m, R)
        assert np.allclose(p1, pe, rtol=1e-5, atol=1e-5)
        scale = np.sqrt(np.outer(np.diag(P0), np.diag(P0))) + 1e-30
        assert np.max(np.abs(P1 - Pe) / scale) < 1e-4
        assert np.min(np.linalg.eigvalsh(P1)) > -1e-6 * np.max(np.diag(P0))   # still PSD


def test_update_limits():
    rng = np.random.default_rng(8)
    W = 16
    blk, hit = _random_blocks(rng, W, zrow=False)
    big = hit.copy()
    big[3 * W:4 * W] = 1e30
    big[5 * W:6 * W] = 1e30
    big[4 * W:5 * W] = 0
    out = O.update_block(blk, big, W)                       # R -> infinity: nothing moves
    assert np.allclose(out[:27 * W], blk[:27 * W], rtol=1e-6, atol=0)
    tiny = hit.copy()
    tiny[3 * W:6 * W] = 0                                   # R -> 0: snap to the hit
    out = O.update_block(blk, tiny, W)
    for l in range(W):
        p1, P1, _ = _unpack(out, W, l)
        _, P0, _ = _unpack(blk, W, l)
        assert abs(p1[0] - hit[l]) <= 1e-4 * (1 + abs(hit[l]))
        assert abs(p1[1] - hit[W + l]) <= 1e-4 * (1 + abs(hit[W + l]))
        assert abs(P1[0, 0]) <= 1e-4 * P0[0, 0] and abs(P1[1, 1]) <= 1e-4 * P0[1, 1]


def _numpy_double_track(p, P, q, hits):
    """One track through its layers in double with numpy: helix step, dense
    Jt P Jt^T, textbook Kalman update."""
    p = p.copy()
    for h in hits:
        hz = float(h[2])
        J = _jac_dense(p, q, hz, KFACT64)
        k = q * KFACT64
        x, y, z, ipt, phi, th = p
        a = (hz - z) * np.sin(th) * ipt / (np.cos(th) * k)
        px, py = np.cos(phi) / ipt, np.sin(phi) / ipt
        p = np.array([x + k * (px * np.sin(a) - py * (1 - np.cos(a))),
                      y + k * (py * np.sin(a) + px * (1 - np.cos(a))), hz, ipt, phi + a, th])
        P = J @ P @ J.T
        R = np.array([[h[3], h[4]], [h[4], h[5]]], np.float64)
        p, P = _textbook_update(p, P, np.array(h[:2], np.float64), R)
    return p, P


def test_fp64_oracle_is_double_precision(golden):
    """The FP64 oracle equals an independent double computation (numpy) to
    double-rounding level over all 20 layers."""
    g = golden["small"]
    nevts, nb, W, nl = (int(g[k]) for k in ("nevts", "nb", "width", "nlayer"))
    o64 = O.run_fp64(g["trk"], g["hit"], nevts, nb, W, nl)
    hits = g["hit"].reshape(nevts * nb, nl, 9, W)
    for t in (0, 5, 31, 40, 95):
        b, l = divmod(t, W)
        p0, P0, q = _unpack(g["trk"][b * 28 * W:(b + 1) * 28 * W], W, l)
        pe, Pe = _numpy_double_track(p0, P0, int(q), hits[b, :, :, l].astype(np.float64))
        assert np.allclose(o64[t, :6], pe, rtol=1e-9, atol=0), t
        P = np.zeros((6, 6))
        k = 0
        for i in range(6):
            for j in range(i + 1):
                P[i, j] = P[j, i] = o64[t, 6 + k]
                k += 1
        scale = np.sqrt(np.outer(np.abs(np.diag(Pe)), np.abs(np.diag(Pe)))) + 1e-300
        assert np.max(np.abs(P - Pe) / scale) < 1e-8, t


def test_float_oracle_error_against_fp64(golden):
    """How far the reference's FP32 arithmetic itself is from double: one layer
    (21 cm propagation + update) and all 20 layers.  These bounds are
    properties of the reference algorithm, measured on its own inputs."""
    g = golden["small"]
    nevts, nb, W, nl = (int(g[k]) for k in ("nevts", "nb", "width", "nlayer"))
    nblk = nevts * nb
    hit1 = g["hit"].reshape(nblk, nl, 9 * W)[:, 1:2].reshape(-1).copy()   # layer 1 only
    for hits, n, par_bound in ((hit1, 1, 3e-5), (g["hit"], nl, 5e-4)):
        o32 = O.run(g["trk"], hits, nevts, nb, W, n)
        o64 = O.run_fp64(g["trk"], hits, nevts, nb, W, n)
        t32 = o32.reshape(nblk, 28, W).transpose(0, 2, 1).reshape(-1, 28).astype(np.float64)
        par_rel = np.abs(t32[:, :6] - o64[:, :6]) / np.maximum(np.abs(o64[:, :6]), 1e-6)
        assert par_rel.max() < par_bound
        q32 = o32.reshape(nblk, 28, W)[:, 27, :].reshape(-1).view(np.int32)
        assert np.array_equal(q32.astype(np.float64), o64[:, 27])


def test_ragged_driver_equals_plain_driver(golden):
    g = golden["small"]
    nevts, nb, W, nl = (int(g[k]) for k in ("nevts", "nb", "width", "nlayer"))
    nblk = nevts * nb
    blk_trk = np.arange(nblk)
    blk_hit = np.arange(nblk) * nl
    blk_nlay = np.full(nblk, nl, np.int32)
    lane = np.full(nblk * W, nl, np.int32)
    out = O.run_ragged(g["trk"], g["hit"], nblk, W, blk_trk, blk_hit, blk_nlay, lane)
    assert np.array_equal(out.view(np.uint32), g["out"].view(np.uint32))
    # lanes that stop early equal a plain run with that many layers
    rng = np.random.default_rng(2)
    lane = rng.integers(1, nl + 1, nblk * W).astype(np.int32)
    out = O.run_ragged(g["trk"], g["hit"], nblk, W, blk_trk, blk_hit, blk_nlay, lane)
    for n in (1, 7, nl):
        sel = np.nonzero(lane == n)[0]
        if sel.size == 0:
            continue
        # plain run with n layers: hits of each block's first n layers
        hit_n = g["hit"].reshape(nblk, nl, 9 * W)[:, :n].reshape(-1).copy()
        plain = O.run(g["trk"], hit_n, nevts, nb, W, n)
        for t in sel:
            b, l = divmod(int(t), W)
            a = out.reshape(nblk, 28, W)[b, :, l]
            e = plain.reshape(nblk, 28, W)[b, :, l]
            assert np.array_equal(a.view(np.uint32), e.view(np.uint32))


# ---------------------------------------------------------------------------
# KalmanUpdate v1 (OMP:375-501), the optional 3-D update
# ---------------------------------------------------------------------------

def _v1_oracle(trk, hit, W, fixed):
    nblk = len(trk) // (28 * W)
    return np.concatenate([O.update_v1_block(np.ascontiguousarray(trk[b * 28 * W:(b + 1) * 28 * W]),
                                              np.ascontiguousarray(hit[b * 9 * W:(b + 1) * 9 * W]), W, fixed)
                           for b in range(nblk)])


def test_update_v1_matches_reference_build(golden):
    """The oracle's v1 as written is the reference's KalmanUpdate word for
    word (fixture written by the reference build, tests/golden/make_golden.py),
    and the reference build itself where it is present."""
    g = golden["update_v1"]
    W = int(g["width"])
    mine = _v1_oracle(g["trk"], g["hit"], W, False)
    assert np.array_equal(mine.view(np.uint32), g["out"].view(np.uint32))
    L = O.ref_lib("small")
    if L is not None:
        from parity import v1_blocks
        rng = np.random.default_rng(8)
        for _ in range(4):
            t, h = v1_blocks(rng, W)
            ref = t.copy()
            L.ref_update_v1_block(ref.ctypes.data, h.ctypes.data)
            assert np.array_equal(O.update_v1_block(t, h, W, False).view(np.uint32), ref.view(np.uint32))


def _sym(c):
    P = np.zeros((6, 6))
    k = 0
    for i in range(6):
        for j in range(i + 1):
            P[i, j] = P[j, i] = c[k]
            k += 1
    return P


def test_update_v1_fixed_is_textbook():
    """The fixed form is the textbook 3-D Kalman update K = P H^T (H P H^T +
    R)^-1, p' = p + K (m - H p), P' = P - K H P with H = [I3 0], computed
    here independently in numpy double."""
    from parity import v1_blocks
    rng = np.random.default_rng(21)
    W = 32
    for _ in range(4):
        trk, hit = v1_blocks(rng, W)
        out = O.update_v1_block(trk, hit, W, True)
        for l in range(W):
            par = np.array([trk[e * W + l] for e in range(6)], np.float64)
            P = _sym([trk[(6 + e) * W + l] for e in range(21)])
            hc = [hit[(3 + e) * W + l] for e in range(6)]
            R = np.array([[hc[0], hc[1], hc[3]], [hc[1], hc[2], hc[4]], [hc[3], hc[4], hc[5]]], np.float64)
            m = np.array([hit[e * W + l] for e in range(3)], np.float64)
            H = np.zeros((3, 6))
            H[0, 0] = H[1, 1] = H[2, 2] = 1
            K = P @ H.T @ np.linalg.inv(H @ P @ H.T + R)
            p2 = par + K @ (m - H @ par)
            P2 = P - K @ H @ P
            gp = np.array([out[e * W + l] for e in range(6)], np.float64)
            gP = _sym([out[(6 + e) * W + l] for e in range(21)])
            assert np.all(np.abs(gp - p2) <= 1e-5 * np.maximum(np.abs(p2), 1.0)), (l, gp, p2)
            d = np.sqrt(np.abs(np.diag(P2)))
            scale = np.maximum(np.abs(P2), 1e-3 * np.outer(d, d))
            assert np.max(np.abs(gP - P2) / scale) < 1e-3, l


def test_update_v1_as_written_writes_pt_into_ipt():
    """The defect the as-written form carries (OMP:447, 494): with an
    uninformative hit (R -> infinity, K -> 0) it leaves x, y, z, phi, theta
    and P as they are but replaces ipt by pt = 1/ipt; the fixed form leaves
    ipt too."""
    from parity import random_blocks
    rng = np.random.default_rng(4)
    W = 16
    trk, hit = random_blocks(rng, W)
    big = np.float32(1e10)                    # the float determinant overflows past ~1e12
    up = hit.copy()                           # upper addressing: xx xy xz yy yz zz
    up[3 * W:9 * W] = 0
    up[3 * W:4 * W] = up[6 * W:7 * W] = up[8 * W:9 * W] = big
    lo = hit.copy()                           # lower packing: xx xy yy xz yz zz
    lo[3 * W:9 * W] = 0
    lo[3 * W:4 * W] = lo[5 * W:6 * W] = lo[8 * W:9 * W] = big
    a = O.update_v1_block(trk, up, W, False)
    f = O.update_v1_block(trk, lo, W, True)
    for out in (a, f):
        for e in (0, 1, 2, 4, 5):
            assert np.allclose(out[e * W:(e + 1) * W], trk[e * W:(e + 1) * W], rtol=1e-6, atol=1e-6)
        assert np.allclose(out[6 * W:27 * W], trk[6 * W:27 * W], rtol=1e-5, atol=1e-12)
    ipt = trk[3 * W:4 * W].astype(np.float64)
    assert np.allclose(a[3 * W:4 * W], 1.0 / ipt, rtol=1e-6)
    assert np.allclose(f[3 * W:4 * W], ipt, rtol=1e-6)
