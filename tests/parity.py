"""Parity metric shared by the GPU tests, smoke() and the precision report.

Per track, relative differences against a reference result b:
  parameters : |a - b| / max(|b|, floor_e), floor 1 cm for x, y, z and 1e-6
               for ipt, phi, theta (positions are compared at a 1 cm scale so
               that a coordinate passing near 0 is not a relative blow-up)
  covariance : |a - b| / max(|b|, floor_ij), floor_ij = COV_FLOOR * sqrt(|b_ii b_jj|)
The covariance floor is a fraction of the entry's correlation scale, so that
entries the update drives to (near) zero by cancellation are measured
against the size of the matrix they belong to, and exact zeros compare equal.
Tracks whose parameters are not finite in b must be non-finite in a too and
are excluded from the relative measures.
"""
from __future__ import annotations

import numpy as np

PAR_TOL = 1e-5
COV_TOL = 1e-4
COV_FLOOR = 1e-3
PAR_FLOOR = np.array([1.0, 1.0, 1.0, 1e-6, 1e-6, 1e-6])
_DIAG = [0, 2, 5, 9, 14, 20]
_IJ = [(i, j) for i in range(6) for j in range(i + 1)]


def track_major(blocks: np.ndarray, width: int) -> np.ndarray:
    """MPTRK blocks (flat float32) -> (ntracks, 28) float64; column 27 = q."""
    nb = blocks.size // (28 * width)
    t = blocks.reshape(nb, 28, width).transpose(0, 2, 1).reshape(-1, 28)
    out = t.astype(np.float64)
    out[:, 27] = t[:, 27].copy().view(np.int32)
    return out


def compare(a: np.ndarray, b: np.ndarray) -> dict:
    """a, b: (ntracks, 28) float64 (track_major or the FP64 oracle's output)."""
    fin_b = np.all(np.isfinite(b[:, :6]), axis=1)
    fin_a = np.all(np.isfinite(a[:, :6]), axis=1)
    a, b = a[fin_b], b[fin_b]
    par = np.abs(a[:, :6] - b[:, :6]) / np.maximum(np.abs(b[:, :6]), PAR_FLOOR)
    Pb = b[:, 6:27]
    d = np.abs(Pb[:, _DIAG])
    scale = np.stack([np.sqrt(d[:, i] * d[:, j]) for i, j in _IJ], 1)
    cov = np.abs(a[:, 6:27] - Pb) / np.maximum(np.maximum(np.abs(Pb), COV_FLOOR * scale), 1e-300)
    return dict(
        ntracks=int(fin_b.size),
        nonfinite_ref=int((~fin_b).sum()),
        nonfinite_mismatch=int((fin_a != fin_b).sum()),
        charge_mismatch=int((a[:, 27] != b[:, 27]).sum()),
        par_max=float(par.max()) if par.size else 0.0,
        par_p99=float(np.quantile(par.max(axis=1), 0.99)) if par.size else 0.0,
        par_median=float(np.median(par)) if par.size else 0.0,
        cov_max=float(cov.max()) if cov.size else 0.0,
        cov_p99=float(np.quantile(cov.max(axis=1), 0.99)) if cov.size else 0.0,
        cov_median=float(np.median(cov)) if cov.size else 0.0,
        par_max_by_elem=[float(v) for v in par.max(axis=0)] if par.size else [],
    )


def ok(stats: dict, par_tol: float = PAR_TOL, cov_tol: float = COV_TOL) -> bool:
    return (stats["nonfinite_mismatch"] == 0 and stats["charge_mismatch"] == 0
            and stats["par_max"] <= par_tol and stats["cov_max"] <= cov_tol)


def same_words(a: np.ndarray, b: np.ndarray) -> bool:
    """Bit-for-bit equality of two float32 arrays."""
    return a.shape == b.shape and bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))


def words_equal(a: np.ndarray, b: np.ndarray) -> float:
    """Fraction of equal 32-bit words (for the log of a failing comparison)."""
    return float(np.mean(a.view(np.uint32) == b.view(np.uint32)))


def track_pass_fraction(a: np.ndarray, b: np.ndarray, par_tol: float = PAR_TOL, cov_tol: float = COV_TOL) -> float:
    """a, b: (ntracks, 28) float64.  Fraction of tracks whose every parameter
    is within par_tol and every covariance entry within cov_tol of b (the
    metric of compare(), per track); non-finite tracks pass when both are
    non-finite."""
    fin_b = np.all(np.isfinite(b[:, :6]), axis=1)
    fin_a = np.all(np.isfinite(a[:, :6]), axis=1)
    par = np.abs(a[:, :6] - b[:, :6]) / np.maximum(np.abs(b[:, :6]), PAR_FLOOR)
    Pb = b[:, 6:27]
    d = np.abs(Pb[:, _DIAG])
    scale = np.stack([np.sqrt(d[:, i] * d[:, j]) for i, j in _IJ], 1)
    cov = np.abs(a[:, 6:27] - Pb) / np.maximum(np.maximum(np.abs(Pb), COV_FLOOR * scale), 1e-300)
    with np.errstate(invalid="ignore"):
        good = (par.max(axis=1) <= par_tol) & (cov.max(axis=1) <= cov_tol) & (a[:, 27] == b[:, 27])
    good = np.where(fin_b, good & fin_a, ~fin_a)
    return float(np.mean(good)) if good.size else 1.0


def random_blocks(rng, W: int, zrow: bool = True):
    """One random but physical track block and hit block of width W: SPD
    covariance (with or without a z row), |cos theta| away from 0, hits near
    the track in x/y and 30 cm around it in z."""
    n = W
    par = np.stack([rng.uniform(-50, 50, n), rng.uniform(-50, 50, n), rng.uniform(-100, 100, n),
                    rng.uniform(0.05, 1.0, n) * rng.choice([-1, 1], n), rng.uniform(-3, 3, n),
                    np.where(rng.random(n) < 0.5, rng.uniform(0.2, 1.2, n), rng.uniform(1.9, 2.9, n))])
    cov = np.zeros((21, n))
    for l in range(n):
        A = rng.normal(size=(6, 6)) * 1e-2
        S = A @ A.T + np.diag(rng.uniform(1e-5, 1e-3, 6))
        if not zrow:
            S[2, :] = S[:, 2] = 0
        cov[:, l] = [S[i, j] for i in range(6) for j in range(i + 1)]
    q = rng.choice([-1, 1], n).astype(np.int32)
    blk = np.concatenate([par.reshape(-1), cov.reshape(-1)]).astype(np.float32)
    blk = np.concatenate([blk, q.view(np.float32)])
    hit = np.concatenate([par[0] + rng.normal(0, 0.5, n), par[1] + rng.normal(0, 0.5, n),
                          par[2] + rng.uniform(-30, 30, n), rng.uniform(1e-4, 1e-1, n),
                          rng.uniform(-1e-5, 1e-5, n), rng.uniform(1e-4, 1e-1, n),
                          np.zeros(3 * n)]).astype(np.float32)
    return blk, hit


def v1_blocks(rng, W: int):
    """random_blocks with all 6 hit covariance words set (xx, xy, yy, xz, yz,
    zz in the path's lower packing), for the 3-D update v1."""
    blk, hit = random_blocks(rng, W)
    n = W
    hit[6 * n:7 * n] = rng.uniform(1e-6, 1e-5, n)   # xz
    hit[7 * n:8 * n] = rng.uniform(1e-6, 1e-5, n)   # yz
    hit[8 * n:9 * n] = rng.uniform(1e-2, 1e-1, n)   # zz
    return blk, hit
