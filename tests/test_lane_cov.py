"""CPU check of the fast modes' covariance propagation after the first layer
(propagate_cov_fast_z0, p2z-tests_amd/csrc/p2z_cov.hpp; reference
MultHelixPropEndcap / MultHelixPropTranspEndcap, OMP:296-373): compiled for
the host and compared with the dense double product Jt P Jt^T, where Jt is the
Jacobian with row 2 zeroed (OMP:316-321), on covariances whose z row and
column are zero."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "p2z-tests_amd", "csrc")
NAMES = ["J02", "J03", "J04", "J05", "J12", "J13", "J14", "J15", "J42", "J43", "J45"]


def packed_index(i, j):
    i, j = max(i, j), min(i, j)
    return i * (i + 1) // 2 + j


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("lane") / "lane_cov_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-I", CSRC,
                    os.path.join(HERE, "lane_cov_check.cpp"), "-o", exe], check=True)
    return exe


def dense(jac, P):
    J = np.eye(6)
    for n, v in zip(NAMES, jac):
        J[int(n[1]), int(n[2])] = v
    J[2, :] = 0.0
    C = np.array([[P[packed_index(i, j)] for j in range(6)] for i in range(6)], np.float64)
    E = J @ C @ J.T
    return np.array([E[i, j] for i in range(6) for j in range(i + 1)])


def test_matches_dense_product(prog):
    rng = np.random.default_rng(3)
    cases = []
    for _ in range(300):
        jac = rng.normal(0, 1, 11).astype(np.float32)
        A = rng.normal(0, 1, (6, 6)) * np.exp(rng.uniform(-3, 3, 6))[:, None]
        C = A @ A.T
        C[2, :] = 0.0
        C[:, 2] = 0.0
        P = np.array([C[i, j] for i in range(6) for j in range(i + 1)], np.float32)
        cases.append((jac, P))
    inp = "\n".join(" ".join(repr(float(x)) for x in list(j) + list(P)) for j, P in cases)
    out = subprocess.run([prog], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for k, (jac, P) in enumerate(cases):
        got = np.array([float(x) for x in out[k].split()])
        exp = dense(jac.astype(np.float64), P.astype(np.float64))
        z = [packed_index(2, j) for j in range(6)]
        assert np.all(got[z] == 0.0), k
        # float rounding of products of O(1) Jacobian entries: relative to the
        # largest entries of the row and column the word belongs to
        scale = np.array([np.sqrt(abs(exp[packed_index(i, i)] * exp[packed_index(j, j)])) + abs(exp).max() * 1e-6
                          for i in range(6) for j in range(i + 1)])
        err = np.abs(got - exp) / scale
        assert err.max() < 2e-5, (k, err.max(), int(err.argmax()))


def test_identity_jacobian_leaves_covariance(prog):
    """With the 11 entries zero Jt is the identity on the live indices: the
    covariance comes back unchanged, bit for bit."""
    rng = np.random.default_rng(4)
    A = rng.normal(0, 1, (6, 6))
    C = A @ A.T
    C[2, :] = 0.0
    C[:, 2] = 0.0
    P = np.array([C[i, j] for i in range(6) for j in range(i + 1)], np.float32)
    inp = " ".join(["0"] * 11 + [repr(float(x)) for x in P])
    got = np.array([float(x) for x in subprocess.run([prog], input=inp, capture_output=True, text=True,
                                                     check=True).stdout.split()], np.float32)
    assert np.array_equal(got.view(np.uint32), P.view(np.uint32))
