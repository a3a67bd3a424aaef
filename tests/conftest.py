"""Shared test setup.

Markers: `gpu` marks tests that need an MI355X (they call the HIP path through
the C ABI); everything else runs on the CPU.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "p2z-tests_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP path through the C ABI)")


@pytest.fixture(scope="session")
def golden():
    """Loads the golden fixtures written by tests/golden/make_golden.py."""
    with open(os.path.join(GOLDEN, "ref_input.json")) as f:
        ref_input = json.load(f)

    def npz(name):
        with np.load(os.path.join(GOLDEN, name)) as z:
            return {k: z[k] for k in z.files}

    def lines(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return f.read().splitlines()

    return dict(
        ref_input=ref_input,
        small=npz("ref_small.npz"),
        w16=npz("ref_w16.npz"),
        layers=npz("ref_track_layers.npz"),
        known_answer=lines("known_answer.txt"),
        ref_summary_full=lines("ref_summary_full.txt"),
        fp64_summary_full=lines("fp64_summary_full.txt"),
        update_v1=npz("ref_update_v1.npz"),
    )


@pytest.fixture(scope="session")
def ctx():
    """One p2z context on GPU 0 for the whole GPU session."""
    import p2z

    c = p2z.Context(0)
    yield c
    c.close()
