"""CPU tests of the sin/cos the reference-typing mode computes on the GPU
(p2z-tests_amd/csrc/p2z_libm.hpp, a restatement of glibc's single-precision
sinf/cosf/sincosf algorithm): the header compiled for the host equals the C
library the reference calls (OMP:633-636, 643-644, 652-653) on every float
input, and its table of 2/pi bits is 2/pi."""
from __future__ import annotations

import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "p2z-tests_amd", "csrc")


@pytest.fixture(scope="module")
def sweep(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("libm") / "libm_sweep")
    # -ffp-contract=off: g++ ignores the header's clang pragma; every product
    # and sum must round as written, as on the device.
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-march=x86-64-v3", "-pthread",
                    "-Wno-unknown-pragmas", "-I", CSRC, os.path.join(REPO, "tests", "libm_sweep.cpp"), "-o", exe],
                   check=True)
    return exe


def _counts(out: str) -> dict:
    m = re.search(r"mismatch sinf (\d+) cosf (\d+) sincosf_s (\d+) sincosf_c (\d+) over (\d+)", out)
    assert m, out
    return dict(zip(("sinf", "cosf", "sincosf_s", "sincosf_c", "n"), map(int, m.groups())))


def test_equals_c_library_on_every_float(sweep):
    """All 2^32 bit patterns (zeros, subnormals, the three reduction ranges,
    infinities, NaNs): sinf, cosf and both outputs of sincosf, bit for bit."""
    out = subprocess.run([sweep], capture_output=True, text=True, check=True, timeout=900).stdout
    print(out)
    c = _counts(out)
    assert c["n"] == 2 ** 32
    assert c["sinf"] == c["cosf"] == c["sincosf_s"] == c["sincosf_c"] == 0, out


def test_two_over_pi_table(sweep):
    """The Payne-Hanek table: word i is the 32 bits of 2/pi starting at bit
    8*(i-3)+1 after the binary point, from pi computed in integers
    (Machin: pi = 16 atan(1/5) - 4 atan(1/239))."""
    bits = 400
    one = 1 << (bits + 32)

    def atan_inv(x: int) -> int:
        total, term, k, x2 = 0, one // x, 1, x * x
        while term:
            total += term // k if (k // 2) % 2 == 0 else -(term // k)
            term //= x2
            k += 2
        return total

    pi = 16 * atan_inv(5) - 4 * atan_inv(239)          # pi * 2^(bits+32)
    two_over_pi = (2 << (2 * (bits + 32))) // pi        # 2/pi * 2^(bits+32)
    frac = two_over_pi >> 32                            # 2/pi * 2^bits, integer part 0
    got = [int(w, 16) for w in subprocess.run([sweep, "table"], capture_output=True, text=True,
                                              check=True).stdout.split()]
    assert len(got) == 24
    for i, w in enumerate(got):
        first = 8 * (i - 3) + 1                          # first bit after the binary point
        shift = bits - (first + 31)
        expect = (frac >> shift) & 0xFFFFFFFF if first >= 1 else (frac >> shift) & ((1 << (first + 31)) - 1)
        assert w == expect, (i, hex(w), hex(expect))
