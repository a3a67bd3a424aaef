"""CPU test of the argument reduction of sincos_accurate
(p2z-tests_amd/csrc/p2z_lane.hpp), which starts the FP64-gain mode's track
angles: k = rint(x * 2/pi) in double, r = fma(-k, c2, fma(-k, c1, x)) with
c1 + c2 = pi/2 to 107 bits, used for |x| < 2^20 (larger arguments go to the
glibc-exact sincosf_glibc, whose reduction is exact at every magnitude).  The
device formula is evaluated here with the same double constants (fma exactly,
through rationals) and held against the exact reduction of each float by a
405-bit pi: the quadrant is the nearest one (or either one when x sits within
rounding of a quadrant boundary), and r is within a few ulp of the exact
remainder, down to the floats nearest to multiples of pi/2."""
from __future__ import annotations

from fractions import Fraction

import numpy as np

C_2PI = 0.63661977236758134308          # 2/pi, rounded to double
C1 = 1.5707963267948965580               # pi/2, rounded to double
C2 = 6.123233995736766036e-17            # pi/2 - C1, rounded to double
LIMIT = 2.0 ** 20


def _pi_fraction(bits: int = 400) -> Fraction:
    one = 1 << (bits + 32)

    def atan_inv(x: int) -> int:
        total, term, k, x2 = 0, one // x, 1, x * x
        while term:
            total += term // k if (k // 2) % 2 == 0 else -(term // k)
            term //= x2
            k += 2
        return total

    return Fraction(16 * atan_inv(5) - 4 * atan_inv(239), one)


PI = _pi_fraction()
HALF_PI = PI / 2


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def device_reduction(x: float):
    k = float(np.rint(np.float64(x) * np.float64(C_2PI)))
    r = fma(-k, C2, fma(-k, C1, x))
    return k, r


def exact_reduction(x: float):
    q = Fraction(x) / HALF_PI
    k = round(q)                           # nearest integer (ties cannot occur: pi is irrational)
    return k, Fraction(x) - k * HALF_PI


def check(x: float) -> None:
    k, r = device_reduction(x)
    ke, re = exact_reduction(x)
    if k != ke:
        # only allowed where x lies within rounding of a quadrant boundary: then
        # both reductions are valid, r just past pi/4
        assert abs(k - ke) == 1 and abs(abs(float(re)) - float(HALF_PI / 2)) < 1e-9, (x, k, ke)
        re = Fraction(x) - int(k) * HALF_PI
    err = abs(Fraction(r) - re)
    ulp = np.spacing(abs(float(re))) if re != 0 else 2.0 ** -1074
    assert err <= 4 * Fraction(float(ulp)) + abs(int(k)) * Fraction(2e-33), (x, float(err), float(ulp))
    assert abs(r) <= float(HALF_PI / 2) * (1 + 1e-12), (x, r)


def test_reduction_random_floats_below_limit():
    rng = np.random.default_rng(7)
    e = rng.uniform(-4.0, 20.0, 3000)
    xs = (np.sign(rng.uniform(-1, 1, 3000)) * 2.0 ** e).astype(np.float32)
    xs = np.concatenate([xs, np.nextafter(np.float32(LIMIT), np.float32(0)) * np.array([1, -1], np.float32)])
    for x in xs:
        assert abs(float(x)) < LIMIT
        check(float(x))


def test_reduction_near_multiples_of_half_pi():
    """The floats nearest to k*pi/2 (the largest cancellation), for k up to
    the largest quadrant below 2^20."""
    kmax = int(LIMIT / float(HALF_PI))
    rng = np.random.default_rng(11)
    ks = np.unique(np.concatenate([np.arange(1, 200), rng.integers(200, kmax, 2000), [kmax]]))
    for k in ks:
        t = Fraction(int(k)) * HALF_PI
        x = float(np.float32(float(t)))
        for y in (x, float(np.nextafter(np.float32(x), np.float32(np.inf))),
                  float(np.nextafter(np.float32(x), np.float32(-np.inf)))):
            if abs(y) < LIMIT:
                check(y)
                check(-y)
