"""Derives the polynomial coefficients of sincos_accurate and sincos_unit
(p2z_lane.hpp).

sincos_accurate (FP64 evaluation), on |r| <= pi/4:
    sin r = r + r^3 * S(r^2),   S of degree ns-1 in r^2
    cos r = 1 + r^2 * C(r^2),   C of degree nc-1 in r^2
S and C are fitted to (sin r - r)/r^3 and (cos r - 1)/r^2 by least squares on
Chebyshev nodes in u = r^2, in numpy's extended precision (80-bit long
double), then rounded to double.  The reported errors are the maximum
relative errors of the double-precision evaluation (Horner, fused
multiply-adds) against math.sin / math.cos on a fine grid.

    python gen_sincos_coeffs.py [ns] [nc]

sincos_unit (FP32 evaluation, no argument reduction), on |x| <= 1: the same
forms with 4 + 4 coefficients fitted in u = x^2 over [0, 1], rounded to
float.  The report evaluates them in float with fused multiply-adds (each
FMA emulated exactly in double, then rounded once) against math.sin/math.cos.

    python gen_sincos_coeffs.py --unit
"""
import math
import sys

import numpy as np

L = np.longdouble


def fit(f, deg, lo, hi, n=4000):
    k = np.arange(n, dtype=L)
    u = (L(lo) + L(hi)) / 2 + (L(hi) - L(lo)) / 2 * np.cos((2 * k + 1) * L(math.pi) / (2 * n))
    A = np.stack([u ** j for j in range(deg + 1)], 1)
    y = f(u)
    # scale columns for conditioning, solve normal equations in long double
    sc = np.max(np.abs(A), 0)
    As = A / sc
    c = np.linalg.solve((As.T @ As).astype(np.float64), (As.T @ y).astype(np.float64))
    # one refinement step in long double
    r = y - As @ c.astype(L)
    dc = np.linalg.solve((As.T @ As).astype(np.float64), (As.T @ r).astype(np.float64))
    return (c.astype(L) + dc.astype(L)) / sc


def sin_tail(u):
    r = np.sqrt(u)
    return (np.sin(r) - r) / (r ** 3)


def cos_tail(u):
    r = np.sqrt(u)
    return (np.cos(r) - 1) / u


def horner(cs, x):
    acc = cs[-1]
    for c in cs[-2::-1]:
        acc = math.fma(acc, x, c) if hasattr(math, "fma") else acc * x + c
    return acc


def fmaf(a, b, c):
    return np.float32(float(a) * float(b) + float(c))


def unit():
    S = [np.float32(v) for v in fit(sin_tail, 3, 1e-12, 1.0001)]
    Cc = [np.float32(v) for v in fit(cos_tail, 3, 1e-12, 1.0001)]
    es = ec = 0.0
    for x in np.linspace(1e-4, 1.0, 200001, dtype=np.float32):
        u = np.float32(x * x)
        t = np.float32(u * x)
        ps = fmaf(fmaf(fmaf(S[3], u, S[2]), u, S[1]), u, S[0])
        s = fmaf(t, ps, x)
        pc = fmaf(fmaf(fmaf(Cc[3], u, Cc[2]), u, Cc[1]), u, Cc[0])
        c = fmaf(u, pc, 1.0)
        ulp_s = float(np.spacing(np.float32(math.sin(float(x)))))
        ulp_c = float(np.spacing(np.float32(math.cos(float(x)))))
        es = max(es, abs(float(s) - math.sin(float(x))) / ulp_s)
        ec = max(ec, abs(float(c) - math.cos(float(x))) / ulp_c)
    print(f"sin (float, |x| <= 1): max error {es:.3f} ulp")
    for v in S:
        print(f"  {float(v).hex()}f  ({float(v)!r})")
    print(f"cos (float, |x| <= 1): max error {ec:.3f} ulp")
    for v in Cc:
        print(f"  {float(v).hex()}f  ({float(v)!r})")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--unit":
        unit()
        return
    ns = int(sys.argv[1]) if len(sys.argv) > 1 else 5
    nc = int(sys.argv[2]) if len(sys.argv) > 2 else 6
    hi = (math.pi / 4) ** 2 * 1.0001
    lo = 1e-12
    S = [float(v) for v in fit(sin_tail, ns - 1, lo, hi)]
    Cc = [float(v) for v in fit(cos_tail, nc - 1, lo, hi)]
    es = ec = 0.0
    for r in np.linspace(1e-6, math.pi / 4, 200001):
        r = float(r)
        u = r * r
        s = r + r * u * horner(S, u)
        c = 1 + u * horner(Cc, u)
        es = max(es, abs(s - math.sin(r)) / math.sin(r))
        ec = max(ec, abs(c - math.cos(r)) / math.cos(r))
    print(f"sin: {ns} coefficients, max rel err {es:.3e}")
    for v in S:
        print(f"  {v!r}")
    print(f"cos: {nc} coefficients, max rel err {ec:.3e}")
    for v in Cc:
        print(f"  {v!r}")


if __name__ == "__main__":
    main()
