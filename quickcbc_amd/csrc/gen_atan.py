#!/usr/bin/env python3
"""Polynomial coefficients of atan(t) / t on |t| <= tan(pi/12), and the error
of the merger-ringdown atan the fused likelihood kernels evaluate with them.

The kernels reduce x to t in [0, tan(pi/12)]:

    a = |x|;  t1 = 1/a if a > 1 else a             (atan a = pi/2 - atan(1/a))
    t2 = (sqrt3 t1 - 1) / (t1 + sqrt3) if t1 > tan(pi/12) else t1
                                                   (atan t1 = pi/6 + atan t2)
    atan(t2) ~ t2 * R(t2^2),  R of degree DEG

and undo the reductions.  The fit is a Chebyshev (near-minimax) fit in
z = t^2 at 40 digits (mpmath.chebyfit), rounded to double.  The script prints
the coefficients as C++ literals and the worst error of the whole evaluation
in double precision (numpy, the kernels' operation order, an exactly rounded
reciprocal standing in for the kernels' Newton-refined one) against mpmath on
a grid of x in [-100, 100], in ulps of the result.

    python gen_atan.py
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40
T = mp.tan(mp.pi / 12)
DEG = 9


def fit():
    c = mp.chebyfit(lambda z: mp.atan(mp.sqrt(z)) / mp.sqrt(z) if z != 0 else mp.mpf(1), [0, T * T], DEG + 1)
    return [float(c[DEG - i]) for i in range(DEG + 1)]  # ascending powers of z


R = fit()
SQRT3 = float(mp.sqrt(3))
TAN12 = float(T)
PI6 = float(mp.pi / 6)
PI2 = float(mp.pi / 2)


def atan_kernel(x):
    a = abs(x)
    inv = a > 1.0
    t1 = 1.0 / a if inv else a
    big = t1 > TAN12
    t2 = (SQRT3 * t1 - 1.0) / (t1 + SQRT3) if big else t1
    z = t2 * t2
    p = R[DEG]
    for k in range(DEG - 1, -1, -1):
        p = p * z + R[k]
    r = t2 * p
    if big:
        r = r + PI6
    if inv:
        r = PI2 - r
    return r if x >= 0 else -r


def ulp_error():
    worst = (0.0, None)
    for x in np.concatenate([np.linspace(-100, 100, 200001), np.geomspace(1e-6, 1e3, 20001)]):
        got = atan_kernel(float(x))
        ref = mp.atan(mp.mpf(float(x)))
        if ref != 0:
            e = float(abs(got - ref)) / np.spacing(abs(float(ref)))
            if e > worst[0]:
                worst = (e, float(x))
    return worst


if __name__ == "__main__":
    for k, c in enumerate(R):
        print("  z^%d: %s" % (k, float(c).hex()))
    print("sqrt3 %s  tan(pi/12) %s  pi/6 %s  pi/2 %s" % (SQRT3.hex(), TAN12.hex(), PI6.hex(), PI2.hex()))
    print("worst error %.3f ulp at x = %r" % ulp_error())
