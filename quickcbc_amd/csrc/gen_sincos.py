#!/usr/bin/env python3
"""Polynomial coefficients of sin(pi t / 2) and cos(pi t / 2) on |t| <= 1/2.

The fused likelihood kernels reduce a phase r (in turns) to t = r - rint(r),
evaluate the quarter angle by

    sin(pi t / 2) ~ t * P(t^2),  P of degree 6
    cos(pi t / 2) ~ Q(t^2),      Q of degree 6

and square (c + i s) twice.  The fits are Chebyshev (near-minimax) fits in
z = t^2 at 40 digits (mpmath.chebyfit), rounded to double.  The script prints
the coefficients as C++ initialisers and the worst error of the quarter-angle
evaluation in double precision, summed in the power basis as the kernels do,
measured against mpmath on a dense grid, in units of the result's ulp.

    python gen_sincos.py
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40
T = mp.mpf(1) / 2
DEG = 6


def fit(fun, scale=1):
    c = mp.chebyfit(fun, [0, T * T], DEG + 1)
    return [float(scale * c[DEG - i]) for i in range(DEG + 1)]  # ascending powers of z


def _sin_q(z):
    return mp.sin(mp.pi / 2 * mp.sqrt(z)) / mp.sqrt(z) if z != 0 else mp.pi / 2


def _cos_q(z):
    return mp.cos(mp.pi / 2 * mp.sqrt(z))


P = fit(_sin_q)
Q = fit(_cos_q)

# sincos_amp_turns: the same fits times 2^(3/4), each rounded once, so that one
# double-angle step leaves sin and cos of the half angle times sqrt(2), whose
# last double-angle step folds with the amplitude:
#   a sin 2y = P (sqrt2 sin y), a cos 2y = P (sqrt2 cos y) - a, P = a (sqrt2 cos y).
SCALE34 = mp.mpf(2) ** (mp.mpf(3) / 4)
PS = fit(_sin_q, SCALE34)
QS = fit(_cos_q, SCALE34)
SQRT2 = float(mp.sqrt(2))


def power_basis(c, z):
    zk = [np.ones_like(z), z]
    for k in range(2, DEG + 1):
        zk.append(zk[-1] * z)
    acc = c[DEG] * zk[DEG]
    for k in range(DEG - 1, 0, -1):
        acc = acc + c[k] * zk[k]
    return acc + c[0]


def ulp_error():
    t = np.linspace(-0.5, 0.5, 20001)
    z = t * t
    s = t * power_basis(P, z)
    c = power_basis(Q, z)
    es = ec = 0.0
    for ti, si, ci in zip(t, s, c):
        a = mp.pi / 2 * mp.mpf(float(ti))
        rs, rc = mp.sin(a), mp.cos(a)
        if rs != 0:
            es = max(es, float(abs(si - rs)) / np.spacing(abs(float(rs))))
        ec = max(ec, float(abs(ci - rc)) / np.spacing(abs(float(rc))))
    return es, ec


if __name__ == "__main__":
    fmt = lambda v: ", ".join(float.hex(x) for x in v)
    print("// sin(pi t / 2) = t * P(t^2): P ascending")
    print("{" + fmt(P) + "}")
    print("// cos(pi t / 2) = Q(t^2): Q ascending")
    print("{" + fmt(Q) + "}")
    es, ec = ulp_error()
    print("// worst error on |t| <= 1/2: sin %.2f ulp, cos %.2f ulp" % (es, ec))
    print("// times 2^(3/4), for sincos_amp_turns: P, Q ascending, then sqrt(2)")
    print("{" + fmt(PS) + "}")
    print("{" + fmt(QS) + "}")
    print(float.hex(SQRT2))
