"""The polynomial fits inside the fused likelihood kernels: the literals in
qcbc_kernels.hip are the generators' fits (gen_sincos.py, gen_atan.py beside
it), and the kernels' evaluation orders, restated here in double precision,
stay within a few ulp of mpmath (sincos of a phase in turns, the
merger-ringdown exp and atan)."""
import importlib.util
import os
import re
from fractions import Fraction

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quickcbc_amd", "csrc")
SRC = open(os.path.join(CSRC, "qcbc_kernels.hip")).read()


def _load(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(CSRC, name + ".py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _body(fn):
    i = SRC.index("__device__ __forceinline__ " + fn)
    return SRC[i:SRC.index("\n}\n", i)]


def _hex_literals(text):
    return [float.fromhex(h) for h in re.findall(r"-?0x1\.[0-9a-f]+p[+-]\d+", text)]


def _kernel_consts(fn):
    """The literals of fn in source order, with the sign each one enters with
    (a literal written -sreg(c) is added as -c)."""
    out = []
    for m in re.finditer(r"(-?)sreg\((-?0x1\.[0-9a-f]+p[+-]\d+)\)", _body(fn)):
        v = float.fromhex(m.group(2))
        out.append(-v if m.group(1) else v)
    return out


def test_sincos_literals_are_the_fit():
    g = _load("gen_sincos")
    k = _kernel_consts("void sincos_turns(double r, double* s, double* c, const SinCosK& k)")
    body = SRC[SRC.index("__device__ __forceinline__ SinCosK sincos_consts()"):]
    p1, q1 = _hex_literals(body[:body.index("}")])[:2]
    # P: z^6 then z^5 (in SinCosK) then z^4..z^0; Q likewise, its z^0 term is 1
    P = [k[5], k[4], k[3], k[2], k[1], p1, k[0]]
    Q = [1.0, k[10], k[9], k[8], k[7], q1, k[6]]
    assert P == list(g.P)
    assert Q == list(g.Q)


def _sincos_kernel(r, P, Q):
    t = r - np.rint(r)
    z = t * t
    p = P[6]
    for c in P[5::-1]:
        p = p * z + c
    q = Q[6]
    for c in Q[5::-1]:
        q = q * z + c
    sn, cs = t * p, q
    for _ in range(2):
        sn, cs = 2.0 * sn * cs, cs * cs - sn * sn
    return sn, cs


def test_sincos_turns_accuracy():
    g = _load("gen_sincos")
    rng = np.random.default_rng(4)
    worst = 0.0
    for r in np.concatenate([rng.uniform(-0.5, 0.5, 1500), rng.uniform(-3e4, 3e4, 1500)]):
        s, c = _sincos_kernel(float(r), g.P, g.Q)
        t = mp.mpf(float(r)) - mp.nint(mp.mpf(float(r)))
        rs, rc = mp.sin(2 * mp.pi * t), mp.cos(2 * mp.pi * t)
        worst = max(worst, float(abs(s - rs)), float(abs(c - rc)))
    assert worst < 8 * 2.0 ** -53, worst  # a few ulp of 1 (the results are in [-1, 1])


def test_sincos_amp_literals_are_the_fit():
    g = _load("gen_sincos")
    k = _kernel_consts("void sincos_amp_turns(double r, double a, double* as, double* ac, const SinCosK& k)")
    body = SRC[SRC.index("__device__ __forceinline__ SinCosK sincos_consts_amp()"):]
    p1, q1, p2, p3, q2, q3, q4 = _hex_literals(body[:body.index("}")])[:7]
    # the literals in source order: P's z^6, z^2, z^1, z^0, Q's z^6, z^1, z^0,
    # then sqrt(2) (subtracted); z^5..z^3 (z^5..z^2 for Q) come from SinCosK
    assert [k[0], p1, p2, p3, k[1], k[2], k[3]] == list(g.PS)[::-1]
    assert [k[4], q1, q2, q3, q4, k[5], k[6]] == list(g.QS)[::-1]
    assert k[7] == -g.SQRT2
    # the scaled fits are the fits times 2^(3/4), to a rounding each
    for a, b in zip(g.PS + g.QS, g.P + g.Q):
        assert abs(a - b * 2 ** 0.75) <= 2 * np.spacing(abs(a))


def _fma(x, y, z):
    """x y + z rounded once, as the GPU's FMA (Fraction to float rounds correctly)."""
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def _sincos_amp_kernel(r, a, PS, QS, sqrt2):
    """sincos_amp_turns' operations in its order, each FMA rounded once."""
    t = r - float(np.rint(r))
    z = t * t
    p = _fma(z, PS[6], PS[5])
    for c in PS[4::-1]:
        p = _fma(p, z, c)
    q = _fma(z, QS[6], QS[5])
    for c in QS[4::-1]:
        q = _fma(q, z, c)
    sn = t * p
    s1, c1 = sn * q, _fma(q, q, -sqrt2)
    pa = a * c1
    return pa * s1, _fma(pa, c1, -a)


def test_sincos_amp_turns_accuracy():
    """a sin and a cos of 2 pi r within 24 ulp of a (the results are in [-a, a]).

    Each double-angle step doubles the error it is handed (the cosine's most:
    d(2 c^2 - 1) = 4 c dc), so the full angle carries ~10 ulp (sine) and ~20
    ulp (cosine) of the quarter angle's 1-2 ulp; the form with doubled
    coefficients and two plain double-angle steps measures 10.3 and 21.5 ulp
    under the same emulation, this one 9.9 and 19.9.  (h|h) does not see it
    (a^2 is formed apart); (d|h) sees it as a relative error of ~2e-15 per bin."""
    g = _load("gen_sincos")
    rng = np.random.default_rng(5)
    worst = 0.0
    for r, a in zip(np.concatenate([rng.uniform(-0.5, 0.5, 1500), rng.uniform(-3e4, 3e4, 1500)]),
                    10.0 ** rng.uniform(-25, 5, 3000)):
        s, c = _sincos_amp_kernel(float(r), float(a), g.PS, g.QS, g.SQRT2)
        t = mp.mpf(float(r)) - mp.nint(mp.mpf(float(r)))
        rs, rc = a * mp.sin(2 * mp.pi * t), a * mp.cos(2 * mp.pi * t)
        worst = max(worst, float(abs(s - rs)) / a, float(abs(c - rc)) / a)
    assert worst < 24 * 2.0 ** -53, worst


def test_atan_literals_are_the_fit():
    g = _load("gen_atan")
    k = _kernel_consts("double atan_mrd(double x)")
    # sqrt3, tan(pi/12), then R from z^9 down to z^1, then pi/6, pi/2
    assert k[0] == g.SQRT3 and k[1] == g.TAN12
    assert k[2:11] == [g.R[9], g.R[8], g.R[7], g.R[6], g.R[5], g.R[4], g.R[3], g.R[2], g.R[1]]
    assert g.R[0] == 1.0
    assert k[11] == g.PI6 and k[12] == g.PI2


def test_atan_accuracy():
    g = _load("gen_atan")
    worst = 0.0
    for x in np.concatenate([np.linspace(-30, 30, 3001), [1e-8, -1.0, 1.0, 0.26794919, 3.7320508]]):
        got = g.atan_kernel(float(x))
        ref = mp.atan(mp.mpf(float(x)))
        if ref != 0:
            worst = max(worst, float(abs(got - ref)) / np.spacing(abs(float(ref))))
    assert worst < 3.0, worst


def _exp_kernel(x, c):
    """exp_mrd with its literals as they enter: c[0] = 1/ln2, c[1] = -ln2_hi,
    c[2] = ln2_hi - ln2, c[3:] = 1/13! .. 1/3!."""
    k = np.rint(x * c[0])
    r = x + k * c[1]        # exact: k ln2_hi has at most 53 significant bits
    r = r + k * c[2]
    p = r * c[3] + c[4]
    for a in c[5:]:
        p = p * r + a
    p = p * r + 0.5
    p = p * r + 1.0
    p = p * r + 1.0
    return np.ldexp(p, int(k))


def test_exp_mrd_accuracy():
    c = _kernel_consts("double exp_mrd(double x)")
    assert len(c) == 14
    assert c[3:] == [float(1 / mp.factorial(n)) for n in range(13, 2, -1)]
    assert -(c[1] + c[2]) == pytest.approx(float(mp.log(2)), rel=1e-16)
    assert c[0] == float(1 / mp.log(2))
    worst = 0.0
    for x in np.linspace(-60.0, 10.0, 3001):
        got = _exp_kernel(float(x), c)
        ref = mp.exp(mp.mpf(float(x)))
        worst = max(worst, float(abs(got - ref)) / np.spacing(float(ref)))
    assert worst < 2.0, worst
