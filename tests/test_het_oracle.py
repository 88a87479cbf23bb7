"""Pins of the oracle's heterodyned likelihood (CPU only).

The oracle restates fullphaseamp (QuickCBC.c:3780-3869), legendre_maker
(:5055-5094), Inverse (Utilities.c:1862-1903), heterodyne (:5112-5299) and
log_likelihood_het (:5303-5444).  Each is checked here against something
that is not its own formula retyped:

- legendre_maker: the closed form of the discrete Chebyshev (Gram)
  polynomials, P_n(k) = sum_j (-1)^j C(n,j) C(n+j,j) k^(j) / U^(j) with
  falling factorials, and their orthogonality on the U + 1 points;
- Inverse: numpy's LAPACK inverse (transposed, as the reference stores it);
- fullphaseamp: fulltemplates (pinned by the full-likelihood tests) is
  A cos(phase), A sin(phase) of the same amplitude and phase above fbegin;
- heterodyne: logLb is the full likelihood of the reference point over
  [MN, MM); the moments are sums of the per-bin products against the closed
  form polynomials;
- log_likelihood_het: logL(reference) = logLb exactly; a brute-force form
  that interpolates each region's samples by a polynomial through its J + 1
  points and sums it against every bin (no Legendre basis, no inverse) gives
  the same logL; and near the reference the heterodyned logL approximates the
  full logL to ~1e-2 (the method's purpose).
"""
from math import comb

import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import workloads as W
from tests import _data


def gram(n, U, k):
    k = np.asarray(k, np.float64)
    out = np.zeros_like(k)
    for j in range(n + 1):
        num = np.ones_like(k)
        den = 1.0
        for i in range(j):
            num = num * (k - i)
            den *= (U - i)
        out += (-1) ** j * comb(n, j) * comb(n + j, j) * num / den
    return out


@pytest.mark.parametrize("J,U", [(1, 1), (2, 2), (2, 7), (3, 12), (5, 40), (2, 400)])
def test_legendre_closed_form_and_orthogonality(J, U):
    P = O.legendre(J, U)
    k = np.arange(U + 1)
    for n in range(J + 1):
        assert np.allclose(P[n], gram(n, U, k), rtol=0, atol=1e-12), n
        assert P[n][0] == 1.0 and abs(P[n][U] - (-1) ** n) < 1e-12
    G = P @ P.T
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() <= 1e-11 * (U + 1)


def test_inverse_matches_lapack_transposed():
    r = np.random.default_rng(3)
    cases = [r.standard_normal((3, 3)) + 3 * np.eye(3), r.standard_normal((5, 5)),
             np.array([[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [4.0, -3.0, 8.0]]),  # needs a row swap
             O.legendre(2, 10)[:, [0, 3, 10]]]
    for M in cases:
        IM = O.inverse(M)
        ref = np.linalg.inv(M).T
        assert np.allclose(IM, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
        assert np.allclose(IM.T @ M, np.eye(M.shape[0]), atol=1e-12)


@pytest.mark.parametrize("which", [0, 1])
def test_fullphaseamp_is_the_template_above_fbegin(which):
    an, inj, D, SN = _data.analysis(which)
    st, A, Ph = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, an.freq(), inj)
    assert st == 0
    st, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj)
    assert st == 0
    f = an.freq()
    fs = O.fbegin(inj)
    i = np.arange(1, an.N // 2)
    on = i[f[i] > fs]
    scale = np.abs(A).max()
    assert np.abs(A[:, on] * np.cos(Ph[:, on]) - h[:, on]).max() <= 1e-12 * scale
    assert np.abs(A[:, on] * np.sin(Ph[:, on]) - h[:, an.N - on]).max() <= 1e-12 * scale
    # below fbegin the amplitude is tapered, not cut: positive, and rising to the untapered value at fs
    below = i[(f[i] < fs) & (A[0, i] > 0)]
    if below.size:
        assert (A[:, below] >= 0).all()
    # a frequency list is evaluated point by point: a subset gives the same numbers
    sub = np.sort(np.random.default_rng(1).choice(np.arange(1, an.N // 2), 50, replace=False))
    st, A2, P2 = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f[sub], inj)
    assert st == 0
    assert np.array_equal(A2, A[:, sub]) and np.array_equal(P2, Ph[:, sub])


def _per_bin(an, ref, D, SN):
    """RC, RS, AA of every bin at the reference point (the heterodyne's products)."""
    st, Ab, Pb = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, an.freq(), ref)
    st2, hb = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, ref)
    assert st == 0 and st2 == 0
    N, n2 = an.N, an.N // 2
    rb = D - hb
    j = np.arange(1, n2)
    RC = np.zeros((an.nifo, n2))
    RS = np.zeros_like(RC)
    AA = np.zeros_like(RC)
    cp, sp = np.cos(Pb[:, j]), np.sin(Pb[:, j])
    x = Ab[:, j] / SN[:, j]
    RC[:, j] = 2 * (rb[:, j] * cp + rb[:, N - j] * sp) * x
    RS[:, j] = 2 * (-rb[:, j] * sp + rb[:, N - j] * cp) * x
    AA[:, j] = Ab[:, j] ** 2 / SN[:, j]
    return RC, RS, AA


@pytest.mark.parametrize("which", [0, 1])
def test_heterodyne_set_up(which):
    an, inj, D, SN = _data.analysis(which)
    h = _data.het(which)
    M, J, NR, g = h["M"], h["J"], h["NR"], h["fgrid"]
    assert M == NR * J + 1
    # logLb: the full likelihood of the reference point over [MN, MM)
    Lf, _, _ = O.loglike_batch(an.labels, an.gmst, an.fref, h["MN"] / an.Tobs, h["MM"] / an.Tobs, D, SN,
                               an.N, an.Tobs, inj[None, :])
    assert abs(h["logLb"] - Lf[0]) <= 1e-9 * abs(Lf[0])
    # reference amplitude and phase on the grid
    st, A, Ph = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, h["freq"], inj)
    assert np.array_equal(A, h["amp"]) and np.array_equal(Ph, h["phase"])
    RC, RS, AA = _per_bin(an, inj, D, SN)
    starts = g[np.arange(NR) * J]
    for arr, ref in ((h["rc"], RC), (h["rs"], RS), (h["aa"], AA)):
        assert np.allclose(arr, ref[:, starts], rtol=1e-12, atol=0)
    for i in (0, 1, NR // 2, NR - 1):
        b0, U = g[i * J], g[(i + 1) * J] - g[i * J]
        n = np.arange(U + 1)
        Pc = np.stack([gram(q, U, n) for q in range(J + 1)])
        PP = Pc[:, g[i * J:i * J + J + 1] - b0]
        assert np.allclose(h["IP"][i].T @ PP, np.eye(J + 1), atol=1e-10)
        for arr, ref in ((h["SL"], AA), (h["lc"], RC), (h["ls"], RS)):
            want = ref[:, b0:b0 + U + 1] @ Pc.T  # nifo x (J+1)
            tol = 1e-11 * (np.abs(ref[:, b0:b0 + U + 1]).sum(1, keepdims=True) + 1e-300)
            assert (np.abs(arr[i] - want) <= tol).all(), (i, arr[i], want)


def brute_force_het(an, h, RC, RS, AA, p):
    """log_likelihood_het by polynomial interpolation of each region's samples."""
    M, J, NR, g = h["M"], h["J"], h["NR"], h["fgrid"]
    st, A, Ph = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, h["freq"], p)
    assert st == 0
    ha, hp = h["amp"], h["phase"]
    hc = 2 * (ha - A * np.cos(hp - Ph)) / ha
    hs = 2 * A * np.sin(hp - Ph) / ha
    cc = hc ** 2 + hs ** 2
    logL = h["logLb"]
    for d in range(an.nifo):
        HH = HR = 0.0
        for i in range(NR):
            s = np.arange(i * J, i * J + J + 1)
            b0, U = g[i * J], g[(i + 1) * J] - g[i * J]
            t = (g[s] - b0) / U
            V = np.vander(t, J + 1, increasing=True)
            n = np.arange(U + 1)
            Vn = np.vander(n / U, J + 1, increasing=True)
            seg = slice(b0, b0 + U + 1)
            HH += AA[d, seg] @ (Vn @ np.linalg.solve(V, cc[d, s]))
            HR += RC[d, seg] @ (Vn @ np.linalg.solve(V, hc[d, s]))
            HR += RS[d, seg] @ (Vn @ np.linalg.solve(V, hs[d, s]))
        # bins shared by two regions are counted twice; the reference takes
        # the second count off at region starts 1 .. NR-2
        for i in range(1, NR - 1):
            m = i * J
            HH -= h["aa"][d, i] * cc[d, m]
            HR -= h["rc"][d, i] * hc[d, m] + h["rs"][d, i] * hs[d, m]
        logL -= HR + 0.5 * HH
    return logL


@pytest.mark.parametrize("which", [0, 1])
def test_loglike_het_brute_force(which):
    an, inj, D, SN = _data.analysis(which)
    h = _data.het(which)
    RC, RS, AA = _per_bin(an, inj, D, SN)
    P = W.ensemble_around(inj, 6, seed=17, scale=0.3)
    L, st = _data.oracle_loglike_het(which, P)
    assert (st == 0).all()
    for b in range(P.shape[0]):
        ref = brute_force_het(an, h, RC, RS, AA, P[b])
        assert abs(L[b] - ref) <= 1e-8 + 1e-9 * abs(h["logLb"] - ref), (b, L[b], ref)


@pytest.mark.parametrize("which", [0, 1])
def test_loglike_het_at_the_reference_is_logLb(which):
    an, inj, D, SN = _data.analysis(which)
    L, st = _data.oracle_loglike_het(which, inj[None, :])
    assert st[0] == 0 and L[0] == _data.het(which)["logLb"]


def test_loglike_het_approximates_the_full_likelihood():
    # config 0 (fbegin below the band, so the taper and the f > fs cut agree)
    an, inj, D, SN = _data.analysis(0)
    h = _data.het(0)
    P = W.ensemble_around(inj, 64, seed=3, scale=0.3)
    L, st = _data.oracle_loglike_het(0, P)
    Lf, _, _ = O.loglike_batch(an.labels, an.gmst, an.fref, h["MN"] / an.Tobs, h["MM"] / an.Tobs, D, SN,
                               an.N, an.Tobs, P)
    assert (st == 0).all()
    assert np.ptp(Lf) > 100.0  # the points span a wide range of logL
    assert np.abs(L - Lf).max() < 0.05
