"""Pins of the oracle's Fisher matrix and eigen-decomposition (CPU only).

The oracle restates Fisher_Het (QuickCBC.c:5446-5617) and FisherEvec
(QuickCBC.c:2449-2546).  FisherEvec's GSL eigen_symmv is replaced by the
cyclic Jacobi method, so it is pinned against LAPACK (numpy.linalg.eigh) and
against the eigen-equation itself.  Fisher_Het is pinned by what the
mathematics fixes:

- the GW phase params[4] enters the phase additively and the amplitude not at
  all, and ln DL (params[6]) scales the amplitude and leaves the phase alone:
  F[4][6] = 0 exactly, and at the heterodyne's reference point F[4][4] and
  F[6][6] both equal sum_d (h_d|h_d) over the heterodyne's bins (the Legendre
  expansion of a constant is exact);
- a brute-force Fisher matrix that evaluates the integrand
  4 (dA_i dA_j + A^2 dphi_i dphi_j) / S at every bin of every region (no
  Legendre basis, no inverse, no interpolation; the region boundaries counted
  as the reference counts them) agrees to the heterodyne's interpolation
  error.  A transposed IP or a shifted overcount index moves entries by far
  more than that.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import workloads as W
from tests import _data


def _sign_fix(V):
    """Rows scaled so each one's largest-magnitude component is positive."""
    V = V.copy()
    for i in range(V.shape[0]):
        k = int(np.argmax(np.abs(V[i])))
        if V[i, k] < 0:
            V[i] = -V[i]
    return V


def _graded_spd(r, d, lo, hi):
    Q, _ = np.linalg.qr(r.standard_normal((d, d)))
    lam = np.exp(r.uniform(np.log(lo), np.log(hi), d))
    A = (Q * lam) @ Q.T
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("seed", range(6))
def test_evec_matches_lapack(seed):
    r = np.random.default_rng(seed)
    A = _graded_spd(r, 11, 20.0, 10.0 ** r.uniform(4, 14))
    ej, ev, sw = O.fisher_evec(A)
    ej, ev = ej[0], ev[0]
    assert 1 <= sw[0] < 60
    lam = 1.0 / ej ** 2  # every eigenvalue is above the floor of 10
    w, V = np.linalg.eigh(A)
    o = np.argsort(np.abs(w), kind="stable")
    w, V = w[o], V[:, o].T
    lmax = np.abs(w).max()
    assert np.allclose(lam, w, rtol=1e-12, atol=1e-13 * lmax)
    # eigenvectors: orthonormal, A v = lambda v, and LAPACK's up to sign
    assert np.abs(ev @ ev.T - np.eye(11)).max() < 1e-13
    res = np.abs(ev @ A - lam[:, None] * ev).max()
    assert res <= 1e-13 * lmax
    gap = np.min(np.abs(np.diff(w)))
    assert np.abs(ev - _sign_fix(V)).max() <= 1e-13 * lmax / gap + 1e-14
    for i in range(11):
        assert ev[i, np.argmax(np.abs(ev[i]))] > 0


def test_evec_edge_cases():
    # a diagonal entry below 1e-16 in magnitude: 1e4 jumps along the axes
    A = np.diag(np.arange(1.0, 12.0))
    A[3, 3] = 1e-17
    ej, ev, _ = O.fisher_evec(A)
    assert (ej[0] == 1e4).all() and np.array_equal(ev[0], np.eye(11))
    # diagonal: exact, ordered by |lambda|, the floor of 10 applied after the sort
    d = np.array([5.0, -300.0, 1e6, 40.0, -2.0, 1e3, 7e8, 11.0, 2e5, 9.0, 3e4])
    ej, ev, sw = O.fisher_evec(np.diag(d))
    o = np.argsort(np.abs(d), kind="stable")
    want = 1.0 / np.sqrt(np.maximum(d[o], 10.0))
    assert np.array_equal(ej[0], want) and sw[0] == 1
    assert np.array_equal(ev[0], np.eye(11)[o])
    # an indefinite matrix: order by |lambda|, negative eigenvalues floored
    r = np.random.default_rng(9)
    Q, _ = np.linalg.qr(r.standard_normal((11, 11)))
    lam = np.array([-1e3, 50.0, -7.0, 2e4, 300.0, -5e5, 1e6, 12.0, 80.0, -2e2, 9e3])
    A = (Q * lam) @ Q.T
    ej, ev, _ = O.fisher_evec(0.5 * (A + A.T))
    o = np.argsort(np.abs(lam))
    assert np.allclose(ej[0], 1.0 / np.sqrt(np.maximum(lam[o], 10.0)), rtol=1e-12)


def _fisher_args(which):
    an, inj, D, SN = _data.analysis(which)
    return an, inj, D, SN, _data.het(which)


def _fisher(which, P):
    an, inj, D, SN, h = _fisher_args(which)
    return O.fisher_het(an.labels, an.gmst, an.fref, an.Tobs, h, P)


def _brute_fisher(an, h, SN, p, eps=1e-5):
    """The Fisher integrand at every bin of every region, counted as the reference counts."""
    g, J, NR = h["fgrid"], h["J"], h["NR"]
    lo, hi = g[0], g[NR * J]
    f = np.arange(lo, hi + 1) / an.Tobs
    st, A, Ph = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f, p)
    assert st == 0
    dA = np.zeros((11,) + A.shape)
    dP = np.zeros_like(dA)
    for i in range(11):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        s1, A1, P1 = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f, pp)
        s2, A2, P2 = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f, pm)
        assert s1 == 0 and s2 == 0
        dA[i] = (A1 - A2) / (2 * eps)
        dP[i] = (P1 - P2) / (2 * eps)
    # each bin once, the region starts 1 .. NR-2 too; the start of region NR-1 twice
    wgt = 4.0 / SN[:, lo:hi + 1]
    wgt[:, g[(NR - 1) * J] - lo] *= 2.0
    F = np.einsum("idk,jdk->ij", dA * wgt, dA) + np.einsum("idk,jdk->ij", dP * (A ** 2 * wgt), dP)
    return F


@pytest.mark.parametrize("which", [0, 1])
def test_fisher_closed_forms_at_the_reference(which):
    an, inj, D, SN, h = _fisher_args(which)
    F, st = _fisher(which, inj[None, :])
    F = F[0]
    assert st[0] == 0
    assert np.array_equal(F, F.T)
    assert F[4, 6] == 0.0 and F[6, 4] == 0.0
    # sum_d (h_d|h_d) over the heterodyne's bins, as the reference counts them
    g, J, NR = h["fgrid"], h["J"], h["NR"]
    lo, hi = g[0], g[NR * J]
    f = np.arange(lo, hi + 1) / an.Tobs
    st, A, _ = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f, inj)
    wgt = 4.0 / SN[:, lo:hi + 1]
    wgt[:, g[(NR - 1) * J] - lo] *= 2.0
    hh = float((A ** 2 * wgt).sum())
    assert abs(F[4, 4] - hh) <= 1e-5 * hh
    assert abs(F[6, 6] - hh) <= 1e-6 * hh


@pytest.mark.parametrize("which", [0, 1])
def test_fisher_matches_brute_force(which):
    an, inj, D, SN, h = _fisher_args(which)
    P = np.vstack([inj, W.ensemble_around(inj, 3, seed=21, scale=0.3)])
    F, st = _fisher(which, P)
    assert (st == 0).all()
    for b in range(P.shape[0]):
        Fb = _brute_fisher(an, h, SN, P[b])
        s = np.sqrt(np.abs(np.diag(Fb)))
        rel = np.abs(F[b] - Fb) / np.outer(s, s)
        # measured <= 1.1e-5; a transposed IP gives 6e-3, a shifted overcount index 7e-4
        assert rel.max() < 5e-5, (b, rel.max(), np.unravel_index(np.argmax(rel), rel.shape))


def test_fisher_spin_limit_and_batch_independence():
    an, inj, D, SN, h = _fisher_args(0)
    p = inj.copy()
    p[3] = 0.9999
    F, st = _fisher(0, p[None, :])
    assert st[0] == 0 and np.array_equal(F[0], 1e4 * np.eye(11))
    P = W.ensemble_around(inj, 5, seed=4, scale=0.3)
    Fa, _ = _fisher(0, P)
    Fb, _ = _fisher(0, P[2:3])
    assert np.array_equal(Fa[2], Fb[0])
