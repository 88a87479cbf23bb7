"""The heterodyned Fisher matrix and its jump directions on the GPU against the oracle.

Fisher_Het (QuickCBC.c:5446-5617) on a heterodyne imported from the oracle's
set-up arrays, so that both sides weigh the same samples; FisherEvec
(QuickCBC.c:2449-2546) of the oracle's matrices; the drop-ins Fisher_Het and
FisherEvec through the C ABI; edge cases (spin limit, points outside the
model, ragged and empty batches, smaller matrices).

Tolerances.  Fisher_Het differentiates fullphaseamp's phase (up to ~1e5 rad,
an ulp ~1.5e-11) by central differences with epsilon 1e-5, so a one-ulp
change of a phase sample moves a derivative by ~1e-6 rad per unit: the
entries are compared as |dF_ij| <= tol sqrt(F_ii F_jj), tol 1e-7 (measured
at most 1.2e-8, median 2e-11 over 64 points of each configuration; the GPU's
sample powers come from planes, within a few ulps of the oracle's pow and
log, and its per-sample weights regroup the reference's region sums).
FisherEvec runs the same rotations in the same order without contraction
(measured: identical to the oracle's bits); eigenvalues to 1e-13 relative,
eigenvectors to 1e-12.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu

FISH_TOL = 1e-7


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    which = request.param
    an, inj, D, SN = _data.analysis(which)
    h = _data.het(which)
    c = Q.Context(an, D, SN)
    g = Q.Het(c, arrays=h)
    yield which, an, inj, h, c, g
    g.close()
    c.close()


def _points(inj, n, seed):
    return np.vstack([inj, W.ensemble_around(inj, n - 1, seed=seed, scale=0.3)])


def _oracle_fisher(an, h, P):
    return O.fisher_het(an.labels, an.gmst, an.fref, an.Tobs, h, P)


def _rel(F, Fo):
    s = np.sqrt(np.abs(np.einsum("bii->bi", Fo)))
    return np.abs(F - Fo) / (s[:, :, None] * s[:, None, :])


def test_fisher_matches_the_oracle(setup):
    which, an, inj, h, c, g = setup
    P = _points(inj, 24, seed=31)
    print("fisher: config %d, %d points, M=%d" % (which, P.shape[0], h["M"]), flush=True)
    F, _, _, st = g.fisher(P, evec=False)
    Fo, so = _oracle_fisher(an, h, P)
    assert (st == 0).all() and (so == 0).all()
    assert np.array_equal(F, np.transpose(F, (0, 2, 1)))
    assert (F[:, 4, 6] == 0).all()
    rel = _rel(F, Fo)
    assert rel.max() <= FISH_TOL, rel.max()


def test_evec_matches_the_oracle(setup):
    which, an, inj, h, c, g = setup
    P = _points(inj, 24, seed=32)
    Fo, _ = _oracle_fisher(an, h, P)
    ej, ev = c.fisher_evec(Fo)
    ejo, evo, _ = O.fisher_evec(Fo)
    assert np.allclose(ej, ejo, rtol=1e-13, atol=0)
    assert np.abs(ev - evo).max() <= 1e-12
    # Fisher and jump directions in one call: the same eigen kernel on its own matrices
    F, ej2, ev2, st = g.fisher(P, evec=True)
    ej3, ev3 = c.fisher_evec(F)
    assert np.array_equal(ej2, ej3) and np.array_equal(ev2, ev3)


@pytest.mark.parametrize("B", [1, 2, 7, 65])
def test_ragged_batches(setup, B):
    which, an, inj, h, c, g = setup
    P = W.ensemble_around(inj, B, seed=40 + B, scale=0.3)
    F, ej, ev, st = g.fisher(P)
    Fa, eja, eva, sta = g.fisher(P[:1])
    # a point's matrix does not depend on the rest of its batch
    assert np.array_equal(F[0], Fa[0]) and np.array_equal(ej[0], eja[0]) and np.array_equal(ev[0], eva[0])
    Fo, _ = _oracle_fisher(an, h, P)
    assert _rel(F, Fo).max() <= FISH_TOL


def test_spin_limit_domain_errors_and_empty_batch(setup):
    which, an, inj, h, c, g = setup
    P = np.vstack([inj] * 3)
    P[0, 2] = 0.99995                              # |chi1| >= 0.9999: 1e4 I, nothing evaluated
    P[1, 1] = np.log(3000.0 * W.MSUN_SI)           # fCut below the grid: outside the model
    F, ej, ev, st = g.fisher(P)
    Fo, so = _oracle_fisher(an, h, P)
    assert list(st) == [0, 33, 0] and list(so) == [0, 33, 0]
    assert np.array_equal(F[0], 1e4 * np.eye(11)) and np.array_equal(F[1], 1e4 * np.eye(11))
    assert np.allclose(ej[0], 1e-2) and np.array_equal(ev[0], np.eye(11))
    assert _rel(F[2:], Fo[2:]).max() <= FISH_TOL
    F0, ej0, ev0, st0 = g.fisher(np.zeros((0, 11)))
    assert F0.shape == (0, 11, 11) and st0.size == 0


@pytest.mark.parametrize("d", [1, 3, 7])
def test_evec_of_smaller_matrices(setup, d):
    # FisherEvec is called with NP = 11, NX = 7 and NS = 7 (ConstCBC.h:67-69)
    which, an, inj, h, c, g = setup
    r = np.random.default_rng(d)
    Q_, _ = np.linalg.qr(r.standard_normal((d, d)))
    lam = np.exp(r.uniform(0, 25, (5, d)))
    A = np.einsum("ij,bj,kj->bik", Q_, lam, Q_)
    A = 0.5 * (A + np.transpose(A, (0, 2, 1)))
    ej, ev = c.fisher_evec(A)
    ejo, evo, _ = O.fisher_evec(A)
    assert np.allclose(ej, ejo, rtol=1e-13, atol=0)
    assert np.abs(ev - evo).max() <= 1e-12


def test_dropins_fisher_het_and_fisher_evec(setup):
    which, an, inj, h, c, g = setup
    # the drop-in heterodyne's struct, then Fisher_Het on it, as MCMC_all calls them (QuickCBC.c:1918-1926)
    v = Q.HetView.from_arrays(h, an.Tobs, an.nifo)
    p = W.ensemble_around(inj, 1, seed=7, scale=0.3)[0]
    F = Q.dropin_fisher_het(an, v, p)
    Fo, _ = _oracle_fisher(an, h, p[None, :])
    assert _rel(F[None], Fo).max() <= FISH_TOL
    ej, ev = Q.dropin_fisher_evec(Fo[0])
    ejo, evo, _ = O.fisher_evec(Fo)
    assert np.allclose(ej, ejo[0], rtol=1e-13, atol=0) and np.abs(ev - evo[0]).max() <= 1e-12
