"""GPU parity of the heterodyne's grid choice against the oracle: limits,
Fisher_All, diff, het_diff, efix (full-data and heterodyne distances) and
het_space, through the batch API and the drop-ins (QuickCBC.c:4275-5053,
2695-2817).

Tolerances: limits' band exact and ratio 1e-10 relative; Fisher_All entries
1e-7 sqrt(F_ii F_jj) (the phase finite differences turn one ulp of a
1e5-rad phase into ~1e-6 of a derivative, as for Fisher_Het); distances
1e-9 relative (the waveforms agree to 1e-12, the sums are regrouped);
efix's scaled jumps 1e-6 relative (the same search on distances equal to
rounding); het_space's grid identical.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def oracle_space(which):
    an, inj, D, SN = _data.analysis(which)
    lo, hi = W.prior_box(an)
    st, h = O.het_space(an.labels, an.gmst, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, inj, lo, hi)
    assert st == 0
    return h


@functools.lru_cache(maxsize=None)
def oracle_het(which):
    an, inj, D, SN = _data.analysis(which)
    sp = oracle_space(which)
    st, h = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, sp["fgrid"], 2, sp["MN"], sp["MM"])
    assert st == 0
    return h


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    which = request.param
    an, inj, D, SN = _data.analysis(which)
    print("grid choice: config %d, %d det, N %d" % (which, an.nifo, an.N), flush=True)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    c = Q.Context(an, D, SN)
    yield which, an, inj, D, SN, c
    c.close()
    Q.lib().qcbc_set_band(16.0, 1024.0)


def test_limits(setup):
    which, an, inj, D, SN, c = setup
    st, band, ratio = O.limits(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj)
    got, r = c.limits(inj)
    assert got == band
    assert np.allclose(r, ratio, rtol=1e-10, atol=0)
    got2, r2 = Q.dropin_limits(an, Q.RowMatrix(SN), inj)
    assert got2 == band and np.array_equal(r2, r)


def test_fisher_all(setup):
    which, an, inj, D, SN, c = setup
    P = np.vstack([inj, W.ensemble_around(inj, 5, seed=61, scale=0.5)])
    F, st = c.fisher_all(P)
    assert (st == 0).all()
    for b in range(P.shape[0]):
        s0, Fo = O.fisher_all(an.labels, an.gmst, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, P[b])
        assert s0 == 0
        scale = np.sqrt(np.outer(np.diag(Fo), np.diag(Fo)))
        err = np.abs(F[b] - Fo) / scale
        print("Fisher_All point %d: max error %.2e of sqrt(F_ii F_jj)" % (b, err.max()), flush=True)
        assert err.max() <= 1e-7, b
    Fd = Q.dropin_fisher_all(an, Q.RowMatrix(SN), inj)
    assert np.array_equal(Fd, F[0])
    # spins at the limit: 1e4 I, nothing evaluated (QuickCBC.c:2709)
    p = inj.copy()
    p[2] = 0.99995
    F1, st1 = c.fisher_all(p)
    assert np.array_equal(F1[0], 1.0e4 * np.eye(11)) and st1[0] == 0


def test_diff(setup):
    which, an, inj, D, SN, c = setup
    sp = oracle_space(which)
    lo, hi = W.prior_box(an)
    P = np.array([O.displace(inj, sp["evec"][i], sp["eval"][i], lo, hi) for i in range(11)])
    starts = np.vstack([inj, W.ensemble_around(inj, 2, seed=62, scale=0.2)])
    of = np.arange(11) % 3
    z = c.diff(starts, P, of)
    zo = np.array([O.diff(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, starts[o], p) for o, p in zip(of, P)])
    assert np.allclose(z, zo, rtol=1e-9, atol=1e-9), np.abs(z - zo).max()
    zd = Q.dropin_diff(an, Q.RowMatrix(SN), starts[of[3]], P[3])
    assert zd == pytest.approx(zo[3], rel=1e-9)


def test_het_diff(setup):
    which, an, inj, D, SN, c = setup
    oh = oracle_het(which)
    g = Q.Het(c, arrays=oh)
    X = W.ensemble_around(inj, 8, seed=63, scale=0.3)
    Y = W.ensemble_around(inj, 24, seed=64, scale=0.3)
    of = np.arange(24) % 8
    z = g.diff(X, Y, of)
    zo = np.array([O.het_diff(an.labels, an.gmst, an.fref, an.Tobs, oh, X[o], y) for o, y in zip(of, Y)])
    assert np.allclose(z, zo, rtol=1e-9, atol=1e-9), np.abs(z - zo).max()
    view = Q.HetView.from_arrays(oh, an.Tobs, an.nifo)
    assert Q.dropin_het_diff(an, view, X[0], Y[0]) == pytest.approx(zo[0], rel=1e-9, abs=1e-9)
    g.close()


def test_efix_full_data(setup):
    which, an, inj, D, SN, c = setup
    lo, hi = W.prior_box(an)
    st, F = O.fisher_all(an.labels, an.gmst, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, inj)
    ej, ev, _ = O.fisher_evec(F)
    for zs in (30.0, 3.0):
        eo, to = O.efix(an.labels, an.gmst, an.fref, an.Tobs, inj, ej, ev, lo, hi, zs, fmin=an.fmin, SN=SN, N=an.N)
        e, t = c.efix(inj, ej, ev, lo, hi, zs)
        assert np.array_equal(t, to), (t, to)
        assert np.allclose(e, eo, rtol=1e-6, atol=0), np.abs(e - eo).max()
    ed = Q.dropin_efix(an, Q.RowMatrix(SN), None, inj, ej[0], ev[0], lo, hi, 30.0, 0)
    eo, _ = O.efix(an.labels, an.gmst, an.fref, an.Tobs, inj, ej, ev, lo, hi, 30.0, fmin=an.fmin, SN=SN, N=an.N)
    assert np.allclose(ed, eo[0], rtol=1e-6, atol=0)


def test_efix_heterodyne(setup):
    """efix with the heterodyne distance at zs = 1, as MCMC_all rescales
    every chain's Fisher jumps (QuickCBC.c:1927, :2025), for 64 points."""
    which, an, inj, D, SN, c = setup
    oh = oracle_het(which)
    g = Q.Het(c, arrays=oh)
    lo, hi = W.prior_box(an)
    X = W.ensemble_in_prior(inj, 64, lo, hi, seed=65, scale=0.3)
    Fo, _ = O.fisher_het(an.labels, an.gmst, an.fref, an.Tobs, oh, X)
    ej, ev, _ = O.fisher_evec(Fo)
    eo, to = O.efix(an.labels, an.gmst, an.fref, an.Tobs, X, ej, ev, lo, hi, 1.0, het=oh)
    e, t = g.efix(X, ej, ev, lo, hi, 1.0)
    same = (t == to).all(1)
    print("heterodyne efix: %d of 64 points take the oracle's trial counts" % same.sum(), flush=True)
    assert same.mean() >= 0.95
    assert np.allclose(e[same], eo[same], rtol=1e-6, atol=0)
    view = Q.HetView.from_arrays(oh, an.Tobs, an.nifo)
    ed = Q.dropin_efix(an, None, view, X[0], ej[0], ev[0], lo, hi, 1.0, 1)
    if same[0]:
        assert np.allclose(ed, eo[0], rtol=1e-6, atol=0)
    g.close()


def test_het_space(setup):
    which, an, inj, D, SN, c = setup
    sp = oracle_space(which)
    lo, hi = W.prior_box(an)
    got = c.het_space(inj, lo, hi)
    print("het_space: GPU M %d (MN %d, MM %d), oracle M %d" % (got["M"], got["MN"], got["MM"], sp["M"]), flush=True)
    for k in ("MN", "MM", "M", "NR"):
        assert got[k] == sp[k], k
    assert np.array_equal(got["fgrid"], sp["fgrid"])
    assert np.allclose(got["eval"], sp["eval"], rtol=1e-6, atol=0)
    # the displaced points are its own jumps applied as the reference applies them
    Pd = np.array([O.displace(inj, got["evec"][i], got["eval"][i], lo, hi) for i in range(11)])
    assert np.array_equal(got["dparams"], Pd)
    d = Q.dropin_het_space(an, Q.RowMatrix(SN), inj, lo, hi)
    assert d["M"] == sp["M"] and d["J"] == 2 and np.array_equal(d["fgrid"], sp["fgrid"])
    assert np.array_equal(d["freq"], sp["fgrid"] / an.Tobs)
    # the heterodyne on it, set up on the GPU, evaluates as the oracle's
    g = Q.Het(c, params=inj, fgrid=got["fgrid"], J=2, MN=got["MN"], MM=got["MM"])
    P = np.array([O.displace(inj, sp["evec"][i], sp["eval"][i], lo, hi) for i in range(11)])
    L, st = g.loglike(P)
    Lo, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, oracle_het(which), P)
    assert (st == 0).all() and np.abs(L - Lo).max() <= 1e-8 + 1e-10 * np.abs(Lo).max()
    g.close()
