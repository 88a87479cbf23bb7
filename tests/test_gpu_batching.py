"""GPU tests of the batched forms: the persistent multi-analysis launch, its
hipGraph replays, the serving loop, the PT step and the drop-in entry points.

Tolerances: FP64 against the oracle 1e-8 absolute in logL; FP64 forms
against each other 1e-9 relative; "alone versus in a mix" and graph replays
versus eager runs are required to be bit-identical.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu


def _an(labels, Tobs, fs, fmin, fmax, scales):
    return W.Analysis(labels=labels, Tobs=Tobs, fs=fs, fmin=fmin, fmax=fmax, psd_scale=scales)


SHAPES = [
    ("cfg0", lambda: W.config0(), lambda: W.injection0(), 64),
    ("cfg1", lambda: W.config1(), lambda: W.injection1(), 32),
    ("H1-8s", lambda: _an([0], 8.0, 2048.0, 20.0, 1024.0, [1.0]),
     lambda: W.make_params(10.0, 8.0, 0.2, 0.1, 1.0, 6.0, 200.0, 1.0, 0.3, 0.4, 0.5), 100),
    ("LV-32s", lambda: _an([1, 2], 32.0, 4096.0, 20.0, 2048.0, [1.0, 2.0]),
     lambda: W.make_params(1.4, 1.3, 0.0, 0.0, 2.0, 30.0, 60.0, 2.0, -0.2, 1.0, -0.3), 16),
]


@functools.lru_cache(maxsize=None)
def shape_data(i):
    name, mk_an, mk_inj, B = SHAPES[i]
    an, inj = mk_an(), mk_inj()

    def tmpl(p):
        st, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, p)
        assert st == 0
        return h

    D = W.make_data(an, tmpl, inj, noise_seed=100 + i)
    P = W.ensemble_around(inj, B, seed=200 + i, scale=3.0)
    return an, inj, D, an.SN(), P


@pytest.fixture(scope="module")
def ctxs():
    out = []
    for i in range(len(SHAPES)):
        an, inj, D, SN, P = shape_data(i)
        out.append(Q.Context(an, D, SN))
    yield out
    for c in out:
        c.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_multi_matches_single_and_oracle(ctxs, prec):
    Ps = [shape_data(i)[4] for i in range(len(SHAPES))]
    m = Q.Multi(ctxs, [p.shape[0] for p in Ps], prec)
    m.set_params(Ps)
    m.run()
    res = m.results()
    for i, (L, R, S) in enumerate(res):
        an, inj, D, SN, P = shape_data(i)
        Ls, Rs, Ss = ctxs[i].loglike(P, prec)
        assert (S == Ss).all()
        # FP64: the two forms split the band differently, so they agree to
        # rounding; mixed precision: both stay within 0.01 of FP64, and the
        # FP32 partial sums of different splits round differently.
        rtol, atol = (1e-9, 1e-9) if prec == Q.FP64 else (0.0, 2e-3)
        assert np.allclose(L, Ls, rtol=rtol, atol=atol), (SHAPES[i][0], np.abs(L - Ls).max())
        Lr, Rr, Sr = _data.oracle_loglike(an, D, SN, P[:8])
        tol = 1e-8 if prec == Q.FP64 else 1e-2
        assert np.abs(L[:8] - Lr).max() <= tol, (SHAPES[i][0], np.abs(L[:8] - Lr).max())
    m.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_alone_equals_mixed(ctxs, prec):
    """Each analysis gives bit-identical results alone and inside the mix."""
    Ps = [shape_data(i)[4] for i in range(len(SHAPES))]
    m = Q.Multi(ctxs, [p.shape[0] for p in Ps], prec)
    m.set_params(Ps)
    m.run()
    mixed = m.results()
    m.close()
    for i in range(len(SHAPES)):
        a = Q.Multi([ctxs[i]], [Ps[i].shape[0]], prec)
        a.set_params([Ps[i]])
        a.run()
        L, R, S = a.results()[0]
        a.close()
        assert np.array_equal(L, mixed[i][0]), SHAPES[i][0]
        assert np.array_equal(R, mixed[i][1])


def test_multi_graph_replays_match_eager(ctxs):
    Ps = [shape_data(i)[4] for i in range(len(SHAPES))]
    m = Q.Multi(ctxs, [p.shape[0] for p in Ps], Q.FP64)
    m.set_params(Ps)
    m.run()
    ref = [r[0].copy() for r in m.results()]
    for rep in range(3):
        m.run(graph=True)
        got = m.results()
        for i in range(len(SHAPES)):
            assert np.array_equal(got[i][0], ref[i]), (rep, SHAPES[i][0])
        m.run()  # eager work between replays
        ctxs[0].loglike(Ps[0][:4])
    # new parameters through the same buffers: the replay sees them
    Ps2 = [p[::-1].copy() for p in Ps]
    m.set_params(Ps2)
    m.run(graph=True)
    got = m.results()
    for i in range(len(SHAPES)):
        Ls, _, _ = ctxs[i].loglike(Ps2[i], Q.FP64)
        assert np.allclose(got[i][0], Ls, rtol=1e-9, atol=1e-9)
    m.close()


def test_multi_autotune_and_shapes(ctxs):
    Ps = [shape_data(i)[4] for i in range(len(SHAPES))]
    m = Q.Multi(ctxs, [p.shape[0] for p in Ps], Q.FP64)
    m.set_params(Ps)
    m.run()
    ref = [r[0].copy() for r in m.results()]
    for shape in [(2, 128, 2048), (8, 512, 16384), (4, 256, 4096)]:
        m.set_shape(*shape)
        m.run()
        for i, r in enumerate(m.results()):
            assert np.allclose(r[0], ref[i], rtol=1e-10, atol=1e-9)
    best, ms = m.autotune(reps=1)
    assert ms > 0 and m.get_shape()[:3] == best
    m.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_ctx_launch_shapes(prec):
    """Every launch shape gives the default shape's logL: one bin per thread
    (the staging that prefetches the next tile) and fewer threads than tile
    bins (the staging loop), 1 to 16 points per block, ragged point counts."""
    an, inj, D, SN = _data.analysis(1)
    c = Q.Context(an, D, SN)
    P = np.vstack([inj, _data.points(1, 100, seed=71)])
    ref, _, st = c.loglike(P, prec)
    assert (st == 0).all()
    tol = (1e-9, 1e-9) if prec == Q.FP64 else (1e-6, 1e-4)
    for cpb, tile, splits in [(1, 128, 0), (2, 512, 0), (4, 512, 3), (8, 256, 0), (12, 512, 0), (16, 512, 7),
                              (16, 128, 0), (3, 256, 1)]:
        c.set_shape(prec, cpb, tile, splits)
        L, _, st = c.loglike(P, prec)
        assert (st == 0).all()
        assert np.allclose(L, ref, rtol=tol[0], atol=tol[1]), (cpb, tile, splits, np.abs(L - ref).max())
    c.close()


def test_server_alone_and_mixed_and_admission():
    budget = 2 * 1024 ** 3
    s = Q.Server(budget, Q.FP64)
    ids = []
    for i in range(3):
        an, inj, D, SN, P = shape_data(i)
        code, aid = s.admit(an, D, SN)
        assert code == 0
        ids.append(aid)
    an0, _, D0, SN0, P0 = shape_data(0)
    alone = s.wait(s.submit(ids[0], P0))
    assert alone[0] == 0
    # the same request inside a mix of concurrent requests from all analyses
    tickets = []
    for rep in range(3):
        for i, aid in enumerate(ids):
            tickets.append((i, s.submit(aid, shape_data(i)[4])))
    results = [(i, s.wait(t)) for i, t in tickets]
    for i, r in results:
        assert r[0] == 0
        if i == 0:
            assert np.array_equal(r[1], alone[1])
            assert np.array_equal(r[2], alone[2])
    Lr, _, _ = _data.oracle_loglike(an0, D0, SN0, P0[:8])
    assert np.abs(alone[1][:8] - Lr).max() <= 1e-8
    st = s.stats()
    assert st["analyses"] == 3 and st["evals"] >= 10 * P0.shape[0]
    # admission is by footprint: room for the batch scratch and one small analysis only
    import ctypes
    fp = [Q.lib().qcbc_ctx_footprint(ctypes.byref(Q.make_desc(shape_data(i)[0]))) for i in range(3)]
    scratch = st["bytes_in_use"] - sum(fp)
    small = Q.Server(scratch + fp[0] + 1000, Q.FP64)
    assert small.admit(an0, D0, SN0)[0] == 0
    an1, _, D1, SN1, _ = shape_data(1)
    assert small.admit(an1, D1, SN1)[0] == 103  # QCBC_EBUSY
    small.close()
    # eviction frees the budget and refuses new requests for that analysis
    assert s.evict(ids[2]) == 0
    with pytest.raises(Q.QcbcError):
        s.submit(ids[2], shape_data(2)[4])
    assert s.stats()["analyses"] == 2
    s.close()


def test_pt_graph_equals_eager():
    an, inj, D, SN = _data.analysis(1)
    c = Q.Context(an, D, SN)
    init = W.ensemble_around(inj, 4 * 16, seed=3)
    sigma = np.array([2e-5, 1e-4, 2e-3, 2e-3, 0.05, 2e-5, 0.01, 0.01, 0.01, 0.02, 0.01])
    a = Q.PT(c, 4, 16, 1.2, init, sigma, seed=11)
    b = Q.PT(c, 4, 16, 1.2, init, sigma, seed=11)
    a.run(5)
    b.run(5, graph=True)
    xa, La, acca, swa = a.get()
    xb, Lb, accb, swb = b.get()
    assert np.array_equal(xa, xb) and np.array_equal(La, Lb) and np.array_equal(acca, accb) and swa == swb
    assert acca.sum() > 0
    # the ensemble's logL is the likelihood of its current points
    Lc, _, _ = c.loglike(xa, Q.FP64)
    assert np.allclose(Lc, La, rtol=1e-12, atol=1e-9)
    a.close()
    b.close()
    c.close()


def test_pt_groups_give_the_same_chains():
    """Two column groups pipelined on two streams step exactly as one group:
    eager, with per-step synchronisation, and replayed as a graph."""
    an, inj, D, SN = _data.analysis(1)
    c = Q.Context(an, D, SN)
    nt, nc = 8, 64
    init = W.ensemble_around(inj, nt * nc, seed=5)
    sigma = np.array([2e-5, 1e-4, 2e-3, 2e-3, 0.05, 2e-5, 0.01, 0.01, 0.01, 0.02, 0.01])
    one = Q.PT(c, nt, nc, 1.2, init, sigma, seed=13, groups=1)
    two = Q.PT(c, nt, nc, 1.2, init, sigma, seed=13, groups=2)
    auto = Q.PT(c, nt, nc, 1.2, init, sigma, seed=13, groups=3)
    assert (one.groups, two.groups, auto.groups) == (1, 2, 3)
    one.run(7)
    two.run(3)
    two.step_timing(True)
    two.run(2)
    two.step_timing(False)
    two.run(2)
    auto.run(7, graph=True)
    r1, r2, r3 = one.get(), two.get(), auto.get()
    for r in (r2, r3):
        assert np.array_equal(r1[0], r[0]) and np.array_equal(r1[1], r[1])
        assert np.array_equal(r1[2], r[2]) and r1[3] == r[3]
    # 67 more steps: the graph path replays the 32-step graph twice and the
    # one-step graph three times, the groups running through each 32 steps
    # without meeting
    one.run(67)
    two.run(67)
    auto.run(67, graph=True)
    auto.run(1, graph=True)
    one.run(1)
    two.run(1)
    r1, r2, r3 = one.get(), two.get(), auto.get()
    for r in (r2, r3):
        assert np.array_equal(r1[0], r[0]) and np.array_equal(r1[1], r[1])
        assert np.array_equal(r1[2], r[2]) and r1[3] == r[3]
    assert r1[2].sum() > 0 and r1[3] > 0
    Lc, _, _ = c.loglike(r2[0], Q.FP64)
    assert np.allclose(Lc, r2[1], rtol=1e-12, atol=1e-9)
    for pt in (one, two, auto):
        pt.close()
    c.close()


def test_dropins_against_oracle():
    an, inj, D, SN = _data.analysis(0)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    Dr, Sr = Q.RowMatrix(D), Q.RowMatrix(SN)
    P = _data.points(0, 6, seed=9)
    for p in P:
        v, rho = Q.dropin_log_likelihood_full(an, Dr, Sr, p)
        vr, rr, st = O.loglike(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, p)
        assert abs(v - vr) <= 1e-8 and np.abs(rho - rr).max() <= 1e-8
        h = Q.dropin_fulltemplates(an, p)
        st, hr = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, p)
        assert np.abs(h - hr).max() <= 1e-9 * np.abs(hr).max()
    # IMRPhenomDGenerateh22FDAmpPhase on an irregular frequency list
    f = np.sort(np.random.default_rng(1).uniform(5.0, 600.0, 333))
    st, a, ph = Q.dropin_h22(f, 0.4, 100.0, 30 * W.MSUN_SI, 20 * W.MSUN_SI, 0.5, -0.3, 300 * W.MPC_SI)
    st2, ar, pr = O.h22(f, 0.4, 100.0, 30 * W.MSUN_SI, 20 * W.MSUN_SI, 0.5, -0.3, 300 * W.MPC_SI)
    assert st == st2 == 0
    assert np.max(np.abs(a - ar) / ar) <= 1e-12
    assert np.max(np.abs(ph - pr) / np.maximum(np.abs(pr), 2 * np.pi)) <= 1e-12
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(an.N), rng.standard_normal(an.N)
    assert Q.dropin_fourier_nwip(x, y, SN[0], an.imin, an.imax, an.N) == pytest.approx(
        O.fourier_nwip(x, y, SN[0], an.imin, an.imax, an.N), rel=1e-12)
    for det in (0, 1, 2):
        fp, fc = Q.dropin_f_ant(0.3, 1.2, -0.4, 0.7, det)
        fpr, fcr = O.f_ant(0.3, 1.2, -0.4, 0.7, det)
        assert fp == pytest.approx(fpr, abs=1e-15) and fc == pytest.approx(fcr, abs=1e-15)
    assert np.allclose(Q.dropin_times(1.2, -0.4, 0.7), O.times(1.2, -0.4, 0.7), atol=1e-18, rtol=1e-14)
    assert Q.dropin_fbegin(inj) == pytest.approx(O.fbegin(inj), rel=1e-13)
    Q.lib().qcbc_set_band(16.0, 1024.0)
