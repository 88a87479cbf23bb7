"""GPU tests of the uneven persistent batch at its full stated shape: twelve
concurrent analyses with Tobs in {4, 8, 16, 32, 64} s, fs in {2048, 4096} Hz,
one to three detectors and 64, 256 or 1024 points each, masses from
1.2 + 1.2 to 60 + 50 Msun, in one persistent queue launch.  One analysis is
the largest shape (64 s at 4096 Hz, three detectors: N/2 = 131072 bins,
129792 in band), two hold 1024 points.

Tolerances: FP64 against the oracle 1e-8 absolute in logL on 32 points per
analysis (the first sixteen and the last sixteen); mixed precision against the
oracle 1e-2; alone versus in the mix bit-identical in both precisions, for
the persistent queue and for the serving loop.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W

pytestmark = pytest.mark.gpu

# (Tobs, fs, detector labels, points, m1, m2)
SPECS = [
    (4.0, 2048.0, [0, 1], 64, 60.0, 50.0),
    (4.0, 4096.0, [0], 256, 30.0, 25.0),
    (8.0, 2048.0, [0, 1, 2], 1024, 36.0, 29.0),
    (8.0, 4096.0, [1, 2], 64, 10.0, 8.0),
    (16.0, 2048.0, [0, 1], 256, 5.0, 4.0),
    (16.0, 4096.0, [0, 1, 2], 64, 1.5, 1.3),
    (32.0, 2048.0, [2], 64, 2.0, 1.5),
    (32.0, 4096.0, [0, 2], 256, 1.4, 1.3),
    (64.0, 4096.0, [0, 1, 2], 64, 1.2, 1.2),
    (64.0, 2048.0, [0, 1], 256, 1.3, 1.25),
    (4.0, 2048.0, [1, 2], 1024, 45.0, 35.0),
    (16.0, 4096.0, [0, 1, 2], 256, 20.0, 15.0),
]


def _analysis(Tobs, fs, labels):
    return W.Analysis(labels=list(labels), Tobs=Tobs, fs=fs, fmin=20.0, fmax=min(2048.0, fs / 2),
                      psd_scale=[2.0 if l == 2 else 1.0 for l in labels])


@functools.lru_cache(maxsize=None)
def instance(i):
    Tobs, fs, labels, npts, m1, m2 = SPECS[i]
    an = _analysis(Tobs, fs, labels)
    r = np.random.default_rng(900 + i)
    dl = 40.0 * (m1 / 1.2) ** 0.8
    inj = W.make_params(m1, m2, r.uniform(-0.3, 0.3), r.uniform(-0.3, 0.3), r.uniform(0, 2 * np.pi), Tobs - 2.0,
                        dl, r.uniform(0, 2 * np.pi), r.uniform(-1, 1), r.uniform(0, np.pi), r.uniform(-1, 1))

    def tmpl(p):
        st, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, p)
        assert st == 0
        return h

    D = W.make_data(an, tmpl, inj, noise_seed=1000 + i)
    P = W.ensemble_around(inj, npts, seed=1100 + i, scale=3.0)
    return an, inj, D, an.SN(), P


def check_rows(B):
    """The checked points of a B-point analysis: the first sixteen and the last sixteen."""
    return np.r_[0:16, B - 16:B]


@pytest.fixture(scope="module")
def ctxs():
    out = []
    for i in range(len(SPECS)):
        an, inj, D, SN, P = instance(i)
        print("uneven batch: analysis %d, Tobs %g s, fs %g Hz, %d det, %d points, %d bins" %
              (i, an.Tobs, an.fs, an.nifo, P.shape[0], an.imax - an.imin), flush=True)
        out.append(Q.Context(an, D, SN))
    yield out
    for c in out:
        c.close()


@pytest.fixture(scope="module")
def oracle_rows():
    """The oracle's logL of the checked points of every analysis."""
    out = []
    for i in range(len(SPECS)):
        an, inj, D, SN, P = instance(i)
        k = check_rows(P.shape[0])
        L, R, S = O.loglike_batch(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P[k])
        out.append((k, L, R, S))
    return out


def test_the_batch_has_the_stated_shape():
    ans = [instance(i)[0] for i in range(len(SPECS))]
    assert len(ans) == 12
    assert {a.Tobs for a in ans} == {4.0, 8.0, 16.0, 32.0, 64.0}
    assert {a.fs for a in ans} == {2048.0, 4096.0}
    assert {a.nifo for a in ans} == {1, 2, 3}
    assert {s[3] for s in SPECS} == {64, 256, 1024}
    big = [a for a in ans if a.N // 2 == 131072 and a.nifo == 3]
    assert big and big[0].imax - big[0].imin == 129792
    masses = [(s[4], s[5]) for s in SPECS]
    assert (1.2, 1.2) in masses and (60.0, 50.0) in masses


def _run_multi(ctx_list, Ps, prec, shape=None, graph=False):
    m = Q.Multi(ctx_list, [p.shape[0] for p in Ps], prec)
    try:
        if shape:
            m.set_shape(*shape)
        m.set_params(Ps)
        m.run(graph=graph)
        return [tuple(np.copy(a) for a in r) for r in m.results()]
    finally:
        m.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_uneven_batch_matches_the_oracle(ctxs, oracle_rows, prec):
    Ps = [instance(i)[4] for i in range(len(SPECS))]
    res = _run_multi(ctxs, Ps, prec)
    tol = 1e-8 if prec == Q.FP64 else 1e-2
    for i, (L, R, S) in enumerate(res):
        k, Lr, Rr, Sr = oracle_rows[i]
        assert (S[k] == Sr).all() and (S == 0).all(), i
        err = np.abs(L[k] - Lr).max()
        print("analysis %d: max |dlogL| %.3e over %d points (|logL| up to %.3e)" %
              (i, err, k.size, np.abs(Lr).max()), flush=True)
        assert err <= tol, (i, err)
        if prec == Q.FP64:
            assert np.abs(R[k] - Rr).max() <= 1e-8, i


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_uneven_alone_equals_mixed(ctxs, prec):
    """Every analysis gives bit-identical logL and rho alone and inside the
    twelve-analysis queue."""
    Ps = [instance(i)[4] for i in range(len(SPECS))]
    mixed = _run_multi(ctxs, Ps, prec)
    for i in range(len(SPECS)):
        L, R, S = _run_multi([ctxs[i]], [Ps[i]], prec)[0]
        assert np.array_equal(L, mixed[i][0]), i
        assert np.array_equal(R, mixed[i][1]), i


def test_uneven_shapes_and_split_cap(ctxs):
    """Launch shapes agree with the default shape, among them a split length
    that gives the largest analysis (129792 band bins) exactly the 64 splits
    the partial-sum layout holds; a shorter split length, which would need
    more, is refused.  Graph replay equals the eager run."""
    Ps = [instance(i)[4] for i in range(len(SPECS))]
    ref = _run_multi(ctxs, Ps, Q.FP64)
    for shape in [(8, 512, 2048), (4, 256, 4096), (1, 128, 16384)]:
        got = _run_multi(ctxs, Ps, Q.FP64, shape=shape)
        for i in range(len(SPECS)):
            assert np.allclose(got[i][0], ref[i][0], rtol=1e-10, atol=1e-9), (shape, i)
    with pytest.raises(Q.QcbcError) as e:
        _run_multi(ctxs, Ps, Q.FP64, shape=(8, 512, 1024))
    assert e.value.code == 1
    got = _run_multi(ctxs, Ps, Q.FP64, graph=True)
    for i in range(len(SPECS)):
        assert np.array_equal(got[i][0], ref[i][0]), i


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_server_alone_equals_mixed_uneven(oracle_rows, prec):
    """The serving loop on the uneven analyses: each request's results are
    the same alone and when it is batched with concurrent requests of every
    other analysis; FP64 results match the oracle."""
    s = Q.Server(8 * 1024 ** 3, prec)
    ids = []
    for i in range(len(SPECS)):
        an, inj, D, SN, P = instance(i)
        code, aid = s.admit(an, D, SN)
        assert code == 0
        ids.append(aid)
    alone = {}
    for i in (2, 8, 10):
        r = s.wait(s.submit(ids[i], instance(i)[4]))
        assert r[0] == 0
        alone[i] = r
    tickets = []
    for rep in range(2):
        for i, aid in enumerate(ids):
            tickets.append((i, s.submit(aid, instance(i)[4])))
    for i, t in tickets:
        r = s.wait(t)
        assert r[0] == 0
        if i in alone:
            assert np.array_equal(r[1], alone[i][1]), (i, prec)
            assert np.array_equal(r[2], alone[i][2]), (i, prec)
    tol = 1e-8 if prec == Q.FP64 else 1e-2
    for i in alone:
        k, Lr, _, _ = oracle_rows[i]
        assert np.abs(alone[i][1][k] - Lr).max() <= tol, i
    s.close()
