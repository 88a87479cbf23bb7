"""The heterodyned likelihood on the GPU against the oracle.

Set-up (heterodyne, QuickCBC.c:5112-5299) on the GPU against the oracle's
set-up arrays, evaluation (log_likelihood_het, :5303-5444) of GPU and
imported set-ups against the oracle's logL, the drop-ins fullphaseamp /
heterodyne / log_likelihood_het through the C ABI, and the edge cases:
ragged and empty batches, points outside the model, other region orders J,
one detector, invalid grids.

Tolerances.  The waveform pieces (fullphaseamp amplitude and phase, the
grid samples of the set-up) 1e-12 relative as the waveform parity tests;
logLb 1e-8 absolute; the set-up sums 1e-10 of their row's magnitude (the
GPU splits each sum over lanes); logL 1e-8 absolute for the evaluation of
the same set-up arrays, and 1e-8 plus 1e-10 |logL - logLb| where the set-up
was computed on the GPU too.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    which = request.param
    an, inj, D, SN = _data.analysis(which)
    h = _data.het(which)
    c = Q.Context(an, D, SN)
    g = Q.Het(c, params=inj, fgrid=h["fgrid"], J=h["J"], MN=h["MN"], MM=h["MM"])
    yield which, an, inj, D, SN, h, c, g
    g.close()
    c.close()


def test_setup_matches_the_oracle(setup):
    which, an, inj, D, SN, h, c, g = setup
    e = g.export()
    assert (e["M"], e["J"], e["NR"]) == (h["M"], h["J"], h["NR"])
    assert np.array_equal(e["freq"], h["freq"])
    assert np.abs(e["amp"] - h["amp"]).max() <= 1e-12 * np.abs(h["amp"]).max()
    assert np.abs(e["phase"] - h["phase"]).max() <= 1e-12 * np.abs(h["phase"]).max()
    assert abs(e["logLb"] - h["logLb"]) <= 1e-8, (e["logLb"], h["logLb"])
    # same recurrence and elimination, no contraction: the inverses agree to rounding
    assert np.allclose(e["IP"], h["IP"], rtol=1e-13, atol=1e-13 * np.abs(h["IP"]).max())
    for k in ("SL", "lc", "ls"):
        for d in range(an.nifo):
            want, got = h[k][:, d, :], e[k][:, d, :]
            tol = 1e-10 * np.abs(want).max(1, keepdims=True) + 1e-300
            assert (np.abs(got - want) <= tol).all(), (k, d, np.abs(got - want).max())
    for k in ("rc", "rs", "aa"):
        for d in range(an.nifo):
            assert np.abs(e[k][d] - h[k][d]).max() <= 1e-10 * np.abs(h[k][d]).max(), (k, d)


def test_weights_are_the_reference_linear_form(setup):
    """wH . cc reproduces HH of log_likelihood_het's region loop for arbitrary sample values."""
    which, an, inj, D, SN, h, c, g = setup
    e = g.export()
    w = g.weights()
    M, J, NR = e["M"], e["J"], e["NR"]
    v = np.random.default_rng(5).standard_normal((an.nifo, M))
    for d in range(an.nifo):
        HH, x = 0.0, 0.0
        for i in range(NR):
            vv = v[d, i * J:i * J + J + 1]
            HH += (e["IP"][i] @ vv) @ e["SL"][i, d]
            if i < NR - 2:
                x += e["aa"][d, i + 1] * vv[J]
        ref = HH - x
        assert abs(w[0, d] @ v[d] - ref) <= 1e-11 * (np.abs(w[0, d]) @ np.abs(v[d])), d


def test_loglike_matches_the_oracle(setup):
    which, an, inj, D, SN, h, c, g = setup
    P = np.vstack([inj, W.ensemble_around(inj, 256, seed=21, scale=0.3)])
    Lo, so = _data.oracle_loglike_het(which, P)
    assert (so == 0).all()
    # the oracle's set-up evaluated on the GPU
    gi = Q.Het(c, arrays=h)
    Li, si = gi.loglike(P)
    gi.close()
    assert (si == 0).all()
    assert np.abs(Li - Lo).max() <= 1e-8, np.abs(Li - Lo).max()
    # set-up and evaluation on the GPU
    L, s = g.loglike(P)
    assert (s == 0).all()
    tol = 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])
    assert (np.abs(L - Lo) <= tol).all(), np.abs(L - Lo).max()
    # at the reference point every heterodyne ratio of the GPU's own set-up is
    # 0: logL is logLb exactly; the oracle's samples differ from the GPU's
    # waveform by rounding
    assert L[0] == g.logLb
    assert abs(Li[0] - h["logLb"]) <= 1e-8


# 1 to 4 waves per point: small batches split each point's samples
@pytest.mark.parametrize("B", [1, 5, 63, 65, 257, 1100, 2048])
def test_ragged_batches(setup, B):
    which, an, inj, D, SN, h, c, g = setup
    P = W.ensemble_around(inj, B, seed=40 + B, scale=0.3)
    Lo, _ = _data.oracle_loglike_het(which, P)
    L, s = g.loglike(P)
    assert L.shape == (B,) and (s == 0).all()
    assert (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])).all()


def test_empty_batch_and_domain_errors(setup):
    which, an, inj, D, SN, h, c, g = setup
    L, s = g.loglike(np.zeros((0, 11)))
    assert L.shape == (0,)
    P = np.vstack([inj, inj, inj])
    P[1, 2] = 1.5  # chi1 > 1: outside the model
    L, s = g.loglike(P)
    Lo, so = _data.oracle_loglike_het(which, P)
    assert list(s) == [0, 33, 0] and list(so) == [0, 33, 0]
    assert L[1] == 0.0 and L[0] == L[2] == g.logLb


@pytest.mark.parametrize("J", [1, 3])
def test_other_region_orders(J):
    an, inj, D, SN = _data.analysis(0)
    h = _data.het(0, J=J)
    c = Q.Context(an, D, SN)
    g = Q.Het(c, params=inj, fgrid=h["fgrid"], J=J, MN=h["MN"], MM=h["MM"])
    P = W.ensemble_around(inj, 64, seed=8, scale=0.3)
    Lo, _ = _data.oracle_loglike_het(0, P, J=J)
    L, s = g.loglike(P)
    Li, _ = Q.Het(c, arrays=h).loglike(P)
    g.close()
    c.close()
    assert (s == 0).all()
    assert np.abs(Li - Lo).max() <= 1e-8
    assert (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])).all()


def test_one_detector():
    base, inj, _, _ = _data.analysis(0)
    an = W.Analysis(labels=[0], Tobs=base.Tobs, fs=base.fs, fmin=base.fmin, fmax=base.fmax, psd_scale=[1.0])

    def tmpl(p):
        st, hw = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, p)
        assert st == 0
        return hw

    D = W.make_data(an, tmpl, inj)
    SN = an.SN()
    MN, MM = _data.het_band(an, inj, SN)
    fg = W.het_grid(MN, MM, an.Tobs)
    st, h = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, fg, 2, MN, MM)
    assert st == 0
    P = W.ensemble_around(inj, 32, seed=4, scale=0.3)
    Lo, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, h, P)
    c = Q.Context(an, D, SN)
    g = Q.Het(c, params=inj, fgrid=fg, J=2, MN=MN, MM=MM)
    L, s = g.loglike(P)
    assert abs(g.logLb - h["logLb"]) <= 1e-8
    g.close()
    c.close()
    assert (s == 0).all()
    assert (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])).all()


def test_invalid_grids_are_rejected(setup):
    which, an, inj, D, SN, h, c, g = setup
    fg = h["fgrid"].copy()
    bad = []
    f1 = fg.copy()
    f1[5] = f1[4]  # not strictly increasing
    bad.append((f1, 2, h["MN"], h["MM"]))
    bad.append((fg[:-1], 2, h["MN"], h["MM"]))  # M != NR J + 1
    f2 = fg.copy()
    f2[-1] = an.N // 2  # past the last bin
    bad.append((f2, 2, h["MN"], h["MM"]))
    bad.append((fg, 2, h["MN"], an.N // 2 + 1))  # logLb band past the data
    bad.append((fg, 9, h["MN"], h["MM"]))  # J above 8
    for f, J, MN, MM in bad:
        with pytest.raises(Q.QcbcError) as e:
            Q.Het(c, params=inj, fgrid=f, J=J, MN=MN, MM=MM)
        assert e.value.code == 1


def test_dropin_fullphaseamp(setup):
    which, an, inj, D, SN, h, c, g = setup
    for freq in (h["freq"], an.freq()):
        st, Ao, Po = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, freq, inj)
        A, P = Q.dropin_fullphaseamp(an, freq, inj)
        assert st == 0
        assert np.abs(A - Ao).max() <= 1e-12 * np.abs(Ao).max()
        assert np.abs(P - Po).max() <= 1e-12 * np.abs(Po).max()


def test_dropin_heterodyne_and_log_likelihood_het(setup):
    which, an, inj, D, SN, h, c, g = setup
    Dr, Sr = Q.RowMatrix(D), Q.RowMatrix(SN)
    v = Q.dropin_heterodyne(an, Dr, Sr, inj, h["fgrid"], h["J"], h["MN"], h["MM"])
    got = v.read(an.nifo)
    e = g.export()
    for k in Q.HET_KEYS:
        assert np.array_equal(got[k], e[k]), k  # the same GPU set-up, copied into the reference's shapes
    assert got["logLb"] == g.logLb and np.array_equal(got["pref"], inj)
    P = W.ensemble_around(inj, 8, seed=12, scale=0.3)
    Lo, _ = _data.oracle_loglike_het(which, P)
    L = np.array([Q.dropin_log_likelihood_het(an, v, p) for p in P])
    assert (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])).all()
    # a struct Het set up elsewhere (the oracle's arrays in the reference's shapes)
    ov = Q.HetView.from_arrays(h, an.Tobs, an.nifo)
    L2 = np.array([Q.dropin_log_likelihood_het(an, ov, p) for p in P])
    assert np.abs(L2 - Lo).max() <= 1e-8


def test_timing_counts_launches(setup):
    which, an, inj, D, SN, h, c, g = setup
    P = W.ensemble_around(inj, 128, seed=2, scale=0.3)
    g.timing(1)
    for _ in range(3):
        g.loglike(P)
    ms, n = g.timing(0)
    assert n == 3 and ms > 0


def test_device_entry_points(setup):
    """qcbc_het_loglike_device on resident points gives qcbc_het_loglike's
    numbers; the kernel-alone timing runs."""
    which, an, inj, D, SN, h, c, g = setup
    P = W.ensemble_around(inj, 300, seed=6, scale=0.3)
    L, _ = g.loglike(P)
    dp = Q.DeviceBuffer.from_array(np.ascontiguousarray(P))
    dl = Q.DeviceBuffer(8 * P.shape[0])
    g.loglike_device(dp.ptr, P.shape[0], dl.ptr)
    g.sync()
    Ld = dl.download((P.shape[0],))
    assert np.array_equal(Ld, L)
    assert g.time_kernel(dp.ptr, P.shape[0], 5) > 0
    dp.free()
    dl.free()


def test_pt_on_the_heterodyned_likelihood(setup):
    """The resident PT step on the heterodyned likelihood: one column group,
    three groups and the graph replay give the same chains, and the
    ensemble's logL is the heterodyned logL of its points (GPU and oracle)."""
    which, an, inj, D, SN, h, c, g = setup
    nt, nc = 8, 48
    init = W.ensemble_around(inj, nt * nc, seed=5, scale=0.3)
    sigma = np.array([2e-5, 1e-4, 2e-3, 2e-3, 0.05, 2e-5, 0.01, 0.01, 0.01, 0.02, 0.01]) * 0.3
    one = Q.PT(c, nt, nc, 1.2, init, sigma, seed=13, groups=1, het=g)
    three = Q.PT(c, nt, nc, 1.2, init, sigma, seed=13, groups=3, het=g)
    gr = Q.PT(c, nt, nc, 1.2, init, sigma, seed=13, groups=3, het=g)
    one.run(40)
    three.run(40)
    gr.run(40, graph=True)
    r1, r3, rg = one.get(), three.get(), gr.get()
    for r in (r3, rg):
        assert np.array_equal(r1[0], r[0]) and np.array_equal(r1[1], r[1])
        assert np.array_equal(r1[2], r[2]) and r1[3] == r[3]
    assert r1[2].sum() > 0
    L, s = g.loglike(r1[0])
    assert (s == 0).all() and np.allclose(L, r1[1], rtol=1e-12, atol=1e-9)
    Lo, _ = _data.oracle_loglike_het(which, r1[0])
    assert (np.abs(r1[1] - Lo) <= 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])).all()
    for pt in (one, three, gr):
        pt.close()
