"""GPU parity of the search likelihoods against the oracle: pbt_shift, PDwave,
log_likelihood_penalized, log_likelihood_APmax and log_likelihood_max
(QuickCBC.c:3448-3525, 5672-6408; Utilities.c:687-713), through the batch
API and the drop-ins, on the noisy data of both workloads.

Tolerances: pbt_shift 1e-11 of the largest sample (hipFFT against the
oracle's radix-2 transform); PDwave 1e-9 of the largest sample (the
waveform agrees to ~1e-12, the merger-time shift multiplies a 1e5-rad
phase); penalized, APmax and log_likelihood_max logL 1e-9 relative and
the moved parameters 1e-8 absolute, with the same bands kept and the same
peaks chosen (every discrete choice identical).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu


def ref_points(which, n, seed):
    """The injection and n points around it, in the reference frame."""
    an, inj, D, SN = _data.analysis(which)
    P = np.vstack([inj, W.ensemble_around(inj, n, seed=seed, scale=0.3)])
    return np.array([_data.ref_frame(an, p) for p in P])


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    which = request.param
    an, inj, D, SN = _data.analysis(which)
    print("search: config %d, %d det, N %d" % (which, an.nifo, an.N), flush=True)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    c = Q.Context(an, D, SN)
    yield which, an, inj, D, SN, c
    c.close()
    Q.lib().qcbc_set_band(16.0, 1024.0)


def test_pbt_shift(setup):
    which, an, inj, D, SN, c = setup
    imin, imax = int(an.fmin * an.Tobs), min(int(an.fmax * an.Tobs), an.N // 2)
    st, w = O.pdwave(an.N, an.Tobs, an.Tobs / 2.0, _data.ref_frame(an, inj)[:7])
    assert st == 0
    for d in range(an.nifo):
        co, cfo = O.pbt_shift(D[d], w, SN[d], imin, imax, an.N)
        cg, cfg = Q.dropin_pbt_shift(D[d], w, SN[d], imin, imax, an.N)
        scale = max(np.abs(co).max(), np.abs(cfo).max())
        print("pbt_shift det %d: max error %.2e of the largest sample" %
              (d, max(np.abs(cg - co).max(), np.abs(cfg - cfo).max()) / scale), flush=True)
        assert np.abs(cg - co).max() <= 1e-11 * scale
        assert np.abs(cfg - cfo).max() <= 1e-11 * scale


def test_pdwave(setup):
    which, an, inj, D, SN, c = setup
    for q in ref_points(which, 3, 71):
        for tcmin in (an.Tobs / 2.0, an.Tobs - 0.5):
            st, wo = O.pdwave(an.N, an.Tobs, tcmin, q[:7], fref=an.fref)
            assert st == 0
            wg = Q.dropin_pdwave(an, q[:7], tcmin)
            assert np.abs(wg - wo).max() <= 1e-9 * np.abs(wo).max()


def test_penalized(setup):
    which, an, inj, D, SN, c = setup
    for q in ref_points(which, 4, 72):
        st, tw = O.pdwave(an.N, an.Tobs, q[5], q[:7], fref=an.fref)
        assert st == 0
        for d in range(an.nifo):
            Lo, Ao, Po = O.penalized(q, D[d], tw, SN[d], an.fmin, an.fmax, an.N, an.Tobs)
            Lg, Ag, Pg = Q.dropin_penalized(q, D[d], tw, SN[d], an.N, an.Tobs)
            assert Lg == pytest.approx(Lo, rel=1e-9, abs=1e-12)
            assert Ag == pytest.approx(Ao, rel=1e-9, abs=1e-15)
            assert Pg == pytest.approx(Po, abs=1e-9)


def test_apmax(setup):
    which, an, inj, D, SN, c = setup
    Q0 = ref_points(which, 7, 73)
    Lo, Po, so = O.apmax(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Q0)
    Lg, Pg, sg = c.apmax(Q0)
    assert np.array_equal(sg, so)
    print("APmax: max logL error %.2e relative, max parameter error %.2e" %
          ((np.abs(Lg - Lo) / np.abs(Lo)).max(), np.abs(Pg - Po).max()), flush=True)
    assert np.allclose(Lg, Lo, rtol=1e-9, atol=1e-12)
    assert np.allclose(Pg, Po, rtol=0, atol=1e-8)
    v, p = Q.dropin_apmax(an, Q.RowMatrix(D), Q.RowMatrix(SN), Q0[1])
    assert v == pytest.approx(Lg[1], rel=1e-12) and np.allclose(p, Pg[1], rtol=0, atol=1e-12)


def test_llmax(setup):
    which, an, inj, D, SN, c = setup
    Q0 = ref_points(which, 5, 74)
    tmin, tmax = max(1.5, an.Tobs - 3.5), an.Tobs - 0.5
    Lo, Po, so = O.llmax(an.labels, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, tmin, tmax, tmax, Q0)
    Lg, Pg, sg = c.llmax(Q0, tmin, tmax, tmax, O.delays(an.labels))
    assert np.array_equal(sg, so)
    print("log_likelihood_max: logL %s, max error %.2e relative, max parameter error %.2e" %
          (np.round(Lo, 2), (np.abs(Lg - Lo) / np.abs(Lo)).max(), np.abs(Pg - Po).max()), flush=True)
    assert np.allclose(Lg, Lo, rtol=1e-9, atol=1e-12)
    assert np.allclose(Pg, Po, rtol=0, atol=1e-8)
    v, p = Q.dropin_llmax(an, Q.RowMatrix(D), Q.RowMatrix(SN), Q0[2], tmin, tmax, tmax, O.delays(an.labels))
    assert v == pytest.approx(Lg[2], rel=1e-12) and np.allclose(p, Pg[2], rtol=0, atol=1e-12)
