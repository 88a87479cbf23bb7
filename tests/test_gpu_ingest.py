"""main's conditioning outputs, CBC_start's re-conditioning and the small
conditioning drop-ins on the GPU against the oracle (QuickCBC.c:195-318,
:807-823; whiten, tukey, tukey_scale: Utilities.c:294-309, :635-683).

Tolerances: D and the whitened series DW 1e-11 of their largest value (two
FFTs' rounding, ~eps log2 N relative); pspec and wcheck 1e-10 of their
largest value (squares of the transform); the likelihood of the bound data
1e-8 absolute plus 1e-10 relative; tukey 4e-16 of each sample (the
window's error is absolute: one ulp of cos, cancelled in 1 + cos near the
start of a taper), whiten 1e-15 relative (sqrt and a division); tukey_scale
1e-13 (a regrouped sum).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from tests import _data

pytestmark = pytest.mark.gpu


def strain(an, seed):
    r = np.random.default_rng(seed)
    t = np.arange(an.N) * (an.Tobs / an.N)
    return r.standard_normal((an.nifo, an.N)) * 1e-21 + 3e-22 * np.sin(2 * np.pi * 123.4 * t)[None, :]


def close(a, b, tol):
    return np.abs(a - b).max() <= tol * np.abs(b).max()


@pytest.mark.parametrize("which", [0, 1])
def test_condition_outputs(which):
    an, inj, D0, SN = _data.analysis(which)
    x = strain(an, 50 + which)
    print("condition outputs: config %d, %d x %d samples" % (which, an.nifo, an.N), flush=True)
    c = Q.Context(an)
    got = c.condition(x, SN, t_rise=0.4)
    want = O.condition_outputs(x, SN, an.Tobs, 0.4)
    for k, tol in (("D", 1e-11), ("DW", 1e-11), ("pspec", 1e-10), ("wcheck", 1e-10)):
        err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
        print("  %s: max error %.2e of the largest value" % (k, err), flush=True)
        assert close(got[k], want[k], tol), k
    # the bound data are D, as qcbc_bind_time_series binds them
    P = _data.points(which, 8, seed=4)
    L, _, st = c.loglike(P, Q.FP64)
    Lo, _, so = _data.oracle_loglike(an, want["D"], SN, P)
    assert (st == 0).all() and (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo)).all()
    # D alone, and bind_time_series, give the same D
    D2 = c.bind_time_series(x, SN, t_rise=0.4)
    assert np.array_equal(D2, got["D"])
    c.close()


@pytest.mark.parametrize("which", [0, 1])
def test_recondition_floors_the_psd(which):
    """CBC_start's re-conditioning sets the bound PSD to 1 below fmin; a floor
    above fmin makes those bins silent in the likelihood."""
    an, inj, D0, SN = _data.analysis(which)
    x = strain(an, 60 + which)
    c = Q.Context(an)
    P = _data.points(which, 8, seed=5)
    fl = 2.0 * an.fmin
    got = c.condition(x, SN, t_rise=0.4, fmin_floor=fl, outputs=False)
    D = O.condition(x, an.Tobs, 0.4)
    assert close(got["D"], D, 1e-11)
    SNf = SN.copy()
    SNf[:, :int(fl * an.Tobs)] = 1.0
    L, _, st = c.loglike(P, Q.FP64)
    Lo, _, so = _data.oracle_loglike(an, D, SNf, P)
    assert (st == 0).all() and (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo)).all()
    # the floor at fmin itself changes no likelihood: the band starts there
    c.condition(x, SN, t_rise=0.4, fmin_floor=an.fmin, outputs=False)
    L1, _, _ = c.loglike(P, Q.FP64)
    Lo1, _, _ = _data.oracle_loglike(an, D, SN, P)
    assert (np.abs(L1 - Lo1) <= 1e-8 + 1e-10 * np.abs(Lo1)).all()
    c.close()


def test_files_to_bound_data(tmp_path):
    """framed_* and spec_* files written as the reference's inputs are, read
    through the C ABI, conditioned and bound: the same D and logL as the
    oracle's reading of the same files."""
    which = 1
    an, inj, D0, SN = _data.analysis(which)
    x = strain(an, 70)
    t = np.arange(an.N) * (an.Tobs / an.N)
    rows_x, rows_s = [], []
    for d, lab in enumerate(an.labels):
        pf = tmp_path / Q.data_file_name("framed", an.Tobs, 1126259462.0, lab)
        pf.write_text("".join("%e %e\n" % (a, b) for a, b in zip(t, x[d])))
        ps = tmp_path / Q.data_file_name("spec", an.Tobs, 1126259462.0, lab)
        f = np.arange(an.N // 2) / an.Tobs
        ps.write_text("".join("%e %e %e %e\n" % (a, b, b, 1.0) for a, b in zip(f, SN[d])))
        tq, xq = Q.read_framed(pf)
        assert xq.size == an.N
        Sq, _ = Q.read_spec(ps, an.N // 2)
        No, to, xo = O.read_framed(pf)
        ro, So, _ = O.read_spec(ps, an.N // 2)
        assert No == an.N and ro == an.N // 2
        assert np.array_equal(xq, xo) and np.array_equal(Sq, So)
        rows_x.append(xq)
        rows_s.append(Sq)
    X, S = np.array(rows_x), np.array(rows_s)
    c = Q.Context(an)
    got = c.condition(X, S, t_rise=0.4, outputs=False)
    want = O.condition(X, an.Tobs, 0.4)
    assert close(got["D"], want, 1e-11)
    P = _data.points(which, 8, seed=6)
    L, _, st = c.loglike(P, Q.FP64)
    Lo, _, so = _data.oracle_loglike(an, want, S, P)
    assert (st == 0).all() and (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo)).all()
    c.close()


def test_dropins():
    r = np.random.default_rng(12)
    for N, alpha in ((1 << 13, 0.1), (1 << 16, 0.05), (1000, 0.5)):
        x = r.standard_normal(N)
        got = Q.dropin_tukey(x, alpha)
        want = x * O.tukey(N, alpha)
        # the window's error is absolute (one ulp of cos near -1 in 1 + cos)
        assert (np.abs(got - want) <= 4e-16 * np.abs(x)).all()
        s1, s2 = Q.dropin_tukey_scale(alpha, N)
        o1, o2 = O.tukey_scale(alpha, N)
        assert s1 == pytest.approx(o1, rel=1e-13) and s2 == pytest.approx(o2, rel=1e-13)
    N = 1 << 14
    d = r.standard_normal(N)
    S = r.uniform(0.5, 3.0, N // 2)
    got = Q.dropin_whiten(d, S)
    want = O.whiten(d, S)
    assert np.allclose(got, want, rtol=1e-15, atol=0)
    assert got[0] == 0.0 and got[N // 2] == 0.0
