"""The reference-frame likelihood on the GPU against the oracle.

templates (QuickCBC.c:3527-3611) through the plain path and
log_likelihood_intrinsic (:4093-4145) through the fused FP64 likelihood with
its own finalize (per-detector snrsq and the two-detector rule), batched and
as drop-ins.  Tolerances as the full likelihood's: templates 1e-9 of the
largest template value (the projected phase reaches ~1e5 rad), logL and
snrsq 1e-8 absolute plus 1e-12 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data
from tests.test_intrinsic_oracle import to_intrinsic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    which = request.param
    an, inj, D, SN = _data.analysis(which)
    c = Q.Context(an, D, SN)
    P = np.vstack([inj, _data.points(which, 31, seed=21)])
    Qp = np.array([to_intrinsic(an, p) for p in P])
    yield which, an, D, SN, c, Qp
    c.close()


def test_templates_match_the_oracle(setup):
    which, an, D, SN, c, Qp = setup
    hw, st = c.templates_intrinsic(Qp[:6])
    assert (st == 0).all()
    for b in range(6):
        so, ho = O.templates(an.nifo, an.fref, an.N, an.Tobs, Qp[b])
        assert so == 0
        assert np.abs(hw[b] - ho).max() <= 1e-9 * np.abs(ho).max(), b


def test_loglike_matches_the_oracle(setup):
    which, an, D, SN, c, Qp = setup
    print("intrinsic: config %d, %d points" % (which, Qp.shape[0]), flush=True)
    L, s2, st = c.loglike_intrinsic(Qp)
    Lo, s2o, so = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Qp)
    assert (st == so).all() and (st == 0).all()
    assert (np.abs(L - Lo) <= 1e-8 + 1e-12 * np.abs(Lo)).all(), np.abs(L - Lo).max()
    assert (np.abs(s2 - s2o) <= 1e-8 + 1e-12 * np.abs(s2o)).all()
    assert (L != 0).any()
    # the two-detector rule: far away every point is quiet and logL is 0
    Qq = Qp.copy()
    Qq[:, 6] += np.log(1.0e4)
    Lq, s2q, _ = c.loglike_intrinsic(Qq)
    Lqo, _, _ = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Qq)
    assert (Lq == 0).all() and (Lqo == 0).all() and (s2q < 25).all()


@pytest.mark.parametrize("B", [1, 3, 65])
def test_ragged_batches_and_domain_errors(setup, B):
    which, an, D, SN, c, Qp = setup
    P = np.resize(Qp, (B, Qp.shape[1])).copy()
    P[0, 2] = 1.5  # chi1 > 1: outside the model
    L, s2, st = c.loglike_intrinsic(P)
    Lo, _, so = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P)
    assert st[0] == 33 and so[0] == 33 and L[0] == 0.0
    assert (np.abs(L - Lo) <= 1e-8 + 1e-12 * np.abs(Lo)).all()
    L0, s20, st0 = c.loglike_intrinsic(np.zeros((0, Qp.shape[1])))
    assert L0.size == 0


def test_dropins_templates_and_log_likelihood_intrinsic(setup):
    which, an, D, SN, c, Qp = setup
    h = Q.dropin_templates(an, Qp[1])
    so, ho = O.templates(an.nifo, an.fref, an.N, an.Tobs, Qp[1])
    assert np.abs(h - ho).max() <= 1e-9 * np.abs(ho).max()
    Dr, Sr = Q.RowMatrix(D), Q.RowMatrix(SN)
    # the drop-ins take the band from qcbc_set_band (the reference's compiled 16-1024 Hz by default)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    try:
        L = Q.dropin_log_likelihood_intrinsic(an, Dr, Sr, Qp[0])
    finally:
        Q.lib().qcbc_set_band(16.0, 1024.0)
    Lo, _, _ = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Qp[:1])
    assert abs(L - Lo[0]) <= 1e-8 + 1e-12 * abs(Lo[0])
