"""Resident PT ensembles in the serving loop (qcbc_server_admit_pt,
qcbc_server_submit_steps): each admitted analysis keeps its states, logL,
proposals, temperatures and random-number counter on the GPU, and step
requests from several analyses share each batch's persistent queue launch
with likelihood requests.

Checks, in FP64 and in mixed precision:
- an ensemble's chain after n steps is bit-identical alone on its server and
  in a mix with two other stepping ensembles and concurrent likelihood
  requests (params, logL, accepted moves, swaps);
- its logL is the oracle's likelihood of its current points (FP64 1e-8
  absolute; mixed 0.02, the mixed path's stated error at these SNRs);
- evicting an analysis with steps pending cancels them and frees its budget.
"""
import numpy as np
import pytest

from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu

SIGMA = np.array([2e-5, 1e-4, 2e-3, 2e-3, 0.05, 2e-5, 0.01, 0.01, 0.01, 0.02, 0.01])


def member(which, nt, nc, seed):
    an, inj, D, SN = _data.analysis(which)
    init = W.ensemble_around(inj, nt * nc, seed=seed)
    return an, D, SN, nt, nc, init


def admit(s, m, seed):
    an, D, SN, nt, nc, init = m
    code, aid = s.admit_pt(an, D, SN, nt, nc, 1.2, init, SIGMA, seed)
    assert code == 0
    return aid


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_chain_alone_equals_in_mix(prec):
    print("serving PT: precision %d" % prec, flush=True)
    X = member(1, 4, 16, 31)
    Y = member(0, 8, 32, 32)
    Z = member(1, 2, 64, 33)
    nsteps = 6
    alone = Q.Server(8 * 1024 ** 3, prec)
    ax = admit(alone, X, 101)
    for _ in range(nsteps):
        code = alone.wait(alone.submit_steps(ax, 1))[0]
        assert code == 0
    ref = alone.get_pt(ax)
    alone.close()
    assert ref["steps"] == nsteps and ref["accepted"].sum() > 0

    mix = Q.Server(8 * 1024 ** 3, prec)
    mx = admit(mix, X, 101)
    my = admit(mix, Y, 202)
    mz = admit(mix, Z, 303)
    # a likelihood-only analysis beside the ensembles
    an1, inj1, D1, SN1 = _data.analysis(1)
    code, ml = mix.admit(an1, D1, SN1)
    assert code == 0
    P = W.ensemble_around(inj1, 48, seed=7)
    tickets = [mix.submit_steps(my, 4), mix.submit_steps(mz, 3)]
    for k in range(nsteps):
        tickets.append(mix.submit_steps(mx, 1))
        tickets.append(mix.submit(ml, P))
    for t in tickets:
        assert mix.wait(t)[0] == 0
    got = mix.get_pt(mx)
    assert got["steps"] == nsteps
    for k in ("params", "logL", "accepted", "swaps"):
        assert np.array_equal(got[k], ref[k]), k
    assert mix.get_pt(my)["steps"] == 4 and mix.get_pt(mz)["steps"] == 3
    # the resident logL is the likelihood of the resident points
    an, D, SN = X[0], X[1], X[2]
    Lo, _, so = _data.oracle_loglike(an, D, SN, got["params"])
    tol = 1e-8 if prec == Q.FP64 else 0.02
    print("  resident logL against the oracle: max |dlogL| %.2e" % np.abs(got["logL"] - Lo).max(), flush=True)
    assert (so == 0).all() and np.abs(got["logL"] - Lo).max() <= tol
    mix.close()


def test_evict_cancels_pending_steps():
    s = Q.Server(8 * 1024 ** 3, Q.FP64)
    a = admit(s, member(0, 4, 16, 41), 7)
    b = admit(s, member(1, 4, 16, 42), 8)
    before = s.stats()["bytes_in_use"]
    t = s.submit_steps(b, 400)
    s.wait(s.submit_steps(a, 2))
    assert s.evict(b) == 0
    code = s.wait(t)[0]
    assert code in (0, 104)  # done already, or QCBC_ECANCELED
    assert s.stats()["bytes_in_use"] < before and s.stats()["analyses"] == 1
    with pytest.raises(Q.QcbcError):
        s.submit_steps(b, 1)
    assert s.get_pt(a)["steps"] == 2
    s.close()
