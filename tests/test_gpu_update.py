"""QuickCBC's own PT step on the GPU against the oracle's replay.

update()'s proposal mix (QuickCBC.c:2238-2405: Fisher, differential
evolution, jiggle and the sky ring) and MCMC_all's step (:2029-2092) on the
heterodyned likelihood, resident on the GPU (qcbc_pt_create_update).  The oracle replays the same chains from the same
counter-based random numbers (its own heterodyned logL, the GPU's Fisher
jumps): proposal types, accepted moves, swaps and history fills must agree
exactly, states to rounding (the GPU's log, cos and pow are within a few ulps
of glibc's, so a proposal differs from the oracle's by ~1e-15 relative; a
sky-ring move's finite-difference density amplifies that to ~1e-9 of its
log on the small ensemble, ~1e-7 on the PT ensemble's full 32 x 64 shape and
~1e-5 where the move maps near-degenerately; logL by <= 1e-8 as the
heterodyned-likelihood tests).  Column groups and graph replay give
bit-identical chains.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu

N_TEMPS, N_CHAINS, N_HIST, EVERY, LADDER, SEED = 4, 6, 16, 2, 1.5, 11


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    an, inj, D, SN = _data.analysis(request.param)
    h = _data.het(request.param)
    c = Q.Context(an, D, SN)
    g = Q.Het(c, arrays=h)
    lo, hi = W.prior_box(an)
    init = W.ensemble_in_prior(inj, N_TEMPS * N_CHAINS, lo, hi, seed=8, scale=0.3)
    yield an, h, c, g, lo, hi, init
    g.close()
    c.close()


def _make(setup, groups, n_hist=N_HIST, c=0.2):
    an, h, ctx, g, lo, hi, init = setup
    desc = Q.update_desc(lo, hi, c=c, n_hist=n_hist, hist_every=EVERY)
    return Q.PTUpdate(ctx, g, N_TEMPS, N_CHAINS, LADDER, init, desc, seed=SEED, groups=groups)


def _uconst(lo, hi, an, c=0.2):
    sky = [an.gmst, an.nifo] + list(an.labels) + [0] * (3 - an.nifo)
    return np.concatenate([lo, hi, [0.5, 0.3, c, 0.05], sky])


def test_start_matches_the_oracle(setup):
    an, h, c, g, lo, hi, init = setup
    pt = _make(setup, 2)
    x, L, _, _ = pt.get()
    ej, ev, y, typ, jac, stats = pt.get_update()
    hist, cnt = pt.get_history()
    assert np.array_equal(x, init)
    Lo, so = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, h, x)
    assert (so == 0).all() and np.abs(L - Lo).max() <= 1e-8
    # the jump sizes are FisherEvec of Fisher_Het at each state (the Fisher tests pin those)
    F, ej2, ev2, _ = g.fisher(x)
    assert np.array_equal(ej, ej2) and np.array_equal(ev, ev2)
    assert np.array_equal(hist, O.pt_update_hist_init(x, N_TEMPS, N_CHAINS, N_HIST, SEED))
    assert (cnt == 0).all() and (stats == 0).all()
    uc = _uconst(lo, hi, an)
    for q in range(N_TEMPS * N_CHAINS):
        t, col = divmod(q, N_CHAINS)
        to, yo, jo = O.update_propose(x[q], ej[q], ev[q], hist[col, t], np.sqrt(LADDER ** t), uc[22:], SEED, q)
        assert typ[q] == to
        assert np.abs(y[q] - yo).max() <= 1e-12 * (1.0 + np.abs(x[q]).max()), (q, typ[q], y[q] - yo)
        assert abs(jac[q] - jo) <= 1e-6 * (1.0 + abs(jo)), (q, jac[q], jo)
    pt.close()


@pytest.mark.parametrize("c", [0.0, 0.2])
def test_chains_match_the_oracle_replay(setup, c):
    an, h, ctx, g, lo, hi, init = setup
    pt = _make(setup, 2, c=c)
    x, L, _, _ = pt.get()
    ej, ev, y, typ, jac, stats = pt.get_update()
    hist, cnt = pt.get_history()
    Lo, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, h, x)
    S = dict(x=x.copy(), L=Lo, ej=ej.copy(), ev=ev.copy(), hist=hist.copy(), hcount=cnt.copy(), y=y.copy(),
             typ=typ.copy(), jac=jac.copy(), stats=stats.copy(), step=0)
    nsteps = 40
    print("update replay: %d det, %d x %d points, %d steps, c = %g" % (an.nifo, N_TEMPS, N_CHAINS, nsteps, c),
          flush=True)
    pt.run(nsteps)
    O.pt_update_run(an.labels, an.gmst, an.fref, an.Tobs, h, _uconst(lo, hi, an, c), N_HIST, EVERY, N_TEMPS,
                    N_CHAINS, LADDER, SEED, nsteps, S)
    x1, L1, _, _ = pt.get()
    ej1, ev1, y1, typ1, jac1, stats1 = pt.get_update()
    hist1, cnt1 = pt.get_history()
    assert np.array_equal(stats1, S["stats"]), (stats1, S["stats"])
    assert np.array_equal(cnt1, S["hcount"]) and np.array_equal(typ1, S["typ"])
    assert stats1[:, 4::2].sum() > 0 and stats1[:, 2].sum() > 0  # moves and swaps were accepted
    assert (stats1[:, 9].sum() > 0) == (c > 0)  # sky-ring proposals with c > 0
    tol = 1e-12 * (1.0 + np.abs(x1))
    assert (np.abs(x1 - S["x"]) <= tol).all()
    assert np.abs(L1 - S["L"]).max() <= 1e-8
    assert np.array_equal(ej1, S["ej"]) and np.array_equal(ev1, S["ev"])  # swapped with their states
    assert (np.abs(hist1 - S["hist"]) <= 1e-12 * (1.0 + np.abs(hist1))).all()
    assert (np.abs(y1 - S["y"]) <= 1e-12 * (1.0 + np.abs(y1))).all()
    assert (np.abs(jac1 - S["jac"]) <= 1e-6 * (1.0 + np.abs(jac1))).all()
    pt.close()


def test_chains_match_the_oracle_replay_full_shape():
    """The PT ensemble's full shape (32 temperatures x 64 chains, Tobs = 16 s, three
    detectors, the default three column groups), a shorter replay."""
    an, inj, D, SN = _data.analysis(1)
    h = _data.het(1)
    ctx = Q.Context(an, D, SN)
    g = Q.Het(ctx, arrays=h)
    lo, hi = W.prior_box(an)
    nt, nc, nsteps = 32, 64, 12
    init = W.ensemble_in_prior(inj, nt * nc, lo, hi, seed=9, scale=0.3)
    desc = Q.update_desc(lo, hi, c=0.2, n_hist=N_HIST, hist_every=EVERY)
    pt = Q.PTUpdate(ctx, g, nt, nc, LADDER, init, desc, seed=SEED, groups=3)
    x, L, _, _ = pt.get()
    ej, ev, y, typ, jac, stats = pt.get_update()
    hist, cnt = pt.get_history()
    Lo, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, h, x)
    S = dict(x=x.copy(), L=Lo, ej=ej.copy(), ev=ev.copy(), hist=hist.copy(), hcount=cnt.copy(), y=y.copy(),
             typ=typ.copy(), jac=jac.copy(), stats=stats.copy(), step=0)
    print("update replay: %d det, %d x %d points, %d steps" % (an.nifo, nt, nc, nsteps), flush=True)
    pt.run(nsteps)
    O.pt_update_run(an.labels, an.gmst, an.fref, an.Tobs, h, _uconst(lo, hi, an, 0.2), N_HIST, EVERY, nt, nc,
                    LADDER, SEED, nsteps, S)
    x1, L1, _, _ = pt.get()
    ej1, ev1, y1, typ1, jac1, stats1 = pt.get_update()
    hist1, cnt1 = pt.get_history()
    assert np.array_equal(stats1, S["stats"])
    assert np.array_equal(cnt1, S["hcount"]) and np.array_equal(typ1, S["typ"])
    assert stats1[:, 4::2].sum() > 0 and stats1[:, 2].sum() > 0
    assert (np.abs(x1 - S["x"]) <= 1e-12 * (1.0 + np.abs(x1))).all()
    assert np.abs(L1 - S["L"]).max() <= 1e-8
    assert np.array_equal(ej1, S["ej"]) and np.array_equal(ev1, S["ev"])
    assert (np.abs(hist1 - S["hist"]) <= 1e-12 * (1.0 + np.abs(hist1))).all()
    assert (np.abs(y1 - S["y"]) <= 1e-12 * (1.0 + np.abs(y1))).all()
    # the proposals' log densities: sky-ring moves carry the finite differences of the sky
    # map, ~1e-7 of their log at this size; where a move maps near-degenerately (log density
    # ~ -15) that grows to ~1e-5, and what enters the acceptance, the density e^jac (2e-7
    # there), still agrees to 3e-11.  Each proposal passes on its log (1e-6 relative) or on
    # its density (1e-9 absolute plus relative).
    d1, d0 = np.exp(jac1), np.exp(S["jac"])
    ok_log = np.abs(jac1 - S["jac"]) <= 1e-6 * (1.0 + np.abs(jac1))
    ok_density = np.abs(d1 - d0) <= 1e-9 * (1.0 + np.maximum(d1, d0))
    assert (ok_log | ok_density).all(), np.flatnonzero(~(ok_log | ok_density))
    pt.close()
    g.close()
    ctx.close()


def test_groups_and_graph_replay_give_the_same_chains(setup):
    a = _make(setup, 1)
    b = _make(setup, 3)
    gph = _make(setup, 3)
    a.run(70)
    b.run(70)
    # a 32-step graph twice, the remainder on the one-step graph, eager work in between
    gph.run(35, graph=True)
    gph.get()
    gph.run(35, graph=True)
    ra, rb, rg = a.get(), b.get(), gph.get()
    ua, ub, ug = a.get_update(), b.get_update(), gph.get_update()
    for k in (0, 1):
        assert np.array_equal(ra[k], rb[k]) and np.array_equal(ra[k], rg[k])
    for k in range(6):
        assert np.array_equal(ua[k], ub[k]) and np.array_equal(ua[k], ug[k])
    for p in (a, b, gph):
        p.close()


def test_refresh_fisher(setup):
    an, h, c, g, lo, hi, init = setup
    pt = _make(setup, 2)
    pt.run(12)
    pt.refresh_fisher()
    x, L, _, _ = pt.get()
    ej, ev, y, typ, jac, stats = pt.get_update()
    hist, cnt = pt.get_history()
    F, ej2, ev2, _ = g.fisher(x)
    assert np.array_equal(ej, ej2) and np.array_equal(ev, ev2)
    uc = _uconst(lo, hi, an)
    P = N_TEMPS * N_CHAINS
    for q in range(P):  # the pending proposals of step 12, remade with the new jumps
        t, col = divmod(q, N_CHAINS)
        to, yo, jo = O.update_propose(x[q], ej[q], ev[q], hist[col, t], np.sqrt(LADDER ** t), uc[22:], SEED,
                                      12 * P + q)
        assert typ[q] == to and np.abs(y[q] - yo).max() <= 1e-12 * (1.0 + np.abs(x[q]).max())
    pt.close()


def test_invalid_descriptions_are_rejected(setup):
    an, h, c, g, lo, hi, init = setup
    for kw in (dict(c=0.4), dict(a=0.2, b=0.3), dict(c=-0.1), dict(n_hist=1), dict(hist_every=0)):
        desc = Q.update_desc(lo, hi, **kw)
        with pytest.raises(RuntimeError):
            Q.PTUpdate(c, g, N_TEMPS, N_CHAINS, LADDER, init, desc, seed=SEED)
    bad_hi = hi.copy()
    bad_hi[3] = lo[3]
    with pytest.raises(RuntimeError):
        Q.PTUpdate(c, g, N_TEMPS, N_CHAINS, LADDER, init, Q.update_desc(lo, bad_hi), seed=SEED)


def test_efix_rescales_the_fisher_jumps(setup):
    """With efix_zs = 1 the ensemble's jumps are FisherEvec's rescaled by
    efix on the heterodyne distance (MCMC_all, QuickCBC.c:1922-1928), at
    creation and on refresh; the oracle's efix gives the same scales."""
    an, h, ctx, g, lo, hi, init = setup
    desc = Q.update_desc(lo, hi, c=0.2, n_hist=N_HIST, hist_every=EVERY, efix_zs=1.0)
    pt = Q.PTUpdate(ctx, g, N_TEMPS, N_CHAINS, LADDER, init, desc, seed=SEED, groups=2)
    x, _, _, _ = pt.get()
    ej, ev, _, _, _, _ = pt.get_update()
    F, ej0, ev0, _ = g.fisher(x)
    assert np.array_equal(ev, ev0)
    want, _ = g.efix(x, ej0, ev0, lo, hi, 1.0)
    assert np.array_equal(ej, want)
    eo, _ = O.efix(an.labels, an.gmst, an.fref, an.Tobs, x, ej0, ev0, lo, hi, 1.0, het=h)
    close = np.isclose(ej, eo, rtol=1e-6, atol=0).all(1)
    assert close.mean() >= 0.9, close
    assert (ej <= ej0 * 1.0e4).all() and (ej >= 0).all()
    pt.run(6)
    pt.refresh_fisher()
    x, _, _, _ = pt.get()
    ej, ev, _, _, _, _ = pt.get_update()
    F, ej0, ev0, _ = g.fisher(x)
    want, _ = g.efix(x, ej0, ev0, lo, hi, 1.0)
    assert np.array_equal(ev, ev0) and np.array_equal(ej, want)
    pt.close()
