"""Pins of the oracle's restatement of QuickCBC's PT step (CPU only).

update() (QuickCBC.c:2238-2405) and MCMC_all's iteration (:2029-2092) with
counter-based random numbers.  Checked against what the mathematics and the
reference's structure fix rather than against the formulas retyped:

- the prior term pMcM = 2 ln Mt + ln eta - ln(1 - 4 eta)/2 is, up to a
  constant, the log of the Jacobian |d(m1, m2)/d(ln Mc, ln Mt)| that makes
  the prior flat in the component masses (finite differences of the mass
  map of QuickCBC.c:3705-3718); the flag rejects eta > 1/4, eta < etamin,
  the box edges and NaN;
- a Fisher proposal moves along one eigendirection, a differential-evolution
  proposal by a multiple of the difference of two distinct history rows, a
  jiggle by one of the three scales 1e-2, 1e-3, 1e-4 times sqrt(T); the
  types come with update()'s frequencies (alpha > a, > b, else), the
  normals are standard, the angles land in range;
- the sky ring: Ring_all keeps the time delay between its two detectors,
  skymap_all gives the new sky location the extrinsics that leave both
  detectors' complex signals unchanged, and the density of the map times
  that of the map back is 1;
- a run keeps its bookkeeping: states inside the prior, logL equal to the
  heterodyned likelihood of the state, the history made of the column's
  states, swaps on about a fifth of the steps.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import workloads as W
from tests import _data


def _uconst(lo, hi, a=0.5, b=0.3, c=0.0, etamin=0.05, an=None):
    sky = [an.gmst, an.nifo] + list(an.labels) + [0] * (3 - an.nifo) if an is not None else [0.0] * 5
    return np.concatenate([lo, hi, [a, b, c, etamin], sky])


def _masses(lnMc, lnMt):
    eta = np.exp((5.0 / 3.0) * (lnMc - lnMt))
    dm = np.sqrt(1.0 - 4.0 * eta)
    Mt = np.exp(lnMt)
    return Mt * (1 + dm) / 2, Mt * (1 - dm) / 2


def test_mass_jacobian_is_flat_in_m1_m2():
    an = W.config1()
    lo, hi = W.prior_box(an)
    uc = _uconst(lo, hi)
    r = np.random.default_rng(1)
    vals = []
    for _ in range(12):
        m1, m2 = np.sort(r.uniform(1.0, 80.0, 2))[::-1]
        if m2 / m1 < 0.08:
            continue
        p = W.make_params(m1, m2, 0.1, 0.1, 1.0, an.Tobs - 2, 100.0, 1.0, 0.2, 1.0, 0.3)
        f, j = O.update_prior(p, uc)
        assert f == 0
        h = 1e-6
        J = np.zeros((2, 2))
        for k in range(2):
            dp = np.zeros(2)
            dp[k] = h
            a = np.array(_masses(p[0] + dp[0], p[1] + dp[1]))
            b = np.array(_masses(p[0] - dp[0], p[1] - dp[1]))
            J[:, k] = (a - b) / (2 * h)
        vals.append(j - np.log(abs(np.linalg.det(J))))
    vals = np.array(vals)
    assert vals.size >= 6 and np.ptp(vals) < 1e-6, vals  # the same constant at every point


def test_prior_flag_cases():
    an = W.config1()
    lo, hi = W.prior_box(an)
    uc = _uconst(lo, hi)
    p = W.make_params(30.0, 20.0, 0.1, -0.1, 1.0, an.Tobs - 2, 400.0, 1.0, 0.2, 1.0, 0.3)
    assert O.update_prior(p, uc)[0] == 0
    q = p.copy()
    q[0] = q[1] + 0.6 * np.log(0.25) + 1e-6  # eta just above 1/4
    assert O.update_prior(q, uc)[0] == 1
    q = W.make_params(100.0, 4.0, 0.1, -0.1, 1.0, an.Tobs - 2, 400.0, 1.0, 0.2, 1.0, 0.3)  # eta 0.037 < 0.05
    assert O.update_prior(q, uc)[0] == 1
    for i in range(11):
        q = p.copy()
        q[i] = hi[i] + 1e-9 * max(1.0, abs(hi[i]))
        assert O.update_prior(q, uc)[0] == 1, i
    q = p.copy()
    q[5] = np.nan
    assert O.update_prior(q, uc)[0] == 1


def _proposal_setup():
    an = W.config1()
    lo, hi = W.prior_box(an)
    inj = W.injection1()
    x = W.ensemble_in_prior(inj, 1, lo, hi, seed=3)[0]
    r = np.random.default_rng(5)
    Q, _ = np.linalg.qr(r.standard_normal((11, 11)))
    ev = Q.T.copy()  # rows: orthonormal directions
    ej = np.exp(r.uniform(np.log(1e-5), np.log(1e-2), 11))
    hist = W.ensemble_in_prior(inj, 50, lo, hi, seed=4)
    return x, ej, ev, hist


def test_proposal_kinds_and_frequencies():
    x, ej, ev, hist = _proposal_setup()
    abc = np.array([0.5, 0.3, 0.0])
    sT = 1.7
    counts = np.zeros(4, int)
    zf = []
    for ctr in range(3000):
        t, y, _ = O.update_propose(x, ej, ev, hist, sT, abc, 77, ctr)
        counts[t] += 1
        d = y - x
        d[[4, 7, 9]] = 0.0  # the wrapped angles aside
        assert 0 <= y[4] <= 2 * np.pi and 0 <= y[7] <= 2 * np.pi and 0 <= y[9] <= np.pi
        if t == 1:
            # along exactly one eigendirection, to the rounding of y - x
            dw = y - x
            for k, per in ((4, 2 * np.pi), (7, 2 * np.pi), (9, np.pi)):
                dw[k] = (dw[k] + per / 2) % per - per / 2
            c = ev @ dw
            i = int(np.argmax(np.abs(c)))
            assert np.abs(np.delete(c, i)).max() <= 1e-9 * abs(c[i]) + 1e-15 * np.abs(x).max()
            zf.append(c[i] / (sT * ej[i]))
        elif t == 2:
            diff = hist[:, None, :] - hist[None, :, :]
            m = np.delete(np.arange(11), [4, 7, 9])
            with np.errstate(invalid="ignore", divide="ignore"):
                g = (y - x)[m][None, None, :] / np.where(diff[:, :, m] == 0, np.nan, diff[:, :, m])
            idx = np.arange(g.shape[0])
            g[idx, idx, :] = np.inf  # i == j is no pair
            with np.errstate(invalid="ignore"):
                spread = np.nanmax(g, axis=2) - np.nanmin(g, axis=2)
            spread[idx, idx] = np.inf
            i, j = np.unravel_index(np.nanargmin(spread), spread.shape)
            assert i != j and spread[i, j] < 1e-6 * (abs(np.nanmean(g[i, j])) + 1e-12)
        else:
            s = np.abs(d).max() / sT
            assert s > 0
    f = counts[1:] / counts.sum()
    assert abs(f[0] - 0.5) < 0.04 and abs(f[1] - 0.2) < 0.04 and abs(f[2] - 0.3) < 0.04, f
    zf = np.array(zf)
    assert abs(zf.mean()) < 0.1 and abs(zf.std() - 1.0) < 0.08


def test_jiggle_scales():
    x, ej, ev, hist = _proposal_setup()
    abc = np.array([1.0, 1.0, 0.0])  # every proposal a jiggle
    sT = 2.0
    n = np.zeros(3)
    for ctr in range(600):
        t, y, _ = O.update_propose(x, ej, ev, hist, sT, abc, 5, ctr)
        assert t == 3
        z = (y - x)[[0, 1, 2, 3, 5, 6, 8, 10]] / sT
        # the per-parameter normals share one scale 10^-(2+k): their rms is that scale times ~1
        k = int(np.round(-np.log10(np.sqrt(np.mean(z ** 2))) - 2.0))
        assert 0 <= k <= 2, k
        n[k] += 1
    assert (np.abs(n / n.sum() - 1.0 / 3.0) < 0.08).all(), n


def test_history_init_rows_are_column_states():
    n_temps, n_chains, n_hist = 4, 3, 10
    x = np.arange(n_temps * n_chains * 11, dtype=np.float64).reshape(-1, 11)
    h = O.pt_update_hist_init(x, n_temps, n_chains, n_hist, 9)
    for c in range(n_chains):
        col = {tuple(x[t * n_chains + c]) for t in range(n_temps)}
        rows = {tuple(r) for r in h[c].reshape(-1, 11)}
        assert rows <= col and len(rows) > 1


@pytest.mark.parametrize("c", [0.0, 0.2])
def test_run_keeps_its_bookkeeping(c):
    an, inj, D, SN = _data.analysis(0)
    h = _data.het(0)
    lo, hi = W.prior_box(an)
    uc = _uconst(lo, hi, c=c, an=an)
    n_temps, n_chains, n_hist, every, ladder, seed = 4, 3, 16, 2, 1.5, 11
    P = n_temps * n_chains
    x = W.ensemble_in_prior(inj, P, lo, hi, seed=8, scale=0.3)
    L, st = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, h, x)
    assert (st == 0).all()
    F, _ = O.fisher_het(an.labels, an.gmst, an.fref, an.Tobs, h, x)
    ej, ev, _ = O.fisher_evec(F)
    hist = O.pt_update_hist_init(x, n_temps, n_chains, n_hist, seed)
    y = np.zeros_like(x)
    typ = np.zeros(P, np.int32)
    jac = np.zeros(P)
    for t in range(n_temps):
        for col in range(n_chains):
            q = t * n_chains + col
            typ[q], y[q], jac[q] = O.update_propose(x[q], ej[q], ev[q], hist[col, t], np.sqrt(ladder ** t), uc[22:],
                                                    seed, q)
    S = dict(x=x, L=L, ej=ej, ev=ev, hist=hist, hcount=np.zeros(n_chains, np.int64), y=y, typ=typ, jac=jac,
             stats=np.zeros((n_chains, 11), np.int64), step=0)
    nsteps = 30
    O.pt_update_run(an.labels, an.gmst, an.fref, an.Tobs, h, uc, n_hist, every, n_temps, n_chains, ladder, seed,
                    nsteps, S)
    st = S["stats"]
    assert S["step"] == nsteps
    assert (st[:, 0] + st[:, 1] == nsteps).all()
    assert (st[:, 3] + st[:, 5] + st[:, 7] + st[:, 9] == n_temps * st[:, 0]).all()
    assert (st[:, 9].sum() > 0) == (c > 0)  # sky-ring proposals only with c > 0
    assert (st[:, 2] <= st[:, 1]).all() and st[:, 1].sum() > 0 and st[:, 4::2].sum() > 0
    for q in range(P):
        assert O.update_prior(S["x"][q], uc)[0] == 0
    Lx, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, h, S["x"])
    assert np.array_equal(Lx, S["L"])
    # histories: the column's rows are states the column held
    for c in range(n_chains):
        assert S["hcount"][c] >= 1
    assert 0.05 < st[:, 1].sum() / (nsteps * n_chains) < 0.4


def _sky_point(an, seed):
    lo, hi = W.prior_box(an)
    inj = W.injection1() if an.nifo == 3 else W.injection0()
    return W.ensemble_in_prior(inj, 1, lo, hi, seed=seed, scale=1.0)[0]


@pytest.mark.parametrize("d1,d2", [(0, 1), (1, 0), (0, 2), (2, 1)])
def test_ring_keeps_the_time_delay(d1, d2):
    an = W.config1()
    for k in range(5):
        x = _sky_point(an, 30 + k)
        y = O.ring_all(x, d1, d2, an.gmst, (k + 0.5) / 5.0)
        tx, ty = O.times(x[7], x[8], an.gmst), O.times(y[7], y[8], an.gmst)
        assert abs((tx[d1] - tx[d2]) - (ty[d1] - ty[d2])) < 1e-12
        assert -1.0 <= y[8] <= 1.0 and abs(y[7] - x[7]) + abs(y[8] - x[8]) > 1e-6
        assert np.array_equal(y[:7], x[:7]) and np.array_equal(y[9:], x[9:])


def test_skymap_keeps_the_two_detector_signals():
    an = W.config1()  # H, L, V
    f = an.freq()[1000:30000:97]
    for k, (d1, d2) in enumerate([(0, 1), (0, 2), (1, 2)]):
        x = _sky_point(an, 40 + k)
        y = O.ring_all(x, d1, d2, an.gmst, 0.37 + 0.2 * k)
        y = O.skymap_all(x, y, an.gmst, d1, d2)
        sx, Ax, Px = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f, x)
        sy, Ay, Py = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, f, y)
        assert sx == 0 and sy == 0
        fs = max(O.fbegin(x), O.fbegin(y)) * 1.05
        m = f > fs
        hx = Ax[:, m] * np.exp(1j * Px[:, m])
        hy = Ay[:, m] * np.exp(1j * Py[:, m])
        scale = np.abs(hx).max()
        for d in (d1, d2):
            assert np.abs(hy[d] - hx[d]).max() <= 1e-6 * scale, (d1, d2, d)
        third = 3 - d1 - d2
        assert np.abs(hy[third] - hx[third]).max() > 1e-3 * scale  # the third detector sees a change


def test_sky_map_round_trip_and_density():
    # the map back onto x's sky location recovers x's extrinsics (up to the
    # other solution branch, phi + pi, psi + pi/2), and where the map is not
    # near singular the two densities multiply to 1 (the reference's
    # one-sided epsilon = 1e-6 Jacobian has truncation error: measured median
    # 2e-5 in log, growing as the density leaves 1e-3 .. 1e3)
    an = W.config1()
    good = 0
    for k in range(30):
        d1, d2 = [(0, 1), (1, 2), (2, 0)][k % 3]
        x = _sky_point(an, 50 + k)
        y = O.ring_all(x, d1, d2, an.gmst, (0.21 + 0.37 * k) % 1.0)
        dxy, y = O.skydensity_all(x, y, an.gmst, d1, d2)
        xb = y.copy()
        xb[7], xb[8] = x[7], x[8]
        dyx, xb = O.skydensity_all(y, xb, an.gmst, d1, d2)
        dpsi = (xb[9] - x[9]) % (np.pi / 2)
        dphi = (xb[4] - x[4]) % np.pi
        assert abs(xb[6] - x[6]) < 1e-8 and abs(xb[10] - x[10]) < 1e-8 and abs(xb[5] - x[5]) < 1e-12
        assert min(dpsi, np.pi / 2 - dpsi) < 1e-8 and min(dphi, np.pi - dphi) < 1e-8
        if 1e-3 < dxy < 1e3:
            good += 1
            assert abs(np.log(dxy * dyx)) < 1e-3, (k, dxy, dyx)
    assert good >= 10
