"""Pins of the oracle's heterodyne grid choice (limits, Fisher_All, diff,
het_diff, efix, het_space; QuickCBC.c:4275-5053, 2695-2817) by facts that do
not restate its formulas:

- limits: the band's SNR^2 is the template inner product (h|h) of
  fulltemplates over the same bins; MN, MM are where the running sum crosses
  0.001 and comes within 0.001 of the total;
- Fisher_All: u^T F u equals the central second difference of the
  template distance (h(p + e u) - h(p - e u) | same) / (4 e^2) along random
  directions u (a band above fbegin, where the two waveforms coincide);
  F[4][6] = 0 exactly; F[4][4] = F[6][6] = (h|h);
- diff: 0 at the same point; (h|h)/4 at half the amplitude (distance
  doubled); 4 (h|h) at the phase shifted by pi;
- het_diff: 0 at the same point; (h|h)/4 over the grid's bins at half the
  amplitude (the Legendre interpolation of a constant is exact); close to
  diff near the reference point;
- efix: its contract, the chosen scale puts the distance within 0.2 of zs
  (or the direction is dropped);
- het_space: a grid of J = 2 equal steps per region from MN, ending at MM
  (or, the last two steps being equal, one bin below),
  and a heterodyne on it reproduces the full logL at the eleven displaced
  points efix chose (distance 30 from the reference).
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import workloads as W
from tests import _data


def _band_inner(an, h, SN, lo, hi):
    k = np.arange(lo, hi)
    return sum(4.0 * np.sum((h[d, k] ** 2 + h[d, an.N - k] ** 2) / SN[d, k]) for d in range(an.nifo))


@functools.lru_cache(maxsize=None)
def _space(which):
    an, inj, D, SN = _data.analysis(which)
    lo, hi = W.prior_box(an)
    st, h = O.het_space(an.labels, an.gmst, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, inj, lo, hi)
    assert st == 0
    return h


@pytest.mark.parametrize("which", [0, 1])
def test_limits_band_is_the_template_snr(which):
    an, inj, D, SN = _data.analysis(which)
    st, (MN, MM), ratio = O.limits(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj)
    assert st == 0
    st, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj)
    fs = O.fbegin(inj)
    jmin = int(an.Tobs * max(an.fmin, fs))
    M = min(int(max(2 * O.fringdown(inj), 64.0) * an.Tobs), an.N // 2)
    # fulltemplates is zero at f <= fbegin and agrees with fullphaseamp above;
    # the first bin (f <= fbegin) carries fullphaseamp's taper only
    x = np.array([_band_inner(an, h, SN, j, j + 1) for j in range(jmin, M)])
    cum = np.cumsum(x)
    total = cum[-1]
    first = jmin + int(np.argmax(cum > 0.001))
    assert abs(first - MN) <= 1
    last = jmin + int(np.argmax(total - cum < 0.001))
    assert abs(last - MM) <= 1 and MN < MM < M
    # ratio: 1/(density / band average) inside the band; the average also
    # holds the tapered first bin, which fulltemplates leaves out, so the
    # shape is compared exactly and the scale to that bin's share
    avg = total / (MM - MN)
    j = np.arange(MN + 2, MM, 97)
    assert np.allclose(ratio[j[0]] / ratio[j], x[j - jmin] / x[j[0] - jmin], rtol=1e-9)
    assert np.allclose(1.0 / ratio[j], x[j - jmin] / avg, rtol=1e-3)


@pytest.mark.parametrize("which", [0, 1])
def test_fisher_all_is_the_template_metric(which):
    an, inj, D, SN = _data.analysis(which)
    # a band above fbegin, where fulltemplates' mask and fullphaseamp's taper agree
    fmin = max(an.fmin, 1.5 * O.fbegin(inj) + 5.0)
    st, F = O.fisher_all(an.labels, an.gmst, an.fref, fmin, an.fmax, SN, an.N, an.Tobs, inj)
    assert st == 0
    assert np.array_equal(F, F.T)
    assert F[4, 6] == 0.0
    lo, hi = int(an.Tobs * fmin), min(int(an.Tobs * an.fmax), an.N // 2)
    st, h0 = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj)
    hh = _band_inner(an, h0, SN, lo, hi)
    assert F[4, 4] == pytest.approx(hh, rel=1e-6)
    assert F[6, 6] == pytest.approx(hh, rel=1e-6)
    rng = np.random.default_rng(3 + which)
    eps = 1e-4
    sig = np.sqrt(1.0 / np.maximum(np.diag(F), 1e-30))
    for _ in range(4):
        u = rng.standard_normal(11) * sig
        u[[4, 7, 9]] *= 0.3
        _, hp = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj + eps * u)
        _, hm = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj - eps * u)
        metric = _band_inner(an, hp - hm, SN, lo, hi) / (4 * eps * eps)
        assert u @ F @ u == pytest.approx(metric, rel=2e-3), (u @ F @ u, metric)


@pytest.mark.parametrize("which", [0, 1])
def test_diff_closed_forms(which):
    an, inj, D, SN = _data.analysis(which)
    z0 = O.diff(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj, inj)
    assert z0 == 0.0
    lo = int(an.Tobs * max(an.fmin, O.fbegin(inj))) + 1
    hi = min(int(max(2 * O.fringdown(inj), 64.0) * an.Tobs), an.N // 2)
    _, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj)
    hh = _band_inner(an, h, SN, lo, hi)
    half = inj.copy()
    half[6] += np.log(2.0)
    z = O.diff(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj, half)
    assert z == pytest.approx(hh / 4.0, rel=2e-3)  # one tapered edge bin apart
    flip = inj.copy()
    flip[4] += np.pi
    z = O.diff(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj, flip)
    assert z == pytest.approx(4.0 * hh, rel=2e-3)


@pytest.mark.parametrize("which", [0, 1])
def test_het_diff_against_closed_form_and_diff(which):
    an, inj, D, SN = _data.analysis(which)
    sp = _space(which)
    st, het = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, sp["fgrid"], 2, sp["MN"],
                           sp["MM"])
    assert st == 0
    assert O.het_diff(an.labels, an.gmst, an.fref, an.Tobs, het, inj, inj) == 0.0
    half = inj.copy()
    half[6] += np.log(2.0)
    z = O.het_diff(an.labels, an.gmst, an.fref, an.Tobs, het, inj, half)
    # |h - h/2|^2 = (h|h)/4 over the grid's bins [MN, last grid bin], the
    # last region boundary counted twice (the reference's overcount
    # correction stops at NR - 2)
    _, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, inj)
    g = sp["fgrid"]
    hh = _band_inner(an, h, SN, g[0], g[-1] + 1) + _band_inner(an, h, SN, g[-3], g[-3] + 1)
    # the grid's first bin may lie at f <= fbegin, where fulltemplates is 0
    # and fullphaseamp tapered: then it adds about the next bin's share
    excess = z - hh / 4.0
    if O.fbegin(inj) < g[0] / an.Tobs:
        assert abs(excess) <= 1e-9 * hh
    else:
        assert 0.0 < excess <= 0.26 * _band_inner(an, h, SN, g[0] + 1, g[0] + 2)
    lo, hi = W.prior_box(an)
    for i in range(11):
        px = O.displace(inj, sp["evec"][i], sp["eval"][i], lo, hi)
        zf = O.diff(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj, px)
        zh = O.het_diff(an.labels, an.gmst, an.fref, an.Tobs, het, inj, px)
        assert abs(zh - zf) <= 0.05 * zf + 0.5, (i, zh, zf)


@pytest.mark.parametrize("which", [0, 1])
def test_efix_meets_its_target(which):
    an, inj, D, SN = _data.analysis(which)
    lo, hi = W.prior_box(an)
    st, F = O.fisher_all(an.labels, an.gmst, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, inj)
    ej, ev, _ = O.fisher_evec(F)
    for zs in (30.0, 5.0):
        e, trials = O.efix(an.labels, an.gmst, an.fref, an.Tobs, inj, ej, ev, lo, hi, zs, fmin=an.fmin, SN=SN,
                           N=an.N)
        assert (trials >= 2).all() and (trials <= 25).all()
        good = 0
        for i in range(11):
            if e[0, i] == 0.0:
                continue
            px = O.displace(inj, ev[0, i], e[0, i], lo, hi)
            z = O.diff(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj, px)
            assert abs(z - zs) <= zs, (i, z)
            good += abs(z - zs) <= 0.2
        assert good >= 8, good


@pytest.mark.parametrize("which", [0, 1])
def test_het_space_grid_and_accuracy(which):
    an, inj, D, SN = _data.analysis(which)
    sp = _space(which)
    g, M, NR, MN, MM = sp["fgrid"], sp["M"], sp["NR"], sp["MN"], sp["MM"]
    st, (lmn, lmm), _ = O.limits(an.labels, an.gmst, an.fref, an.fmin, SN, an.N, an.Tobs, inj)
    assert (MN, MM) == (lmn, lmm)
    # the last region refitted to two equal steps ending within a step of 1 below MM
    assert M == 2 * NR + 1 and g[0] == MN and MM - 2 < g[-1] <= MM
    assert (np.diff(g) > 0).all()
    steps = np.diff(g).reshape(NR, 2)
    assert (steps[:, 0] == steps[:, 1]).all()
    assert steps.max() <= int(100.0 * an.Tobs) + 1
    st, het = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, g, 2, MN, MM)
    lo, hi = W.prior_box(an)
    P = np.array([O.displace(inj, sp["evec"][i], sp["eval"][i], lo, hi) for i in range(11)])
    Lh, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, het, P)
    Lf, _, _ = O.loglike_batch(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P)
    assert np.abs(Lh - Lf).max() <= 1.0, np.abs(Lh - Lf)
    # a grid of every other sample (one region in two dropped) is worse
    g2 = np.concatenate([g[:1], g[2::2]])
    g2 = g2[: 2 * ((g2.size - 1) // 2) + 1]
    st, het2 = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, g2, 2, MN, MM)
    L2, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, het2, P)
    assert np.abs(L2 - Lf).max() > np.abs(Lh - Lf).max()
