"""Pins of the oracle's Fisher_Full (QuickCBC.c:2551-2688, AmpPhase :3395-3444)
by facts that do not restate its formulas.  AmpPhase is the bare h22 of
templates() (QuickCBC.c:3527-3611) without the 1/sqrt(Tobs) and the mask,
so on a band above fbegin, with h_l the templates of detector l and
(a|b) = sum_l 4 Re(a_l conj(b_l)) / S_l over the band:

- F[6][6] (ln DL) = Tobs (h|h): the amplitude falls as 1/DL;
- F[4][4] = 4 Tobs (h|h) and F[4][6] = 0 exactly: phi0 enters the model
  phase twice (templates hands the model 2 phi0 / 2) and not the amplitude;
- F[5][5] (tc) = Tobs sum 4 (2 pi f)^2 |h|^2 / S;
- u^T F u = Tobs |h(p + e u) - h(p - e u)|^2 / (4 e^2) along random u;
- a far source (every diagonal entry below 1) gives diagonal entries of 1;
- a spin at 0.9999 gives 1e4 I; d < 7 is the top-left block of d = 7.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _data


def band_inner(an, a, b, SN, lo, hi):
    k = np.arange(lo, hi)
    return sum(4.0 * np.sum((a[d, k] * b[d, k] + a[d, an.N - k] * b[d, an.N - k]) / SN[d, k]) for d in range(an.nifo))


def setup(which):
    an, inj, D, SN = _data.analysis(which)
    q = _data.ref_frame(an, inj)
    # a band above fbegin, where the templates' mask is inactive
    fmin = max(an.fmin, 1.5 * O.fbegin(q) + 5.0)
    lo, hi = int(an.Tobs * fmin), min(int(an.Tobs * an.fmax), an.N // 2)
    return an, q, SN, fmin, lo, hi


@pytest.mark.parametrize("which", [0, 1])
def test_fisher_full_closed_forms(which):
    an, q, SN, fmin, lo, hi = setup(which)
    st, F = O.fisher_full(an.nifo, an.fref, fmin, an.fmax, SN, an.N, an.Tobs, q)
    assert st == 0 and np.array_equal(F, F.T)
    st, h = O.templates(an.nifo, an.fref, an.N, an.Tobs, q)
    assert st == 0
    hh = band_inner(an, h, h, SN, lo, hi)
    assert F[6, 6] == pytest.approx(an.Tobs * hh, rel=1e-7)
    assert F[4, 4] == pytest.approx(4.0 * an.Tobs * hh, rel=1e-7)
    assert F[4, 6] == 0.0
    f = np.arange(an.N // 2 + 1) / an.Tobs
    w = np.zeros(an.N)
    w[:an.N // 2 + 1] = 2.0 * np.pi * f
    w[an.N // 2 + 1:] = (2.0 * np.pi * f[1:an.N // 2])[::-1]
    hw = h * w[None, :]
    assert F[5, 5] == pytest.approx(an.Tobs * band_inner(an, hw, hw, SN, lo, hi), rel=1e-6)


@pytest.mark.parametrize("which", [0, 1])
def test_fisher_full_is_the_template_metric(which):
    an, q, SN, fmin, lo, hi = setup(which)
    st, F = O.fisher_full(an.nifo, an.fref, fmin, an.fmax, SN, an.N, an.Tobs, q)
    rng = np.random.default_rng(20 + which)
    eps = 1e-4
    sig = np.sqrt(1.0 / np.diag(F))
    for _ in range(4):
        u = np.zeros(q.size)
        u[:7] = rng.standard_normal(7) * sig
        _, hp = O.templates(an.nifo, an.fref, an.N, an.Tobs, q + eps * u)
        _, hm = O.templates(an.nifo, an.fref, an.N, an.Tobs, q - eps * u)
        dh = hp - hm
        metric = an.Tobs * band_inner(an, dh, dh, SN, lo, hi) / (4 * eps * eps)
        assert u[:7] @ F @ u[:7] == pytest.approx(metric, rel=2e-3), (u[:7] @ F @ u[:7], metric)


def test_fisher_full_floor_limits_and_blocks():
    an, q, SN, fmin, lo, hi = setup(1)
    far = q.copy()
    far[6] += 40.0  # the amplitude times e^-40: every entry far below 1
    st, F = O.fisher_full(an.nifo, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, far)
    assert st == 0 and np.array_equal(np.diag(F), np.ones(7))
    assert np.abs(F - np.diag(np.diag(F))).max() < 1e-20
    spin = q.copy()
    spin[3] = -0.9999
    st, F = O.fisher_full(an.nifo, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, spin)
    assert st == 0 and np.array_equal(F, 1.0e4 * np.eye(7))
    st, F7 = O.fisher_full(an.nifo, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, q)
    for d in (1, 3, 5):
        st, Fd = O.fisher_full(an.nifo, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, q, d=d)
        assert np.array_equal(Fd, F7[:d, :d])
