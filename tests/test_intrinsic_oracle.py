"""Pins of the oracle's reference-frame likelihood (CPU only).

templates (QuickCBC.c:3527-3611) and log_likelihood_intrinsic (:4093-4145)
describe the signal in detector 0's frame: phase, time and amplitude of
detectors 1, 2 relative to it instead of a sky position.  A full point p
maps onto that form exactly: phi0 = (lambda_0 + p[4]) / 2, tc = p[5] + td_0,
ln DL - ln Fs_0, and per detector d the offsets lambda_d - lambda_0,
td_d - td_0 and the ratio Fs_d / Fs_0 (fulltemplates' Fs, lambda, td).  So:

- templates of the mapped point is fulltemplates of p in every detector
  (above fbegin of both, which differ by the shift of tc);
- log_likelihood_intrinsic of the mapped point is log_likelihood_full of p
  when the signal is loud in two detectors (a band where fbegin cuts no
  bin), and 0 when it is not (the two-detector SNR rule), except with one
  detector.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import workloads as W
from tests import _data


def to_intrinsic(an, p):
    """The reference-frame parameters of the full point p."""
    ciota = p[10]
    Ap, Ac = (1 + ciota ** 2) / 2, -ciota
    dt = O.times(p[7], p[8], an.gmst)
    Fs, lam, td = [], [], []
    for d in an.labels:
        Fp, Fc = O.f_ant(p[9], p[7], p[8], an.gmst, d)
        Fs.append(np.sqrt(Ap ** 2 * Fp ** 2 + Ac ** 2 * Fc ** 2))
        l = np.arctan2(Ac * Fc, Ap * Fp)
        lam.append(l + 2 * np.pi if l < 0 else l)
        td.append(dt[d])
    q = [p[0], p[1], p[2], p[3], 0.5 * (lam[0] + p[4]), p[5] + td[0], p[6] - np.log(Fs[0])]
    for d in range(1, an.nifo):
        q += [lam[d] - lam[0], td[d] - td[0], Fs[d] / Fs[0]]
    return np.array(q)


@pytest.mark.parametrize("which", [0, 1])
def test_templates_are_fulltemplates_in_the_reference_frame(which):
    an, inj, D, SN = _data.analysis(which)
    for p in np.vstack([inj, _data.points(which, 3, seed=12)]):
        q = to_intrinsic(an, p)
        s1, hf = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, p)
        s2, hi = O.templates(an.nifo, an.fref, an.N, an.Tobs, q)
        assert s1 == 0 and s2 == 0
        f = np.arange(an.N // 2) / an.Tobs
        k = np.nonzero(f > 1.01 * max(O.fbegin(p), O.fbegin(q)))[0]
        k = k[k > 0]
        scale = np.abs(hf).max()
        for d in range(an.nifo):
            assert np.abs(hi[d, k] - hf[d, k]).max() <= 1e-9 * scale
            assert np.abs(hi[d, an.N - k] - hf[d, an.N - k]).max() <= 1e-9 * scale


def test_intrinsic_likelihood_is_the_full_likelihood_and_the_snr_rule():
    an, inj, D, SN = _data.analysis(0)  # fbegin below the band: no bin is cut
    P = np.vstack([inj, W.ensemble_around(inj, 5, seed=14, scale=0.3)])
    Lf, _, sf = O.loglike_batch(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P)
    Q = np.array([to_intrinsic(an, p) for p in P])
    Li, s2, si = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Q)
    assert (sf == 0).all() and (si == 0).all()
    loud = (s2 > 25).sum(1) >= 2
    assert loud[0]  # the injection is loud in both detectors
    assert np.allclose(Li[loud], Lf[loud], rtol=1e-10, atol=1e-8)
    # far away: quiet in both, logL 0 whatever the sum
    Qq = Q.copy()
    Qq[:, 6] += np.log(1.0e4)
    Lq, s2q, _ = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Qq)
    assert (s2q < 25).all() and (Lq == 0).all()
    # loud in one detector only: 0 as well
    Q1 = Q[:1].copy()
    Q1[0, 9] = 1e-3  # detector 1 nearly silent
    L1, s21, _ = O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Q1)
    assert s21[0, 0] > 25 and s21[0, 1] < 25 and L1[0] == 0.0
