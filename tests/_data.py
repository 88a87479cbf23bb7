"""Shared synthetic workloads for the tests, built with the CPU oracle."""
import functools

import numpy as np

from oracle import oracle as O
from quickcbc_amd import workloads as W


@functools.lru_cache(maxsize=None)
def analysis(which, noise=True):
    an = W.config0() if which == 0 else W.config1()
    inj = W.injection0() if which == 0 else W.injection1()

    def tmpl(p):
        st, h = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, p)
        assert st == 0
        return h

    D = W.make_data(an, tmpl, inj, with_noise=noise)
    SN = an.SN()
    return an, inj, D, SN


def oracle_loglike(an, D, SN, P, nthreads=0):
    return O.loglike_batch(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P, nthreads)


def points(which, n, seed=5):
    an, inj, D, SN = analysis(which)
    if which == 0:
        return W.prior_draws0(n, seed=seed)
    return W.ensemble_around(inj, n, seed=seed)


def het_band(an, ref, SN):
    """[MN, MM] of limits() (QuickCBC.c:4275-4385) from the oracle's fullphaseamp."""
    st, A, _ = O.fullphaseamp(an.labels, an.gmst, an.fref, an.Tobs, an.freq(), ref)
    assert st == 0
    return W.snr_band(an, A, SN, O.fbegin(ref))


@functools.lru_cache(maxsize=None)
def het(which, J=2, frac=1.0 / 64.0):
    """The oracle's heterodyne at the injection, on het_grid over het_band."""
    an, inj, D, SN = analysis(which)
    MN, MM = het_band(an, inj, SN)
    g = W.het_grid(MN, MM, an.Tobs, J=J, frac=frac)
    st, h = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, g, J, MN, MM)
    assert st == 0
    return h


def oracle_loglike_het(which, P, J=2, frac=1.0 / 64.0):
    an, inj, D, SN = analysis(which)
    return O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, het(which, J, frac), P)


def ref_frame(an, p):
    """The full point's reference-frame parameters (detector 0's frame)."""
    ciota = p[10]
    Ap, Ac = (1 + ciota ** 2) / 2, -ciota
    dt = O.times(p[7], p[8], an.gmst)
    Fs, lam = [], []
    for d in an.labels:
        Fp, Fc = O.f_ant(p[9], p[7], p[8], an.gmst, d)
        Fs.append(np.sqrt(Ap ** 2 * Fp ** 2 + Ac ** 2 * Fc ** 2))
        l = np.arctan2(Ac * Fc, Ap * Fp)
        lam.append(l + 2 * np.pi if l < 0 else l)
    q = [p[0], p[1], p[2], p[3], 0.5 * (lam[0] + p[4]), p[5] + dt[an.labels[0]], p[6] - np.log(Fs[0])]
    for d in range(1, an.nifo):
        q += [lam[d] - lam[0], dt[an.labels[d]] - dt[an.labels[0]], Fs[d] / Fs[0]]
    return np.array(q)
