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
