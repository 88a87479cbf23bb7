"""The drop-in entry points under QuickCBC's threading: QuickCBC calls
log_likelihood_full, fulltemplates and log_likelihood_het from up to 12
OpenMP threads at once, one chain per thread (QuickCBC.c:880-955,
:2063-2067).  Twelve Python threads call them through the C ABI at the same
time (ctypes releases the interpreter lock during a foreign call), each with
points of its own, and every result is checked against the oracle.  Also:
qcbc_invalidate() after D changes in place, and heterodyne() rebinding an SN
that was redrawn in place (noise marginalisation, QuickCBC.c:2001-2006).

Tolerances: logL 1e-8 absolute; templates 1e-9 of their largest value;
heterodyned logL 1e-8 + 1e-10 |logL - logLb|.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu

N_THREADS = 12


def _run_threads(fn):
    errors = []
    barrier = threading.Barrier(N_THREADS)

    def body(k):
        try:
            barrier.wait()
            fn(k)
        except BaseException as e:  # noqa: BLE001
            errors.append((k, e))

    ts = [threading.Thread(target=body, args=(k,)) for k in range(N_THREADS)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


def test_twelve_threads_full_likelihood_and_templates():
    an, inj, D, SN = _data.analysis(0)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    try:
        Dr, Sr = Q.RowMatrix(D), Q.RowMatrix(SN)
        P = _data.points(0, 4 * N_THREADS, seed=31)
        Lr, Rr, Sr_ = _data.oracle_loglike(an, D, SN, P)
        got = np.zeros(P.shape[0])
        rho = np.zeros((P.shape[0], an.nifo))
        tmpl = {}

        def work(k):
            for rep in range(3):
                for j in range(4 * k, 4 * k + 4):
                    got[j], rho[j] = Q.dropin_log_likelihood_full(an, Dr, Sr, P[j])
                tmpl[k] = Q.dropin_fulltemplates(an, P[4 * k])

        _run_threads(work)
        assert np.abs(got - Lr).max() <= 1e-8, np.abs(got - Lr).max()
        assert np.abs(rho - Rr).max() <= 1e-8
        for k, h in tmpl.items():
            st, hr = O.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, P[4 * k])
            assert st == 0 and np.abs(h - hr).max() <= 1e-9 * np.abs(hr).max(), k
    finally:
        Q.lib().qcbc_set_band(16.0, 1024.0)


def test_twelve_threads_heterodyned_likelihood():
    an, inj, D, SN = _data.analysis(0)
    h = _data.het(0)
    view = Q.HetView.from_arrays(h, an.Tobs, an.nifo)
    P = W.ensemble_around(inj, 4 * N_THREADS, seed=32, scale=0.3)
    Lo, _ = _data.oracle_loglike_het(0, P)
    got = np.zeros(P.shape[0])

    def work(k):
        for rep in range(3):
            for j in range(4 * k, 4 * k + 4):
                got[j] = Q.dropin_log_likelihood_het(an, view, P[j])

    _run_threads(work)
    assert (np.abs(got - Lo) <= 1e-8 + 1e-10 * np.abs(Lo - h["logLb"])).all(), np.abs(got - Lo).max()


def test_invalidate_after_data_change_in_place():
    an, inj, D, SN = _data.analysis(0)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    try:
        Dm = np.array(D)
        Dr, Sr = Q.RowMatrix(Dm), Q.RowMatrix(SN)
        p = _data.points(0, 1, seed=33)[0]
        v0, _ = Q.dropin_log_likelihood_full(an, Dr, Sr, p)
        Dr.arr[:] = Dr.arr + W.noise(an, 77)  # the rows keep their addresses
        Q.lib().qcbc_invalidate()
        v1, _ = Q.dropin_log_likelihood_full(an, Dr, Sr, p)
        r0 = O.loglike(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, p)[0]
        r1 = O.loglike(an.labels, an.gmst, an.fref, an.fmin, an.fmax, Dr.arr, SN, an.N, an.Tobs, p)[0]
        assert abs(v0 - r0) <= 1e-8 and abs(v1 - r1) <= 1e-8 and abs(r1 - r0) > 1.0
    finally:
        Q.lib().qcbc_set_band(16.0, 1024.0)


def test_heterodyne_rebinds_a_psd_redrawn_in_place():
    """No qcbc_invalidate(): heterodyne() binds D and SN afresh on every call."""
    an, inj, D, SN = _data.analysis(0)
    h = _data.het(0)
    Dr, Sr = Q.RowMatrix(np.array(D)), Q.RowMatrix(np.array(SN))
    v0 = Q.dropin_heterodyne(an, Dr, Sr, inj, h["fgrid"], h["J"], h["MN"], h["MM"])
    assert abs(v0.s.logLb - h["logLb"]) <= 1e-8
    Sr.arr[:] = Sr.arr * 1.3  # a new PSD draw written into the same rows
    v1 = Q.dropin_heterodyne(an, Dr, Sr, inj, h["fgrid"], h["J"], h["MN"], h["MM"])
    st, h1 = O.heterodyne(an.labels, an.gmst, an.fref, Dr.arr, Sr.arr, an.N, an.Tobs, inj, h["fgrid"], h["J"],
                          h["MN"], h["MM"])
    assert st == 0 and abs(v1.s.logLb - h1["logLb"]) <= 1e-8 and abs(h1["logLb"] - h["logLb"]) > 1.0
