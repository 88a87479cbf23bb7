"""Data conditioning on the GPU (QuickCBC.c:268-318) against the oracle.

qcbc_bind_time_series windows each detector's time series (tukey,
Utilities.c:635-654), transforms it with hipFFT into GSL's half-complex
layout, scales it by sqrt(Tobs)/N and binds it.  The conditioned data agree
with the oracle's (numpy's FFT) to 1e-11 of their largest value (two FFTs'
rounding, ~eps log2 N relative), and the likelihood of the bound analysis
agrees with the oracle's on the oracle's conditioned data to 1e-8 absolute
plus 1e-10 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from tests import _data

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", [0, 1])
def test_conditioned_data_and_likelihood(which):
    an, inj, D0, SN = _data.analysis(which)
    r = np.random.default_rng(40 + which)
    t = np.arange(an.N) * (an.Tobs / an.N)
    x = r.standard_normal((an.nifo, an.N)) * 1e-21 + 3e-22 * np.sin(2 * np.pi * 123.4 * t)[None, :]
    print("condition: config %d, %d x %d samples" % (which, an.nifo, an.N), flush=True)
    c = Q.Context(an)
    D = c.bind_time_series(x, SN, t_rise=0.4)
    Do = O.condition(x, an.Tobs, 0.4)
    assert np.abs(D - Do).max() <= 1e-11 * np.abs(Do).max()
    P = _data.points(which, 16, seed=3)
    L, _, st = c.loglike(P, Q.FP64)
    Lo, _, so = _data.oracle_loglike(an, Do, SN, P)
    assert (st == 0).all() and (so == 0).all()
    assert (np.abs(L - Lo) <= 1e-8 + 1e-10 * np.abs(Lo)).all(), np.abs(L - Lo).max()
    with pytest.raises(RuntimeError):
        c.bind_time_series(x, SN, t_rise=-1.0)
    c.close()
