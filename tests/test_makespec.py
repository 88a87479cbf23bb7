"""makespec (Utilities.c:1908-1959): the oracle pinned, and the GPU against it.

The oracle's spline is the natural cubic spline scipy computes
(CubicSpline(bc_type='natural')) to 1e-12; without lines SN = exp(spline);
at a line's centre the factor is 1 + height exactly (z = 1, the Lorentzian
denominator 1); beyond a line's half-width it adds nothing.  The GPU runs the
same solve and evaluation (1e-13 relative: exp may differ by an ulp).
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from quickcbc_amd import workloads as W


def _knots(Tobs, N, n=24, seed=5):
    r = np.random.default_rng(seed)
    fmax = (N // 2) / Tobs
    x = np.concatenate([[0.5 / Tobs], np.sort(r.uniform(2.0 / Tobs, fmax - 2.0 / Tobs, n - 2)), [fmax]])
    f = np.maximum(x, 1.0)
    y = np.log(W.psd(np.maximum(f, 10.0))) + 0.3 * r.standard_normal(n)
    return x, y


def _lines():
    return (np.array([60.0, 331.5]), np.array([50.0, 8.0]), np.array([1e3, 4e2]), np.array([2.0, 1.5]),
            np.array([0.3, 0.2]))


def test_oracle_makespec_pins():
    Tobs, N = 4.0, 8192
    x, y = _knots(Tobs, N)
    st, SN = O.makespec(x, y, None, Tobs, N)
    assert st == 0
    f = np.arange(1, N // 2 - 1) / Tobs
    assert np.allclose(np.log(SN[1:N // 2 - 1]), CubicSpline(x, y, bc_type="natural")(f), rtol=0, atol=1e-12)
    assert SN[0] == np.exp(y[0]) and SN[N // 2 - 1] == np.exp(y[-1])
    lines = _lines()
    st, SL = O.makespec(x, y, lines, Tobs, N)
    k = int(round(lines[0][0] * Tobs))  # 60 Hz is bin 240
    assert abs(SL[k] / SN[k] - (1.0 + lines[1][0])) < 1e-14 * (1.0 + lines[1][0])
    far = np.abs(np.arange(N // 2) / Tobs - 60.0) >= 2.0
    far &= np.abs(np.arange(N // 2) / Tobs - 331.5) >= 1.5
    assert np.array_equal(SL[far], SN[far])
    # knots that do not cover the bins: the reference's spline would abort
    assert O.makespec(x[1:], y[1:], None, Tobs, N)[0] != 0


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1])
def test_gpu_makespec(which):
    from quickcbc_amd import qcbc as Q
    an = W.config0() if which == 0 else W.config1()
    x, y = _knots(an.Tobs, an.N, seed=7 + which)
    c = Q.Context(an)
    for lines in (None, _lines()):
        SN = c.makespec(x, y, lines)
        st, So = O.makespec(x, y, lines, an.Tobs, an.N)
        assert st == 0
        assert np.abs(SN / So - 1.0).max() <= 1e-13
    with pytest.raises(RuntimeError):
        c.makespec(x[1:], y[1:], None)
    c.close()
