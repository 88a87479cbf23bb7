"""Pins of the oracle's data conditioning (CPU only): QuickCBC.c:268-313,
tukey() (Utilities.c:635-654), the real FFT in GSL's half-complex layout and
the sqrt(Tobs)/N scale.

- the window: 0 at the first sample, 1 on the flat part, the two cosine
  tapers rising monotonically, the flat part (1 - alpha) of the samples;
- the transform: the half-complex layout of a direct O(N^2) DFT with the
  forward sign on small N (a swapped real/imaginary slot or a conjugate sign
  fails it), and Parseval's relation on a large N;
- the scale: without a window, Parseval gives sum |D|^2 = (Tobs / N) sum x^2,
  so unit-variance white noise has a mean |D_i|^2 of Tobs / N per bin.
"""
import numpy as np

from oracle import oracle as O


def test_tukey_window():
    for N, alpha in ((8192, 0.2), (65536, 0.05), (1000, 0.5)):
        w = O.tukey(N, alpha)
        imin = int(alpha * (N - 1) / 2.0)
        imax = int((N - 1) * (1.0 - alpha / 2.0))
        assert w[0] == 0.0 and (w[imin:imax + 1] == 1.0).all()
        assert (np.diff(w[:imin + 1]) > 0).all() and (np.diff(w[imax:]) < 0).all()
        assert abs((imax - imin + 1) / N - (1 - alpha)) < 3.0 / N


def test_halfcomplex_fft_is_the_dft():
    r = np.random.default_rng(2)
    for N in (8, 16, 64):
        x = r.standard_normal(N)
        k = np.arange(N)
        X = np.array([(x * np.exp(-2j * np.pi * j * k / N)).sum() for j in range(N // 2 + 1)])
        hc = O.halfcomplex(np.fft.rfft(x), N)
        assert np.allclose(hc[:N // 2 + 1], X.real, atol=1e-12)
        assert np.allclose(hc[N - np.arange(1, N // 2)], X[1:N // 2].imag, atol=1e-12)
        D = O.condition(x[None, :], 2.0, t_rise=0.0)[0]
        assert np.allclose(D, hc * np.sqrt(2.0) / N, atol=1e-12)


def test_parseval_and_scale():
    r = np.random.default_rng(3)
    N, Tobs = 1 << 16, 16.0
    x = r.standard_normal(N)
    D = O.condition(x[None, :], Tobs, t_rise=0.0)[0]  # no window: Parseval exactly
    h = N // 2
    power = D[0] ** 2 + D[h] ** 2 + 2 * (D[1:h] ** 2 + D[N - np.arange(1, h)] ** 2).sum()
    assert abs(power - (x ** 2).sum() * Tobs / N) < 1e-9 * power
    # unit-variance white noise: mean |D_i|^2 = Tobs / N per bin
    m = (D[1:h] ** 2 + D[N - np.arange(1, h)] ** 2).mean()
    assert abs(m / (Tobs / N) - 1.0) < 0.02
