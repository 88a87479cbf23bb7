"""The input and output text formats and main's conditioning outputs
(QuickCBC.c:195-313; whiten, tukey_scale: Utilities.c:294-309, :656-683),
on the CPU: the oracle pinned, and the library's host readers and writer
(C ABI, no GPU needed) against it.

- framed_* count: the oracle runs the reference's loop through glibc's
  stdio; its count is the closed rule (two numbers per call, the end seen by
  the call that finds no number or whose last number ends the file), and
  the library's reader gives the same count and values bit for bit, on files
  with and without a final newline, odd token counts, blank lines and tabs;
- spec_* rows: columns 2 and 3, as numpy's loadtxt reads them;
- the writer: "%.{d}e " per value and a newline per row, byte for byte;
- DW: without a window, white noise whitened by its own PSD 2 Tobs/N comes
  back as itself less its DC and Nyquist components (fixes whiten's
  1/sqrt(S) and the sqrt(N/2)/sqrt(Tobs) scale);
- pspec: Parseval's relation over bins 1 .. N/2 - 1, and bin 0 the DC term;
- tukey_scale: 1 and 1 without a window; 1 - alpha/2 and sqrt(1 - 5 alpha/8)
  (the tapers' mean 1/2 and mean square 3/8) to O(1/N).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q


TEXTS = [
    "",
    "\n",
    "1.0 2.0\n3.0 4.0\n",
    "1.0 2.0\n3.0 4.0",
    "1.0 2.0\n3.0\n",
    "1.0 2.0\n3.0",
    "  1.5e-3\t-2.25e-21\n\n7 8   \n\n",
    "1 2 3 4 5 6",
    "1 2 3 4 5 6 ",
]


@pytest.mark.parametrize("k", range(len(TEXTS)))
def test_framed_count_and_values(tmp_path, k):
    text = TEXTS[k]
    p = tmp_path / "framed_8_100_0.dat"
    p.write_text(text)
    N, t, x = O.read_framed(p)
    assert N == O.framed_count(text), (repr(text), N)
    tq, xq = Q.read_framed(p)
    assert tq.size == N
    assert np.array_equal(tq, t) and np.array_equal(xq, x)
    vals = np.array(text.split(), float)
    assert np.array_equal(t, vals[0:2 * N:2]) and np.array_equal(x, vals[1:2 * N:2])


def test_framed_as_written(tmp_path):
    """A file written "%e %e\\n" per sample, as the framed_* files are: every sample counts."""
    r = np.random.default_rng(7)
    N = 4096
    t = np.arange(N) / 512.0
    s = r.standard_normal(N) * 1e-21
    p = tmp_path / Q.data_file_name("framed", 8.0, 1126259462.4, 1)
    p.write_text("".join("%e %e\n" % (a, b) for a, b in zip(t, s)))
    tq, xq = Q.read_framed(p)
    want = np.loadtxt(p)
    assert tq.size == N and np.array_equal(tq, want[:, 0]) and np.array_equal(xq, want[:, 1])


def test_file_name():
    assert Q.data_file_name("framed", 8.0, 1126259462.4, 0) == "framed_8_1126259462_0.dat"
    assert Q.data_file_name("spec", 4.0, 1187008882.43, 2) == "spec_4_1187008882_2.dat"


def test_read_spec(tmp_path):
    r = np.random.default_rng(8)
    n2 = 2048
    f = np.arange(n2) / 4.0
    SN, SM = r.uniform(1, 2, n2) * 1e-46, r.uniform(1, 2, n2) * 1e-46
    p = tmp_path / "spec_4_100_0.dat"
    p.write_text("".join("%e %e %e %e\n" % (a, b, c, b / c) for a, b, c in zip(f, SN, SM)))
    rows, So, Mo = O.read_spec(p, n2)
    assert rows == n2
    want = np.loadtxt(p)
    assert np.array_equal(So, want[:, 1]) and np.array_equal(Mo, want[:, 2])
    Sq, Mq = Q.read_spec(p, n2)
    assert np.array_equal(Sq, So) and np.array_equal(Mq, Mo)
    # fewer rows than asked: an error, not values left unset
    with pytest.raises(RuntimeError):
        Q.read_spec(p, n2 + 1)


def test_reader_errors(tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text("1.0 2.0\nabc 4.0\n")
    assert O.read_framed(p)[0] == -2
    with pytest.raises(RuntimeError, match="QCBC_EINVAL"):
        Q.read_framed(p)
    with pytest.raises(RuntimeError, match="QCBC_EIO"):
        Q.read_framed(tmp_path / "missing.dat")


def test_write_table_format(tmp_path):
    r = np.random.default_rng(9)
    cols = [np.arange(5) / 3.0, r.standard_normal(5) * 1e-20, -r.standard_normal(5)]
    for digits in (6, 15):
        p = tmp_path / ("t%d.dat" % digits)
        Q.write_table(p, cols, digits)
        fmt = "%%.%de " % digits
        want = "".join("".join(fmt % c[i] for c in cols) + "\n" for i in range(5))
        assert p.read_text() == want
        back = np.loadtxt(p)
        if digits == 15:
            assert np.allclose(back.T, cols, rtol=1e-15, atol=0)


def test_whitened_series_of_white_noise():
    r = np.random.default_rng(10)
    N, Tobs = 1 << 14, 8.0
    x = r.standard_normal((2, N))
    out = O.condition_outputs(x, np.full((2, N // 2), 2.0 * Tobs / N), Tobs, t_rise=0.0)
    n = np.arange(N)
    alt = (-1.0) ** n
    want = x - x.mean(1, keepdims=True) - ((x * alt).sum(1) / N)[:, None] * alt
    assert np.abs(out["DW"] - want).max() <= 1e-12
    # with the window, whitened white noise keeps unit variance on the flat part
    out = O.condition_outputs(x, np.full((2, N // 2), 2.0 * Tobs / N), Tobs, t_rise=0.4)
    mid = out["DW"][:, N // 4:3 * N // 4]
    assert abs(mid.var() - 1.0) < 0.05


def test_pspec_parseval():
    r = np.random.default_rng(11)
    N, Tobs = 1 << 13, 4.0
    h = N // 2
    x = r.standard_normal((3, N)) * 2.0
    SN = r.uniform(1.0, 3.0, (3, h))
    out = O.condition_outputs(x, SN, Tobs, t_rise=0.0)
    X = np.fft.rfft(x)
    want = Tobs / N * ((x ** 2).sum(1) - (np.abs(X[:, 0]) ** 2 + np.abs(X[:, h]) ** 2) / N)
    assert np.allclose(out["pspec"][:, 1:].sum(1), want, rtol=1e-12)
    assert np.allclose(out["pspec"][:, 0], 2.0 * Tobs / N ** 2 * x.sum(1) ** 2, rtol=1e-12)
    assert np.array_equal(out["wcheck"], out["pspec"] / SN)
    # the mean periodogram of white noise is its one-sided PSD, 2 sigma^2 Tobs / N
    assert abs(out["pspec"][:, 1:].mean() / (2 * 4.0 * Tobs / N) - 1.0) < 0.02


def test_tukey_scale_closed_forms():
    assert O.tukey_scale(0.0, 4096) == (1.0, 1.0)
    for N, alpha in ((1 << 14, 0.1), (1 << 16, 0.05), (1 << 12, 0.4)):
        s1, s2 = O.tukey_scale(alpha, N)
        assert abs(s1 - (1.0 - alpha / 2.0)) < 3.0 / N
        assert abs(s2 - np.sqrt(1.0 - 5.0 * alpha / 8.0)) < 3.0 / N
