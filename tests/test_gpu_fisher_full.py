"""GPU parity of Fisher_Full (QuickCBC.c:2551-2688) against the oracle,
through the batch call and the drop-in, on both workloads.

Tolerance: entries 1e-7 sqrt(F_ii F_jj), as for Fisher_All: the central
differences of a 1e5-rad phase turn one rounding of the waveform into ~1e-6
of a derivative, and the GPU's waveform agrees with the oracle's to ~1e-12.
The floor, the 1e4 I of a spin at the limit and the d < 7 blocks are exact.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W
from tests import _data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0, 1])
def setup(request):
    which = request.param
    an, inj, D, SN = _data.analysis(which)
    print("Fisher_Full: config %d, %d det, N %d" % (which, an.nifo, an.N), flush=True)
    Q.lib().qcbc_set_band(an.fmin, an.fmax)
    c = Q.Context(an, D, SN)
    P = np.vstack([inj, W.ensemble_around(inj, 7, seed=81, scale=0.5)])
    R = np.array([_data.ref_frame(an, p) for p in P])
    yield which, an, SN, c, R
    c.close()
    Q.lib().qcbc_set_band(16.0, 1024.0)


def test_fisher_full_batch(setup):
    which, an, SN, c, R = setup
    F, st = c.fisher_full(R)
    assert (st == 0).all()
    for b in range(R.shape[0]):
        s0, Fo = O.fisher_full(an.nifo, an.fref, an.fmin, an.fmax, SN, an.N, an.Tobs, R[b])
        assert s0 == 0
        scale = np.sqrt(np.outer(np.diag(Fo), np.diag(Fo)))
        err = np.abs(F[b] - Fo) / scale
        print("Fisher_Full point %d: max error %.2e of sqrt(F_ii F_jj)" % (b, err.max()), flush=True)
        assert err.max() <= 1e-7, b
        assert np.array_equal(F[b], F[b].T)


def test_fisher_full_dropin_and_blocks(setup):
    which, an, SN, c, R = setup
    F, _ = c.fisher_full(R[:2])
    Fd = Q.dropin_fisher_full(an, Q.RowMatrix(SN), R[0])
    assert np.array_equal(Fd, F[0])
    F3, _ = c.fisher_full(R[:2], d=3)
    assert np.array_equal(F3, F[:, :3, :3])
    F3d = Q.dropin_fisher_full(an, Q.RowMatrix(SN), R[1], d=3)
    assert np.array_equal(F3d, F[1, :3, :3])


def test_fisher_full_floor_and_spin_limit(setup):
    which, an, SN, c, R = setup
    far = R[0].copy()
    far[6] += 40.0
    spin = R[0].copy()
    spin[2] = 0.99995
    F, st = c.fisher_full(np.vstack([far, spin]))
    assert (st == 0).all()
    assert np.array_equal(np.diag(F[0]), np.ones(7))
    assert np.array_equal(F[1], 1.0e4 * np.eye(7))
