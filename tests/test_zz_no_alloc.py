"""qcbc_loglike_device_noalloc, the likelihood entry point that allocates
nothing (its point records and partial sums live in the library's fixed
device array), against qcbc_loglike_device, the entry point it is a variant
of: two calls in a row, the second with fewer points, on a context that never
ran the allocating entry point; both precisions.  The same kernels and the
same split layout give bit-identical logL and rho; FP64 also stays within
1e-8 of the oracle.  Batches above QCBC_NOALLOC_MAX_POINTS are refused with
QCBC_ERANGE.
"""
import ctypes

import numpy as np
import pytest

from quickcbc_amd import qcbc as Q
from tests import _data

pytestmark = pytest.mark.gpu

QCBC_ERANGE = 34
MAX_POINTS = 4096


def _noalloc(ctx, dp, B, prec, dl, dr, ds):
    return Q.lib().qcbc_loglike_device_noalloc(ctx.h, dp.ptr, B, prec, dl.ptr, dr.ptr, ds.ptr)


@pytest.mark.parametrize("prec", [Q.FP64, Q.MIXED])
def test_noalloc_matches_the_allocating_entry_point(prec):
    an, inj, D, SN = _data.analysis(1)
    fresh = Q.Context(an, D, SN)   # its batch scratch is never allocated
    ref_ctx = Q.Context(an, D, SN)
    nd = an.nifo
    for B, seed in ((2048, 41), (700, 42)):
        P = np.vstack([inj, _data.points(1, B - 1, seed=seed)])
        print("noalloc: precision %d, %d points" % (prec, B), flush=True)
        dp = Q.DeviceBuffer.from_array(P)
        bufs = [Q.DeviceBuffer(8 * B), Q.DeviceBuffer(8 * nd * B), Q.DeviceBuffer(4 * B)]
        ref = [Q.DeviceBuffer(8 * B), Q.DeviceBuffer(8 * nd * B), Q.DeviceBuffer(4 * B)]
        assert _noalloc(fresh, dp, B, prec, *bufs) == 0
        fresh.sync()
        ref_ctx.loglike_device(dp.ptr, B, prec, ref[0].ptr, ref[1].ptr, ref[2].ptr)
        ref_ctx.sync()
        L, R, S = bufs[0].download((B,)), bufs[1].download((B, nd)), bufs[2].download((B,), np.int32)
        Lr, Rr, Sr = ref[0].download((B,)), ref[1].download((B, nd)), ref[2].download((B,), np.int32)
        assert (S == 0).all() and np.array_equal(S, Sr)
        assert np.array_equal(L, Lr), np.abs(L - Lr).max()
        assert np.array_equal(R, Rr)
        k = np.r_[0:8, B - 8:B]
        Lo, _, So = _data.oracle_loglike(an, D, SN, P[k])
        tol = 1e-8 if prec == Q.FP64 else 1e-2
        assert (So == 0).all() and np.abs(L[k] - Lo).max() <= tol
        for b in bufs + ref + [dp]:
            b.free()
    fresh.close()
    ref_ctx.close()


def test_noalloc_refuses_batches_above_its_capacity():
    an, inj, D, SN = _data.analysis(0)
    c = Q.Context(an, D, SN)
    B = MAX_POINTS + 1
    dp = Q.DeviceBuffer(8 * 11 * B)
    bufs = [Q.DeviceBuffer(8 * B), Q.DeviceBuffer(8 * 3 * B), Q.DeviceBuffer(4 * B)]
    assert _noalloc(c, dp, B, Q.FP64, *bufs) == QCBC_ERANGE
    assert Q.lib().qcbc_loglike_device_noalloc(None, dp.ptr, 4, Q.FP64, bufs[0].ptr, bufs[1].ptr,
                                               bufs[2].ptr) == 1
    # an unbound analysis
    c2 = Q.Context(an)
    assert _noalloc(c2, dp, 4, Q.FP64, *bufs) == 102
    for b in bufs + [dp]:
        b.free()
    c.close()
    c2.close()
