"""GPU path against the CPU oracle on identical seeded inputs.

Tolerances (FP64 path): waveform amplitude and phase 1e-12 relative
(|dA|/A and |dPsi|/max(|Psi|, 2 pi), measured before projection because the
projected argument reaches ~1e5 rad); logL 1e-8 absolute plus 1e-12 of
|logL| for the fused launch (1e-8 absolute for the plain path), rho 1e-8.
Config 1 is the PT ensemble's full 2048 points.  Mixed precision:
|dlogL| <= 0.01.
"""
import numpy as np
import pytest

from tests import _data
from quickcbc_amd import qcbc as Q
from quickcbc_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0, 1])
def ctx_cfg(request):
    an, inj, D, SN = _data.analysis(request.param)
    c = Q.Context(an, D, SN)
    yield request.param, an, inj, D, SN, c
    c.close()


def test_waveform_amp_phase(ctx_cfg, oracle):
    which, an, inj, D, SN, c = ctx_cfg
    P = np.vstack([inj, _data.points(which, 15, seed=21)])
    amp, ph, st = c.waveform(P)
    assert (st == 0).all()
    for b in range(P.shape[0]):
        s, ra, rp = oracle.waveform_params(P[b], an.N, an.Tobs, an.fref)
        assert s == 0
        assert np.array_equal(ra == 0, amp[b] == 0), "fCut mask differs"
        nz = ra > 0
        rel_a = np.abs(amp[b][nz] - ra[nz]) / ra[nz]
        rel_p = np.abs(ph[b] - rp) / np.maximum(np.abs(rp), 2 * np.pi)
        k = int(np.argmax(rel_p))
        f = k / an.Tobs if k else 1.0 / an.Tobs
        assert rel_a.max() <= 1e-12, (b, rel_a.max())
        assert rel_p.max() <= 1e-12, ("point", b, "bin", k, "f", f, "gpu", ph[b][k], "oracle", rp[k],
                                      "rel", rel_p.max(), "params", list(P[b]))


def test_templates(ctx_cfg, oracle):
    which, an, inj, D, SN, c = ctx_cfg
    P = np.vstack([inj, _data.points(which, 7, seed=22)])
    hw, st = c.templates(P)
    assert (st == 0).all()
    for b in range(P.shape[0]):
        s, ref = oracle.fulltemplates(an.labels, an.gmst, an.fref, an.N, an.Tobs, P[b])
        nh = an.N // 2
        for d in range(an.nifo):
            A = np.hypot(ref[d, 1:nh], ref[d, an.N - 1:nh:-1])
            Ag = np.hypot(hw[b, d, 1:nh], hw[b, d, an.N - 1:nh:-1])
            nz = A > 0
            assert np.array_equal(nz, Ag > 0), "fbegin / fCut mask differs"
            assert (np.abs(Ag[nz] - A[nz]) / A[nz]).max() <= 1e-12
            # the projected argument is ~1e5 rad: allow its ulp-level shift
            err = np.abs(hw[b, d] - ref[d]).max()
            assert err <= 1e-9 * A.max(), (b, d, err / A.max())


def test_loglike_fp64(ctx_cfg):
    which, an, inj, D, SN, c = ctx_cfg
    n = 1024 if which == 0 else 2048  # config 1: the PT ensemble's full 2048 points
    P = np.vstack([inj, _data.points(which, n - 1, seed=23)])
    L, R, S = c.loglike(P, Q.FP64)
    Lr, Rr, Sr = _data.oracle_loglike(an, D, SN, P)
    assert (S == Sr).all()
    print("fused FP64 |dlogL| max %.3e (|logL| up to %.0f), |drho| max %.3e" %
          (np.abs(L - Lr).max(), np.abs(Lr).max(), np.abs(R - Rr).max()))
    # 1e-8 absolute, plus 1e-12 of |logL|: two FP64 sums over ~1e5 terms in
    # different orders differ by ~1e-12 relative, which passes 1e-8 at
    # |logL| ~ 1e4 (2 of the 2048 config-1 points, at logL = -1.04e4, reach
    # 1.2e-8 with either fused kernel)
    assert (np.abs(L - Lr) <= 1e-8 + 1e-12 * np.abs(Lr)).all(), np.abs(L - Lr).max()
    assert np.abs(R - Rr).max() <= 1e-8


def test_loglike_plain(ctx_cfg):
    which, an, inj, D, SN, c = ctx_cfg
    P = np.vstack([inj, _data.points(which, 31, seed=24)])
    L, R, S = c.loglike_plain(P)
    Lr, Rr, Sr = _data.oracle_loglike(an, D, SN, P)
    assert (S == Sr).all()
    assert np.abs(L - Lr).max() <= 1e-8, np.abs(L - Lr).max()


def test_loglike_mixed(ctx_cfg):
    which, an, inj, D, SN, c = ctx_cfg
    n = 1024 if which == 0 else 2048  # config 1: the PT ensemble's full 2048 points
    P = np.vstack([inj, _data.points(which, n - 1, seed=25)])
    L, R, S = c.loglike(P, Q.MIXED)
    Lr, Rr, Sr = _data.oracle_loglike(an, D, SN, P)
    err = np.abs(L - Lr)
    print("mixed |dlogL| max %.3e p99 %.3e median %.3e" % (err.max(), np.percentile(err, 99), np.median(err)))
    assert err.max() <= 0.01


def test_domain_errors_are_per_point(ctx_cfg):
    which, an, inj, D, SN, c = ctx_cfg
    P = np.vstack([inj, inj, inj])
    P[1, 2] = 1.5  # |chi1| > 1: the reference exits (IMRPhenomD.c:224-225)
    L, R, S = c.loglike(P, Q.FP64)
    assert list(S) == [0, 33, 0]
    assert L[1] == 0.0 and L[0] == L[2] and L[0] != 0.0


def test_noise_free_identity():
    """With D = h(theta), logL(theta) = (h|h)/2 and rho_d = sqrt((h_d|h_d))."""
    for which in (0, 1):
        an, inj, D0, SN = _data.analysis(which, noise=False)
        c = Q.Context(an, D0, SN)
        L, R, S = c.loglike(inj[None, :], Q.FP64)
        assert S[0] == 0
        assert abs(L[0] - 0.5 * (R[0] ** 2).sum()) <= 1e-9 * L[0]
        c.close()


ORACLE_FIELDS = ["Msec", "amp0", "fCut", "eta", "finspin", "fRD", "fDM", "fmaxCalc", "ai_23", "ai_1", "ai_43",
                 "ai_53", "ai_2", "ai_73", "ai_83", "ai_3", "pi_init", "pi_23", "pi_13", "pi_13L", "pi_L", "pi_m13",
                 "pi_m23", "pi_m1", "pi_m53", "pi_1", "pi_43", "pi_53", "pi_2", "beta1", "beta2", "beta3",
                 "alpha1", "alpha2", "alpha3", "alpha4", "C1Int", "C2Int", "C1MRD", "C2MRD", "t0", "MfRef",
                 "phi_precalc"]


PHASE_CONSTANTS = {"C1Int", "C2Int", "C1MRD", "C2MRD", "phi_precalc", "pi_init"}


def test_setup_records_vs_oracle(ctx_cfg, oracle):
    """Per-point IMRPhenomD set-up on the GPU against the oracle's, field by field."""
    which, an, inj, D, SN, c = ctx_cfg
    P = np.vstack([inj, _data.points(which, 31, seed=21)])
    recs = Q.point_records(c, P)
    worst = (0.0, None)
    for b, r in enumerate(recs):
        p = P[b]
        # the mass conversion of fulltemplates (QuickCBC.c:3705-3718), correctly rounded as in
        # the oracle and the GPU set-up
        mc, mt = oracle.exp_cr(p[0]), oracle.exp_cr(p[1])
        eta = oracle.pow_cr(mc / mt, 5.0 / 3.0)
        dm = 0.0 if eta > 0.25 else np.sqrt(1 - 4 * eta)
        m1, m2 = mt * (1 + dm) / 2 / W.MSUN_SI, mt * (1 - dm) / 2 / W.MSUN_SI
        st, ref = oracle.setup(m1, m2, p[2], p[3], oracle.exp_cr(p[6]), 0.5 * p[4], an.fref)
        assert st == 0 and r.status == 0
        for name in ORACLE_FIELDS:
            g = getattr(r, name)
            o = ref["rho1" if name == "ai_73" else "rho2" if name == "ai_83" else "rho3" if name == "ai_3"
                    else name]
            # the phase connection constants and slopes (radians, and radians per unit Mf with
            # Mf < 0.2) come out of cancelling differences of connection values: they are judged
            # as the phase is, absolute error over 2 pi
            scale = 2 * np.pi if name in PHASE_CONSTANTS else max(abs(o), 1e-300)
            err = abs(g - o) / scale
            if err > worst[0]:
                worst = (err, (b, name, g, o))
    print("worst set-up field:", worst)
    assert worst[0] <= 1e-12, worst
