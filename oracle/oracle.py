"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: imported by tests/, the smoke check and the bench's
CPU-baseline leg only.  The product (quickcbc_amd) never imports this.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.qo_fbegin.restype = ctypes.c_double
        L.qo_fbegin.argtypes = [_dp]
        L.qo_times.argtypes = [ctypes.c_double] * 3 + [_dp]
        L.qo_f_ant.argtypes = [ctypes.c_double] * 4 + [_dp, _dp, ctypes.c_int]
        L.qo_fourier_nwip.restype = ctypes.c_double
        L.qo_fourier_nwip.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.qo_loglike_flat.restype = ctypes.c_double
        L.qo_loglike_flat.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int,
                                      ctypes.c_double, _dp, _dp, _ip]
        L.qo_loglike_batch_flat.restype = None
        L.qo_loglike_batch_flat.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int,
                                            ctypes.c_double, _dp, ctypes.c_int, _dp, _dp, _ip,
                                            ctypes.c_int]
        L.qo_fulltemplates_flat.restype = ctypes.c_int
        L.qo_fulltemplates_flat.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double,
                                            _dp, ctypes.c_int, ctypes.c_double, _dp]
        L.qo_h22_flat.restype = ctypes.c_int
        L.qo_h22_flat.argtypes = [_dp, ctypes.c_size_t] + [ctypes.c_double] * 7 + [_dp, _dp]
        L.qo_setup_flat.restype = ctypes.c_int
        L.qo_setup_flat.argtypes = [ctypes.c_double] * 7 + [_dp, ctypes.c_int]
        L.qo_spline_eval_flat.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.qo_final_spin.restype = ctypes.c_double
        L.qo_final_spin.argtypes = [ctypes.c_double] * 3
        L.qo_erad.restype = ctypes.c_double
        L.qo_erad.argtypes = [ctypes.c_double] * 3
        L.qo_fit.restype = ctypes.c_double
        L.qo_fit.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.qo_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


# Field order of qo_phenomd (qcbc_oracle.h); all doubles.
SETUP_FIELDS = (
    "m1 m2 chi1 chi2 M eta Msec amp0 fCut finspin chi fRD fDM "
    "gamma1 gamma2 gamma3 fmaxCalc rho1 rho2 rho3 f1 f2 f3 v1 v2 v3 d1 d2 "
    "delta0 delta1 delta2 delta3 delta4 "
    "ai_amp0 ai_23 ai_1 ai_43 ai_53 ai_2 ai_73 ai_83 ai_3 "
    "sigma1 sigma2 sigma3 sigma4 beta1 beta2 beta3 alpha1 alpha2 alpha3 alpha4 alpha5 "
    "C1Int C2Int C1MRD C2MRD fPhiIns fPhiMRD").split()
SETUP_FIELDS += ["pn_v%d" % i for i in range(13)] + ["pn_vlogv%d" % i for i in range(13)]
SETUP_FIELDS += ("pi_init pi_23 pi_13 pi_13L pi_L pi_m13 pi_m23 pi_m1 pi_m53 "
                 "pi_1 pi_43 pi_53 pi_2 t0 MfRef phifRef phi_precalc").split()


def setup(m1, m2, chi1, chi2, distance=1.0, phi0=0.0, fref=100.0):
    out = np.zeros(len(SETUP_FIELDS))
    st = lib().qo_setup_flat(m1, m2, chi1, chi2, distance, phi0, fref, _d(out), len(out))
    return st, dict(zip(SETUP_FIELDS, out))


def fbegin(params):
    p = np.ascontiguousarray(params, dtype=np.float64)
    return lib().qo_fbegin(_d(p))


def times(alpha, sindelta, gmst):
    out = np.zeros(3)
    lib().qo_times(alpha, sindelta, gmst, _d(out))
    return out


def f_ant(psi, alpha, sindelta, gmst, det):
    fp = ctypes.c_double(0.0)
    fc = ctypes.c_double(0.0)
    lib().qo_f_ant(psi, alpha, sindelta, gmst, ctypes.byref(fp), ctypes.byref(fc), det)
    return fp.value, fc.value


def fourier_nwip(a, b, Sn, imin, imax, N):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    Sn = np.ascontiguousarray(Sn, np.float64)
    return lib().qo_fourier_nwip(_d(a), _d(b), _d(Sn), imin, imax, N)


def h22(freq, phi0, fref, m1_SI, m2_SI, chi1, chi2, distance):
    freq = np.ascontiguousarray(freq, np.float64)
    amp = np.zeros_like(freq)
    ph = np.zeros_like(freq)
    st = lib().qo_h22_flat(_d(freq), freq.size, phi0, fref, m1_SI, m2_SI, chi1, chi2, distance,
                           _d(amp), _d(ph))
    return st, amp, ph


def fulltemplates(labels, gmst, fref, N, Tobs, params):
    labels = np.ascontiguousarray(labels, np.int32)
    p = np.ascontiguousarray(params, np.float64)
    h = np.zeros((labels.size, N))
    st = lib().qo_fulltemplates_flat(labels.size, _i(labels), gmst, fref, _d(h), N, Tobs, _d(p))
    return st, h


def loglike(labels, gmst, fref, fmin, fmax, D, SN, N, Tobs, params):
    labels = np.ascontiguousarray(labels, np.int32)
    D = np.ascontiguousarray(D, np.float64)
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    rho = np.zeros(labels.size)
    st = ctypes.c_int(0)
    v = lib().qo_loglike_flat(labels.size, _i(labels), gmst, fref, fmin, fmax, _d(D), _d(SN), N,
                              Tobs, _d(p), _d(rho), ctypes.byref(st))
    return v, rho, st.value


def loglike_batch(labels, gmst, fref, fmin, fmax, D, SN, N, Tobs, params, nthreads=0):
    labels = np.ascontiguousarray(labels, np.int32)
    D = np.ascontiguousarray(D, np.float64)
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
    B = p.shape[0]
    logL = np.zeros(B)
    rho = np.zeros((B, labels.size))
    st = np.zeros(B, np.int32)
    lib().qo_loglike_batch_flat(labels.size, _i(labels), gmst, fref, fmin, fmax, _d(D), _d(SN), N,
                                Tobs, _d(p), B, _d(logL), _d(rho), _i(st), nthreads)
    return logL, rho, st


def spline_eval(which, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    lib().qo_spline_eval_flat(which, _d(x), x.size, _d(y))
    return y


def final_spin(eta, chi1, chi2):
    return lib().qo_final_spin(eta, chi1, chi2)


def erad(eta, chi1, chi2):
    return lib().qo_erad(eta, chi1, chi2)


def fit(which, eta, chi):
    return lib().qo_fit(which, eta, chi)


def waveform_params(params, N, Tobs, fref=100.0):
    L = lib()
    L.qo_waveform_params.restype = ctypes.c_int
    L.qo_waveform_params.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
    p = np.ascontiguousarray(params, np.float64)
    amp = np.zeros(N // 2)
    ph = np.zeros(N // 2)
    st = L.qo_waveform_params(_d(p), N, Tobs, fref, _d(amp), _d(ph))
    return st, amp, ph


def collocation(m1, m2, chi1, chi2):
    """The intermediate amplitude's collocation (IMRPhenomD_internals.c:848-1063) of one system: the
    closed-form deltas, the elimination's deltas, (f1, f2, f3, v1, v2, v3, d1, d2) and the term-wise
    derivative of the inspiral ansatz at f1."""
    L = lib()
    L.qo_collocation_probe.restype = ctypes.c_int
    L.qo_collocation_probe.argtypes = [ctypes.c_double] * 4 + [_dp]
    out = np.zeros(19)
    st = L.qo_collocation_probe(m1, m2, chi1, chi2, _d(out))
    return st, {"closed": out[0:5].copy(), "elim": out[5:10].copy(), "f": out[10:13].copy(),
                "v": out[13:16].copy(), "d": out[16:18].copy(), "d1_termwise": float(out[18])}


def probe(which, m1, m2, chi1, chi2, Mf):
    L = lib()
    L.qo_probe.restype = ctypes.c_double
    L.qo_probe.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5
    return L.qo_probe(which, m1, m2, chi1, chi2, Mf)


def exp_cr(x):
    L = lib()
    L.qo_exp_cr.restype = ctypes.c_double
    L.qo_exp_cr.argtypes = [ctypes.c_double]
    return L.qo_exp_cr(x)


def pow_cr(x, y):
    L = lib()
    L.qo_pow_cr.restype = ctypes.c_double
    L.qo_pow_cr.argtypes = [ctypes.c_double, ctypes.c_double]
    return L.qo_pow_cr(x, y)


# ------------------------------------------------- heterodyned likelihood
def legendre(J, U):
    """legendre_maker (QuickCBC.c:5055): P[j][k], j <= J, k <= U."""
    L = lib()
    L.qo_legendre.restype = None
    L.qo_legendre.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    P = np.zeros((J + 1, U + 1))
    L.qo_legendre(J, U, _d(P))
    return P


def inverse(M):
    """Inverse (Utilities.c:1862): the transpose of inv(M)."""
    L = lib()
    L.qo_inverse.restype = None
    L.qo_inverse.argtypes = [_dp, _dp, ctypes.c_int]
    M = np.ascontiguousarray(M, np.float64)
    IM = np.zeros_like(M)
    L.qo_inverse(_d(M), _d(IM), M.shape[0])
    return IM


def fullphaseamp(labels, gmst, fref, Tobs, freq, params):
    """fullphaseamp (QuickCBC.c:3780) on a frequency list: (status, amp, phase), nifo x n."""
    L = lib()
    L.qo_fullphaseamp_flat.restype = ctypes.c_int
    L.qo_fullphaseamp_flat.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       _dp, ctypes.c_int, _dp, _dp, _dp]
    labels = np.ascontiguousarray(labels, np.int32)
    freq = np.ascontiguousarray(freq, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    amp = np.zeros((labels.size, freq.size))
    ph = np.zeros((labels.size, freq.size))
    st = L.qo_fullphaseamp_flat(labels.size, _i(labels), gmst, fref, Tobs, _d(freq), freq.size, _d(p),
                                _d(amp), _d(ph))
    return st, amp, ph


def heterodyne(labels, gmst, fref, D, SN, N, Tobs, params, fgrid, J, MN, MM):
    """heterodyne (QuickCBC.c:5112) on a given grid: (status, het dict)."""
    L = lib()
    L.qo_heterodyne_flat.restype = ctypes.c_int
    L.qo_heterodyne_flat.argtypes = ([ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                      ctypes.c_int, ctypes.c_double, _ip] + [ctypes.c_int] * 4 + [_dp] * 10)
    labels = np.ascontiguousarray(labels, np.int32)
    nd = labels.size
    fgrid = np.ascontiguousarray(fgrid, np.int32)
    M = fgrid.size
    NR = (M - 1) // J
    D = np.ascontiguousarray(D, np.float64)
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    h = dict(M=M, J=J, NR=NR, MN=MN, MM=MM, fgrid=fgrid, freq=fgrid / Tobs,
             amp=np.zeros((nd, M)), phase=np.zeros((nd, M)), IP=np.zeros((NR, J + 1, J + 1)),
             SL=np.zeros((NR, nd, J + 1)), lc=np.zeros((NR, nd, J + 1)), ls=np.zeros((NR, nd, J + 1)),
             rc=np.zeros((nd, NR)), rs=np.zeros((nd, NR)), aa=np.zeros((nd, NR)))
    lb = np.zeros(1)
    st = L.qo_heterodyne_flat(nd, _i(labels), gmst, fref, _d(D), _d(SN), _d(p), N, Tobs, _i(fgrid), M, J, MN, MM,
                              *[_d(h[k]) for k in ("amp", "phase", "IP", "SL", "lc", "ls", "rc", "rs", "aa")],
                              _d(lb))
    h["logLb"] = float(lb[0])
    h["pref"] = p.copy()
    return st, h


def loglike_het(labels, gmst, fref, Tobs, het, params, nthreads=0):
    """log_likelihood_het (QuickCBC.c:5303) for a batch: (logL, status)."""
    L = lib()
    L.qo_loglike_het_batch_flat.restype = None
    L.qo_loglike_het_batch_flat.argtypes = ([ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int] + [_dp] * 10 +
                                            [ctypes.c_double, _dp, ctypes.c_int, _dp, _ip, ctypes.c_int])
    labels = np.ascontiguousarray(labels, np.int32)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
    B = p.shape[0]
    logL = np.zeros(B)
    st = np.zeros(B, np.int32)
    arr = [np.ascontiguousarray(het[k], np.float64)
           for k in ("freq", "amp", "phase", "IP", "SL", "lc", "ls", "rc", "rs", "aa")]
    L.qo_loglike_het_batch_flat(labels.size, _i(labels), gmst, fref, Tobs, het["M"], het["J"],
                                *[_d(a) for a in arr], het["logLb"], _d(p), B, _d(logL), _i(st), nthreads)
    return logL, st


def fisher_het(labels, gmst, fref, Tobs, het, params, nthreads=0):
    """Fisher_Het (QuickCBC.c:5446) for a batch: (fish B x 11 x 11, status)."""
    L = lib()
    L.qo_fisher_het_batch_flat.restype = None
    L.qo_fisher_het_batch_flat.argtypes = ([ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_int] + [_dp] * 7 +
                                           [ctypes.c_int, _dp, _ip, ctypes.c_int])
    labels = np.ascontiguousarray(labels, np.int32)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
    B = p.shape[0]
    fish = np.zeros((B, 11, 11))
    st = np.zeros(B, np.int32)
    arr = [np.ascontiguousarray(het[k], np.float64) for k in ("freq", "amp", "phase", "IP", "SL", "aa")]
    L.qo_fisher_het_batch_flat(labels.size, _i(labels), gmst, fref, Tobs, het["M"], het["J"],
                               *[_d(a) for a in arr[:5]], _d(arr[5]), _d(p), B, _d(fish), _i(st), nthreads)
    return fish, st


def fisher_evec(fish):
    """FisherEvec (QuickCBC.c:2449) for a batch of d x d matrices: (ej B x d, ev B x d x d, sweeps)."""
    L = lib()
    L.qo_fisher_evec_batch.restype = None
    L.qo_fisher_evec_batch.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _ip]
    F = np.ascontiguousarray(fish, np.float64)
    d = F.shape[-1]
    F = F.reshape(-1, d, d)
    B = F.shape[0]
    ej = np.zeros((B, d))
    ev = np.zeros((B, d, d))
    sw = np.zeros(B, np.int32)
    L.qo_fisher_evec_batch(_d(F), B, d, _d(ej), _d(ev), _i(sw))
    return ej, ev, sw


def pt_update_hist_init(x, n_temps, n_chains, n_hist, seed):
    """The initial histories of QuickCBC's PT step (n_chains x n_temps x n_hist x 11)."""
    L = lib()
    L.qo_update_hist_init.restype = None
    L.qo_update_hist_init.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _dp]
    x = np.ascontiguousarray(x, np.float64)
    hist = np.zeros((n_chains, n_temps, n_hist, 11))
    L.qo_update_hist_init(_d(x), n_temps, n_chains, n_hist, seed, _d(hist))
    return hist


def update_propose(x, ej, ev, hist, sT, mix, seed, ctr):
    """update()'s proposal from x with the counters (seed, ctr); mix = a, b, c, etamin, gmst, n_det,
    labels[3]: (type, y, log density of a sky-ring map or 0)."""
    L = lib()
    L.qo_update_propose.restype = ctypes.c_int
    L.qo_update_propose.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_uint64,
                                    ctypes.c_uint64, _dp, _dp]
    m = np.zeros(9)
    mm = np.asarray(mix, np.float64)
    m[:mm.size] = mm
    a = [np.ascontiguousarray(v, np.float64) for v in (x, ej, ev, hist, m)]
    y = np.zeros(11)
    j = np.zeros(1)
    t = L.qo_update_propose(_d(a[0]), _d(a[1]), _d(a[2]), _d(a[3]), a[3].size // 11, sT, _d(a[4]), seed, ctr, _d(y),
                            _d(j))
    return t, y, float(j[0])


def ring_all(px, d1, d2, gmst, u_rot):
    """Ring_all (Utilities.c:1445): y = x with the sky location moved on the d1-d2 ring."""
    L = lib()
    L.qo_ring_all.restype = None
    L.qo_ring_all.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    x = np.ascontiguousarray(px, np.float64)
    y = x.copy()
    L.qo_ring_all(_d(x), _d(y), d1, d2, gmst, u_rot)
    return y


def skymap_all(px, py, gmst, ifo1, ifo2):
    """skymap_all (QuickCBC.c:7085): y's extrinsics for its sky location."""
    L = lib()
    L.qo_skymap_all.restype = None
    L.qo_skymap_all.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    x = np.ascontiguousarray(px, np.float64)
    y = np.ascontiguousarray(py, np.float64).copy()
    L.qo_skymap_all(_d(x), _d(y), gmst, ifo1, ifo2)
    return y


def skydensity_all(px, py, gmst, ifo1, ifo2):
    """skydensity_all (QuickCBC.c:7411): (density, y after skymap_all)."""
    L = lib()
    L.qo_skydensity_all.restype = ctypes.c_double
    L.qo_skydensity_all.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    x = np.ascontiguousarray(px, np.float64)
    y = np.ascontiguousarray(py, np.float64).copy()
    d = L.qo_skydensity_all(_d(x), _d(y), gmst, ifo1, ifo2)
    return d, y


def pt_update_run(labels, gmst, fref, Tobs, het, uconst, n_hist, hist_every, n_temps, n_chains, ladder, seed,
                  nsteps, state):
    """QuickCBC's PT step, nsteps times, on the state dict (x, L, ej, ev, hist, hcount, y, typ, jac, stats,
    step), updated in place; uconst = min[11] + max[11] + [a, b, c, etamin, gmst, n_det, labels[3]]."""
    L = lib()
    L.qo_pt_update_run.restype = None
    L.qo_pt_update_run.argtypes = ([ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int] + [_dp] * 10 +
                                   [ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_longlong), _dp, _dp, _dp, _dp, _dp,
                                    ctypes.POINTER(ctypes.c_longlong), _dp, _ip, _dp,
                                    ctypes.POINTER(ctypes.c_longlong)])
    labels = np.ascontiguousarray(labels, np.int32)
    arr = [np.ascontiguousarray(het[k], np.float64)
           for k in ("freq", "amp", "phase", "IP", "SL", "lc", "ls", "rc", "rs", "aa")]
    uc = np.ascontiguousarray(uconst, np.float64)
    for k in ("x", "L", "ej", "ev", "hist", "y", "jac"):
        state[k] = np.ascontiguousarray(state[k], np.float64)
    state["typ"] = np.ascontiguousarray(state["typ"], np.int32)
    hc = (ctypes.c_longlong * n_chains)(*[int(v) for v in state["hcount"]])
    st = (ctypes.c_longlong * (11 * n_chains))(*[int(v) for v in np.asarray(state["stats"]).ravel()])
    step = ctypes.c_longlong(int(state["step"]))
    L.qo_pt_update_run(labels.size, _i(labels), gmst, fref, Tobs, het["M"], het["J"], *[_d(a) for a in arr],
                       het["logLb"], _d(uc), n_hist, hist_every, n_temps, n_chains, ladder, seed, nsteps,
                       ctypes.byref(step), *[_d(state[k]) for k in ("x", "L", "ej", "ev", "hist")], hc,
                       _d(state["y"]), _i(state["typ"]), _d(state["jac"]), st)
    state["hcount"] = np.array(hc[:], np.int64)
    state["stats"] = np.array(st[:], np.int64).reshape(n_chains, 11)
    state["step"] = step.value
    return state


def update_prior(y, uconst):
    """update()'s prior check: (flag, mass Jacobian term)."""
    L = lib()
    L.qo_update_prior.restype = ctypes.c_int
    L.qo_update_prior.argtypes = [_dp, _dp, _dp]
    y = np.ascontiguousarray(y, np.float64)
    uc = np.ascontiguousarray(uconst, np.float64)
    j = np.zeros(1)
    f = L.qo_update_prior(_d(y), _d(uc), _d(j))
    return f, float(j[0])


def templates(nifo, fref, N, Tobs, params):
    """templates (QuickCBC.c:3527): (status, hwave nifo x N) for the reference-frame parameters."""
    L = lib()
    L.qo_templates_flat.restype = ctypes.c_int
    L.qo_templates_flat.argtypes = [ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double, _dp]
    p = np.ascontiguousarray(params, np.float64)
    h = np.zeros((nifo, N))
    st = L.qo_templates_flat(nifo, fref, _d(h), N, Tobs, _d(p))
    return st, h


def loglike_intrinsic(nifo, fref, fmin, fmax, D, SN, N, Tobs, params):
    """log_likelihood_intrinsic (QuickCBC.c:4093) for a batch: (logL, snrsq B x nifo, status)."""
    L = lib()
    L.qo_loglike_intrinsic_batch_flat.restype = None
    L.qo_loglike_intrinsic_batch_flat.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                  _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, _dp,
                                                  _dp, _ip]
    ips = 7 + 3 * (nifo - 1)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, ips)
    B = p.shape[0]
    logL = np.zeros(B)
    snr = np.zeros((B, nifo))
    st = np.zeros(B, np.int32)
    D = np.ascontiguousarray(D, np.float64)
    SN = np.ascontiguousarray(SN, np.float64)
    L.qo_loglike_intrinsic_batch_flat(nifo, fref, fmin, fmax, _d(D), _d(SN), N, Tobs, _d(p), B, _d(logL), _d(snr),
                                      _i(st))
    return logL, snr, st


def tukey(N, alpha):
    """tukey() (Utilities.c:635-654): the window values, flat for (1 - alpha) of the N samples."""
    i = np.arange(N, dtype=np.float64)
    imin = int(alpha * (N - 1) / 2.0)
    imax = int((N - 1) * (1.0 - alpha / 2.0))
    Nwin = N - imax
    w = np.ones(N)
    lo = i < imin
    hi = i > imax
    w[lo] = 0.5 * (1.0 + np.cos(np.pi * (i[lo] / imin - 1.0)))
    w[hi] = 0.5 * (1.0 + np.cos(np.pi * ((i[hi] - imax) / Nwin)))
    return w


def halfcomplex(X, N):
    """GSL's half-complex layout of the N/2 + 1 coefficients of a real FFT."""
    out = np.zeros(N)
    h = N // 2
    out[0] = X[0].real
    out[h] = X[h].real
    out[1:h] = X[1:h].real
    out[N - np.arange(1, h)] = X[1:h].imag
    return out


def condition(x, Tobs, t_rise=0.4):
    """QuickCBC.c:268-313: tukey(alpha = 2 t_rise / Tobs), the real FFT in the half-complex layout
    (gsl_fft_real_radix2_transform, forward sign), times sqrt(Tobs) / N; x is n_det x N."""
    x = np.atleast_2d(np.asarray(x, np.float64))
    N = x.shape[1]
    w = tukey(N, 2.0 * t_rise / Tobs)
    return np.array([halfcomplex(np.fft.rfft(row * w), N) for row in x]) * (np.sqrt(Tobs) / N)


def tukey_scale(alpha, N):
    """tukey_scale() (Utilities.c:656-683): (mean, root mean square) of the window."""
    w = tukey(N, alpha)
    return w.sum() / N, np.sqrt((w * w).sum() / N)


def whiten(data, Sn):
    """whiten() (Utilities.c:294-309) of a half-complex row: bins 1 .. N/2 - 1 times 1/sqrt(S), DC and Nyquist 0."""
    d = np.array(data, np.float64)
    N = d.size
    h = N // 2
    x = 1.0 / np.sqrt(np.asarray(Sn, np.float64)[1:h])
    d[1:h] *= x
    d[N - np.arange(1, h)] *= x
    d[0] = d[h] = 0.0
    return d


def condition_outputs(x, SN, Tobs, t_rise=0.4):
    """main's conditioning (QuickCBC.c:250-313) of x (n_det x N) with the PSD SN (n_det x N/2):
    D (the scaled half-complex data), DW (dataw.dat's whitened series), pspec and wcheck (the
    spectra of pspec_*.dat and wcheck_*.dat; bin 0 takes the DC term only)."""
    x = np.atleast_2d(np.asarray(x, np.float64))
    SN = np.atleast_2d(np.asarray(SN, np.float64))
    nd, N = x.shape
    h = N // 2
    w = tukey(N, 2.0 * t_rise / Tobs)
    raw = np.array([halfcomplex(np.fft.rfft(row * w), N) for row in x])  # unscaled, GSL's layout
    fac = np.sqrt(N // 2) / np.sqrt(Tobs)
    DW = np.zeros((nd, N))
    for d in range(nd):
        W = whiten(raw[d], SN[d])
        # GSL's half-complex inverse: the backward sum times 1/N
        Z = np.zeros(h + 1, complex)
        Z[:h + 1] = W[:h + 1]
        Z[1:h] += 1j * W[N - np.arange(1, h)]
        DW[d] = np.fft.irfft(Z, N) / fac
    pfac = Tobs / (float(N) * float(N))
    sq = raw[:, :h] ** 2
    sq[:, 1:] += raw[:, N - np.arange(1, h)] ** 2
    pspec = pfac * 2.0 * sq
    return {"D": raw * (np.sqrt(Tobs) / N), "DW": DW, "pspec": pspec, "wcheck": pspec / SN}


def framed_count(text):
    """The count of QuickCBC.c:204-210 for a file of T whitespace-separated numbers: each fscanf takes two,
    and the end of the file is seen by the one that finds no further number, or whose last number ends
    the file (no character after it); the count is the calls made less one."""
    T = len(text.split())
    if T == 0:
        return 0
    if text[-1].isspace() or T % 2 == 1:
        return T // 2
    return T // 2 - 1


def read_framed(path, cap=1 << 24):
    """A framed_* file through stdio as QuickCBC.c:204-220 reads it: (N, time, strain)."""
    L = lib()
    L.qo_read_framed.restype = ctypes.c_int
    L.qo_read_framed.argtypes = [ctypes.c_char_p, ctypes.c_int, _dp, _dp]
    N = L.qo_read_framed(str(path).encode(), 0, None, None)
    if N < 0:
        return N, None, None
    t, x = np.zeros(max(N, 1)), np.zeros(max(N, 1))
    L.qo_read_framed(str(path).encode(), min(N, cap), _d(t), _d(x))
    return N, t[:N], x[:N]


def read_spec(path, n2):
    """A spec_* file through stdio as QuickCBC.c:242-248 reads it: (rows read, SN, SM)."""
    L = lib()
    L.qo_read_spec.restype = ctypes.c_int
    L.qo_read_spec.argtypes = [ctypes.c_char_p, ctypes.c_int, _dp, _dp]
    SN, SM = np.zeros(n2), np.zeros(n2)
    r = L.qo_read_spec(str(path).encode(), n2, _d(SN), _d(SM))
    return r, SN, SM


def makespec(ffit, Sspline, lines, Tobs, N):
    """makespec (Utilities.c:1908): (status, SN of N/2); lines = (linef, lineh, lineQ, linew, deltafmax) or None."""
    L = lib()
    L.qo_makespec.restype = ctypes.c_int
    L.qo_makespec.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 8 + [ctypes.c_double, ctypes.c_int]
    x = np.ascontiguousarray(ffit, np.float64)
    y = np.ascontiguousarray(Sspline, np.float64)
    Ls = [np.ascontiguousarray(v, np.float64) for v in lines] if lines is not None else [np.zeros(1)] * 5
    nl = Ls[0].size if lines is not None else 0
    SN = np.zeros(N // 2)
    st = L.qo_makespec(x.size, nl, _d(x), _d(y), *[_d(v) for v in Ls], _d(SN), Tobs, N)
    return st, SN


# ----------------------------------------------- heterodyne grid choice
ETAMIN = 0.05  # ConstCBC.h:57


def fringdown(params):
    """fringdown (Utilities.c:567): ringdown frequency (Hz)."""
    L = lib()
    L.qo_fringdown.restype = ctypes.c_double
    L.qo_fringdown.argtypes = [_dp]
    p = np.ascontiguousarray(params, np.float64)
    return L.qo_fringdown(_d(p))


def limits(labels, gmst, fref, fmin, SN, N, Tobs, params):
    """limits (QuickCBC.c:4275): (status, [MN, MM], ratio (N/2))."""
    L = lib()
    L.qo_limits_flat.restype = ctypes.c_int
    L.qo_limits_flat.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                 ctypes.c_int, ctypes.c_double, _ip, _dp]
    labels = np.ascontiguousarray(labels, np.int32)
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    stst = np.zeros(2, np.int32)
    ratio = np.ones(N // 2)
    st = L.qo_limits_flat(labels.size, _i(labels), gmst, fref, fmin, _d(SN), _d(p), N, Tobs, _i(stst), _d(ratio))
    return st, (int(stst[0]), int(stst[1])), ratio


def fisher_all(labels, gmst, fref, fmin, fmax, SN, N, Tobs, params):
    """Fisher_All (QuickCBC.c:2695): (status, fish 11 x 11)."""
    L = lib()
    L.qo_fisher_all_flat.restype = ctypes.c_int
    L.qo_fisher_all_flat.argtypes = [ctypes.c_int, _ip] + [ctypes.c_double] * 4 + [_dp, _dp, ctypes.c_int,
                                                                                   ctypes.c_double, _dp]
    labels = np.ascontiguousarray(labels, np.int32)
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    fish = np.zeros((11, 11))
    st = L.qo_fisher_all_flat(labels.size, _i(labels), gmst, fref, fmin, fmax, _d(SN), _d(p), N, Tobs, _d(fish))
    return st, fish


def fisher_full(nifo, fref, fmin, fmax, SN, N, Tobs, params, d=7):
    """Fisher_Full (QuickCBC.c:2551) of a reference-frame point: (status, fish d x d)."""
    L = lib()
    L.qo_fisher_full_flat.restype = ctypes.c_int
    L.qo_fisher_full_flat.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + [_dp, _dp, ctypes.c_int, ctypes.c_int,
                                                                              ctypes.c_double, _dp]
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    fish = np.zeros((d, d))
    st = L.qo_fisher_full_flat(nifo, fref, fmin, fmax, _d(SN), _d(p), d, N, Tobs, _d(fish))
    return st, fish


def diff(labels, gmst, fref, fmin, SN, N, Tobs, params, dparams):
    """diff (QuickCBC.c:4913)."""
    L = lib()
    L.qo_diff_flat.restype = ctypes.c_double
    L.qo_diff_flat.argtypes = [ctypes.c_int, _ip] + [ctypes.c_double] * 3 + [_dp, _dp, _dp, ctypes.c_int,
                                                                             ctypes.c_double]
    labels = np.ascontiguousarray(labels, np.int32)
    SN = np.ascontiguousarray(SN, np.float64)
    a = np.ascontiguousarray(params, np.float64)
    b = np.ascontiguousarray(dparams, np.float64)
    return L.qo_diff_flat(labels.size, _i(labels), gmst, fref, fmin, _d(SN), _d(a), _d(b), N, Tobs)


def het_diff(labels, gmst, fref, Tobs, het, params, dparams):
    """het_diff (QuickCBC.c:4966)."""
    L = lib()
    L.qo_het_diff_flat.restype = ctypes.c_double
    L.qo_het_diff_flat.argtypes = ([ctypes.c_int, _ip] + [ctypes.c_double] * 3 + [ctypes.c_int] * 2 +
                                   [_dp] * 7)
    labels = np.ascontiguousarray(labels, np.int32)
    arr = [np.ascontiguousarray(het[k], np.float64) for k in ("freq", "amp", "IP", "SL", "aa")]
    a = np.ascontiguousarray(params, np.float64)
    b = np.ascontiguousarray(dparams, np.float64)
    return L.qo_het_diff_flat(labels.size, _i(labels), gmst, fref, Tobs, het["M"], het["J"], *[_d(x) for x in arr],
                              _d(a), _d(b))


def displace(params, direction, step, pmin, pmax, etamin=ETAMIN):
    """The displaced point of efix and het_space (QuickCBC.c:4418-4450)."""
    L = lib()
    L.qo_displace.restype = None
    L.qo_displace.argtypes = [_dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_double, _dp]
    a = [np.ascontiguousarray(x, np.float64) for x in (params, direction, pmin, pmax)]
    out = np.zeros(11)
    L.qo_displace(_d(a[0]), _d(a[1]), step, _d(a[2]), _d(a[3]), etamin, _d(out))
    return out


def efix(labels, gmst, fref, Tobs, params, ej, ev, pmin, pmax, zs, het=None, fmin=None, SN=None, N=None,
         etamin=ETAMIN):
    """efix (QuickCBC.c:4387) for a batch of points: the scaled jump sizes (B x 11) and the
    distances evaluated per direction (B x 11).  het given: het_diff on it (hflag 1), else diff
    on the full data SN (hflag 0)."""
    L = lib()
    L.qo_efix_batch_flat.restype = None
    L.qo_efix_batch_flat.argtypes = ([ctypes.c_int, _ip] + [ctypes.c_double] * 3 + [_dp, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int] + [_dp] * 6 +
                                     [ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_double, _ip])
    labels = np.ascontiguousarray(labels, np.int32)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
    B = p.shape[0]
    e = np.array(ej, np.float64).reshape(B, 11)
    v = np.ascontiguousarray(ev, np.float64).reshape(B, 11, 11)
    lo = np.ascontiguousarray(pmin, np.float64)
    hi = np.ascontiguousarray(pmax, np.float64)
    trials = np.zeros((B, 11), np.int32)
    if het is not None:
        arr = [np.ascontiguousarray(het[k], np.float64) for k in ("freq", "amp", "IP", "SL", "aa")]
        M, J, ptrs = het["M"], het["J"], [_d(x) for x in arr]
        S, NN, fm = None, 0, 0.0
    else:
        S = np.ascontiguousarray(SN, np.float64)
        M, J, ptrs = 0, 1, [None] * 5
        NN, fm = N, fmin
    L.qo_efix_batch_flat(labels.size, _i(labels), gmst, fref, fm, _d(S) if S is not None else None, NN, Tobs, M, J,
                         *ptrs, _d(p), B, _d(lo), _d(hi), etamin, _d(e), _d(v), zs, _i(trials))
    return e, trials


def het_space(labels, gmst, fref, fmin, fmax, SN, N, Tobs, params, pmin, pmax, etamin=ETAMIN):
    """het_space (QuickCBC.c:4570): (status, dict with MN, MM, M, NR, J, fgrid, eval, evec)."""
    L = lib()
    L.qo_het_space_flat.restype = ctypes.c_int
    L.qo_het_space_flat.argtypes = ([ctypes.c_int, _ip] + [ctypes.c_double] * 4 + [_dp, _dp, _dp, _dp,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_double, _ip, _ip, _dp, _dp])
    labels = np.ascontiguousarray(labels, np.int32)
    SN = np.ascontiguousarray(SN, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    lo = np.ascontiguousarray(pmin, np.float64)
    hi = np.ascontiguousarray(pmax, np.float64)
    out4 = np.zeros(4, np.int32)
    fgrid = np.zeros(N // 2, np.int32)
    ev = np.zeros(11)
    evec = np.zeros((11, 11))
    st = L.qo_het_space_flat(labels.size, _i(labels), gmst, fref, fmin, fmax, _d(SN), _d(p), _d(lo), _d(hi),
                             etamin, N, Tobs, _i(out4), _i(fgrid), _d(ev), _d(evec))
    MN, MM, M, NR = (int(x) for x in out4)
    return st, dict(MN=MN, MM=MM, M=M, NR=NR, J=2, fgrid=fgrid[:M].copy(), eval=ev, evec=evec)


# ------------------------------------------------- search likelihoods
LIGHT_TIMES = {(0, 1): 0.0106, (0, 2): 0.0278, (1, 2): 0.0257}  # ConstCBC.h:19-21


def delays(labels):
    """net->delays as QuickCBC's set-up fills it (QuickCBC.c:7776-7790): nifo x nifo."""
    n = len(labels)
    d = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                d[i, j] = LIGHT_TIMES[tuple(sorted((labels[i], labels[j])))]
    return d


def penalized(params, D, twave, SN, fmin, fmax, N, Tobs):
    """log_likelihood_penalized (QuickCBC.c:6233) of one detector: (logL, A, P)."""
    L = lib()
    L.qo_penalized_flat.restype = ctypes.c_double
    L.qo_penalized_flat.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double, _dp]
    a = [np.ascontiguousarray(x, np.float64) for x in (params, D, twave, SN)]
    AP = np.zeros(2)
    v = L.qo_penalized_flat(*[_d(x) for x in a], fmin, fmax, N, Tobs, _d(AP))
    return v, AP[0], AP[1]


def pdwave(N, Tobs, tcmin, params, fref=100.0):
    """PDwave (QuickCBC.c:3448): (status, half-complex N)."""
    L = lib()
    L.qo_pdwave_flat.restype = ctypes.c_int
    L.qo_pdwave_flat.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_double]
    w = np.zeros(N)
    p = np.ascontiguousarray(params, np.float64)
    st = L.qo_pdwave_flat(_d(w), N, Tobs, tcmin, _d(p), fref)
    return st, w


def pbt_shift(data1, data2, Sn, imin, imax, N):
    """pbt_shift (Utilities.c:687): (corr, corrf)."""
    L = lib()
    L.qo_pbt_shift.restype = None
    L.qo_pbt_shift.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    a = [np.ascontiguousarray(x, np.float64) for x in (data1, data2, Sn)]
    c, cf = np.zeros(N), np.zeros(N)
    L.qo_pbt_shift(_d(c), _d(cf), *[_d(x) for x in a], imin, imax, N)
    return c, cf


def apmax(nifo, fref, fmin, fmax, D, SN, N, Tobs, params):
    """log_likelihood_APmax (QuickCBC.c:5672) for a batch: (logL, moved params, status)."""
    L = lib()
    L.qo_apmax_batch_flat.restype = None
    L.qo_apmax_batch_flat.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + [_dp, _dp, ctypes.c_int,
                                                                             ctypes.c_double, _dp, ctypes.c_int, _dp,
                                                                             _ip]
    ips = 7 + 3 * (nifo - 1)
    p = np.array(params, np.float64).reshape(-1, ips)
    B = p.shape[0]
    logL = np.zeros(B)
    st = np.zeros(B, np.int32)
    L.qo_apmax_batch_flat(nifo, fref, fmin, fmax, _d(np.ascontiguousarray(D, np.float64)),
                          _d(np.ascontiguousarray(SN, np.float64)), N, Tobs, _d(p), B, _d(logL), _i(st))
    return logL, p, st


def llmax(labels, fref, fmin, fmax, D, SN, N, Tobs, tmin, tmax, net_tmax, params):
    """log_likelihood_max (QuickCBC.c:5765) for a batch: (logL, moved params, status)."""
    L = lib()
    L.qo_llmax_batch_flat.restype = None
    L.qo_llmax_batch_flat.argtypes = ([ctypes.c_int] + [ctypes.c_double] * 3 + [_dp, _dp, ctypes.c_int] +
                                      [ctypes.c_double] * 4 + [_dp, _dp, ctypes.c_int, _dp, _ip])
    nifo = len(labels)
    ips = 7 + 3 * (nifo - 1)
    p = np.array(params, np.float64).reshape(-1, ips)
    B = p.shape[0]
    logL = np.zeros(B)
    st = np.zeros(B, np.int32)
    dl = np.ascontiguousarray(delays(labels), np.float64)
    L.qo_llmax_batch_flat(nifo, fref, fmin, fmax, _d(np.ascontiguousarray(D, np.float64)),
                          _d(np.ascontiguousarray(SN, np.float64)), N, Tobs, tmin, tmax, net_tmax, _d(dl), _d(p), B,
                          _d(logL), _i(st))
    return logL, p, st
