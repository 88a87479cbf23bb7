"""ctypes binding of libqcbc (include/qcbc.h).

The library is the product: every call below goes through its C ABI and runs
on the GPU.  There is no CPU fallback; if libqcbc.so is missing or no gfx950
device is usable, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqcbc.so")

FP64, MIXED = 0, 1
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

ERRORS = {0: "QCBC_OK", 1: "QCBC_EINVAL", 12: "QCBC_ENOMEM", 14: "QCBC_EFAULT", 33: "QCBC_EDOM",
          34: "QCBC_ERANGE", 5: "QCBC_EIO", 100: "QCBC_EHIP", 101: "QCBC_ENODEV", 102: "QCBC_ENOTBOUND", 103: "QCBC_EBUSY",
          104: "QCBC_ECANCELED", 105: "QCBC_ETIMEDOUT"}


class QcbcError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("%s: %s (%s)" % (what, ERRORS.get(code, code), strerror(code)))


class Desc(ctypes.Structure):
    _fields_ = [("n_det", ctypes.c_int), ("labels", ctypes.c_int * 3), ("gmst", ctypes.c_double),
                ("N", ctypes.c_int), ("Tobs", ctypes.c_double), ("fmin", ctypes.c_double),
                ("fmax", ctypes.c_double), ("fref", ctypes.c_double)]


class Net(ctypes.Structure):
    """struct Net of QuickCBC.h:13-25."""
    _fields_ = [("Nifo", ctypes.c_int), ("Tobs", ctypes.c_double), ("tmax", ctypes.c_double),
                ("tmin", ctypes.c_double), ("GMST", ctypes.c_double), ("labels", _ip),
                ("tds", _dp), ("delays", ctypes.POINTER(_dp)), ("lstart", _ip), ("lstop", _ip)]


class RealVector(ctypes.Structure):
    _fields_ = [("data", _dp), ("length", ctypes.c_size_t)]


class AmpPhaseFDWaveform(ctypes.Structure):
    _fields_ = [("freq", _dp), ("amp", _dp), ("phase", _dp), ("length", ctypes.c_size_t)]


_dpp = ctypes.POINTER(_dp)


class HetStruct(ctypes.Structure):
    """struct Het of QuickCBC.h:27-50."""
    _fields_ = [("M", ctypes.c_int), ("MM", ctypes.c_int), ("MN", ctypes.c_int), ("ell", ctypes.c_int),
                ("J", ctypes.c_int), ("NR", ctypes.c_int), ("hh", ctypes.c_double), ("hr", ctypes.c_double),
                ("logLb", ctypes.c_double), ("IP", ctypes.POINTER(_dpp)), ("SL", ctypes.POINTER(_dpp)),
                ("amp", _dpp), ("phase", _dpp), ("rc", _dpp), ("rs", _dpp), ("aa", _dpp),
                ("lc", ctypes.POINTER(_dpp)), ("ls", ctypes.POINTER(_dpp)), ("pref", _dp), ("fgrid", _ip),
                ("freq", ctypes.POINTER(RealVector))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqcbc.so is not built (run python -m quickcbc_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        L.qcbc_strerror.restype = ctypes.c_char_p
        L.qcbc_strerror.argtypes = [ctypes.c_int]
        L.qcbc_ctx_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(_vp)]
        L.qcbc_ctx_destroy.argtypes = [_vp]
        L.qcbc_ctx_footprint.restype = ctypes.c_size_t
        L.qcbc_ctx_footprint.argtypes = [ctypes.POINTER(Desc)]
        L.qcbc_bind_data.argtypes = [_vp, _dp, _dp]
        L.qcbc_loglike.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _ip]
        L.qcbc_loglike_device.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
        L.qcbc_loglike_device_noalloc.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
        L.qcbc_loglike_plain.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _ip]
        L.qcbc_waveform.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _ip]
        L.qcbc_templates.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip]
        L.qcbc_ctx_set_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.qcbc_ctx_get_shape.argtypes = [_vp, ctypes.c_int, _ip, _ip, _ip]
        L.qcbc_autotune.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.qcbc_ctx_timing.argtypes = [_vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_ctx_sync.argtypes = [_vp]
        L.qcbc_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, _ip, _dp]
        # drop-ins
        L.log_likelihood_full.restype = ctypes.c_double
        L.log_likelihood_full.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(_dp), _dp, ctypes.POINTER(RealVector),
                                          ctypes.POINTER(_dp), _dp, ctypes.c_int, ctypes.c_double]
        L.fulltemplates.restype = None
        L.fulltemplates.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(_dp), ctypes.POINTER(RealVector), _dp,
                                    ctypes.c_int]
        L.IMRPhenomDGenerateh22FDAmpPhase.restype = ctypes.c_int
        L.IMRPhenomDGenerateh22FDAmpPhase.argtypes = [ctypes.POINTER(ctypes.POINTER(AmpPhaseFDWaveform)),
                                                      ctypes.POINTER(RealVector)] + [ctypes.c_double] * 7
        L.DestroyAmpPhaseFDWaveform.argtypes = [ctypes.POINTER(AmpPhaseFDWaveform)]
        L.fourier_nwip.restype = ctypes.c_double
        L.fourier_nwip.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.F_ant.restype = None
        L.F_ant.argtypes = [ctypes.c_double] * 4 + [_dp, _dp, ctypes.c_int]
        L.Times.restype = None
        L.Times.argtypes = [ctypes.c_double] * 3 + [_dp]
        L.fbegin.restype = ctypes.c_double
        L.fbegin.argtypes = [_dp]
        L.qcbc_set_band.argtypes = [ctypes.c_double, ctypes.c_double]
        L.qcbc_set_fref.argtypes = [ctypes.c_double]
        L.qcbc_set_exit_on_error.argtypes = [ctypes.c_int]
        L.fullphaseamp.restype = None
        L.fullphaseamp.argtypes = [ctypes.POINTER(Net), _dpp, _dpp, ctypes.POINTER(RealVector), _dp, ctypes.c_int]
        L.heterodyne.restype = None
        L.heterodyne.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(HetStruct), _dpp, _dp,
                                 ctypes.POINTER(RealVector), _dpp, ctypes.c_int, ctypes.c_double]
        L.log_likelihood_het.restype = ctypes.c_double
        L.log_likelihood_het.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(HetStruct), _dp, ctypes.c_double]
        # heterodyned likelihood
        L.qcbc_het_create.argtypes = [_vp, _dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_vp)]
        L.qcbc_het_import.argtypes = [_vp, ctypes.c_int, ctypes.c_int] + [_dp] * 10 + [ctypes.c_double,
                                                                                        ctypes.POINTER(_vp)]
        L.qcbc_het_info.argtypes = [_vp, _ip, _ip, _ip, _ip, _dp]
        L.qcbc_het_export.argtypes = [_vp] + [_dp] * 10
        L.qcbc_het_weights.argtypes = [_vp, _dp]
        L.qcbc_het_loglike.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip]
        L.qcbc_het_loglike_device.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
        L.qcbc_het_sync.argtypes = [_vp]
        L.qcbc_het_timing.argtypes = [_vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_het_time_kernel.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _dp]
        L.qcbc_het_destroy.argtypes = [_vp]
        L.qcbc_bind_time_series.argtypes = [_vp, _dp, _dp, ctypes.c_double, _dp]
        L.qcbc_makespec.argtypes = [_vp, ctypes.c_int, ctypes.c_int] + [_dp] * 8
        L.qcbc_condition.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.qcbc_data_file_name.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int]
        L.qcbc_read_framed.argtypes = [ctypes.c_char_p, ctypes.c_int, _dp, _dp, _ip]
        L.qcbc_read_spec.argtypes = [ctypes.c_char_p, ctypes.c_int, _dp, _dp]
        L.qcbc_write_table.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _dpp, ctypes.c_int]
        L.tukey.restype = None
        L.tukey.argtypes = [_dp, ctypes.c_double, ctypes.c_int]
        L.tukey_scale.restype = None
        L.tukey_scale.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int]
        L.whiten.restype = None
        L.whiten.argtypes = [_dp, _dp, ctypes.c_int]
        # reference-frame likelihood
        L.qcbc_loglike_intrinsic.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _ip]
        L.qcbc_templates_intrinsic.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip]
        L.templates.restype = None
        L.templates.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(_dp), ctypes.POINTER(RealVector), _dp,
                                ctypes.c_int]
        L.log_likelihood_intrinsic.restype = ctypes.c_double
        L.log_likelihood_intrinsic.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(_dp), _dp,
                                               ctypes.POINTER(RealVector), ctypes.POINTER(_dp), ctypes.c_int,
                                               ctypes.c_double]
        # Fisher matrix and jump directions
        L.qcbc_het_fisher.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, _ip]
        L.qcbc_het_fisher_device.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.qcbc_het_fisher_time.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _dp]
        L.qcbc_fisher_evec.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.Fisher_Het.restype = None
        L.Fisher_Het.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(HetStruct), _dpp, _dp, ctypes.c_double]
        L.FisherEvec.restype = None
        L.FisherEvec.argtypes = [_dpp, _dp, _dpp, ctypes.c_int]
        # search likelihoods
        L.qcbc_loglike_apmax.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip]
        L.qcbc_loglike_max.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       _dp, _dp, _ip]
        L.log_likelihood_APmax.restype = ctypes.c_double
        L.log_likelihood_APmax.argtypes = [ctypes.POINTER(Net), _dpp, _dp, ctypes.POINTER(RealVector), _dpp,
                                           ctypes.c_int, ctypes.c_double]
        L.log_likelihood_max.restype = ctypes.c_double
        L.log_likelihood_max.argtypes = [ctypes.POINTER(Net), _dpp, _dp, ctypes.POINTER(RealVector), _dpp,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int]
        L.log_likelihood_penalized.restype = ctypes.c_double
        L.log_likelihood_penalized.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_int]
        L.PDwave.restype = None
        L.PDwave.argtypes = [_dp, ctypes.POINTER(RealVector), ctypes.c_double, _dp, ctypes.c_int]
        L.pbt_shift.restype = None
        L.pbt_shift.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        # the heterodyne's grid choice
        L.qcbc_limits.argtypes = [_vp, _dp, _ip, _dp]
        L.qcbc_fisher_all.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip]
        L.qcbc_fisher_full.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_int, _dp, _ip]
        L.Fisher_Full.restype = None
        L.Fisher_Full.argtypes = [ctypes.POINTER(Net), _dpp, ctypes.c_int, _dp, ctypes.POINTER(RealVector), _dpp,
                                  ctypes.c_int, ctypes.c_double]
        L.qcbc_diff.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip, ctypes.c_int, _dp]
        L.qcbc_efix.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _ip]
        L.qcbc_het_space.argtypes = [_vp, _dp, _dp, _dp, ctypes.c_double, _ip, _ip, _ip, _ip, _ip, ctypes.c_int,
                                     _dp, _dp, _dp]
        L.qcbc_het_diff.argtypes = [_vp, _dp, ctypes.c_int, _dp, _ip, ctypes.c_int, _dp]
        L.qcbc_het_efix.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                    _ip]
        L.limits.restype = None
        L.limits.argtypes = [ctypes.POINTER(Net), _dp, ctypes.POINTER(RealVector), _dpp, _ip, _dp, ctypes.c_int,
                             ctypes.c_double]
        L.Fisher_All.restype = None
        L.Fisher_All.argtypes = [ctypes.POINTER(Net), _dpp, _dp, ctypes.POINTER(RealVector), _dpp, ctypes.c_int,
                                 ctypes.c_double]
        L.diff.restype = ctypes.c_double
        L.diff.argtypes = [ctypes.POINTER(Net), _dp, _dp, ctypes.POINTER(RealVector), _dpp, ctypes.c_int,
                           ctypes.c_double]
        L.het_diff.restype = ctypes.c_double
        L.het_diff.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(HetStruct), _dp, _dp]
        L.efix.restype = None
        L.efix.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(HetStruct), _dp, _dp, _dp, ctypes.POINTER(RealVector),
                           _dpp, ctypes.c_int, ctypes.c_double, _dp, _dpp, ctypes.c_double, ctypes.c_int]
        L.het_space.restype = None
        L.het_space.argtypes = [ctypes.POINTER(Net), ctypes.POINTER(HetStruct), _dp, _dp, _dp,
                                ctypes.POINTER(RealVector), _dpp, ctypes.c_int, ctypes.c_double]
        _lib = L
    return _lib


def strerror(code):
    try:
        return lib().qcbc_strerror(code).decode()
    except Exception:  # noqa: BLE001 - error text only
        return "?"


def _check(code, what):
    if code != 0:
        raise QcbcError(code, what)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int(0)
    mem = ctypes.c_double(0)
    _check(lib().qcbc_device_info(name, 256, ctypes.byref(cu), ctypes.byref(mem)), "qcbc_device_info")
    return name.value.decode(), cu.value, mem.value


def make_desc(an):
    d = Desc()
    d.n_det = an.nifo
    for i, l in enumerate(an.labels):
        d.labels[i] = l
    d.gmst = an.gmst
    d.N = an.N
    d.Tobs = an.Tobs
    d.fmin = an.fmin
    d.fmax = an.fmax
    d.fref = an.fref
    return d


class _Handle:
    """Lifetime links between handles: a handle made on another (an ensemble
    or a heterodyne on an analysis) must be destroyed first.  Owners hold
    their dependents strongly, so even when the garbage collector finalises
    both in one cycle (in any order), closing the owner closes the dependents
    before it."""

    def _depend_on(self, *owners):
        self._owners = [o for o in owners if o is not None]
        for o in self._owners:
            o.__dict__.setdefault("_deps", []).append(self)

    def _release(self):
        for o in self.__dict__.get("_owners", []):
            deps = o.__dict__.get("_deps")
            if deps is not None:
                o._deps = [d for d in deps if d is not self]
        self._owners = []

    def _close_deps(self):
        deps = self.__dict__.get("_deps", [])
        self._deps = []
        for d in reversed(deps):
            d.close()


class Context(_Handle):
    """One resident analysis on the GPU (qcbc_ctx)."""

    def __init__(self, an, D=None, SN=None):
        self.an = an
        self.desc = make_desc(an)
        h = _vp()
        _check(lib().qcbc_ctx_create(ctypes.byref(self.desc), ctypes.byref(h)), "qcbc_ctx_create")
        self.h = h
        if D is not None:
            self.bind(D, SN if SN is not None else an.SN())

    def bind_time_series(self, x, SN, t_rise=0.4):
        """Tukey window, FFT and scaling on the GPU, then bind: returns the conditioned D (n_det x N)."""
        x = np.ascontiguousarray(x, np.float64)
        SN = np.ascontiguousarray(SN, np.float64)
        D = np.zeros((self.an.nifo, self.an.N))
        _check(lib().qcbc_bind_time_series(self.h, _d(x), _d(SN), t_rise, _d(D)), "qcbc_bind_time_series")
        return D

    def condition(self, x, SN, t_rise=0.4, fmin_floor=0.0, outputs=True):
        """main's conditioning (and CBC_start's with fmin_floor) on the GPU, then bind: a dict of D and,
        with outputs, DW (n_det x N), pspec and wcheck (n_det x N/2)."""
        x = np.ascontiguousarray(x, np.float64)
        SN = np.ascontiguousarray(SN, np.float64)
        nd, N = self.an.nifo, self.an.N
        out = {"D": np.zeros((nd, N))}
        if outputs:
            out.update(DW=np.zeros((nd, N)), pspec=np.zeros((nd, N // 2)), wcheck=np.zeros((nd, N // 2)))
        g = [_d(out[k]) if k in out else None for k in ("D", "DW", "pspec", "wcheck")]
        _check(lib().qcbc_condition(self.h, _d(x), _d(SN), t_rise, fmin_floor, *g), "qcbc_condition")
        return out

    def makespec(self, ffit, Sspline, lines=None):
        """makespec on the GPU: SN (N/2); lines = (linef, lineh, lineQ, linew, deltafmax) arrays or None."""
        x = np.ascontiguousarray(ffit, np.float64)
        y = np.ascontiguousarray(Sspline, np.float64)
        L = [np.ascontiguousarray(v, np.float64) for v in lines] if lines is not None else [np.zeros(1)] * 5
        nl = L[0].size if lines is not None else 0
        SN = np.zeros(self.an.N // 2)
        _check(lib().qcbc_makespec(self.h, x.size, nl, _d(x), _d(y), *[_d(v) for v in L], _d(SN)), "qcbc_makespec")
        return SN

    def loglike_intrinsic(self, params):
        """log_likelihood_intrinsic for a batch of templates' parameters: (logL, snrsq B x n_det, status)."""
        ips = 7 + 3 * (self.an.nifo - 1)
        p = np.ascontiguousarray(params, np.float64).reshape(-1, ips)
        B = p.shape[0]
        logL = np.zeros(B)
        snr = np.zeros((B, self.an.nifo))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_loglike_intrinsic(self.h, _d(p), B, _d(logL), _d(snr), _i(st)), "qcbc_loglike_intrinsic")
        return logL, snr, st

    def templates_intrinsic(self, params):
        """templates() for a batch: (hwave B x n_det x N, status)."""
        ips = 7 + 3 * (self.an.nifo - 1)
        p = np.ascontiguousarray(params, np.float64).reshape(-1, ips)
        B = p.shape[0]
        hw = np.zeros((B, self.an.nifo, self.an.N))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_templates_intrinsic(self.h, _d(p), B, _d(hw), _i(st)), "qcbc_templates_intrinsic")
        return hw, st

    def fisher_evec(self, fish):
        """FisherEvec of a batch of d x d matrices on this analysis's GPU: (ej B x d, ev B x d x d)."""
        F = np.ascontiguousarray(fish, np.float64)
        d = F.shape[-1]
        F = F.reshape(-1, d, d)
        B = F.shape[0]
        ej = np.zeros((B, d))
        ev = np.zeros((B, d, d))
        _check(lib().qcbc_fisher_evec(self.h, _d(F), B, d, _d(ej), _d(ev)), "qcbc_fisher_evec")
        return ej, ev

    # ---- search likelihoods (QuickCBC.c:5672-6408)
    def apmax(self, params):
        """log_likelihood_APmax for a batch of reference-frame points: (logL, moved params, status)."""
        ips = 7 + 3 * (self.an.nifo - 1)
        p = np.array(params, np.float64).reshape(-1, ips)
        B = p.shape[0]
        L = np.zeros(B)
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_loglike_apmax(self.h, _d(p), B, _d(L), _i(st)), "qcbc_loglike_apmax")
        return L, p, st

    def llmax(self, params, tmin, tmax, net_tmax, delays):
        """log_likelihood_max for a batch of reference-frame points: (logL, moved params, status)."""
        ips = 7 + 3 * (self.an.nifo - 1)
        p = np.array(params, np.float64).reshape(-1, ips)
        B = p.shape[0]
        L = np.zeros(B)
        st = np.zeros(B, np.int32)
        dl = np.ascontiguousarray(delays, np.float64)
        _check(lib().qcbc_loglike_max(self.h, _d(p), B, tmin, tmax, net_tmax, _d(dl), _d(L), _i(st)),
               "qcbc_loglike_max")
        return L, p, st

    # ---- the heterodyne's grid choice (QuickCBC.c:4275-5053, 2695-2817)
    def limits(self, params):
        """limits: ((MN, MM), ratio N/2)."""
        p = np.ascontiguousarray(params, np.float64)
        band = np.zeros(2, np.int32)
        ratio = np.ones(self.an.N // 2)
        _check(lib().qcbc_limits(self.h, _d(p), _i(band), _d(ratio)), "qcbc_limits")
        return (int(band[0]), int(band[1])), ratio

    def fisher_all(self, params):
        """Fisher_All for a batch: (fish B x 11 x 11, status)."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        F = np.zeros((B, 11, 11))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_fisher_all(self.h, _d(p), B, _d(F), _i(st)), "qcbc_fisher_all")
        return F, st

    def fisher_full(self, params, d=7):
        """Fisher_Full for a batch of reference-frame points: (fish B x d x d, status)."""
        ips = 7 + 3 * (self.an.nifo - 1)
        p = np.ascontiguousarray(params, np.float64).reshape(-1, ips)
        B = p.shape[0]
        F = np.zeros((B, d, d))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_fisher_full(self.h, _d(p), B, d, _d(F), _i(st)), "qcbc_fisher_full")
        return F, st

    def diff(self, params, px, of=None):
        """diff between start points params[of[k]] and px[k]."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        x = np.ascontiguousarray(px, np.float64).reshape(-1, 11)
        o = np.ascontiguousarray(np.zeros(x.shape[0]) if of is None else of, np.int32)
        z = np.zeros(x.shape[0])
        _check(lib().qcbc_diff(self.h, _d(p), p.shape[0], _d(x), _i(o), x.shape[0], _d(z)), "qcbc_diff")
        return z

    def efix(self, params, ej, ev, pmin, pmax, zs, etamin=0.05):
        """efix on the full data for a batch: (scaled ej B x 11, trials B x 11)."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        e = np.array(ej, np.float64).reshape(B, 11)
        v = np.ascontiguousarray(ev, np.float64).reshape(B, 11, 11)
        lo = np.ascontiguousarray(pmin, np.float64)
        hi = np.ascontiguousarray(pmax, np.float64)
        tr = np.zeros((B, 11), np.int32)
        _check(lib().qcbc_efix(self.h, _d(p), B, _d(e), _d(v), _d(lo), _d(hi), etamin, zs, _i(tr)), "qcbc_efix")
        return e, tr

    def het_space(self, params, pmin, pmax, etamin=0.05):
        """het_space: dict with MN, MM, M, NR, J, fgrid, eval, evec and the displaced points dparams."""
        p = np.ascontiguousarray(params, np.float64)
        lo = np.ascontiguousarray(pmin, np.float64)
        hi = np.ascontiguousarray(pmax, np.float64)
        cap = self.an.N // 2
        g = np.zeros(cap, np.int32)
        ev = np.zeros(11)
        evec = np.zeros((11, 11))
        dpar = np.zeros((11, 11))
        MN, MM, M, NR = (ctypes.c_int() for _ in range(4))
        _check(lib().qcbc_het_space(self.h, _d(p), _d(lo), _d(hi), etamin, ctypes.byref(MN), ctypes.byref(MM),
                                    ctypes.byref(M), ctypes.byref(NR), _i(g), cap, _d(ev), _d(evec), _d(dpar)),
               "qcbc_het_space")
        return dict(MN=MN.value, MM=MM.value, M=M.value, NR=NR.value, J=2, fgrid=g[:M.value].copy(), eval=ev,
                    evec=evec, dparams=dpar)

    def close(self):
        self._close_deps()
        if self.h:
            lib().qcbc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def bind(self, D, SN):
        D = np.ascontiguousarray(D, np.float64)
        SN = np.ascontiguousarray(SN, np.float64)
        assert D.shape == (self.an.nifo, self.an.N) and SN.shape == (self.an.nifo, self.an.N // 2)
        self._D, self._SN = D, SN
        _check(lib().qcbc_bind_data(self.h, _d(D), _d(SN)), "qcbc_bind_data")

    def loglike(self, params, precision=FP64):
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        logL = np.zeros(B)
        rho = np.zeros((B, self.an.nifo))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_loglike(self.h, _d(p), B, precision, _d(logL), _d(rho), _i(st)), "qcbc_loglike")
        return logL, rho, st

    def loglike_plain(self, params):
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        logL = np.zeros(B)
        rho = np.zeros((B, self.an.nifo))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_loglike_plain(self.h, _d(p), B, _d(logL), _d(rho), _i(st)), "qcbc_loglike_plain")
        return logL, rho, st

    def loglike_device(self, d_params, B, precision, d_logL, d_rho, d_status=0):
        _check(lib().qcbc_loglike_device(self.h, d_params, B, precision, d_logL, d_rho, d_status),
               "qcbc_loglike_device")

    def waveform(self, params):
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        nh = self.an.N // 2
        amp = np.zeros((B, nh))
        ph = np.zeros((B, nh))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_waveform(self.h, _d(p), B, _d(amp), _d(ph), _i(st)), "qcbc_waveform")
        return amp, ph, st

    def templates(self, params):
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        hw = np.zeros((B, self.an.nifo, self.an.N))
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_templates(self.h, _d(p), B, _d(hw), _i(st)), "qcbc_templates")
        return hw, st

    def set_shape(self, precision, cpb, tile, splits):
        _check(lib().qcbc_ctx_set_shape(self.h, precision, cpb, tile, splits), "qcbc_ctx_set_shape")

    def get_shape(self, precision):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().qcbc_ctx_get_shape(self.h, precision, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
               "qcbc_ctx_get_shape")
        return a.value, b.value, c.value

    def autotune(self, precision, B):
        _check(lib().qcbc_autotune(self.h, precision, B), "qcbc_autotune")
        return self.get_shape(precision)

    def timing(self, enable):
        ms = ctypes.c_double(0)
        n = ctypes.c_longlong(0)
        _check(lib().qcbc_ctx_timing(self.h, enable, ctypes.byref(ms), ctypes.byref(n)), "qcbc_ctx_timing")
        return ms.value, n.value

    def time_kernel(self, d_params, B, precision, reps):
        """Average duration (ms) of the fused likelihood kernel over reps
        back-to-back launches on B points (device memory)."""
        L = lib()
        L.qcbc_time_loglike_kernel.argtypes = [_vp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double)]
        ms = ctypes.c_double(0)
        _check(L.qcbc_time_loglike_kernel(self.h, d_params, B, precision, reps, ctypes.byref(ms)),
               "qcbc_time_loglike_kernel")
        return ms.value

    def sync(self):
        _check(lib().qcbc_ctx_sync(self.h), "qcbc_ctx_sync")


# ------------------------------------------------------------ drop-ins
class RowMatrix:
    """double** rows as QuickCBC's double_matrix builds them (Utilities.c:2981-2992)."""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, np.float64)
        n = self.arr.shape[0]
        self.ptrs = (_dp * n)(*[self.arr[i].ctypes.data_as(_dp) for i in range(n)])

    @property
    def p(self):
        return ctypes.cast(self.ptrs, ctypes.POINTER(_dp))


def make_net(an):
    labels = (ctypes.c_int * 3)(*an.labels, *([0] * (3 - an.nifo)))
    net = Net()
    net.Nifo = an.nifo
    net.Tobs = an.Tobs
    net.GMST = an.gmst
    net.labels = ctypes.cast(labels, _ip)
    net._keep = labels
    return net


def make_freq(an):
    f = np.ascontiguousarray(an.freq())
    rv = RealVector()
    rv.data = _d(f)
    rv.length = f.size
    rv._keep = f
    return rv


def dropin_log_likelihood_full(an, Drows, SNrows, params):
    net = make_net(an)
    fr = make_freq(an)
    rho = np.zeros(an.nifo)
    p = np.ascontiguousarray(params, np.float64)
    v = lib().log_likelihood_full(ctypes.byref(net), Drows.p, _d(p), ctypes.byref(fr), SNrows.p, _d(rho),
                                  an.N, an.Tobs)
    return v, rho


def dropin_fulltemplates(an, params):
    net = make_net(an)
    fr = make_freq(an)
    out = RowMatrix(np.zeros((an.nifo, an.N)))
    p = np.ascontiguousarray(params, np.float64)
    lib().fulltemplates(ctypes.byref(net), out.p, ctypes.byref(fr), _d(p), an.N)
    return out.arr


def dropin_h22(freq, phi0, fref, m1_SI, m2_SI, chi1, chi2, distance):
    f = np.ascontiguousarray(freq, np.float64)
    rv = RealVector()
    rv.data = _d(f)
    rv.length = f.size
    h = ctypes.POINTER(AmpPhaseFDWaveform)()
    st = lib().IMRPhenomDGenerateh22FDAmpPhase(ctypes.byref(h), ctypes.byref(rv), phi0, fref, m1_SI, m2_SI,
                                                chi1, chi2, distance)
    if st != 0:
        return st, None, None
    n = h.contents.length
    amp = np.ctypeslib.as_array(h.contents.amp, shape=(n,)).copy()
    ph = np.ctypeslib.as_array(h.contents.phase, shape=(n,)).copy()
    lib().DestroyAmpPhaseFDWaveform(h)
    return st, amp, ph


def dropin_fourier_nwip(a, b, Sn, imin, imax, N):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    Sn = np.ascontiguousarray(Sn, np.float64)
    return lib().fourier_nwip(_d(a), _d(b), _d(Sn), imin, imax, N)


def dropin_f_ant(psi, alpha, sindelta, gmst, det):
    fp = ctypes.c_double(0)
    fc = ctypes.c_double(0)
    lib().F_ant(psi, alpha, sindelta, gmst, ctypes.byref(fp), ctypes.byref(fc), det)
    return fp.value, fc.value


def dropin_times(alpha, sindelta, gmst):
    out = np.zeros(3)
    lib().Times(alpha, sindelta, gmst, _d(out))
    return out


def dropin_fbegin(params):
    p = np.ascontiguousarray(params, np.float64)
    return lib().fbegin(_d(p))


class PT(_Handle):
    """Resident parallel-tempered ensemble (qcbc_pt)."""

    def __init__(self, ctx, n_temps, n_chains, ladder, init, sigma, seed=1, precision=FP64, groups=0, het=None):
        L = lib()
        L.qcbc_pt_create_groups.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                            ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.qcbc_pt_create_het.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                         ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_vp)]
        L.qcbc_pt_groups.argtypes = [_vp, _ip]
        L.qcbc_pt_run.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.qcbc_pt_get.argtypes = [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_pt_latencies.argtypes = [_vp, _dp, ctypes.c_int, _ip]
        L.qcbc_pt_set_step_timing.argtypes = [_vp, ctypes.c_int]
        L.qcbc_pt_destroy.argtypes = [_vp]
        self.ctx = ctx
        self.P = n_temps * n_chains
        init = np.ascontiguousarray(init, np.float64).reshape(self.P, 11)
        sigma = np.ascontiguousarray(sigma, np.float64).reshape(11)
        h = _vp()
        self.het = het
        if het is None:
            _check(L.qcbc_pt_create_groups(ctx.h, n_temps, n_chains, ladder, _d(init), _d(sigma), seed, precision,
                                           groups, ctypes.byref(h)), "qcbc_pt_create_groups")
        else:
            _check(L.qcbc_pt_create_het(ctx.h, het.h, n_temps, n_chains, ladder, _d(init), _d(sigma), seed, groups,
                                        ctypes.byref(h)), "qcbc_pt_create_het")
        self.h = h
        self._depend_on(ctx, het)
        g = ctypes.c_int()
        _check(L.qcbc_pt_groups(h, ctypes.byref(g)), "qcbc_pt_groups")
        self.groups = g.value
        self.steps_run = 0

    def run(self, nsteps, graph=False):
        _check(lib().qcbc_pt_run(self.h, nsteps, 1 if graph else 0), "qcbc_pt_run")
        self.steps_run += nsteps

    def step_timing(self, on):
        _check(lib().qcbc_pt_set_step_timing(self.h, 1 if on else 0), "qcbc_pt_set_step_timing")

    def latencies(self):
        n = ctypes.c_int(0)
        _check(lib().qcbc_pt_latencies(self.h, None, 0, ctypes.byref(n)), "qcbc_pt_latencies")
        out = np.zeros(max(n.value, 1))
        _check(lib().qcbc_pt_latencies(self.h, _d(out), n.value, ctypes.byref(n)), "qcbc_pt_latencies")
        return out[:n.value]

    def get(self):
        x = np.zeros((self.P, 11))
        L = np.zeros(self.P)
        acc = (ctypes.c_longlong * self.P)()
        sw = ctypes.c_longlong(0)
        _check(lib().qcbc_pt_get(self.h, _d(x), _d(L), acc, ctypes.byref(sw)), "qcbc_pt_get")
        return x, L, np.array(acc[:]), sw.value

    def close(self):
        if self.h:
            lib().qcbc_pt_destroy(self.h)
            self.h = None
        self._release()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class UpdateDesc(ctypes.Structure):
    """qcbc_update_desc: prior box and proposal mix of QuickCBC's own PT step."""
    _fields_ = [("min", ctypes.c_double * 11), ("max", ctypes.c_double * 11), ("a", ctypes.c_double),
                ("b", ctypes.c_double), ("c", ctypes.c_double), ("etamin", ctypes.c_double),
                ("n_hist", ctypes.c_int), ("hist_every", ctypes.c_int), ("efix_zs", ctypes.c_double)]


def update_desc(pmin, pmax, a=0.5, b=0.3, c=0.2, etamin=0.05, n_hist=1000, hist_every=20, efix_zs=0.0):
    """efix_zs = 1.0 rescales the Fisher jumps as MCMC_all does (QuickCBC.c:1927)."""
    d = UpdateDesc()
    d.min[:] = [float(v) for v in pmin]
    d.max[:] = [float(v) for v in pmax]
    d.a, d.b, d.c, d.etamin, d.n_hist, d.hist_every = a, b, c, etamin, n_hist, hist_every
    d.efix_zs = efix_zs
    return d


class PTUpdate(PT):
    """A resident ensemble stepped by QuickCBC's own update() and MCMC_all
    step on the heterodyned likelihood (qcbc_pt_create_update)."""

    def __init__(self, ctx, het, n_temps, n_chains, ladder, init, desc, seed=1, groups=0):
        L = lib()
        L.qcbc_pt_create_update.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                                            ctypes.POINTER(UpdateDesc), ctypes.c_uint64, ctypes.c_int,
                                            ctypes.POINTER(_vp)]
        L.qcbc_pt_refresh_fisher.argtypes = [_vp]
        L.qcbc_pt_get_update.argtypes = [_vp, _dp, _dp, _dp, _ip, _dp, ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_pt_get_history.argtypes = [_vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_pt_groups.argtypes = [_vp, _ip]
        L.qcbc_pt_run.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.qcbc_pt_get.argtypes = [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_pt_latencies.argtypes = [_vp, _dp, ctypes.c_int, _ip]
        L.qcbc_pt_set_step_timing.argtypes = [_vp, ctypes.c_int]
        L.qcbc_pt_destroy.argtypes = [_vp]
        self.ctx = ctx
        self.het = het
        self.n_temps, self.n_chains = n_temps, n_chains
        self.P = n_temps * n_chains
        self.desc = desc
        init = np.ascontiguousarray(init, np.float64).reshape(self.P, 11)
        h = _vp()
        _check(L.qcbc_pt_create_update(ctx.h, het.h, n_temps, n_chains, ladder, _d(init), ctypes.byref(desc), seed,
                                       groups, ctypes.byref(h)), "qcbc_pt_create_update")
        self.h = h
        self._depend_on(ctx, het)
        self.steps_run = 0
        g = ctypes.c_int()
        _check(L.qcbc_pt_groups(h, ctypes.byref(g)), "qcbc_pt_groups")
        self.groups = g.value

    def refresh_fisher(self):
        _check(lib().qcbc_pt_refresh_fisher(self.h), "qcbc_pt_refresh_fisher")

    def get_update(self):
        """(ej P x 11, ev P x 11 x 11, proposals P x 11, types P, sky-ring log densities P,
        stats n_chains x 11), point order."""
        ej = np.zeros((self.P, 11))
        ev = np.zeros((self.P, 11, 11))
        y = np.zeros((self.P, 11))
        ty = np.zeros(self.P, np.int32)
        jc = np.zeros(self.P)
        st = (ctypes.c_longlong * (11 * self.n_chains))()
        _check(lib().qcbc_pt_get_update(self.h, _d(ej), _d(ev), _d(y), _i(ty), _d(jc), st), "qcbc_pt_get_update")
        return ej, ev, y, ty, jc, np.array(st[:], np.int64).reshape(self.n_chains, 11)

    def get_history(self):
        hist = np.zeros((self.n_chains, self.n_temps, self.desc.n_hist, 11))
        cnt = (ctypes.c_longlong * self.n_chains)()
        _check(lib().qcbc_pt_get_history(self.h, _d(hist), cnt), "qcbc_pt_get_history")
        return hist, np.array(cnt[:], np.int64)


# ------------------------------------------------------- device buffers
class DeviceBuffer:
    """Device memory from qcbc_dev_alloc (for the device-pointer entry points)."""

    def __init__(self, nbytes):
        L = lib()
        L.qcbc_dev_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.qcbc_dev_free.argtypes = [_vp]
        L.qcbc_memcpy_h2d.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.qcbc_memcpy_d2h.argtypes = [_vp, _vp, ctypes.c_size_t]
        self.nbytes = int(nbytes)
        p = _vp()
        _check(L.qcbc_dev_alloc(max(self.nbytes, 8), ctypes.byref(p)), "qcbc_dev_alloc")
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().qcbc_memcpy_h2d(self.ptr, a.ctypes.data_as(_vp), a.nbytes), "qcbc_memcpy_h2d")

    def download(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype)
        _check(lib().qcbc_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, out.nbytes), "qcbc_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().qcbc_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001
            pass


def _ptr_array(ptrs):
    arr = (_vp * len(ptrs))(*[p.ptr if isinstance(p, DeviceBuffer) else p for p in ptrs])
    return arr


class Multi(_Handle):
    """Persistent multi-analysis plan (qcbc_multi)."""

    def __init__(self, ctxs, B, precision=FP64):
        L = lib()
        L.qcbc_multi_create.argtypes = [ctypes.POINTER(_vp), _ip, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.qcbc_multi_set_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.qcbc_multi_get_shape.argtypes = [_vp, _ip, _ip, _ip, _ip]
        L.qcbc_multi_run.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                     ctypes.POINTER(_vp), ctypes.c_int]
        L.qcbc_multi_autotune.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                          ctypes.c_int, _ip, _ip, _ip, _dp]
        L.qcbc_multi_last_ms.argtypes = [_vp, _dp]
        L.qcbc_multi_destroy.argtypes = [_vp]
        self.ctxs = list(ctxs)
        self.B = [int(b) for b in B]
        hs = (_vp * len(ctxs))(*[c.h for c in ctxs])
        bb = (ctypes.c_int * len(B))(*self.B)
        h = _vp()
        _check(L.qcbc_multi_create(hs, bb, len(ctxs), precision, ctypes.byref(h)), "qcbc_multi_create")
        self.h = h
        self._depend_on(*self.ctxs)
        self.params = [DeviceBuffer(8 * 11 * b) for b in self.B]
        self.logL = [DeviceBuffer(8 * b) for b in self.B]
        self.rho = [DeviceBuffer(8 * 3 * b) for b in self.B]
        self.status = [DeviceBuffer(4 * b) for b in self.B]

    def set_params(self, params_list):
        for buf, p in zip(self.params, params_list):
            buf.upload(np.ascontiguousarray(p, np.float64))

    def set_shape(self, cpb, tile, split_bins):
        _check(lib().qcbc_multi_set_shape(self.h, cpb, tile, split_bins), "qcbc_multi_set_shape")

    def get_shape(self):
        a, b, c, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().qcbc_multi_get_shape(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                          ctypes.byref(n)), "qcbc_multi_get_shape")
        return a.value, b.value, c.value, n.value

    def run(self, graph=False):
        _check(lib().qcbc_multi_run(self.h, _ptr_array(self.params), _ptr_array(self.logL), _ptr_array(self.rho),
                                    _ptr_array(self.status), 1 if graph else 0), "qcbc_multi_run")
        ms = ctypes.c_double()
        lib().qcbc_multi_last_ms(self.h, ctypes.byref(ms))
        return ms.value

    def autotune(self, reps=3):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        _check(lib().qcbc_multi_autotune(self.h, _ptr_array(self.params), _ptr_array(self.logL),
                                         _ptr_array(self.rho), reps, ctypes.byref(a), ctypes.byref(b),
                                         ctypes.byref(c), ctypes.byref(ms)), "qcbc_multi_autotune")
        return (a.value, b.value, c.value), ms.value

    def results(self):
        out = []
        for c, b, l, r, s in zip(self.ctxs, self.B, self.logL, self.rho, self.status):
            out.append((l.download((b,)), r.download((b, c.an.nifo)), s.download((b,), np.int32)))
        return out

    def close(self):
        if self.h:
            lib().qcbc_multi_destroy(self.h)
            self.h = None
        self._release()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class Server:
    """Continuous-batching likelihood server (qcbc_server)."""

    def __init__(self, budget_bytes, precision=FP64):
        L = lib()
        L.qcbc_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_vp)]
        L.qcbc_server_admit.argtypes = [_vp, ctypes.POINTER(Desc), _dp, _dp, _ip]
        L.qcbc_server_evict.argtypes = [_vp, ctypes.c_int]
        L.qcbc_server_submit.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.qcbc_server_wait.argtypes = [_vp, ctypes.c_int64, _dp, _dp, _ip, _dp, ctypes.c_int]
        L.qcbc_server_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_size_t), _ip,
                                        ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
        L.qcbc_server_destroy.argtypes = [_vp]
        L.qcbc_server_admit_pt.argtypes = [_vp, ctypes.POINTER(Desc), _dp, _dp, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _dp, _dp, ctypes.c_uint64, _ip]
        L.qcbc_server_submit_steps.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.qcbc_server_get_pt.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_longlong),
                                         ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
        h = _vp()
        _check(L.qcbc_server_create(budget_bytes, precision, ctypes.byref(h)), "qcbc_server_create")
        self.h = h
        self.nd = {}

    def admit(self, an, D, SN):
        d = make_desc(an)
        D = np.ascontiguousarray(D, np.float64)
        SN = np.ascontiguousarray(SN, np.float64)
        aid = ctypes.c_int(0)
        code = lib().qcbc_server_admit(self.h, ctypes.byref(d), _d(D), _d(SN), ctypes.byref(aid))
        if code != 0:
            return code, None
        self.nd[aid.value] = an.nifo
        return 0, aid.value

    def admit_pt(self, an, D, SN, n_temps, n_chains, ladder, init, sigma, seed):
        """Admit an analysis with a resident PT ensemble: (code, analysis id or None)."""
        d = make_desc(an)
        D = np.ascontiguousarray(D, np.float64)
        SN = np.ascontiguousarray(SN, np.float64)
        x0 = np.ascontiguousarray(init, np.float64).reshape(n_temps * n_chains, 11)
        sg = np.ascontiguousarray(sigma, np.float64)
        aid = ctypes.c_int(0)
        code = lib().qcbc_server_admit_pt(self.h, ctypes.byref(d), _d(D), _d(SN), n_temps, n_chains, ladder, _d(x0),
                                          _d(sg), seed, ctypes.byref(aid))
        if code != 0:
            return code, None
        self.nd[aid.value] = an.nifo
        self.P = getattr(self, "P", {})
        self.P[aid.value] = (n_temps * n_chains, n_chains)
        return 0, aid.value

    def submit_steps(self, aid, nsteps=1):
        rid = ctypes.c_int64(0)
        _check(lib().qcbc_server_submit_steps(self.h, aid, nsteps, ctypes.byref(rid)), "qcbc_server_submit_steps")
        return rid.value, 0, aid

    def get_pt(self, aid):
        """An ensemble's state: dict of params, logL, accepted, swaps, steps."""
        P, nc = self.P[aid]
        x = np.zeros((P, 11))
        L = np.zeros(P)
        acc = (ctypes.c_longlong * P)()
        sw = (ctypes.c_longlong * nc)()
        n = ctypes.c_longlong(0)
        _check(lib().qcbc_server_get_pt(self.h, aid, _d(x), _d(L), acc, sw, ctypes.byref(n)), "qcbc_server_get_pt")
        return {"params": x, "logL": L, "accepted": np.array(acc[:]), "swaps": np.array(sw[:]), "steps": n.value}

    def evict(self, aid):
        return lib().qcbc_server_evict(self.h, aid)

    def submit(self, aid, params):
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        rid = ctypes.c_int64(0)
        _check(lib().qcbc_server_submit(self.h, aid, _d(p), p.shape[0], ctypes.byref(rid)), "qcbc_server_submit")
        return rid.value, p.shape[0], aid

    def wait(self, ticket, timeout_ms=-1):
        rid, B, aid = ticket
        nd = self.nd.get(aid, 3)
        logL = np.zeros(max(B, 1))
        rho = np.zeros((max(B, 1), nd))
        st = np.zeros(max(B, 1), np.int32)
        lat = ctypes.c_double(0)
        code = lib().qcbc_server_wait(self.h, rid, _d(logL), _d(rho), _i(st), ctypes.byref(lat), timeout_ms)
        return code, logL, rho, st, lat.value

    def stats(self):
        b = ctypes.c_size_t(0)
        n = ctypes.c_int(0)
        e = ctypes.c_longlong(0)
        k = ctypes.c_longlong(0)
        _check(lib().qcbc_server_stats(self.h, ctypes.byref(b), ctypes.byref(n), ctypes.byref(e), ctypes.byref(k)),
               "qcbc_server_stats")
        return {"bytes_in_use": b.value, "analyses": n.value, "evals": e.value, "batches": k.value}

    def close(self):
        if self.h:
            lib().qcbc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class PointRecord(ctypes.Structure):
    """Mirror of qcbc::Point (quickcbc_amd/csrc/qcbc_types.h)."""
    _fields_ = ([(n, ctypes.c_int) for n in ("status", "kstart", "kend", "region_flags",
                                             "seg_a1", "seg_a2", "seg_p1", "seg_p2")] +
                [(n, ctypes.c_double) for n in ("Msec", "s6_scale", "is6_scale", "v_scale")] +
                [(n, ctypes.c_double * 3) for n in ("tshift", "wc", "ws")] +
                [(n, ctypes.c_double) for n in (
                    "ins_q0 ins_q1 ins_q13 ins_q13L ins_qL ins_q23 ins_q43 ins_q53 ins_q2 ins_qm13 ins_qm23 "
                    "ins_qm1 ins_qm53 ia0 ia23 ia1 ia43 ia53 ia2 ia73 ia83 ia3 int_q0 int_q1 int_qL int_qm3 "
                    "col_f1 col_f2 col_f3 ib0 ib1 ib2 ib3 ib4 nf1 im2 im3 fRD mrd_rate mnum mg2 "
                    "mrd_q0 mrd_q1 mrd_qm1 mrd_q34 mrd_qat mrd_a5 mrd_ifDM").split()] +
                [(n, ctypes.c_double * 3) for n in ("dt", "dl")] +
                [(n, ctypes.c_double) for n in (
                    "fs fCut amp0 logMsec inv_eta amp_pre ai_23 ai_1 ai_43 ai_53 ai_2 ai_73 ai_83 ai_3 fmaxCalc "
                    "col_c0 col_c1 col_c2 col_c3 col_c4 fDM fDMg3 mrd_num pi_init pi_23 pi_13 pi_13L pi_L "
                    "pi_m13 pi_m23 pi_m1 pi_m53 pi_1 pi_43 pi_53 pi_2 fPhiMRD beta1 beta2 beta3 C1Int C2Int "
                    "alpha1 alpha2 alpha3 alpha4 a5fRD C1MRD C2MRD t0 MfRef phi_precalc").split()] +
                [(n, ctypes.c_double * 3) for n in ("Fs", "lambda_", "Ascale")] +
                [("eta", ctypes.c_double), ("finspin", ctypes.c_double)])


def point_records(ctx, params):
    L = lib()
    L.qcbc_point_record_size.restype = ctypes.c_size_t
    L.qcbc_point_records.argtypes = [_vp, _dp, ctypes.c_int, _vp]
    assert L.qcbc_point_record_size() == ctypes.sizeof(PointRecord), "PointRecord layout out of date"
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
    recs = (PointRecord * p.shape[0])()
    _check(L.qcbc_point_records(ctx.h, _d(p), p.shape[0], recs), "qcbc_point_records")
    return list(recs)


# ------------------------------------------------- heterodyned likelihood
HET_KEYS = ("amp", "phase", "IP", "SL", "lc", "ls", "rc", "rs", "aa")


class Het(_Handle):
    """A heterodyne on the GPU (qcbc_het): made by heterodyne() on a bound
    analysis (``Het(ctx, params=..., fgrid=..., J=..., MN=..., MM=...)``) or
    imported from set-up arrays (``Het(ctx, arrays=dict)``, the layout of
    the oracle's heterodyne)."""

    def __init__(self, ctx, params=None, fgrid=None, J=2, MN=None, MM=None, arrays=None):
        self.ctx = ctx
        nd = ctx.an.nifo
        h = _vp()
        if arrays is None:
            p = np.ascontiguousarray(params, np.float64)
            g = np.ascontiguousarray(fgrid, np.int32)
            _check(lib().qcbc_het_create(ctx.h, _d(p), _i(g), g.size, J, MN, MM, ctypes.byref(h)),
                   "qcbc_het_create")
        else:
            a = [np.ascontiguousarray(arrays[k], np.float64) for k in ("freq",) + HET_KEYS]
            _check(lib().qcbc_het_import(ctx.h, int(arrays["M"]), int(arrays["J"]), *[_d(x) for x in a],
                                         float(arrays["logLb"]), ctypes.byref(h)), "qcbc_het_import")
        self.h = h
        self._depend_on(ctx)
        M, Jo, NR, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lb = ctypes.c_double()
        _check(lib().qcbc_het_info(h, ctypes.byref(M), ctypes.byref(Jo), ctypes.byref(NR), ctypes.byref(n),
                                   ctypes.byref(lb)), "qcbc_het_info")
        self.M, self.J, self.NR, self.nd, self.logLb = M.value, Jo.value, NR.value, nd, lb.value

    def export(self):
        M, J1, NR, nd = self.M, self.J + 1, self.NR, self.nd
        out = dict(M=M, J=self.J, NR=NR, logLb=self.logLb, freq=np.zeros(M), amp=np.zeros((nd, M)),
                   phase=np.zeros((nd, M)), IP=np.zeros((NR, J1, J1)), SL=np.zeros((NR, nd, J1)),
                   lc=np.zeros((NR, nd, J1)), ls=np.zeros((NR, nd, J1)), rc=np.zeros((nd, NR)),
                   rs=np.zeros((nd, NR)), aa=np.zeros((nd, NR)))
        _check(lib().qcbc_het_export(self.h, _d(out["freq"]), *[_d(out[k]) for k in HET_KEYS]), "qcbc_het_export")
        return out

    def weights(self):
        w = np.zeros((3, self.nd, self.M))
        _check(lib().qcbc_het_weights(self.h, _d(w)), "qcbc_het_weights")
        return w

    def loglike(self, params):
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        logL = np.zeros(B)
        st = np.zeros(B, np.int32)
        _check(lib().qcbc_het_loglike(self.h, _d(p), B, _d(logL), _i(st)), "qcbc_het_loglike")
        return logL, st

    def loglike_device(self, d_params, B, d_logL, d_status=0):
        _check(lib().qcbc_het_loglike_device(self.h, d_params, B, d_logL, d_status), "qcbc_het_loglike_device")

    def sync(self):
        _check(lib().qcbc_het_sync(self.h), "qcbc_het_sync")

    def fisher(self, params, evec=True):
        """Fisher_Het at each point: (fish B x 11 x 11, ej B x 11, ev B x 11 x 11, status); ej and ev are
        None unless evec."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        F = np.zeros((B, 11, 11))
        st = np.zeros(B, np.int32)
        ej = np.zeros((B, 11)) if evec else None
        ev = np.zeros((B, 11, 11)) if evec else None
        _check(lib().qcbc_het_fisher(self.h, _d(p), B, _d(F), _d(ej) if evec else None, _d(ev) if evec else None,
                                     _i(st)), "qcbc_het_fisher")
        return F, ej, ev, st

    def fisher_device(self, d_params, B, d_fish, d_ej=0, d_ev=0, d_status=0):
        _check(lib().qcbc_het_fisher_device(self.h, d_params, B, d_fish, d_ej, d_ev, d_status),
               "qcbc_het_fisher_device")

    def fisher_time(self, d_params, B, reps):
        """Mean ms per launch: (set-up of the 23 B perturbed points, Fisher kernel, eigen kernel)."""
        ms = np.zeros(3)
        _check(lib().qcbc_het_fisher_time(self.h, d_params, B, reps, _d(ms)), "qcbc_het_fisher_time")
        return tuple(float(x) for x in ms)

    def timing(self, enable):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(lib().qcbc_het_timing(self.h, enable, ctypes.byref(ms), ctypes.byref(n)), "qcbc_het_timing")
        return ms.value, n.value

    def time_kernel(self, d_params, B, reps):
        ms = ctypes.c_double()
        _check(lib().qcbc_het_time_kernel(self.h, d_params, B, reps, ctypes.byref(ms)), "qcbc_het_time_kernel")
        return ms.value

    def diff(self, params, px, of=None):
        """het_diff between start points params[of[k]] and px[k] on this heterodyne."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        x = np.ascontiguousarray(px, np.float64).reshape(-1, 11)
        o = np.ascontiguousarray(np.zeros(x.shape[0]) if of is None else of, np.int32)
        z = np.zeros(x.shape[0])
        _check(lib().qcbc_het_diff(self.h, _d(p), p.shape[0], _d(x), _i(o), x.shape[0], _d(z)), "qcbc_het_diff")
        return z

    def efix(self, params, ej, ev, pmin, pmax, zs, etamin=0.05):
        """efix with the heterodyne distance for a batch: (scaled ej B x 11, trials B x 11)."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1, 11)
        B = p.shape[0]
        e = np.array(ej, np.float64).reshape(B, 11)
        v = np.ascontiguousarray(ev, np.float64).reshape(B, 11, 11)
        lo = np.ascontiguousarray(pmin, np.float64)
        hi = np.ascontiguousarray(pmax, np.float64)
        tr = np.zeros((B, 11), np.int32)
        _check(lib().qcbc_het_efix(self.h, _d(p), B, _d(e), _d(v), _d(lo), _d(hi), etamin, zs, _i(tr)),
               "qcbc_het_efix")
        return e, tr

    def close(self):
        self._close_deps()
        if self.h:
            lib().qcbc_het_destroy(self.h)
            self.h = None
        self._release()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def dropin_fullphaseamp(an, freq, params):
    net = make_net(an)
    f = np.ascontiguousarray(freq, np.float64)
    rv = RealVector()
    rv.data = _d(f)
    rv.length = f.size
    amp = RowMatrix(np.zeros((an.nifo, f.size)))
    ph = RowMatrix(np.zeros((an.nifo, f.size)))
    p = np.ascontiguousarray(params, np.float64)
    lib().fullphaseamp(ctypes.byref(net), amp.p, ph.p, ctypes.byref(rv), _d(p), 2 * f.size)
    return amp.arr, ph.arr


class HetView:
    """A struct Het as QuickCBC holds it: either filled by the drop-in
    heterodyne() (``read()`` copies its arrays out) or built here from
    set-up arrays in the reference's double_matrix / double_tensor shapes."""

    def __init__(self, M, J, NR, MN, MM, fgrid, Tobs):
        self.s = HetStruct()
        self.s.M, self.s.J, self.s.NR, self.s.MN, self.s.MM = M, J, NR, MN, MM
        self._g = np.ascontiguousarray(fgrid, np.int32)
        self._f = np.ascontiguousarray(self._g / Tobs, np.float64)
        self._rv = RealVector()
        self._rv.data = _d(self._f)
        self._rv.length = self._f.size
        self.s.fgrid = _i(self._g)
        self.s.freq = ctypes.pointer(self._rv)
        self._keep = []

    @classmethod
    def from_arrays(cls, h, Tobs, nd):
        v = cls(h["M"], h["J"], h["NR"], h.get("MN", 0), h.get("MM", 0), h["fgrid"], Tobs)
        v.s.logLb = h["logLb"]
        for k in ("amp", "phase", "rc", "rs", "aa"):
            m = RowMatrix(h[k])
            v._keep.append(m)
            setattr(v.s, k, m.p)
        for k in ("IP", "SL", "lc", "ls"):
            rows = [RowMatrix(h[k][i]) for i in range(h[k].shape[0])]
            v._keep += rows
            t = (_dpp * len(rows))(*[r.p for r in rows])
            v._keep.append(t)
            setattr(v.s, k, ctypes.cast(t, ctypes.POINTER(_dpp)))
        if "pref" in h:
            pr = np.ascontiguousarray(h["pref"], np.float64)
            v._keep.append(pr)
            v.s.pref = _d(pr)
        return v

    def read(self, nd):
        s = self.s
        M, J1, NR = s.M, s.J + 1, s.NR
        mat = lambda p, n, m: np.array([[p[i][j] for j in range(m)] for i in range(n)])
        return dict(M=M, J=s.J, NR=NR, logLb=s.logLb, freq=self._f.copy(), fgrid=self._g.copy(),
                    amp=mat(s.amp, nd, M), phase=mat(s.phase, nd, M), rc=mat(s.rc, nd, NR), rs=mat(s.rs, nd, NR),
                    aa=mat(s.aa, nd, NR), IP=np.array([mat(s.IP[i], J1, J1) for i in range(NR)]),
                    SL=np.array([mat(s.SL[i], nd, J1) for i in range(NR)]),
                    lc=np.array([mat(s.lc[i], nd, J1) for i in range(NR)]),
                    ls=np.array([mat(s.ls[i], nd, J1) for i in range(NR)]),
                    pref=np.array([s.pref[i] for i in range(11)]))


def dropin_heterodyne(an, Drows, SNrows, params, fgrid, J, MN, MM):
    g = np.ascontiguousarray(fgrid, np.int32)
    v = HetView(g.size, J, (g.size - 1) // J, MN, MM, g, an.Tobs)
    net = make_net(an)
    fr = make_freq(an)
    p = np.ascontiguousarray(params, np.float64)
    lib().heterodyne(ctypes.byref(net), ctypes.byref(v.s), Drows.p, _d(p), ctypes.byref(fr), SNrows.p, an.N,
                     an.Tobs)
    return v


def dropin_log_likelihood_het(an, view, params):
    net = make_net(an)
    p = np.ascontiguousarray(params, np.float64)
    return lib().log_likelihood_het(ctypes.byref(net), ctypes.byref(view.s), _d(p), an.Tobs)


def dropin_fisher_het(an, view, params):
    net = make_net(an)
    p = np.ascontiguousarray(params, np.float64)
    F = RowMatrix(np.zeros((11, 11)))
    lib().Fisher_Het(ctypes.byref(net), ctypes.byref(view.s), F.p, _d(p), an.Tobs)
    return F.arr


def dropin_fisher_evec(fish):
    F = RowMatrix(np.array(fish, np.float64))
    d = F.arr.shape[0]
    ej = np.zeros(d)
    ev = RowMatrix(np.zeros((d, d)))
    lib().FisherEvec(F.p, _d(ej), ev.p, d)
    return ej, ev.arr


def dropin_templates(an, params):
    net = make_net(an)
    fr = make_freq(an)
    h = RowMatrix(np.zeros((an.nifo, an.N)))
    p = np.ascontiguousarray(params, np.float64)
    lib().templates(ctypes.byref(net), h.p, ctypes.byref(fr), _d(p), an.N)
    return h.arr


def dropin_log_likelihood_intrinsic(an, Drows, SNrows, params):
    net = make_net(an)
    fr = make_freq(an)
    p = np.ascontiguousarray(params, np.float64)
    return lib().log_likelihood_intrinsic(ctypes.byref(net), Drows.p, _d(p), ctypes.byref(fr), SNrows.p, an.N,
                                          an.Tobs)


# ------------------------------------------- drop-ins of the grid choice
def dropin_limits(an, SNrows, params):
    net = make_net(an)
    fr = make_freq(an)
    p = np.ascontiguousarray(params, np.float64)
    stst = np.zeros(2, np.int32)
    ratio = np.ones(an.N // 2)
    lib().limits(ctypes.byref(net), _d(p), ctypes.byref(fr), SNrows.p, _i(stst), _d(ratio), an.N, an.Tobs)
    return (int(stst[0]), int(stst[1])), ratio


def dropin_fisher_all(an, SNrows, params):
    net = make_net(an)
    fr = make_freq(an)
    p = np.ascontiguousarray(params, np.float64)
    F = RowMatrix(np.zeros((11, 11)))
    lib().Fisher_All(ctypes.byref(net), F.p, _d(p), ctypes.byref(fr), SNrows.p, an.N, an.Tobs)
    return F.arr


def dropin_fisher_full(an, SNrows, params, d=7):
    net = make_net(an)
    fr = make_freq(an)
    p = np.ascontiguousarray(params, np.float64)
    F = RowMatrix(np.zeros((d, d)))
    lib().Fisher_Full(ctypes.byref(net), F.p, d, _d(p), ctypes.byref(fr), SNrows.p, an.N, an.Tobs)
    return F.arr


def dropin_diff(an, SNrows, params, dparams):
    net = make_net(an)
    fr = make_freq(an)
    a = np.ascontiguousarray(params, np.float64)
    b = np.ascontiguousarray(dparams, np.float64)
    return lib().diff(ctypes.byref(net), _d(a), _d(b), ctypes.byref(fr), SNrows.p, an.N, an.Tobs)


def dropin_het_diff(an, view, params, dparams):
    net = make_net(an)
    a = np.ascontiguousarray(params, np.float64)
    b = np.ascontiguousarray(dparams, np.float64)
    return lib().het_diff(ctypes.byref(net), ctypes.byref(view.s), _d(a), _d(b))


def dropin_efix(an, SNrows, view, params, ej, ev, pmin, pmax, zs, hflag):
    net = make_net(an)
    fr = make_freq(an)
    a = np.ascontiguousarray(params, np.float64)
    e = np.array(ej, np.float64)
    V = RowMatrix(np.array(ev, np.float64))
    lo = np.ascontiguousarray(pmin, np.float64)
    hi = np.ascontiguousarray(pmax, np.float64)
    lib().efix(ctypes.byref(net), ctypes.byref(view.s) if view is not None else None, _d(a), _d(lo), _d(hi),
               ctypes.byref(fr), SNrows.p if SNrows is not None else None, an.N, an.Tobs, _d(e), V.p, zs, hflag)
    return e


def dropin_het_space(an, SNrows, params, pmin, pmax):
    """het_space() into a fresh struct Het: (MN, MM, M, NR, J, fgrid, freq)."""
    net = make_net(an)
    fr = make_freq(an)
    a = np.ascontiguousarray(params, np.float64)
    lo = np.ascontiguousarray(pmin, np.float64)
    hi = np.ascontiguousarray(pmax, np.float64)
    h = HetStruct()
    lib().het_space(ctypes.byref(net), ctypes.byref(h), _d(a), _d(lo), _d(hi), ctypes.byref(fr), SNrows.p, an.N,
                    an.Tobs)
    M = h.M
    g = np.array([h.fgrid[i] for i in range(M)], np.int32)
    f = np.array([h.freq.contents.data[i] for i in range(M)])
    return dict(MN=h.MN, MM=h.MM, M=M, NR=h.NR, J=h.J, fgrid=g, freq=f)


# ------------------------------------------- drop-ins of the search likelihoods
def dropin_apmax(an, Drows, SNrows, params):
    net = make_net(an)
    fr = make_freq(an)
    p = np.array(params, np.float64)
    v = lib().log_likelihood_APmax(ctypes.byref(net), Drows.p, _d(p), ctypes.byref(fr), SNrows.p, an.N, an.Tobs)
    return v, p


def dropin_llmax(an, Drows, SNrows, params, tmin, tmax, net_tmax, delays):
    net = make_net(an)
    net.tmax = net_tmax
    dl = RowMatrix(np.array(delays, np.float64))
    net.delays = dl.p
    net._keep2 = dl
    fr = make_freq(an)
    p = np.array(params, np.float64)
    v = lib().log_likelihood_max(ctypes.byref(net), Drows.p, _d(p), ctypes.byref(fr), SNrows.p, an.N, an.Tobs, tmin,
                                 tmax, 0)
    return v, p


def dropin_penalized(params, D, twave, SN, N, Tobs):
    a = [np.ascontiguousarray(x, np.float64) for x in (params, D, twave, SN)]
    A = ctypes.c_double(0)
    P = ctypes.c_double(0)
    v = lib().log_likelihood_penalized(*[_d(x) for x in a], ctypes.byref(A), ctypes.byref(P), N, Tobs, 0)
    return v, A.value, P.value


def dropin_pdwave(an, params, tcmin):
    fr = make_freq(an)
    w = np.zeros(an.N)
    p = np.ascontiguousarray(params, np.float64)
    lib().PDwave(_d(w), ctypes.byref(fr), tcmin, _d(p), an.N)
    return w


def dropin_tukey(data, alpha):
    x = np.array(data, np.float64)
    lib().tukey(_d(x), alpha, x.size)
    return x


def dropin_tukey_scale(alpha, N):
    s1, s2 = ctypes.c_double(0), ctypes.c_double(0)
    lib().tukey_scale(ctypes.byref(s1), ctypes.byref(s2), alpha, N)
    return s1.value, s2.value


def dropin_whiten(data, Sn):
    x = np.array(data, np.float64)
    S = np.ascontiguousarray(Sn, np.float64)
    lib().whiten(_d(x), _d(S), x.size)
    return x


def data_file_name(kind, Tobs, ttrig, label):
    buf = ctypes.create_string_buffer(256)
    _check(lib().qcbc_data_file_name(buf, 256, kind.encode(), Tobs, ttrig, label), "qcbc_data_file_name")
    return buf.value.decode()


def read_framed(path):
    """A framed_* file: (time, strain) of the N pairs the reference counts."""
    n = ctypes.c_int(0)
    p = str(path).encode()
    _check(lib().qcbc_read_framed(p, 0, None, None, ctypes.byref(n)), "qcbc_read_framed")
    t, x = np.zeros(max(n.value, 1)), np.zeros(max(n.value, 1))
    _check(lib().qcbc_read_framed(p, n.value, _d(t), _d(x), ctypes.byref(n)), "qcbc_read_framed")
    return t[:n.value], x[:n.value]


def read_spec(path, n2):
    """A spec_* file: (SN, SM) of its first n2 rows."""
    SN, SM = np.zeros(n2), np.zeros(n2)
    _check(lib().qcbc_read_spec(str(path).encode(), n2, _d(SN), _d(SM)), "qcbc_read_spec")
    return SN, SM


def write_table(path, cols, digits):
    """Columns written as main writes dataw.dat and the spectra: "%.{digits}e " per value, a newline per row."""
    cs = [np.ascontiguousarray(c, np.float64) for c in cols]
    nrows = cs[0].size if cs else 0
    arr = (_dp * max(len(cs), 1))(*[c.ctypes.data_as(_dp) for c in cs])
    _check(lib().qcbc_write_table(str(path).encode(), nrows, len(cs), ctypes.cast(arr, _dpp), digits),
           "qcbc_write_table")


def dropin_pbt_shift(data1, data2, Sn, imin, imax, N):
    a = [np.ascontiguousarray(x, np.float64) for x in (data1, data2, Sn)]
    c, cf = np.zeros(N), np.zeros(N)
    lib().pbt_shift(_d(c), _d(cf), *[_d(x) for x in a], imin, imax, N)
    return c, cf
