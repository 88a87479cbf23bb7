"""Synthetic analyses for the waveform + likelihood path.

Parameter vectors follow QuickCBC's 11-entry layout (QuickCBC.c:1867):
  [0] ln(Mc/kg)  [1] ln(Mt/kg)  [2] chi1  [3] chi2  [4] GW phase (2 phi_orb)
  [5] tc (s from segment start)  [6] ln(DL/m)  [7] alpha  [8] sin(delta)
  [9] psi  [10] cos(iota)
Frequency-domain data use the GSL half-complex layout of QuickCBC
(QuickCBC.c:257-262): D[i] = Re, D[N-i] = Im of bin i, scaled so that
<|D_k|^2> = S_k/2 (QuickCBC.c:309-313).

Everything here is host-side numpy; the injected signal is produced by a
caller-supplied template function so that the same workload can be built with
the GPU path (bench) or with the CPU oracle (tests).
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

MSUN_SI = 1.988546954961461467461011951140572744e30
PC_SI = 3.085677581491367278913937957796471611e16
MPC_SI = 1.0e6 * PC_SI
GMST = 1.7318  # fixed sidereal angle (rad) of every synthetic analysis

# detector labels as in F_ant / Times (Utilities.c:995-1109)
H1, L1, V1 = 0, 1, 2


def psd(f):
    """Analytic one-sided PSD S(f) = 1e-49 [x^-4.14 - 5x^-2 + 111(1-x^2+x^4/2)/(1+x^2/2)], x=f/215."""
    x = np.asarray(f, dtype=np.float64) / 215.0
    with np.errstate(divide="ignore", invalid="ignore"):
        s = 1e-49 * (x ** -4.14 - 5.0 * x ** -2 + 111.0 * (1.0 - x * x + 0.5 * x ** 4) / (1.0 + 0.5 * x * x))
    return s


@dataclass
class Analysis:
    """One analysis: detector network, segment and band."""
    labels: List[int]
    Tobs: float
    fs: float                 # sample rate (Hz)
    fmin: float
    fmax: float
    fref: float = 100.0       # ConstCBC.h:31
    gmst: float = GMST
    psd_scale: List[float] = field(default_factory=list)

    @property
    def N(self):
        return int(round(self.Tobs * self.fs))

    @property
    def nifo(self):
        return len(self.labels)

    @property
    def imin(self):
        return int(self.Tobs * self.fmin)     # QuickCBC.c:4250

    @property
    def imax(self):
        return min(int(self.Tobs * self.fmax), self.N // 2)  # QuickCBC.c:4251-4252

    def freq(self):
        f = np.arange(self.N // 2, dtype=np.float64) / self.Tobs
        f[0] = 1.0 / self.Tobs                # QuickCBC.c:418
        return f

    def SN(self):
        """PSD rows (nifo x N/2); bin 0 carries the bin-1 value (never in band)."""
        f = self.freq()
        f[0] = f[1]
        base = psd(f)
        scales = self.psd_scale or [1.0] * self.nifo
        return np.stack([base * s for s in scales])


def config0():
    """2 detectors (H, L), Tobs=4 s, fs=2048 Hz, band 16-1024 Hz."""
    return Analysis(labels=[H1, L1], Tobs=4.0, fs=2048.0, fmin=16.0, fmax=1024.0, psd_scale=[1.0, 1.0])


def config1():
    """3 detectors (H, L, Virgo-like with 2x PSD), Tobs=16 s, fs=4096 Hz, band 20-2048 Hz."""
    return Analysis(labels=[H1, L1, V1], Tobs=16.0, fs=4096.0, fmin=20.0, fmax=2048.0,
                    psd_scale=[1.0, 1.0, 2.0])


def chirp_mass(m1, m2):
    return (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2


def make_params(m1, m2, chi1, chi2, phase, tc, dl_mpc, alpha, sindelta, psi, ciota):
    """11-vector from component masses (Msun) and distance (Mpc)."""
    mc = chirp_mass(m1, m2)
    return np.array([np.log(mc * MSUN_SI), np.log((m1 + m2) * MSUN_SI), chi1, chi2, phase, tc,
                     np.log(dl_mpc * MPC_SI), alpha, sindelta, psi, ciota], dtype=np.float64)


def injection0(seed=7):
    """m1=36, m2=29, chi=(0.3,-0.2), DL=400 Mpc, tc=2 s, random extrinsics."""
    r = np.random.default_rng(seed)
    return make_params(36.0, 29.0, 0.3, -0.2, r.uniform(0, 2 * np.pi), 2.0, 400.0,
                       r.uniform(0, 2 * np.pi), r.uniform(-1, 1), r.uniform(0, np.pi), r.uniform(-1, 1))


def injection1(seed=11, dl_mpc=40.0):
    """m1=1.5, m2=1.3, chi=(0.05,-0.05), DL=40 Mpc, tc=Tobs-2=14 s."""
    r = np.random.default_rng(seed)
    return make_params(1.5, 1.3, 0.05, -0.05, r.uniform(0, 2 * np.pi), 14.0, dl_mpc,
                       r.uniform(0, 2 * np.pi), r.uniform(-1, 1), r.uniform(0, np.pi), r.uniform(-1, 1))


def prior_draws0(n, seed=1234, tc0=2.0):
    """Prior draws of the small CPU-runnable workload: Mc~U[10,60] Msun,
    q=m2/m1~U[0.1,1], chi~U[-0.9,0.9], tc=tc0+-0.1 s, p(DL)~DL^2 on
    [100,2000] Mpc, isotropic sky, psi~U[0,pi], cos(iota)~U[-1,1]."""
    r = np.random.default_rng(seed)
    mc = r.uniform(10.0, 60.0, n)
    q = r.uniform(0.1, 1.0, n)
    # m1 from Mc and q: Mc = m1 q^{3/5} / (1+q)^{1/5}
    m1 = mc * (1.0 + q) ** 0.2 / q ** 0.6
    m2 = q * m1
    dl = (r.uniform(100.0 ** 3, 2000.0 ** 3, n)) ** (1.0 / 3.0)
    out = np.empty((n, 11))
    out[:, 0] = np.log(mc * MSUN_SI)
    out[:, 1] = np.log((m1 + m2) * MSUN_SI)
    out[:, 2] = r.uniform(-0.9, 0.9, n)
    out[:, 3] = r.uniform(-0.9, 0.9, n)
    out[:, 4] = r.uniform(0, 2 * np.pi, n)
    out[:, 5] = tc0 + r.uniform(-0.1, 0.1, n)
    out[:, 6] = np.log(dl * MPC_SI)
    out[:, 7] = r.uniform(0, 2 * np.pi, n)
    out[:, 8] = r.uniform(-1, 1, n)
    out[:, 9] = r.uniform(0, np.pi, n)
    out[:, 10] = r.uniform(-1, 1, n)
    return out


def ensemble_around(center, n, seed=99, scale=1.0):
    """Points scattered around a parameter vector (a warm PT ensemble):
    small Gaussian offsets in the intrinsic parameters, broad extrinsics."""
    r = np.random.default_rng(seed)
    out = np.tile(np.asarray(center, np.float64), (n, 1))
    sig = np.array([2e-4, 1e-3, 0.02, 0.02, 0.5, 2e-4, 0.1, 0.2, 0.2, 0.3, 0.2]) * scale
    out += r.standard_normal((n, 11)) * sig
    out[:, 2:4] = np.clip(out[:, 2:4], -0.95, 0.95)
    out[:, 8] = np.clip(out[:, 8], -0.999, 0.999)
    out[:, 10] = np.clip(out[:, 10], -0.999, 0.999)
    # keep eta <= 1/4: Mc <= 0.4352752816480621 Mt
    lim = np.log(0.25 ** 0.6) + out[:, 1]
    out[:, 0] = np.minimum(out[:, 0], lim - 1e-9)
    return out


def noise(an: Analysis, seed=2024):
    """Gaussian noise rows (nifo x N), half-complex, Re and Im ~ N(0, S/4)."""
    r = np.random.default_rng(seed)
    N = an.N
    S = an.SN()
    out = np.zeros((an.nifo, N))
    k = np.arange(1, N // 2)
    sig = np.sqrt(S[:, 1:N // 2] / 4.0)
    out[:, k] = r.standard_normal((an.nifo, k.size)) * sig
    out[:, N - k] = r.standard_normal((an.nifo, k.size)) * sig
    return out


def make_data(an: Analysis, template_fn, inj, noise_seed=2024, with_noise=True):
    """D = h(inj) + n.  ``template_fn(params) -> (nifo x N)`` half-complex."""
    h = np.asarray(template_fn(inj), dtype=np.float64).reshape(an.nifo, an.N)
    if with_noise:
        h = h + noise(an, noise_seed)
    return np.ascontiguousarray(h)


def snr_band(an: Analysis, amp, SN, fs):
    """[MN, MM] as limits() chooses them (QuickCBC.c:4275-4385): from bin
    Tobs max(fmin, fbegin) up, the first bin where the reference point's
    cumulative network SNR^2 passes 0.001 and the first where it comes within
    0.001 of its total.  amp: fullphaseamp's amplitude rows (nifo x N/2) at
    the reference point, fs its fbegin."""
    jmin = int(an.Tobs * max(an.fmin, fs))
    x = (4.0 * np.asarray(amp) ** 2 / SN).sum(0)
    x[:jmin] = 0.0
    c = np.cumsum(x)
    j0 = int(np.argmax(c > 0.001))
    j1 = int(np.argmax(c[-1] - c < 0.001))
    return max(j0, jmin), j1


def het_grid(MN, MM, Tobs, J=2, frac=1.0 / 64.0, max_hz=100.0):
    """A heterodyne grid of the shape het_space builds (QuickCBC.c:4784-4908):
    bin indices from MN in regions of J equal steps, M = NR J + 1 of them,
    the last one below MM.  het_space sizes the steps from the parameter
    space's Fisher directions; here a region starting at bin k steps by
    frac * k bins (at least 1, at most max_hz * Tobs, het_space's own cap)."""
    cap = max(1, int(max_hz * Tobs))
    g = [int(MN)]
    k = int(MN)
    while True:
        s = max(1, min(cap, int(frac * k)))
        if k + J * s > MM - 1:
            s = (MM - 1 - k) // J
            if s >= 1:
                g += [k + s * (i + 1) for i in range(J)]
            break
        g += [k + s * (i + 1) for i in range(J)]
        k += J * s
    return np.asarray(g, dtype=np.int32)


def prior_box(an: Analysis):
    """MCMC_all's prior box (QuickCBC.c:1869-1892) with ConstCBC.h's limits:
    chirp mass 0.5-160 Msun, total mass 0.5-300 Msun, |chi| <= smax = 0.85,
    DL 1e6-1e10 pc, and the peak-time window [Tobs - 3.5 (at least 1.5),
    Tobs - 0.5] of a trigger at Tobs - 2 (QuickCBC.c:182-184)."""
    smax = 0.85
    tmax = an.Tobs - 0.5
    tmin = max(an.Tobs - 3.5, 1.5)
    lo = np.array([np.log(0.5 * MSUN_SI), np.log(0.5 * MSUN_SI), -smax, -smax, 0.0, tmin, np.log(1e6 * PC_SI),
                   0.0, -1.0, 0.0, -1.0])
    hi = np.array([np.log(160.0 * MSUN_SI), np.log(300.0 * MSUN_SI), smax, smax, 2 * np.pi, tmax,
                   np.log(1e10 * PC_SI), 2 * np.pi, 1.0, np.pi, 1.0])
    return lo, hi


def ensemble_in_prior(center, n, lo, hi, seed=99, scale=1.0):
    """ensemble_around, with every point moved inside the prior box (angles
    wrapped, the rest clipped just inside)."""
    P = ensemble_around(center, n, seed=seed, scale=scale)
    P[:, 4] = np.mod(P[:, 4], 2 * np.pi)
    P[:, 7] = np.mod(P[:, 7], 2 * np.pi)
    P[:, 9] = np.mod(P[:, 9], np.pi)
    eps = 1e-6 * (hi - lo)
    return np.clip(P, lo + eps, hi - eps)
