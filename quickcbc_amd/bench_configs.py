"""Benchmarks of the batched, mixed-precision and serving forms.

Each function builds its synthetic workload on the GPU (templates from the
library, seeded Gaussian noise coloured by the analytic PSD), runs it, and
returns a dict of measurements.  bench.py calls them with --config.

  uneven(...)     12 analyses with Tobs in {4..64} s, fs in {2048, 4096} Hz,
                  1-3 detectors and 64/256/1024 points each, in one persistent
                  queue launch: fixed launch shape vs autotuned vs hipGraph
                  replay vs one launch per analysis.
  precision(...)  mixed precision against FP64 at Tobs = 16 s, 3 detectors:
                  |dlogL| over many draws and over injections at network SNR
                  10, 25 and 80, and the throughput ratio.
  serving(...)    Poisson arrivals of analyses (sizes as in uneven), each
                  admitted against a device-memory budget with a resident PT
                  ensemble of 512-4096 points and issuing one-step requests,
                  continuously batched; p50/p99 step latency and aggregate
                  logL evaluations per second.
"""
import threading
import time

import numpy as np

from . import qcbc as Q
from . import workloads as W

SIGMA = np.array([2e-5, 1e-4, 2e-3, 2e-3, 0.05, 2e-5, 0.01, 0.01, 0.01, 0.02, 0.01])


def _analysis(Tobs, fs, labels):
    scales = [2.0 if l == 2 else 1.0 for l in labels]
    fmax = min(2048.0, fs / 2)
    return W.Analysis(labels=list(labels), Tobs=float(Tobs), fs=float(fs), fmin=20.0, fmax=fmax, psd_scale=scales)


def uneven_specs(seed=2026, n=12):
    """n analyses of mixed shape: (an, injection, n_points)."""
    r = np.random.default_rng(seed)
    det_sets = [[0], [1], [0, 1], [1, 2], [0, 1, 2]]
    out = []
    for i in range(n):
        Tobs = float(r.choice([4, 8, 16, 32, 64]))
        fs = float(r.choice([2048, 4096]))
        labels = det_sets[r.integers(len(det_sets))]
        an = _analysis(Tobs, fs, labels)
        # masses from 1.2+1.2 to 60+50 Msun; heavier systems need shorter segments to merge inside
        u = r.uniform()
        m1 = 1.2 * (60.0 / 1.2) ** u
        m2 = m1 * r.uniform(0.5, 1.0)
        m2 = max(m2, 1.2)
        tc = Tobs - 2.0
        dl = 40.0 * (m1 / 1.2) ** 0.8
        inj = W.make_params(m1, m2, r.uniform(-0.3, 0.3), r.uniform(-0.3, 0.3), r.uniform(0, 2 * np.pi), tc, dl,
                            r.uniform(0, 2 * np.pi), r.uniform(-1, 1), r.uniform(0, np.pi), r.uniform(-1, 1))
        npts = int(r.choice([64, 256, 1024]))
        out.append((an, inj, npts))
    return out


def build_context(an, inj, noise_seed):
    ctx = Q.Context(an)
    h, st = ctx.templates(inj[None, :])
    D = np.ascontiguousarray(h[0] + W.noise(an, noise_seed))
    ctx.bind(D, an.SN())
    return ctx, D


def uneven(reps=20, precision=Q.FP64, log=print):
    specs = uneven_specs()
    ctxs, Ps, bins = [], [], []
    for i, (an, inj, npts) in enumerate(specs):
        ctx, D = build_context(an, inj, 500 + i)
        ctxs.append(ctx)
        Ps.append(W.ensemble_around(inj, npts, seed=600 + i, scale=3.0))
        bins.append(an.imax - an.imin)
    total = sum(p.shape[0] for p in Ps)
    res = {"analyses": len(specs), "points": total,
           "shapes": [{"Tobs": a.Tobs, "fs": a.fs, "det": a.nifo, "points": n, "bins": b}
                      for (a, _, n), b in zip(specs, bins)]}
    m = Q.Multi(ctxs, [p.shape[0] for p in Ps], precision)
    m.set_params(Ps)

    def timeit(graph=False):
        m.run(graph=graph)
        ts = [m.run(graph=graph) for _ in range(reps)]
        return float(np.median(ts))

    m.set_shape(4, 256, 4096)
    t_fixed = timeit()
    res["fixed_shape"] = {"shape": m.get_shape()[:3], "ms": t_fixed, "logl_per_s": total / t_fixed * 1e3}
    best, _ = m.autotune(reps=3)
    t_tuned = timeit()
    res["autotuned"] = {"shape": best, "ms": t_tuned, "logl_per_s": total / t_tuned * 1e3}
    t_graph = timeit(graph=True)
    res["autotuned_graph"] = {"ms": t_graph, "logl_per_s": total / t_graph * 1e3}
    ref = [r[0].copy() for r in m.results()]
    # one launch per analysis, same points, through the single-analysis API
    for c, p in zip(ctxs, Ps):
        c.loglike(p, precision)
    t0 = time.perf_counter()
    for _ in range(reps):
        outs = [c.loglike(p, precision)[0] for c, p in zip(ctxs, Ps)]
    t_sep = (time.perf_counter() - t0) / reps * 1e3
    res["per_analysis_launches"] = {"ms": t_sep, "logl_per_s": total / t_sep * 1e3}
    dev = max(float(np.max(np.abs(a - b))) for a, b in zip(ref, outs))
    res["max_abs_dlogl_queue_vs_per_analysis"] = dev
    m.close()
    for c in ctxs:
        c.close()
    return res


def precision(n_draws=100000, log=print):
    an = W.config1()
    res = {"draws": n_draws}
    ctx = Q.Context(an)
    out = {}
    for snr_target in (None, 10.0, 25.0, 80.0):
        inj = W.injection1()
        if snr_target is not None:
            # rescale the distance so the injection has the requested network SNR
            h, _ = ctx.templates(inj[None, :])
            SN = an.SN()
            hh = 0.0
            for d in range(an.nifo):
                k = np.arange(an.imin, an.imax)
                hh += 4 * np.sum((h[0, d, k] ** 2 + h[0, d, an.N - k] ** 2) / SN[d, k])
            inj = inj.copy()
            inj[6] += np.log(np.sqrt(hh) / snr_target)
        ctx2, D = build_context(an, inj, 4242)
        r = np.random.default_rng(17)
        n = n_draws if snr_target is None else max(4096, n_draws // 10)
        near = W.ensemble_around(inj, n // 2, seed=18, scale=5.0)
        far = W.ensemble_around(inj, n - n // 2, seed=19, scale=1.0)
        far[:, 0] += r.uniform(-0.05, 0.05, far.shape[0])
        far[:, 1] = np.maximum(far[:, 1], far[:, 0] - np.log(0.25 ** 0.6) + 1e-9)
        P = np.vstack([near, far])
        L64 = np.concatenate([ctx2.loglike(P[i:i + 8192], Q.FP64)[0] for i in range(0, n, 8192)])
        L32 = np.concatenate([ctx2.loglike(P[i:i + 8192], Q.MIXED)[0] for i in range(0, n, 8192)])
        err = np.abs(L32 - L64)
        key = "prior_and_posterior_draws" if snr_target is None else "snr_%g" % snr_target
        out[key] = {"n": int(n), "max": float(err.max()), "p99": float(np.percentile(err, 99)),
                    "median": float(np.median(err)), "logl_range": [float(L64.min()), float(L64.max())]}
        log(key, out[key])
        ctx2.close()
    res["abs_dlogl_mixed_vs_fp64"] = out
    ctx.close()
    return res


def serving(seconds=60.0, rate=20.0, budget_gb=200.0, lifetime_steps=20, precision=Q.MIXED, seed=7, log=print):
    """Poisson arrivals of analyses; each is admitted with a resident PT
    ensemble of 512-4096 points (8 temperatures), issues lifetime_steps
    one-step requests against it and leaves.  Returns step-latency
    percentiles and the aggregate evaluation rate."""
    specs = uneven_specs(seed=seed, n=16)
    pool = []
    for i, (an, inj, _) in enumerate(specs):
        ctx, D = build_context(an, inj, 900 + i)
        ctx.close()
        pool.append((an, inj, D, an.SN()))
    srv = Q.Server(int(budget_gb * 1e9), precision)
    rng = np.random.default_rng(seed)
    lat, lock = [], threading.Lock()
    counters = {"admitted": 0, "refused": 0, "requests": 0, "points": 0, "bin_points": 0, "errors": 0}
    stop = threading.Event()
    threads = []

    def client(aid, B, nbd):
        for _ in range(lifetime_steps):
            if stop.is_set():
                break
            code, L, R, S, ms = srv.wait(srv.submit_steps(aid, 1))
            with lock:
                if code == 0:
                    lat.append(ms)
                    counters["requests"] += 1
                    counters["points"] += B
                    counters["bin_points"] += B * nbd
                else:
                    counters["errors"] += 1
        srv.evict(aid)

    t_start = time.perf_counter()
    next_arrival = t_start
    k = 0
    while time.perf_counter() - t_start < seconds:
        now = time.perf_counter()
        if now < next_arrival:
            time.sleep(min(0.005, next_arrival - now))
            continue
        next_arrival += rng.exponential(1.0 / rate)
        an, inj, D, SN = pool[k % len(pool)]
        B = int(rng.choice([512, 1024, 2048, 4096]))
        init = W.ensemble_around(inj, B, seed=1000 + k)
        code, aid = srv.admit_pt(an, D, SN, 8, B // 8, 1.2, init, SIGMA * 0.3, 1000 + k)
        if code != 0:
            counters["refused"] += 1
        else:
            counters["admitted"] += 1
            th = threading.Thread(target=client, args=(aid, B, (an.imax - an.imin) * an.nifo), daemon=True)
            th.start()
            threads.append(th)
        k += 1
    t_end = time.perf_counter()
    n_done = counters["points"]
    stop.set()
    for th in threads:
        th.join(timeout=30)
    st = srv.stats()
    srv.close()
    lat = np.asarray(lat)
    return {"seconds": t_end - t_start, "arrival_rate_per_s": rate, "budget_gb": budget_gb,
            "precision": "mixed" if precision == Q.MIXED else "fp64",
            "form": "resident PT ensembles (8 temperatures, 512-4096 points), one-step requests",
            "step_latency_ms_p50": float(np.percentile(lat, 50)) if lat.size else None,
            "step_latency_ms_p99": float(np.percentile(lat, 99)) if lat.size else None,
            "logl_per_s": n_done / (t_end - t_start),
            "bin_detector_points_per_s": counters["bin_points"] / (t_end - t_start), "batches": st["batches"],
            **counters}


# FP64 operations of one (point, bin) in the fused FP64 kernel, counted from
# its source (qcbc_kernels.hip: amp_turns, sincos_turns, F64Seg; an FMA
# counts 2, exp, atan, sqrt and a division 1 each).  The waveform part by
# (amplitude region, phase region), 0 inspiral, 1 intermediate, 2
# merger-ringdown: f^(-1/3), f^(-2/3) and the f tshift_0 term (4), f^(-1)
# (1, intermediate), Mf = Msec f, Mf^(-1/6), Mf^(-1/3) and 1/Mf (5, merger),
# and each region's amplitude with its f^(-7/6) or Mf^(-7/6) scale and phase:
_FUSED_AMP = (17, 13, 9)     # Horner in f^(1/3) and f^(-2/3); Newton form in f; exp / Lorentzian
_FUSED_PHASE = (24, 6, 15)   # Horner in f^(+-1/3) + log terms; + log, f^(-3); + sqrt, atan (the Mf-linear
                             # term rides in the f coefficient's FMA)


def _fused_waveform_flops(a, p):
    return (4 + (1 if a == 1 or p == 1 else 0) + (5 if a == 2 or p == 2 else 0)
            + _FUSED_AMP[a] + _FUSED_PHASE[p])


FUSED_WAVEFORM_FLOPS = {(a, p): _fused_waveform_flops(a, p) for a in range(3) for p in range(3)}
# reduction to t in [-1/2, 1/2] (2), t^2 (1), the two degree-6 polynomials
# (12 + 12), t P (1), two double-angle steps of twice sin and cos (3 each)
FUSED_SINCOS_FLOPS = 34


def fused_detector_flops(nd):
    """a^2, a cos x0, a sin x0 (3); per detector the a^2 / S sum (2); detector
    0's Re(D e^{-i x0}) sum (4); every further detector's Horner step
    H e^{i phi} + a D e^{-i x0} (14)."""
    return 3 + 2 * nd + 4 + 14 * (nd - 1)


def fused_work(ctx, an, params):
    """(bin-points, FP64 flops) of one fused FP64 launch over params: each
    point's bins [max(kstart, imin), min(kend, imax)) split by region pair."""
    recs = Q.point_records(ctx, params)
    per_det = fused_detector_flops(an.nifo) + FUSED_SINCOS_FLOPS
    bins = 0
    flops = 0
    for r in recs:
        if r.status != 0:
            continue
        k = np.arange(max(r.kstart, an.imin), min(r.kend, an.imax))
        ar = np.where(k < r.seg_a1, 0, np.where(k >= r.seg_a2, 2, 1))
        pr = np.where(k < r.seg_p1, 0, np.where(k >= r.seg_p2, 2, 1))
        cnt = np.bincount(ar * 3 + pr, minlength=9)
        bins += int(cnt.sum())
        flops += sum(int(cnt[a * 3 + p]) * (FUSED_WAVEFORM_FLOPS[(a, p)] + per_det)
                     for a in range(3) for p in range(3))
    return bins, flops
