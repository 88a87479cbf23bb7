#!/usr/bin/env python3
"""Benchmark of the waveform + likelihood path on one MI355X.

Workload: the batched PT ensemble at Tobs = 16 s, fs = 4096 Hz, 3 detectors
(H, L, Virgo-like with 2x PSD), band 20-2048 Hz (32448 bins), a low-mass
injection (1.5 + 1.3 Msun at 40 Mpc) in seeded Gaussian noise, and 32
temperatures x 64 chains = 2048 parameter points per step.  A step is one PT
step resident on the GPU: propose 2048 points, set up IMRPhenomD for each,
evaluate amplitude, phase, projection and the PSD-weighted inner products over
all bins and detectors (one fused launch), Metropolis accept, temperature
swaps.  The metric is logL evaluations per second over the timed steps.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line on stdout; diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TF = 78.6        # MI355X FP64, vector ALU and matrix cores alike
FP32_VALU_PEAK_TF = 157.3  # MI355X FP32 vector (packed issue)

# HBM bytes per launch of the fused FP64 kernel on this workload (one launch
# over the 2048 points), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (profiles/k_ll_fp64_pt_step_hbm_traffic.md), each scaled by what the counter
# reads for a 1 GiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).
K_LL_FP64_TRAFFIC_BYTES = 30_700_000
# The same for the mixed-precision kernel (profiles/k_ll_mixed_pt_step_valu.md).
K_LL_MIXED_TRAFFIC_BYTES = 20_670_000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def warm_clocks(pt, seconds):
    """Untimed PT steps for `seconds` of wall time.  After even half a second
    of idle the GPU steps ~20% slower for its first tens of milliseconds of
    work (its clocks ramp), so the W warm-up steps alone leave a 20-step timed
    region below the steady rate."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        pt.run(50)


def percentile(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def cpu_baseline(an, D, SN, P, seconds):
    """The CPU oracle (the reference's per-call structure, OpenMP over points
    as QuickCBC.c:2063) on a bounded sample of the same workload."""
    from oracle import oracle as O
    O.lib()
    n = 256
    t0 = time.perf_counter()
    O.loglike_batch(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P[:n])
    dt = time.perf_counter() - t0
    n2 = int(min(P.shape[0] * 64, max(n, seconds / max(dt / n, 1e-6))))
    idx = np.arange(n2) % P.shape[0]
    t0 = time.perf_counter()
    O.loglike_batch(an.labels, an.gmst, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, P[idx])
    dt = time.perf_counter() - t0
    return n2 / dt, n2, dt, O.lib().qo_num_threads()


def het_bench(Q, W, an, ctx, inj, D, SN, P, launches, cpu_seconds, no_cpu):
    """The heterodyned likelihood (log_likelihood_het, QuickCBC.c:5303-5444)
    for the same 2048-point ensemble: heterodyne at the injection on a grid
    over limits()' SNR band, then batched evaluation with the points resident
    on the GPU (per launch: IMRPhenomD set-up of every point, then one wave per
    point over the grid samples)."""
    lo, hi = W.prior_box(an)
    ctx.het_space(inj, lo, hi)
    t0 = time.perf_counter()
    sp = ctx.het_space(inj, lo, hi)  # the reference's grid choice (QuickCBC.c:4570-4910)
    space_ms = (time.perf_counter() - t0) * 1e3
    g, MN, MM = sp["fgrid"], sp["MN"], sp["MM"]
    t0 = time.perf_counter()
    het = Q.Het(ctx, params=inj, fgrid=g, J=2, MN=MN, MM=MM)
    setup_ms = (time.perf_counter() - t0) * 1e3
    B = P.shape[0]
    dp = Q.DeviceBuffer.from_array(np.ascontiguousarray(P, np.float64))
    dl = Q.DeviceBuffer(8 * B)
    for _ in range(20):
        het.loglike_device(dp.ptr, B, dl.ptr)
    het.sync()
    t0 = time.perf_counter()
    for _ in range(launches):
        het.loglike_device(dp.ptr, B, dl.ptr)
    het.sync()
    T = time.perf_counter() - t0
    kernel_ms = het.time_kernel(dp.ptr, B, launches)
    Lh, sh = het.loglike(P)
    Lf, _, _ = ctx.loglike(P, Q.FP64)
    # het_space's own accuracy target: the eleven points efix put at distance 30
    Lhd, _ = het.loglike(sp["dparams"])
    Lfd, _, _ = ctx.loglike(sp["dparams"], Q.FP64)
    near = float(np.abs(Lhd - Lfd).max())
    out = {"logl_per_s": B * launches / T, "launch_ms": T / launches * 1e3, "kernel_ms": kernel_ms,
           "kernel_ms_source": "HIP events around %d back-to-back launches: the kernel's duration plus the "
                               "launch gaps, which show at this size (rocprof's duration: "
                               "profiles/k_het_ll_fp64_2048_points.md)" % launches,
           "kernel": "k_het_ll<%d>" % an.nifo, "points": B, "grid_samples": het.M, "regions": het.NR,
           "grid": "het_space (limits, Fisher_All, FisherEvec, efix to 30, tolerance 0.003)",
           "band_bins": [MN, MM], "het_space_ms": space_ms, "setup_ms": setup_ms,
           "het_vs_full_abs_dlogl_at_efix_points_max": near,
           "het_vs_full_abs_dlogl_median": percentile(np.abs(Lh - Lf), 50)}
    out["fisher"] = fisher_bench(Q, het, P, dp, launches, no_cpu, an, D, SN, inj, g, MN, MM)
    out["pt_update"] = update_bench(Q, W, an, ctx, het, inj, P.shape[0], no_cpu, D, SN, g, MN, MM)
    het.close()
    dp.free()
    dl.free()
    if not no_cpu:
        from oracle import oracle as O
        st, oh = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, g, 2, MN, MM)
        n0 = 2048
        t0 = time.perf_counter()
        O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, oh, P[:n0])
        dt = time.perf_counter() - t0
        n2 = int(min(P.shape[0] * 512, max(n0, cpu_seconds / max(dt / n0, 1e-9))))
        idx = np.arange(n2) % P.shape[0]
        t0 = time.perf_counter()
        O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, oh, P[idx])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n2 / dt, "unit": "logL evals/s", "cores": O.lib().qo_num_threads(),
                               "kind": "port", "sample": "%d points of the same ensemble, %.1f s, gcc -O2 -fopenmp, "
                                                         "OpenMP over points" % (n2, dt)}
    return out


def fisher_bench(Q, het, P, dp, launches, no_cpu, an, D, SN, inj, g, MN, MM):
    """Fisher_Het + FisherEvec (QuickCBC.c:5446-5617, :2449-2546) at every
    point of the ensemble on the same heterodyne: the set-up of the 23
    perturbed parameter sets per point, the Fisher kernel and the eigen kernel,
    each timed alone (HIP events around back-to-back launches), and the whole
    host call."""
    B = P.shape[0]
    reps = max(5, launches // 50)
    het.fisher_time(dp.ptr, B, 2)
    setup_ms, fisher_ms, evec_ms = het.fisher_time(dp.ptr, B, reps)
    het.fisher(P)
    t0 = time.perf_counter()
    for _ in range(5):
        het.fisher(P)
    call_ms = (time.perf_counter() - t0) / 5 * 1e3
    out = {"points": B, "setup_ms": setup_ms, "fisher_kernel_ms": fisher_ms, "evec_kernel_ms": evec_ms,
           "call_ms": call_ms, "points_per_s": B / (call_ms * 1e-3),
           "kernels": ["k_setup x 23 points", "k_het_fisher<%d>" % an.nifo, "k_fisher_evec"]}
    if not no_cpu:
        from oracle import oracle as O
        oh = _oracle_het(O, an, D, SN, inj, g, MN, MM)
        n = 1024
        t0 = time.perf_counter()
        F, _ = O.fisher_het(an.labels, an.gmst, an.fref, an.Tobs, oh, P[:n])
        O.fisher_evec(F)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "points/s (Fisher_Het + FisherEvec)",
                               "cores": O.lib().qo_num_threads(), "kind": "port",
                               "sample": "%d points of the same ensemble, %.2f s, OpenMP over points (Fisher), "
                                         "one thread (eigen)" % (n, dt)}
    return out


def _oracle_het(O, an, D, SN, inj, g, MN, MM):
    st, oh = O.heterodyne(an.labels, an.gmst, an.fref, D, SN, an.N, an.Tobs, inj, g, 2, MN, MM)
    return oh


def update_bench(Q, W, an, ctx, het, inj, B, no_cpu, D, SN, g, MN, MM, steps=1000, lat_steps=200, groups=0):
    """QuickCBC's own PT step on the heterodyned likelihood (update()'s
    Fisher / differential-evolution / jiggle / sky-ring proposals, 0.5 / 0.2 /
    0.1 / 0.2, with the reference's priors, MCMC_all's swap-or-update step,
    QuickCBC.c:2029-2092 and :2238-2405) for the 32 x 64 ensemble: steps per second over `steps`
    resident steps (eager and graph replay), the synchronised step latency,
    the acceptance per proposal type, and the Fisher refresh."""
    lo, hi = W.prior_box(an)
    n_temps, n_chains = 32, B // 32
    init = W.ensemble_in_prior(inj, n_temps * n_chains, lo, hi, seed=99)
    desc = Q.update_desc(lo, hi, efix_zs=1.0)  # efix of the Fisher jumps, QuickCBC.c:1927
    t0 = time.perf_counter()
    pt = Q.PTUpdate(ctx, het, n_temps, n_chains, 1.15, init, desc, seed=5, groups=groups)
    create_ms = (time.perf_counter() - t0) * 1e3
    pt.run(20)
    t0 = time.perf_counter()
    pt.run(steps)
    T = time.perf_counter() - t0
    pt.run(steps, graph=True)  # captures the graphs of this step count
    t0 = time.perf_counter()
    pt.run(steps, graph=True)
    Tg = time.perf_counter() - t0
    pt.step_timing(True)
    pt.run(lat_steps)
    lat = pt.latencies()
    pt.step_timing(False)
    t0 = time.perf_counter()
    pt.refresh_fisher()
    refresh_ms = (time.perf_counter() - t0) * 1e3
    _, _, _, _, _, st = pt.get_update()
    tot = st.sum(0)
    out = {"points": n_temps * n_chains, "groups": pt.groups, "steps_per_s": steps / T,
           "step_ms": T / steps * 1e3, "graph_steps_per_s": steps / Tg,
           "logl_per_s": n_temps * n_chains * steps / T,
           "step_ms_p50": percentile(lat * 1.0, 50), "step_ms_p99": percentile(lat * 1.0, 99),
           "swap_step_fraction": float(tot[1] / max(1, tot[0] + tot[1])),
           "swap_accept": float(tot[2] / max(1, tot[1])),
           "accept": {k: float(tot[4 + 2 * i] / max(1, tot[3 + 2 * i])) for i, k in
                      enumerate(("fisher", "de", "jiggle", "sky_ring"))},
           "proposed": {k: int(tot[3 + 2 * i]) for i, k in enumerate(("fisher", "de", "jiggle", "sky_ring"))},
           "fisher_refresh_ms": refresh_ms, "create_ms": create_ms}
    pt.close()
    if not no_cpu:
        from oracle import oracle as O
        oh = _oracle_het(O, an, D, SN, inj, g, MN, MM)
        P = n_temps * n_chains
        x = init.copy()
        L, _ = O.loglike_het(an.labels, an.gmst, an.fref, an.Tobs, oh, x)
        F, _ = O.fisher_het(an.labels, an.gmst, an.fref, an.Tobs, oh, x)
        ej, ev, _ = O.fisher_evec(F)
        hist = O.pt_update_hist_init(x, n_temps, n_chains, 1000, 5)
        y = x.copy()
        S = dict(x=x, L=L, ej=ej, ev=ev, hist=hist, hcount=np.zeros(n_chains, np.int64), y=y,
                 typ=np.full(P, 3, np.int32), jac=np.zeros(P), stats=np.zeros((n_chains, 11), np.int64), step=0)
        uc = np.concatenate([lo, hi, [0.5, 0.3, 0.2, 0.05, an.gmst, an.nifo], list(an.labels) + [0] * (3 - an.nifo)])
        n = 100
        t0 = time.perf_counter()
        O.pt_update_run(an.labels, an.gmst, an.fref, an.Tobs, oh, uc, 1000, 20, n_temps, n_chains, 1.15, 5, n, S)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": P * n / dt, "unit": "logL evals/s (whole PT steps)",
                               "cores": O.lib().qo_num_threads(), "kind": "port",
                               "sample": "%d steps of the same %d-point ensemble, %.2f s, OpenMP over points for "
                                         "the likelihood" % (n, P, dt)}
    return out


def intrinsic_bench(Q, an, ctx, P, D, SN, no_cpu, reps=50):
    """log_likelihood_intrinsic (QuickCBC.c:4093-4145), the reference-frame
    likelihood of MCMC_intrinsic, for the same 2048 points mapped to detector
    0's frame (phase, time and amplitude offsets of the other detectors),
    through the fused FP64 kernel; host calls (parameters in, logL out)."""
    from oracle import oracle as O
    Qp = []
    for p in P:
        ciota = p[10]
        Ap, Ac = (1 + ciota ** 2) / 2, -ciota
        dt = O.times(p[7], p[8], an.gmst)
        Fs, lam = [], []
        for d in an.labels:
            Fp, Fc = O.f_ant(p[9], p[7], p[8], an.gmst, d)
            Fs.append(np.sqrt(Ap ** 2 * Fp ** 2 + Ac ** 2 * Fc ** 2))
            l = np.arctan2(Ac * Fc, Ap * Fp)
            lam.append(l + 2 * np.pi if l < 0 else l)
        q = [p[0], p[1], p[2], p[3], 0.5 * (lam[0] + p[4]), p[5] + dt[an.labels[0]], p[6] - np.log(Fs[0])]
        for d in range(1, an.nifo):
            q += [lam[d] - lam[0], dt[an.labels[d]] - dt[an.labels[0]], Fs[d] / Fs[0]]
        Qp.append(q)
    Qp = np.array(Qp)
    ctx.loglike_intrinsic(Qp)
    t0 = time.perf_counter()
    for _ in range(reps):
        L, s2, st = ctx.loglike_intrinsic(Qp)
    T = (time.perf_counter() - t0) / reps
    Lf, _, _ = ctx.loglike(P, Q.FP64)
    loud = (s2 > 25).sum(1) >= 2
    out = {"points": Qp.shape[0], "logl_per_s": Qp.shape[0] / T, "call_ms": T * 1e3,
           "loud_fraction": float(loud.mean()),
           "vs_full_max_abs_dlogl_when_loud": float(np.abs(L - Lf)[loud].max()) if loud.any() else None}
    if not no_cpu:
        n = 256
        t0 = time.perf_counter()
        O.loglike_intrinsic(an.nifo, an.fref, an.fmin, an.fmax, D, SN, an.N, an.Tobs, Qp[:n])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "logL evals/s", "cores": O.lib().qo_num_threads(),
                               "kind": "port", "sample": "%d points, %.2f s, OpenMP over points" % (n, dt)}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--warm-seconds", type=float, default=0.5,
                    help="untimed PT steps before the warm-up steps, so the GPU leaves its idle clock state")
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--kernel-steps", type=int, default=500, help="PT steps of the kernel-alone timing")
    ap.add_argument("--plain-points", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="autotune the launch shape first")
    ap.add_argument("--only-main", action="store_true", help="time the FP64 PT steps only (for profiling)")
    ap.add_argument("--main-precision", choices=("fp64", "mixed"), default="fp64",
                    help="with --only-main: the precision of the timed PT steps (mixed: to profile that kernel)")
    ap.add_argument("--pt-groups", type=int, default=0, help="PT column groups (0: the library's choice)")
    ap.add_argument("--config", type=int, default=1,
                    help="1: the PT-ensemble headline (default); 2: uneven persistent batch; "
                         "3: mixed-precision accuracy; 4: serving stream")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-steps", type=int, default=200, help="PT-step requests per analysis")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="analysis arrivals per second")
    ap.add_argument("--serve-bench-seconds", type=float, default=8.0,
                    help="seconds of the serving stream inside the default run (0: none)")
    ap.add_argument("--draws", type=int, default=100000)
    ap.add_argument("--het-launches", type=int, default=500)
    ap.add_argument("--only-het", action="store_true", help="time the heterodyned likelihood only (for profiling)")
    ap.add_argument("--only-aux", action="store_true",
                    help="time the reference-frame likelihood, data conditioning and makespec only (for profiling)")
    ap.add_argument("--only-update", action="store_true",
                    help="time QuickCBC's own PT step on the heterodyned likelihood only (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the points of a PT step "
              "would shard over GPUs by column groups, each GPU holding the band planes); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from quickcbc_amd import qcbc as Q
    from quickcbc_amd import workloads as W

    name, cus, mem = Q.device_info()
    log("device:", name, cus, "CUs", "%.0f GB" % mem)
    if args.config != 1:
        from quickcbc_amd import bench_configs as C
        if args.config == 2:
            out = {"config": "uneven persistent batch", "fp64": C.uneven(precision=Q.FP64, log=log),
                   "mixed": C.uneven(precision=Q.MIXED, log=log)}
        elif args.config == 3:
            out = {"config": "mixed precision vs FP64", **C.precision(n_draws=args.draws, log=log)}
        elif args.config == 4:
            out = {"config": "serving stream", **C.serving(seconds=args.serve_seconds, rate=args.serve_rate,
                                                           lifetime_steps=args.serve_steps, log=log)}
        else:
            raise SystemExit("unknown --config %d" % args.config)
        print(json.dumps(out), flush=True)
        return

    an = W.config1()
    inj = W.injection1()
    ctx = Q.Context(an)
    h, st = ctx.templates(inj[None, :])
    assert st[0] == 0
    D = np.ascontiguousarray(h[0] + W.noise(an, 2024))
    SN = an.SN()
    ctx.bind(D, SN)
    n_temps, n_chains = 32, 64
    B = n_temps * n_chains
    init = W.ensemble_around(inj, B, seed=99)
    sigma = np.array([5e-6, 2.5e-5, 5e-4, 5e-4, 0.0125, 5e-6, 0.0025, 0.0025, 0.0025, 0.005, 0.0025])
    nb = an.imax - an.imin
    plane_bytes = (4 + 3 * an.nifo) * nb * 8    # f, f^(1/3), f^(-1/6), log f; Re d/S, Im d/S, 1/S per detector

    if args.only_aux:
        r = {"intrinsic": intrinsic_bench(Q, an, ctx, init, D, SN, True)}
        rng = np.random.default_rng(1)
        x = rng.standard_normal((an.nifo, an.N)) * 1e-21
        c2 = Q.Context(an)
        c2.bind_time_series(x, SN)
        t0 = time.perf_counter()
        for _ in range(10):
            c2.bind_time_series(x, SN)
        r["bind_time_series_ms"] = (time.perf_counter() - t0) / 10 * 1e3
        fk = np.concatenate([[0.5 / an.Tobs], np.linspace(5.0, an.N / 2 / an.Tobs - 1.0, 40),
                             [an.N / 2 / an.Tobs]])
        yk = np.log(W.psd(np.maximum(fk, 10.0)))
        lines = (np.array([60.0, 331.5, 1000.0]), np.array([50.0, 8.0, 3.0]), np.array([1e3, 4e2, 2e2]),
                 np.array([2.0, 1.5, 3.0]), np.array([0.3, 0.2, 0.5]))
        c2.makespec(fk, yk, lines)
        t0 = time.perf_counter()
        for _ in range(20):
            c2.makespec(fk, yk, lines)
        r["makespec_ms"] = (time.perf_counter() - t0) / 20 * 1e3
        c2.close()
        print(json.dumps({"config": "reference-frame likelihood, data conditioning, makespec", **r}), flush=True)
        return

    if args.only_update:
        lo, hi = W.prior_box(an)
        sp = ctx.het_space(inj, lo, hi)
        g, MN, MM = sp["fgrid"], sp["MN"], sp["MM"]
        het = Q.Het(ctx, params=inj, fgrid=g, J=2, MN=MN, MM=MM)
        dp = Q.DeviceBuffer.from_array(np.ascontiguousarray(init, np.float64))
        r = {"fisher": fisher_bench(Q, het, init, dp, 100, True, an, D, SN, inj, g, MN, MM),
             "pt_update": update_bench(Q, W, an, ctx, het, inj, B, True, D, SN, g, MN, MM, steps=args.steps,
                                       lat_steps=20, groups=args.pt_groups)}
        print(json.dumps({"config": "QuickCBC PT step, heterodyned likelihood", **r}), flush=True)
        het.close()
        dp.free()
        return

    if args.only_het:
        r = het_bench(Q, W, an, ctx, inj, D, SN, init, args.het_launches, 5.0, True)
        print(json.dumps({"config": "heterodyned likelihood", **r}), flush=True)
        return

    if args.autotune:
        for prec in (Q.FP64, Q.MIXED):
            log("autotune", prec, ctx.autotune(prec, B))

    if args.only_main:
        # a copy of known size beside the timed kernel, for scaling byte counters
        import ctypes
        nbytes = 1 << 30
        a, b = Q.DeviceBuffer(nbytes), Q.DeviceBuffer(nbytes)
        Q.lib().qcbc_dev_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        Q._check(Q.lib().qcbc_dev_copy(b.ptr, a.ptr, nbytes), "qcbc_dev_copy")
        a.free()
        b.free()

    # ---- headline: FP64 fused path, PT steps (the library's column groups,
    # overlapping on their own streams), wall clock over the timed steps
    main_prec = Q.MIXED if (args.only_main and args.main_precision == "mixed") else Q.FP64
    pt = Q.PT(ctx, n_temps, n_chains, 1.15, init, sigma, seed=7, precision=main_prec, groups=args.pt_groups)
    warm_clocks(pt, args.warm_seconds)
    pt.run(args.warmup)
    ctx.sync()
    t0 = time.perf_counter()
    pt.run(args.steps)
    ctx.sync()
    T = time.perf_counter() - t0
    value = B * args.steps / T
    ms_per_step = 1000.0 * T / args.steps
    log("%s fused: %.1f logL/s, %.3f ms/step (%d column groups)" %
        ("FP64" if main_prec == Q.FP64 else "mixed", value, ms_per_step, pt.groups))

    # ---- the likelihood kernel alone: one group, each launch over all 2048
    # points, timed with HIP events on the stream it runs on
    pt1 = Q.PT(ctx, n_temps, n_chains, 1.15, init, sigma, seed=7, precision=main_prec, groups=1)
    pt1.run(args.warmup)
    ctx.timing(1)
    ctx.sync()
    pt1.run(args.kernel_steps)
    ctx.sync()
    kms, nl = ctx.timing(0)
    kernel_ms_in_step = kms / max(nl, 1)  # each launch inside the PT steps between its own two events
    # the same kernel on the ensemble as the steps left it, launched back to
    # back between two events on its stream: the kernel's own duration
    x_end = np.ascontiguousarray(pt1.get()[0], np.float64)
    dx = Q.DeviceBuffer.from_array(x_end)
    kernel_ms = ctx.time_kernel(dx.ptr, B, main_prec, args.kernel_steps)
    dx.free()
    cpb, tile, _ = ctx.get_shape(main_prec)
    log("likelihood kernel alone: %.4f ms/launch over %d back-to-back launches, %.4f ms inside the PT steps "
        "(%d points per block, %d-bin tiles)" % (kernel_ms, args.kernel_steps, kernel_ms_in_step, cpb, tile))
    # algorithmic FP64 work of one launch, on the ensemble as the timed steps left it
    from quickcbc_amd import bench_configs as C
    bin_points, src_flops = C.fused_work(ctx, an, x_end)
    pt1.close()
    # the path's algorithmic FP64 work per bin-point: the waveform's 60 flops
    # (Mf 1, powers 7, amplitude 20, inspiral phase 27, combine 5) and per
    # detector 19 (projection 9, the two inner products 10), times the
    # bin-points the launch evaluates
    flops_per_bp = 60 + 19 * an.nifo
    flops = flops_per_bp * bin_points
    achieved_tf = flops / (kernel_ms * 1e-3) / 1e12
    extra = {}
    if not args.only_main:
        # graph-replayed steps (the same step count once untimed: its graphs are captured then)
        pt.run(args.steps, graph=True)
        ctx.sync()
        t0 = time.perf_counter()
        pt.run(args.steps, graph=True)
        Tg = time.perf_counter() - t0
        extra["fp64_graph_logl_per_s"] = B * args.steps / Tg
        # step latency
        pt.step_timing(True)
        pt.run(args.latency_steps)
        lat = pt.latencies()
        extra["fp64_step_ms_p50"] = percentile(lat, 50)
        extra["fp64_step_ms_p99"] = percentile(lat, 99)
        # the same, each step one replay of the captured one-step graph
        pt.run(args.latency_steps, graph=True)
        lat = pt.latencies()
        pt.step_timing(False)
        extra["fp64_graph_step_ms_p50"] = percentile(lat, 50)
        extra["fp64_graph_step_ms_p99"] = percentile(lat, 99)
        x, Lx, acc, sw = pt.get()
        extra["pt_accept_rate"] = float(acc.sum()) / (B * pt.steps_run)
        extra["pt_groups"] = pt.groups
        pt.close()
        # latency-first configuration: one column group (three launches per
        # step instead of nine, one stream to wait for)
        ptl = Q.PT(ctx, n_temps, n_chains, 1.15, init, sigma, seed=7, precision=Q.FP64, groups=1)
        ptl.run(args.warmup)
        ptl.step_timing(True)
        ptl.run(args.latency_steps)
        lat = ptl.latencies()
        ptl.step_timing(False)
        ptl.close()
        extra["fp64_one_group_step_ms_p50"] = percentile(lat, 50)
        extra["fp64_one_group_step_ms_p99"] = percentile(lat, 99)
        # mixed precision on the same shape
        ptm = Q.PT(ctx, n_temps, n_chains, 1.15, init, sigma, seed=7, precision=Q.MIXED, groups=args.pt_groups)
        warm_clocks(ptm, 0.2)
        ptm.run(args.warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ptm.run(args.steps)
        Tm = time.perf_counter() - t0
        extra["mixed_logl_per_s"] = B * args.steps / Tm
        pm1 = Q.PT(ctx, n_temps, n_chains, 1.15, init, sigma, seed=7, precision=Q.MIXED, groups=1)
        pm1.run(args.warmup)
        ctx.timing(1)
        ctx.sync()
        pm1.run(args.kernel_steps)
        mms, mnl = ctx.timing(0)
        dxm = Q.DeviceBuffer.from_array(np.ascontiguousarray(pm1.get()[0], np.float64))
        extra["mixed_kernel_ms"] = ctx.time_kernel(dxm.ptr, B, Q.MIXED, args.kernel_steps)
        extra["mixed_kernel_ms_in_pt_step"] = mms / max(mnl, 1)
        dxm.free()
        pm1.close()
        ptm.step_timing(True)
        ptm.run(args.latency_steps)
        lat = ptm.latencies()
        extra["mixed_step_ms_p50"] = percentile(lat, 50)
        extra["mixed_step_ms_p99"] = percentile(lat, 99)
        ptm.run(args.latency_steps, graph=True)
        lat = ptm.latencies()
        extra["mixed_graph_step_ms_p50"] = percentile(lat, 50)
        extra["mixed_graph_step_ms_p99"] = percentile(lat, 99)
        ptm.close()
        L64, _, _ = ctx.loglike(init, Q.FP64)
        L32, _, _ = ctx.loglike(init, Q.MIXED)
        err = np.abs(L32 - L64)
        extra["mixed_abs_dlogl_vs_fp64"] = {"max": float(err.max()), "p99": percentile(err, 99),
                                            "median": percentile(err, 50)}
        # plain FP64 path: the reference's three loops, one point after another
        npl = args.plain_points
        ctx.loglike_plain(init[:8])
        t0 = time.perf_counter()
        Lp, _, _ = ctx.loglike_plain(init[:npl])
        Tp = time.perf_counter() - t0
        extra["plain_fp64_logl_per_s"] = npl / Tp
        extra["plain_vs_fused_max_abs_dlogl"] = float(np.abs(Lp - L64[:npl]).max())
        extra["mixed_over_plain"] = extra["mixed_logl_per_s"] / extra["plain_fp64_logl_per_s"]
        extra["mixed_over_fp64_fused"] = extra["mixed_logl_per_s"] / value
        extra["fp64_fused_over_plain"] = value / extra["plain_fp64_logl_per_s"]
        extra["launch_shape_fp64"] = ctx.get_shape(Q.FP64)
        extra["launch_shape_mixed"] = ctx.get_shape(Q.MIXED)
        extra["het"] = het_bench(Q, W, an, ctx, inj, D, SN, init, args.het_launches, args.cpu_seconds / 3,
                                 args.no_cpu)
        extra["intrinsic"] = intrinsic_bench(Q, an, ctx, init, D, SN, args.no_cpu)
        if args.serve_bench_seconds > 0:
            # the serving form, briefly: Poisson arrivals of analyses with resident PT ensembles
            from quickcbc_amd import bench_configs as C
            sv = C.serving(seconds=args.serve_bench_seconds, rate=args.serve_rate, lifetime_steps=args.serve_steps,
                           log=log)
            extra["serving_mixed_logl_per_s"] = sv["logl_per_s"]
            extra["serving_step_ms_p50"] = sv["step_latency_ms_p50"]
            extra["serving_step_ms_p99"] = sv["step_latency_ms_p99"]
            extra["serving"] = sv
            extra["serving_mixed_over_plain"] = sv["logl_per_s"] / extra["plain_fp64_logl_per_s"]
        log(json.dumps(extra))

    cpu = None
    if not args.no_cpu and not args.only_main:
        rate, n, dt, threads = cpu_baseline(an, D, SN, init, args.cpu_seconds)
        cpu = {"value": rate, "unit": "logL evals/s", "cores": threads, "kind": "port",
               "sample": "%d points of the same 2048-point ensemble (Tobs=16 s, 3 det, 32448 bins), "
                         "%.1f s, gcc -O2 -fopenmp, OpenMP over points" % (n, dt)}
        log("cpu baseline:", cpu)

    traffic = K_LL_FP64_TRAFFIC_BYTES if main_prec == Q.FP64 else K_LL_MIXED_TRAFFIC_BYTES
    out = {
        "metric": "logL evals/sec (waveform+logL) at Tobs=16s 3-det; p99 PT-step latency",
        "value": value,
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == Q.FP64 else "mixed (f64 waveform and sums, f32 sin/cos and data products)",
        "data": "synthetic: seeded Gaussian noise coloured by the analytic PSD plus an IMRPhenomD injection",
        "config": {"workload": "batched PT ensemble, Tobs=16 s, fs=4096 Hz, 3 det (H/L/V), band 20-2048 Hz "
                               "(32448 bins), 32 temps x 64 chains = 2048 points/step, %s fused launch"
                               % ("FP64" if main_prec == Q.FP64 else "mixed-precision"),
                   "points_per_step": B, "bins": nb, "detectors": an.nifo},
        "roofline": {"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TF, "unit": "TFLOP/s",
                     "frac": achieved_tf / FP64_PEAK_TF, "traffic": traffic,
                     "traffic_source": "constant: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel on this "
                                       "workload, scaled by a 1 GiB copy (profiles/%s); not measured in this run"
                                       % ("k_ll_fp64_pt_step_hbm_traffic.md" if main_prec == Q.FP64
                                          else "k_ll_mixed_pt_step_valu.md"),
                     "ceiling": ("FP64 vector ALU issue: the kernel has no MFMA; MI355X's FP64 vector peak "
                                 "(78.6 TFLOP/s, an FMA counted as 2)" if main_prec == Q.FP64 else
                                 "mixed kernel (about 70% of its executed flops in FP64, the rest FP32): the path's "
                                 "flops against the FP64 vector peak, a comparison with the FP64 kernel only"),
                     "kernel": "k_ll<3,%s,%d>" % ("FP64" if main_prec == Q.FP64 else "MIXED", tile),
                     "kernel_ms": kernel_ms, "kernel_ms_source": "HIP events around %d back-to-back launches on "
                     "the kernel's stream, over the ensemble the timed PT steps left" % args.kernel_steps,
                     "kernel_ms_in_pt_step": kernel_ms_in_step, "points_per_block": cpb,
                     "flops_per_launch": flops, "bin_points_per_launch": bin_points,
                     "flops_per_bin_point": flops_per_bp,
                     "kernel_source_flops_per_bin_point": src_flops / max(bin_points, 1)},
        "hbm": {"data_stream_bytes_per_launch": plane_bytes,
                "traffic_bytes_per_launch": traffic,
                "achieved_gbs": traffic / (kernel_ms * 1e-3) / 1e9 if traffic else None,
                "peak_gbs": HBM_PEAK_GBS},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
