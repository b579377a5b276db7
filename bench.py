#!/usr/bin/env python3
"""Benchmark of the MI355X grid-propagation path (QDng Chebyshev / GSPO).

Workload (one GPU): 2D grid of 512 x 512 points on [-10, 10]^2, two electronic
states with masses (1836, 12000) per dimension, V1 a 2D Morse surface
(De = 0.2, a = 1), V2 a displaced harmonic surface (k = 0.1, shift 1.5,
+0.05 Eh), coupling V12 = 0.01 exp(-r^2), both surfaces capped at 1 Eh;
Gaussian psi0 (sigma 0.5) on state 1 centred at a seed-1234 random point in
[-1, 1]^2.  Chebyshev propagation, dt = 5 a.u., QDng's automatic order and
default break test (prec 1e-10), plain FP64 path.  One step = one Chebyshev
time step of the two-state wave function, inputs resident in HBM.

metric: grid-pt * H*psi / s = sum over steps of N_points * states * executed
H applications (1 + recurrence iterations actually run), divided by the wall
time of the timed region.

Also reported (object "extra"), one record per benchmark configuration
("configs": the 2D two-state Chebyshev and GSPO workload, the 3D 128^3
precision variant, the 512-job batch and the serving stream), each with the
FP64, compact (FP64 arithmetic, 40-bit stored terms) and FP32 paths, the
relative L2 error against FP64, a roofline figure and a CPU baseline: the
FP64 oracle with QDng's FFT library interface (FFTW3 API, served by MKL's
libmkl_rt.so when it loads; else the oracle's own transform, labelled) on the
host cores and on one thread, medians over repeats.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "grid-pt H·ψ/s + %HBM roofline, FP32 vs FP64; p50/p99 batched-job latency"
UNIT = "grid-pt*Hpsi/s"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per grid point per state per executed order of the row
# kernel (streaming model): phi_{k-1} + phi_{k-2} + phi_k + psi read/write +
# V_ii + V_ij (coupling); the prologue (order 1) reads no phi_{k-2}
B_ROW = {"fp64": (16 + 16 + 16 + 32 + 8 + 8), "fp32": (8 + 8 + 8 + 32 + 4 + 4),
         "compact": (10 + 10 + 10 + 32 + 8 + 8)}
B_ROW_PRO = {"fp64": B_ROW["fp64"] - 16, "fp32": B_ROW["fp32"] - 8, "compact": B_ROW["compact"] - 10}
# GSPO per step and grid point of one state: psi read and write + V
B_GSPO = {"fp64": 32 + 8, "fp32": 16 + 4}
PRECISIONS = ("fp64", "compact", "fp32")


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "cheby_row_2d_512x512_2state_fp64_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per row-kernel launch from the committed FETCH_SIZE /
    WRITE_SIZE profile of this workload (None if it is absent)."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def sync(ctx):
    ctx.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def build_problem(N=512):
    import cases
    return cases.morse_2state_2d(N)


def run_cheby(ctx, p, precision, steps, warmup, profile):
    import qdng_amd as Q
    H = p.gpu_ham(ctx)
    U = Q.Cheby(H, 5.0, prec=1e-10, precision=precision)
    w = p.gpu_wf(ctx)
    if warmup:
        U.run(w, 0, warmup, observables=False)
    sync(ctx)
    if profile:
        ctx.profile(True)
    t0 = time.perf_counter()
    _, _, nh, st = U.run(w, warmup, steps, observables=False)
    sync(ctx)
    t1 = time.perf_counter()
    rowk = ctx.kernel_time(Q.K_CHEBY_ROW) if profile else None
    linek = ctx.kernel_time(Q.K_LINE_PASS) if profile else None
    if profile:
        ctx.profile(False)
    return {"seconds": t1 - t0, "n_hpsi": nh, "order": U.order, "row": rowk, "line": linek}


def run_gspo(ctx, p, precision, steps, warmup, profile=False):
    """GSPO steps of the two-state workload through qdx_prop_run (one row pass
    between steps); with profile, the per-kernel times of a second pass."""
    import qdng_amd as Q
    G = Q.GSPO([p.gpu_split_op(ctx, ["kin"]), p.gpu_split_op(ctx, ["pot"])], 5.0, precision=precision)
    w = p.gpu_wf(ctx)
    if warmup:
        G.run(w, 0, warmup, observables=False)
    sync(ctx)
    t0 = time.perf_counter()
    G.run(w, warmup, steps, observables=False)
    sync(ctx)
    t = time.perf_counter() - t0
    if not profile:
        return t
    ctx.profile(True)
    G.run(w, warmup + steps, steps, observables=False)
    sync(ctx)
    rowk, colk = ctx.kernel_time(Q.K_GSPO_ROW), ctx.kernel_time(Q.K_LINE_PASS)
    ctx.profile(False)
    return t, rowk, colk


# batched jobs, single state with a laser dipole: phi_{k-1}, phi_{k-2}, phi_k,
# psi read/write, V and mu per grid point and executed order
B_BATCH = {"fp64": 16 * 3 + 32 + 8 + 8, "fp32": 8 * 3 + 32 + 4 + 4, "compact": 10 * 3 + 32 + 8 + 8}
B_BATCH_M2 = {"fp64": 16, "fp32": 8, "compact": 10}   # the prologue reads no phi_{k-2}


def run_batched(ctx, precision, steps, prof_ticks, seed=2026, tune=False):
    """512 laser-driven jobs (256 1D, 256 2D, sizes drawn from seed 2026),
    dt = 1; one tick = one step of every job.  tune: run the autotuner on
    every 2D shape before the jobs are submitted.  Returns throughput, kernel
    timings and the final states."""
    import cases
    import qdng_amd as Q
    eager_ticks = max(1, min(20, steps))
    total = steps + eager_ticks + prof_ticks + 2 + (13 if tune else 0)
    jobs = cases.batch_jobs(steps=total, seed=seed)
    B = Q.Batch(ctx, precision, 16 << 30)
    tuning = {}
    tune_s = 0.0
    if tune:
        t0 = time.perf_counter()
        for n in sorted({p.N[0] for p in jobs if p.nd == 2}):
            tuning[f"{n}x{n}"] = B.autotune(n, n)
        tune_s = time.perf_counter() - t0
    ids = [B.submit(p.gpu_ham(None), 1.0, total, p.psi0) for p in jobs]
    lds = None
    if tune:
        B.run(1)
        t0 = time.perf_counter()
        ms = B.autotune_lds()      # 12 ticks of the jobs themselves
        tune_s += time.perf_counter() - t0
        lds = {"ms_per_tick": dict(zip(["2,2", "1,2", "2,1", "1,1"], ms)), "chosen": B.lds_schedule()}
    B.run(2)                       # tables built, tick graph captured
    sync(ctx)
    h0 = np.array([B.state(j)[1] for j in ids])
    t0 = time.perf_counter()
    B.run(steps)                   # graph replays
    sync(ctx)
    t1 = time.perf_counter()
    h1 = np.array([B.state(j)[1] for j in ids])
    sizes = np.array([p.size for p in jobs])
    work = float((sizes * (h1 - h0)).sum())
    graphs = B.graph_stats()
    B.set_graphs(False)            # the same ticks launched kernel by kernel
    te0 = time.perf_counter()
    B.run(eager_ticks)
    sync(ctx)
    eager_ms = (time.perf_counter() - te0) * 1e3 / eager_ticks
    h1 = np.array([B.state(j)[1] for j in ids])
    ctx.profile(True)
    B.run(prof_ticks)
    sync(ctx)
    h2 = np.array([B.state(j)[1] for j in ids])
    rowk = ctx.kernel_time(Q.K_BATCH_ROW)
    colk = ctx.kernel_time(Q.K_BATCH_COL)
    linek = ctx.kernel_time(Q.K_BATCH_LINE)
    ctx.profile(False)
    two_d = np.array([p.nd == 2 for p in jobs])
    alg = sizes * (B_BATCH[precision] * (h2 - h1) - B_BATCH_M2[precision] * prof_ticks)
    return {"seconds": t1 - t0, "work": work, "row": rowk, "col": colk, "line": linek,
            "row_bytes": float(alg[two_d].sum()), "line_bytes": float(alg[~two_d].sum()),
            "work_2d": float((sizes * (h1 - h0))[two_d].sum()),
            "jobs": jobs, "ids": ids, "batch": B, "points": int(sizes.sum()), "ticks_total": int(B.state(ids[0])[2]),
            "graphs": graphs, "eager_ms_per_tick": eager_ms, "tuning": tuning, "tune_s": tune_s, "lds": lds}


TUNING_FILES = {}   # precision -> tuning file written by the batched run (read by the serving run)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def batched_extra(ctx, steps, prof_ticks):
    """The 512-job batch (seed 2026) in every precision: throughput, the row
    kernel's roofline, FP64-relative error per job, and the oracle on the host
    over one step of a sample of the jobs as CPU baseline."""
    import tempfile
    out = {"metric": "grid-pt*Hpsi/s, 512 laser-driven jobs (256 1D, 256 2D) per tick", "unit": UNIT,
           "precisions": {}}
    res = {}
    for prec in PRECISIONS:
        r = run_batched(ctx, prec, steps, prof_ticks, tune=True)
        res[prec] = r
        fd, path = tempfile.mkstemp(prefix=f"qdx_tuning_{prec}_", suffix=".txt")
        os.close(fd)
        r["batch"].save_tuning(path)
        TUNING_FILES[prec] = path
        n, ms = r["row"]
        nc, msc = r["col"]
        nl, msl = r["line"]
        ach = r["row_bytes"] / (ms * 1e-3) / 1e9
        value = r["work"] / r["seconds"]
        out["precisions"][prec] = {
            "value": value, "unit": UNIT, "ms_per_tick": r["seconds"] * 1e3 / steps,
            "effective_GBs": B_BATCH[prec] * value / 1e9,
            "effective_frac": B_BATCH[prec] * value / 1e9 / HBM_PEAK_GBS,
            "value_2d_jobs": r["work_2d"] / r["seconds"],
            "tick_graphs": {"captured": r["graphs"][0], "replays": r["graphs"][1],
                            "ms_per_tick_kernel_by_kernel": r["eager_ms_per_tick"]},
            "autotune": {"seconds": r["tune_s"], "lds_schedule": r["lds"],
                         "shapes": {k: {"radix": v["radix0"], "tiling": v["tiling"], "row_us": v["row_us"],
                                        "col_us": v["col_us"]} for k, v in r["tuning"].items()}},
            "roofline": {"bound": "hbm", "kernel": "k_brow (2D jobs)", "achieved": ach, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "launches": n,
                         "avg_launch_us": ms / max(n, 1) * 1e3,
                         "bytes_model": f"{B_BATCH[prec]} B per grid point per executed order"},
            "col_kernel": {"launches": nc, "avg_launch_us": msc / max(nc, 1) * 1e3},
            "line_kernel": {"launches": nl, "avg_launch_us": msl / max(nl, 1) * 1e3, "jobs": "1D",
                            "note": "whole step per block with the line in registers"},
        }
        log(f"batched {prec}: {value:.4g} grid-pt*Hpsi/s, {r['seconds'] * 1e3 / steps:.2f} ms/tick, "
            f"row kernel {ach:.0f} GB/s")
    ref = [res["fp64"]["batch"].fetch(j) for j in res["fp64"]["ids"]]
    jobs = res["fp64"]["jobs"]
    full = full_series_finals(ctx, jobs, res["fp64"]["ticks_total"])
    for prec in ("compact", "fp32"):
        b = res[prec]["batch"]
        got = [b.fetch(j) for j in res[prec]["ids"]]
        errs = [rel_l2(g, a) for g, a in zip(got, ref)]
        # against the FP64 run of the same series rule: the whole series for
        # the jobs this precision runs at fixed length, QDng's break test for
        # the others (compact 1D jobs are the FP64 line kernel)
        errs_same = [rel_l2(g, f if fixed_series(prec, p.nd) else a) for g, a, f, p in zip(got, ref, full, jobs)]
        w = int(np.argmax(errs))
        ws = int(np.argmax(errs_same))
        out["precisions"][prec].update({
            "vs_fp64": out["precisions"][prec]["value"] / out["precisions"]["fp64"]["value"],
            "l2_vs_fp64_worst": float(errs[w]), "l2_worst_job": jobs[w].N,
            "l2_vs_fp64_median": float(np.median(errs)),
            "l2_vs_fp64_same_series_worst": float(errs_same[ws]), "l2_same_series_worst_job": jobs[ws].N,
            "l2_vs_fp64_same_series_median": float(np.median(errs_same)),
            "l2_vs_fp64_same_series_above_1e-6": int(np.sum(np.array(errs_same) > 1e-6))})
    ebreak = [rel_l2(a, b) for a, b in zip(ref, full)]
    out["fp64_break_test_truncation"] = {
        "worst": float(np.max(ebreak)), "median": float(np.median(ebreak)), "worst_job": jobs[int(np.argmax(ebreak))].N,
        "note": "relative L2 distance of the FP64 run (QDng's norm break test, prec 1e-10) from the FP64 run of "
                "the whole series (prec 0): the break test can end a series while it is still this far from "
                "converged; the compact and FP32 2D jobs run a fixed-length series (coefficients down to 1e-10), "
                "so l2_vs_fp64 carries this truncation; l2_vs_fp64_same_series compares each job with the FP64 run "
                "of its own series rule (whole series for the fixed-length jobs, the break test for the others)"}
    out["value"] = out["precisions"]["fp64"]["value"]
    out["roofline"] = out["precisions"]["fp64"]["roofline"]
    out["workload"] = {"jobs": 512, "points": res["fp64"]["points"], "dt": 1.0, "timed_ticks": steps,
                       "profiled_ticks": prof_ticks, "steps_compared": res["fp64"]["ticks_total"]}
    out["cpu_baseline"] = cpu_batched(jobs[::4])
    for r in res.values():
        r["batch"] = None
    return out


def fixed_series(prec, nd):
    """the job runs a fixed-length series (no norm break test) in this
    precision: every FP32 job, the compact 2D jobs"""
    return prec == "fp32" or (prec == "compact" and nd == 2)


def full_series_finals(ctx, jobs, ticks):
    """Final states of the jobs propagated in FP64 with the whole Chebyshev
    series (prec 0, no break test): the accuracy reference."""
    import qdng_amd as Q
    B = Q.Batch(ctx, "fp64", 16 << 30)
    ids = [B.submit(p.gpu_ham(None), 1.0, ticks, p.psi0, prec=0.0) for p in jobs]
    B.run(ticks)
    sync(ctx)
    return [B.fetch(j) for j in ids]


def serve_requests(lam, n_jobs, seed=7):
    """Poisson arrivals (rate lam, seed 7): 90 % laser-driven Chebyshev jobs of
    the batched distribution (half 1D, half 2D), 10 % two-state 256 x 256 GSPO
    jobs (coupled Morse / harmonic surfaces); steps ~ U[200, 5000], dt = 1."""
    import cases
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_jobs))
    reqs, meta = [], []
    for i in range(n_jobs):
        steps = int(rng.integers(200, 5001))
        if rng.uniform() < 0.1:
            p = cases.morse_2state_2d(256, seed=int(rng.integers(1 << 30)))
            ops = [p.gpu_split_op(None, ["kin"]), p.gpu_split_op(None, ["pot"])]
            reqs.append(dict(dt=1.0, nsteps=steps, psi0=p.psi0, ops=ops))
            meta.append(("gspo", p.size * p.S, p.nd))
        else:
            if rng.uniform() < 0.5:
                p = cases.driven_job([int(rng.choice(cases.SIZES_1D_BATCH))], rng, steps)
            else:
                n = int(rng.choice(cases.SIZES_2D_BATCH))
                p = cases.driven_job([n, n], rng, steps)
            reqs.append(dict(H=p.gpu_ham(None), dt=1.0, nsteps=steps, psi0=p.psi0))
            meta.append(("cheby", p.size, p.nd))
    return arrivals, reqs, meta


def serving_extra(ctx, lams, n_jobs):
    import qdng_amd as Q
    out = {"metric": "grid-pt*Hpsi/s (GSPO: grid-pt*step) and p50/p99 job latency of a Poisson stream",
           "mix": "90% batched-distribution Chebyshev jobs, 10% two-state 256x256 GSPO jobs, "
                  "steps U[200, 5000], dt 1, Poisson arrivals (seed 7)", "jobs_per_rate": n_jobs, "rates": {}}
    full = None
    for lam in lams:
        arrivals, reqs, meta = serve_requests(lam, n_jobs)
        pts = np.array([m[1] for m in meta], dtype=float)
        r = {}
        finals = {}
        if full is None:
            # the jobs (seed 7) are the same at every rate: one FP64 run of the
            # whole Chebyshev series (prec 0) is the accuracy reference
            B = Q.Batch(ctx, "fp64", 32 << 30)
            freqs = [dict(q, prec=0.0) if "H" in q else q for q in reqs]
            full = B.serve(freqs, arrivals, want_psi=True)[3]
            B = None
        for prec in PRECISIONS:
            B = Q.Batch(ctx, prec, 32 << 30)
            if prec in TUNING_FILES:
                B.load_tuning(TUNING_FILES[prec])   # the batched run's schedules
            t0 = time.perf_counter()
            admit, finish, nh, psis = B.serve(reqs, arrivals, want_psi=True)
            wall = time.perf_counter() - t0
            lat = finish - arrivals
            span = finish.max() - arrivals.min()
            r[prec] = {"throughput": float((pts * nh).sum() / span), "unit": "grid-pt*Hpsi/s (GSPO: grid-pt*step)",
                       "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                       "mean_queue_wait_s": float((admit - arrivals).mean()), "span_s": float(span),
                       "wall_s": wall}
            finals[prec] = psis
            log(f"serve lam={lam} {prec}: {r[prec]['throughput']:.3g}, p50 {r[prec]['p50_latency_s']:.3f} s, "
                f"p99 {r[prec]['p99_latency_s']:.3f} s")
        for prec in ("compact", "fp32"):
            errs = [rel_l2(a, b) for a, b in zip(finals[prec], finals["fp64"])]
            r[prec]["l2_vs_fp64_median"] = float(np.median(errs))
            r[prec]["l2_vs_fp64_worst"] = float(np.max(errs))
            # against the FP64 run of the same series rule (see batched_extra)
            errs = [rel_l2(a, f if (m[0] == "cheby" and fixed_series(prec, m[2])) else b)
                    for a, b, f, m in zip(finals[prec], finals["fp64"], full, meta)]
            r[prec]["l2_vs_fp64_same_series_median"] = float(np.median(errs))
            r[prec]["l2_vs_fp64_same_series_worst"] = float(np.max(errs))
            r[prec]["l2_vs_fp64_same_series_above_1e-6"] = int(np.sum(np.array(errs) > 1e-6))
            r[prec]["vs_fp64"] = r[prec]["throughput"] / r["fp64"]["throughput"]
        out["rates"][str(lam)] = r
    return out


# ---------------------------------------------------------------------------
# CPU baselines: the FP64 oracle (QDng's algorithm) with the FFTW3 API through
# MKL where it loads (QDng calls FFTW), on the host cores and on one thread

def host_threads():
    return min(os.cpu_count() or 1, 16)


def cpu_fft_backend():
    from oracle import oracle as O
    return O.set_fft_backend("fftw")


def timed_median(run, repeats):
    """run() -> (seconds, work); median of the per-repeat rates after one
    untimed warm-up (plans, first touch)."""
    run()
    rates, secs = [], 0.0
    for _ in range(repeats):
        t, work = run()
        rates.append(work / t)
        secs += t
    return float(np.median(rates)), secs


def cpu_cheby(p, dt, steps, repeats, threads, prec=1e-10):
    """Chebyshev steps of problem p on the oracle: (rate, seconds, H applications)."""
    from oracle import oracle as O
    O.set_threads(threads)
    U = O.Cheby(p.oracle_ham(), dt, prec=prec)
    psi = p.psi0.copy()
    state = {"step": 0, "nh": 0}

    def run():
        t0 = time.perf_counter()
        nh = 0
        for _ in range(steps):
            nh += U.step(psi, state["step"])
            state["step"] += 1
        state["nh"] = nh
        return time.perf_counter() - t0, p.size * p.S * nh
    rate, secs = timed_median(run, repeats)
    return rate, secs, state["nh"]


def cpu_record(rate, rate1, backend, cores, sample, unit=UNIT):
    return {"value": rate, "unit": unit, "cores": cores, "kind": "port",
            "fft": "FFTW3 API (MKL libmkl_rt.so)" if backend == "fftw" else "oracle's own FFT (FFTW3/MKL absent)",
            "value_1_thread": rate1, "sample": sample}


def cpu_cheby_2d(p):
    backend = cpu_fft_backend()
    th = host_threads()
    rate, secs, nh = cpu_cheby(p, 5.0, 2, 5, th)
    rate1, secs1, _ = cpu_cheby(p, 5.0, 1, 3, 1)
    from oracle import oracle as O
    O.set_threads(th)
    return cpu_record(rate, rate1, backend, th,
                      f"Chebyshev steps of the 512x512 two-state workload on the FP64 oracle: median of 5 x 2 steps "
                      f"on {th} threads ({secs:.1f} s), of 3 x 1 step on 1 thread ({secs1:.1f} s), after one warm-up "
                      f"step each; {nh} H applications per 2 steps")


def cpu_gspo(p):
    from oracle import oracle as O
    backend = cpu_fft_backend()
    res = {}
    for th, steps, rep in ((host_threads(), 10, 5), (1, 4, 3)):
        O.set_threads(th)
        G = O.GSPO([p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])], 5.0, p.S)
        psi = p.psi0.copy()

        def run():
            t0 = time.perf_counter()
            for i in range(steps):
                G.step(psi, i)
            return time.perf_counter() - t0, p.size * p.S * steps
        res[th] = timed_median(run, rep)
    O.set_threads(host_threads())
    th = host_threads()
    return cpu_record(res[th][0], res[1][0], backend, th,
                      f"GSPO steps of the same workload on the FP64 oracle: median of 5 x 10 steps on {th} threads, "
                      f"of 3 x 4 steps on 1 thread", unit="grid-pt*step/s")


def cpu_batched(jobs):
    """One Chebyshev step of each job of a sample, one after the other (the
    batch has no CPU counterpart: QDng runs jobs one at a time)."""
    from oracle import oracle as O
    backend = cpu_fft_backend()
    th = host_threads()
    O.set_threads(th)
    t = 0.0
    work = 0
    for p in jobs:
        U = O.Cheby(p.oracle_ham(), 1.0, prec=1e-10)
        psi = p.psi0.copy()
        U.step(psi, 0)
        t0 = time.perf_counter()
        nh = U.step(psi, 1)
        t += time.perf_counter() - t0
        work += p.size * nh
    return cpu_record(work / t, None, backend, th,
                      f"second Chebyshev step of every 4th job of the 512-job batch ({len(jobs)} jobs) on the FP64 "
                      f"oracle, jobs one after the other on {th} threads, {t:.1f} s")


def precision_3d(ctx, steps=5):
    """The 3D 128^3 precision workload (anharmonic oscillator on [-6, 6]^3,
    complex64 Gaussian) at dt 3.5 (order ~64) and 37 (order ~256): each path's
    throughput, the row kernel's roofline and the relative L2 to FP64 after
    `steps` steps from the same state; the oracle on the host at dt 3.5."""
    import cases
    import qdng_amd as Q
    p = cases.anharmonic_3d(128)
    p.psi0 = p.psi0.astype(np.complex64).astype(np.complex128)
    npts = p.size
    out = {"metric": "grid-pt*Hpsi/s and relative L2 to FP64, 3D 128^3 single state", "unit": UNIT, "dt": {}}
    for dt in (3.5, 37.0):
        rec = {}
        ref = None
        for prec in PRECISIONS:
            U = Q.Cheby(p.gpu_ham(ctx), dt, prec=1e-10, precision=prec)
            w = p.gpu_wf(ctx)
            sync(ctx)
            t0 = time.perf_counter()
            _, _, nh, _ = U.run(w, 0, steps, observables=False)
            sync(ctx)
            t = time.perf_counter() - t0
            hpsi = int(nh.sum())
            v = w.download()
            ctx.profile(True)
            w2 = p.gpu_wf(ctx)
            _, _, nh2, _ = U.run(w2, 0, 1, observables=False)
            sync(ctx)
            nrow, msrow = ctx.kernel_time(Q.K_CHEBY_ROW)
            ctx.profile(False)
            h2 = int(nh2.sum())
            cpl = 4 if prec == "fp32" else 8   # one state: no coupling field
            rb = npts * ((B_ROW[prec] - cpl) * (h2 - 1) + (B_ROW_PRO[prec] - cpl))
            ach = rb / (msrow * 1e-3) / 1e9 if msrow > 0 else 0.0
            rec[prec] = {"value": npts * hpsi / t, "unit": UNIT, "ms_per_step": t * 1e3 / steps,
                         "hpsi_per_step": [int(nh.min()), int(nh.max())], "order": U.order,
                         "roofline": {"bound": "hbm", "kernel": "k_cheby_row", "achieved": ach,
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                                      "avg_launch_us": msrow / max(nrow, 1) * 1e3}}
            if prec == "fp64":
                ref = v
            else:
                rec[prec]["l2_vs_fp64"] = rel_l2(v, ref)
                rec[prec]["vs_fp64"] = rec[prec]["value"] / rec["fp64"]["value"]
            log(f"3D 128^3 dt={dt} {prec}: {rec[prec]['value']:.4g} grid-pt*Hpsi/s"
                + (f", L2 vs fp64 {rec[prec]['l2_vs_fp64']:.2e}" if prec != "fp64" else ""))
        out["dt"][str(dt)] = rec
    out["value"] = out["dt"]["3.5"]["fp64"]["value"]
    out["roofline"] = out["dt"]["3.5"]["fp64"]["roofline"]
    backend = cpu_fft_backend()
    th = host_threads()
    rate, secs, nh = cpu_cheby(p, 3.5, 1, 3, th)
    out["cpu_baseline"] = cpu_record(rate, None, backend, th,
                                     f"Chebyshev steps at dt 3.5 on the FP64 oracle: median of 3 x 1 step on {th} "
                                     f"threads after one warm-up step ({secs:.1f} s, {nh} H applications per step)")
    out["workload"] = {"grid": [128, 128, 128], "box": [-6.0, 6.0], "mass": 1836.0, "steps": steps}
    return out


def cheby_1d_small(ctx, steps=1000):
    """The small 1D case (256 points on [-8, 8], m = 1836, V = 0.05 x^2,
    Gaussian psi0 at x = 1, dt = 10, prec = 1e-12, 1000 steps): a latency-bound
    workload (a few kernels of 256 points per order), reported but not judged
    on the roofline; parity of the final psi, the norm and the
    autocorrelation series against the oracle."""
    import cases
    import qdng_amd as Q
    from oracle import oracle as O
    p = cases.harmonic_1d(256, L=8.0, mass=1836.0, k=0.1, x0=1.0, sigma=0.5)
    U = Q.Cheby(p.gpu_ham(ctx), 10.0, prec=1e-12)
    w, w0 = p.gpu_wf(ctx), p.gpu_wf(ctx)
    sync(ctx)
    t0 = time.perf_counter()
    n2, ac, nh, _ = U.run(w, 0, steps, w0)
    sync(ctx)
    gpu_s = time.perf_counter() - t0
    O.set_threads(1)   # 256 points: threads would only add overhead
    Uo = O.Cheby(p.oracle_ham(), 10.0, prec=1e-12)
    g = p.oracle_grid()
    psi = p.psi0.copy()
    n2o, aco = np.zeros(steps), np.zeros(steps, dtype=complex)
    cpu_s = 0.0
    nho = 0
    for i in range(steps):
        n2o[i] = O.norm2(g, psi)
        aco[i] = O.overlap(g, psi, p.psi0)
        t1 = time.perf_counter()
        nho += Uo.step(psi, i)
        cpu_s += time.perf_counter() - t1
    O.set_threads(min(os.cpu_count() or 1, 16))
    gv = w.download()
    return {"value": 256 * int(nh.sum()) / gpu_s, "unit": UNIT, "ms_per_step": gpu_s * 1e3 / steps,
            "hpsi_per_step": [int(nh.min()), int(nh.max())], "hpsi_oracle": int(nho), "hpsi_gpu": int(nh.sum()),
            "cpu_value": 256 * nho / cpu_s, "cpu_threads": 1,
            "l2_vs_oracle": float(np.linalg.norm(gv - psi) / np.linalg.norm(psi)),
            "norm_drift_max": float(np.abs(n2 - n2[0]).max()),
            "autocorr_max_abs_err": float(np.abs(ac - aco).max()),
            "note": "latency bound (the whole step runs in one block, 4 points per thread: one wave "
                    "holds the line): not judged on the roofline"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=10, help="oracle steps timed for cpu_baseline")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=100, help="timed ticks of the batched-jobs workload")
    ap.add_argument("--serve-jobs", type=int, default=96, help="requests per arrival rate of the serving run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB device buffer 4 times (reference for FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--serve-rates", default="50,500,2000", help="Poisson arrival rates (jobs/s), comma separated")
    ap.add_argument("--batched-only", choices=["fp64", "fp32", "compact"],
                    help="only run the batched-jobs workload in this precision (default schedules, "
                         "--batch-steps ticks) and print its tick time: a short program for counter runs")
    ap.add_argument("--gspo-only", action="store_true",
                    help="only run the GSPO leg (FP64 and FP32, --steps steps after --warmup) and print its step "
                         "time: a short program for traces")
    ap.add_argument("--precision-3d-only", action="store_true",
                    help="only run the 3D 128^3 precision workload (FP64, compact, FP32 at dt 3.5 and 37) and print "
                         "its record: a short program for traces")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    import qdng_amd as Q
    ctx = Q.Context(0)
    if args.calibrate_copy:
        a = Q.WaveFunction(ctx, [4096, 4096], [-1.0, -1.0], [1.0, 1.0], 1)
        b = Q.WaveFunction(ctx, [4096, 4096], [-1.0, -1.0], [1.0, 1.0], 1)
        for _ in range(4):
            b.copy_from(a)
        log("copied 4 x 256 MiB device to device")
        return
    if args.batched_only:
        r = run_batched(ctx, args.batched_only, args.batch_steps, 1)
        log(f"batched {args.batched_only}: {r['seconds'] * 1e3 / args.batch_steps:.3f} ms/tick "
            f"over {args.batch_steps} ticks")
        return
    if args.precision_3d_only:
        rec = precision_3d(ctx)
        for dt, r in rec["dt"].items():
            for prec in PRECISIONS:
                log(f"3D dt {dt} {prec}: {r[prec]['value']:.4g} grid-pt*Hpsi/s, "
                    f"row kernel {r[prec]['roofline']['avg_launch_us']:.2f} us ({r[prec]['roofline']['frac']:.3f})")
        return
    p = build_problem(512)
    npts = p.size * p.S
    if args.gspo_only:
        for prec in ("fp64", "fp32"):
            t = run_gspo(ctx, p, prec, args.steps, args.warmup)
            log(f"gspo {prec}: {npts * args.steps / t:.4g} grid-pt*step/s, {t * 1e6 / args.steps:.2f} us/step")
        return
    log(f"workload 512x512 x 2 states, {args.warmup} warmup + {args.steps} timed steps")

    # ---- headline: FP64 Chebyshev.  Pass 1 times the steps; pass 2 repeats
    # them with HIP events on every launch (attached to the dispatch packets)
    # for the per-kernel roofline.
    r = run_cheby(ctx, p, "fp64", args.steps, args.warmup, profile=False)
    nh = r["n_hpsi"]
    hpsi = int(nh.sum())
    value = npts * hpsi / r["seconds"]
    rp = run_cheby(ctx, p, "fp64", args.steps, args.warmup, profile=True)
    assert int(rp["n_hpsi"].sum()) == hpsi
    launches, row_ms = rp["row"]
    row_bytes = p.size * p.S * (B_ROW["fp64"] * (hpsi - args.steps) + B_ROW_PRO["fp64"] * args.steps)
    achieved = row_bytes / (row_ms * 1e-3) / 1e9
    log(f"fp64: {value:.4g} grid-pt*Hpsi/s, {r['seconds'] * 1e3 / args.steps:.3f} ms/step, "
        f"orders/step {nh.min()}-{nh.max()}, row kernel {launches} launches avg {row_ms / launches * 1e3:.2f} us")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "chebyshev_2d_512x512_2state_fp64",
            "grid": [512, 512], "box": [-10.0, 10.0], "states": 2, "masses": [1836.0, 12000.0],
            "dt": 5.0, "prec": 1e-10, "order": r["order"],
            "hpsi_per_step": [int(nh.min()), int(nh.max())],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "traffic_source": "profiles/cheby_row_2d_512x512_2state_fp64_hbm_traffic.md (rocprofv3 FETCH_SIZE x 2.000 "
                              "+ WRITE_SIZE x 1.000, calibrated on a 256 MiB device copy)",
            "kernel": "qdx::k_cheby_row<double, 2, false, false, 256, 4, false, false>",
            "points_per_thread": 4,
            "launches": launches,
            "avg_launch_us": row_ms / launches * 1e3,
            "algorithmic_bytes_per_launch": row_bytes / launches,
            "bytes_model": "96 B per grid point and state per executed order (80 B for order 1); "
                           "early-exit launches after the break count 0 B",
            "measured": "HIP events attached to every launch (hipExtLaunchKernelGGL) over a second "
                        "timed pass of the same steps",
            "ms_per_step_profiled_pass": rp["seconds"] * 1e3 / args.steps,
            "step_frac": B_ROW["fp64"] * value / 1e9 / HBM_PEAK_GBS,
            "step_frac_note": "the whole step (row kernel, column pass, launch gaps) at the same 96 B per grid "
                              "point and state per executed order",
            "column_pass": {"kernel": "qdx::k_col<double, 4, 4, false>", "launches": rp["line"][0],
                            "avg_launch_us": rp["line"][1] / max(rp["line"][0], 1) * 1e3,
                            "algorithmic_bytes_per_launch": 32.0 * p.size * p.S,
                            "note": "phi_{k-1} read + Y written, 16 B each per point and state"},
        },
    }

    # ---- parity against the oracle (its own FFT: the checker) over the first
    # steps, and the CPU baseline (the oracle with the FFTW3 API, timed) ----
    ex = out.setdefault("extra", {})
    if args.cpu_steps > 0:
        from oracle import oracle as O
        O.set_fft_backend("own")
        O.set_threads(host_threads())
        Uo = O.Cheby(p.oracle_ham(), 5.0, prec=1e-10)
        g = p.oracle_grid()
        psi_cpu = p.psi0.copy()
        n2_cpu, ac_cpu = [], []
        for i in range(args.cpu_steps):
            n2_cpu.append(O.norm2(g, psi_cpu, p.S))
            ac_cpu.append(O.overlap(g, psi_cpu, p.psi0, p.S))
            Uo.step(psi_cpu, i)
        n2_cpu, ac_cpu = np.array(n2_cpu), np.array(ac_cpu)

        def parity(precision):
            U = Q.Cheby(p.gpu_ham(ctx), 5.0, prec=1e-10, precision=precision)
            w, w0 = p.gpu_wf(ctx), p.gpu_wf(ctx)
            n2, ac, _, _ = U.run(w, 0, args.cpu_steps, w0)
            return {"l2_vs_oracle": rel_l2(w.download(), psi_cpu),
                    "norm_drift_max": float(np.abs(n2 - n2[0]).max()),
                    "norm_drift_max_oracle": float(np.abs(n2_cpu - n2_cpu[0]).max()),
                    "autocorr_max_abs_err": float(np.abs(ac - ac_cpu).max()),
                    "steps": args.cpu_steps}
        ex["parity"] = {prec: parity(prec) for prec in (PRECISIONS if not args.no_extra else ("fp64",))}
        ex["l2_fp64_vs_oracle"] = ex["parity"]["fp64"]["l2_vs_oracle"]
        out["cpu_baseline"] = cpu_cheby_2d(p)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.4g} grid-pt*Hpsi/s on {out['cpu_baseline']['cores']} "
            f"threads ({out['cpu_baseline']['fft']}), {out['cpu_baseline']['value_1_thread']:.4g} on 1")

    # ---- the benchmark configurations, every path --------------------------
    if not args.no_extra:
        configs = {}
        rec2d = {"metric": "grid-pt*Hpsi/s, 2D 512x512 two states, Chebyshev dt 5", "unit": UNIT,
                "value": value, "roofline": out["roofline"], "cpu_baseline": out.get("cpu_baseline"),
                "precisions": {"fp64": {"value": value, "ms_per_step": out["ms_per_step"],
                                        "hpsi_per_step": out["config"]["hpsi_per_step"]}}}
        for prec in ("compact", "fp32"):
            # timed without events (as the FP64 headline), kernels from a second pass
            rr = run_cheby(ctx, p, prec, args.steps, args.warmup, profile=False)
            hh = int(rr["n_hpsi"].sum())
            rq = run_cheby(ctx, p, prec, args.steps, args.warmup, profile=True)
            ln, lms = rq["row"]
            bb = p.size * p.S * (B_ROW[prec] * (hh - args.steps) + B_ROW_PRO[prec] * args.steps)
            v = npts * hh / rr["seconds"]
            rec2d["precisions"][prec] = {
                "value": v, "unit": UNIT, "ms_per_step": rr["seconds"] * 1e3 / args.steps,
                "hpsi_per_step": [int(rr["n_hpsi"].min()), int(rr["n_hpsi"].max())], "vs_fp64": v / value,
                "l2_vs_oracle": ex.get("parity", {}).get(prec, {}).get("l2_vs_oracle"),
                "roofline": {"bound": "hbm", "kernel": "k_cheby_row", "achieved": bb / (lms * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": bb / (lms * 1e-3) / 1e9 / HBM_PEAK_GBS, "avg_launch_us": lms / ln * 1e3,
                             "bytes_model": f"{B_ROW[prec]} B per grid point and state per executed order"},
                "col_kernel_avg_launch_us": rq["line"][1] / max(rq["line"][0], 1) * 1e3}
            log(f"{prec}: {v:.4g} grid-pt*Hpsi/s")
        configs["cheby_2d_512x512_2state"] = rec2d
        # GSPO on the same workload: at least 100 timed steps
        gsteps = max(100, args.steps)
        gs = {"metric": "grid-pt*step/s, 2D 512x512 two coupled states, GSPO dt 5", "unit": "grid-pt*step/s",
              "precisions": {}}
        for prec in ("fp64", "fp32"):
            t, rowk, colk = run_gspo(ctx, p, prec, gsteps, min(args.warmup, 10), profile=True)
            v = npts * gsteps / t
            eff = B_GSPO[prec] * v / 1e9
            gs["precisions"][prec] = {
                "value": v, "unit": "grid-pt*step/s", "ms_per_step": t * 1e3 / gsteps, "steps": gsteps,
                "roofline": {"bound": "hbm", "kernel": "whole step (row pass + column pass)", "achieved": eff,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": eff / HBM_PEAK_GBS,
                             "bytes_model": f"{B_GSPO[prec]} B per grid point and state per step"},
                "row_kernel_avg_launch_us": rowk[1] / max(rowk[0], 1) * 1e3, "row_launches": rowk[0],
                "col_kernel_avg_launch_us": colk[1] / max(colk[0], 1) * 1e3, "col_launches": colk[0]}
            log(f"gspo {prec}: {v:.4g} grid-pt*step/s, {t * 1e3 / gsteps:.4f} ms/step")
        gs["value"] = gs["precisions"]["fp64"]["value"]
        gs["roofline"] = gs["precisions"]["fp64"]["roofline"]
        gs["cpu_baseline"] = cpu_gspo(p)
        configs["gspo_2d_512x512_2state"] = gs
        configs["precision_3d_128"] = precision_3d(ctx)
        if args.batch_steps > 0:
            configs["batched_512_jobs"] = batched_extra(ctx, args.batch_steps, 5)
        if args.serve_jobs > 0:
            configs["serving"] = serving_extra(ctx, [float(x) for x in args.serve_rates.split(",")], args.serve_jobs)
        ex["configs"] = configs
        ex["cheby_1d_256_latency"] = cheby_1d_small(ctx)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
