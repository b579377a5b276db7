"""Schedules of the batched kernels and the autotuner (qdx_batch_autotune).

A schedule only changes how the device computes a job, not what: the LDS
layout (one or two buffers per line) leaves every bit of the result alone;
the tile shape changes only the grouping of the break-test partial sums; a
radix plan changes the transform's rounding.  The autotuner picks among
them by timing and must leave jobs within the usual parity bars.  QDng's
counterpart is FFTW's measured planning with its wisdom file
(src/fft/FFTGlobal.cpp:41-80).
"""
import numpy as np
import pytest

import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def note(msg):
    print("[gpu case]", msg, flush=True)


SHAPES = [[64, 64], [96, 64], [64, 192], [128, 128], [384, 384]]


def jobs_2d(steps, seed=51):
    rng = np.random.default_rng(seed)
    return [cases.driven_job(s, rng, steps) for s in SHAPES]


def run_batch(ctx, precision, jobs, steps, tilings=None, lds=None):
    import qdng_amd as Q
    B = Q.Batch(ctx, precision, 1 << 30)
    if lds:
        B.set_lds_schedule(*lds)
    for (n0, n1), t in (tilings or {}).items():
        B.set_tiling(n0, n1, **t)
    ids = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in jobs]
    B.run(steps)
    return [B.fetch(j) for j in ids], [B.state(j)[1] for j in ids], B


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lds_schedules_are_bit_identical(ctx, precision):
    steps = 5
    jobs = jobs_2d(steps)
    res = {}
    for lds in [(2, 2), (1, 2), (2, 1), (1, 1)]:
        note(f"LDS buffers per line (rows, columns) = {lds}, {precision}")
        res[lds] = run_batch(ctx, precision, jobs, steps, lds=lds)
    ref, nh_ref, _ = res[(2, 2)]
    for lds, (got, nh, _) in res.items():
        for p, a, b, h, hr in zip(jobs, got, ref, nh, nh_ref):
            assert h == hr
            assert np.array_equal(a, b), (lds, p.N)


def test_lds_schedule_tuning_on_the_live_mix(ctx):
    """qdx_batch_autotune_lds times the schedules on the batch's own jobs; as
    the schedule leaves results alone, the jobs end exactly as without it."""
    import qdng_amd as Q
    steps = 16
    jobs = jobs_2d(steps, seed=56)
    ref, nh_ref, _ = run_batch(ctx, "fp32", jobs, steps)
    B = Q.Batch(ctx, "fp32", 1 << 30)
    ids = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in jobs]
    B.run(1)
    ms = B.autotune_lds()                    # 12 ticks
    note(f"LDS schedule tick times {ms}, chosen {B.lds_schedule()}")
    assert len(ms) == 4 and all(m > 0 for m in ms)
    best = [(2, 2), (1, 2), (2, 1), (1, 1)][int(np.argmin(ms))]
    assert B.lds_schedule() == best
    B.run(steps - 13)
    for jid, a, h in zip(ids, ref, nh_ref):
        assert B.state(jid)[0] == 0
        assert np.array_equal(B.fetch(jid), a)
    with pytest.raises(Q.QdxError) as e:     # nothing left to time on
        B.autotune_lds()
    assert e.value.code == -6


def test_tile_shapes_agree(ctx):
    steps = 5
    jobs = jobs_2d(steps, seed=52)
    note("512-thread row tiles, other column widths, fp64")
    ref, nh_ref, B = run_batch(ctx, "fp64", jobs, steps)
    tilings = {}
    for p in jobs:
        n0, n1 = p.N
        t = B.tiling(n0, n1)
        t["row_threads"] = 512
        t["row_lines"] = min(n1, 512 // (n0 // 8))
        t["col_width"] = max(1, t["col_width"] // 2)
        tilings[(n0, n1)] = t
    got, nh, _ = run_batch(ctx, "fp64", jobs, steps, tilings)
    for a, b, h, hr in zip(got, ref, nh, nh_ref):
        assert h == hr
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", cases.BULK_BAR)])
def test_radix_plans_agree(ctx, precision, tol):
    import qdng_amd as Q
    steps = 4
    jobs = jobs_2d(steps, seed=53)
    note(f"other radix plans, {precision}")
    ref, nh_ref, _ = run_batch(ctx, precision, jobs, steps)
    c2 = Q.Context(0)
    for L, r in [(64, [4, 4, 4]), (96, [4, 3, 8]), (128, [2, 8, 8]), (192, [4, 3, 2, 8]), (384, [8, 3, 2, 8])]:
        c2.set_radix_plan(L, r, precision)
        assert c2.radix_plan(L, precision) == r
    got, nh, _ = run_batch(c2, precision, jobs, steps)
    if precision == "fp32":
        # the two FP32 runs differ by their rounding (other transform
        # orders); each is held to the bar against the FP64 batch
        ref, _, _ = run_batch(ctx, "fp64", jobs, steps)
    errs = []
    for p, a, b, h, hr in zip(jobs, got, ref, nh, nh_ref):
        note(f"  job {p.N} rel {rel(a, b):.2e}")
        if precision == "fp64":
            assert h == hr
            assert rel(a, b) < tol
        errs.append(rel(a, b))
    if precision == "fp32":
        cases.bulk_bar(errs, "fp32", " (other radix plans vs the FP64 batch)")


def test_schedule_arguments_are_checked(ctx):
    import qdng_amd as Q
    B = Q.Batch(ctx, "fp64", 1 << 24)
    bad = [dict(row_threads=128, row_lines=4, col_width=4),
           dict(row_threads=256, row_lines=33, col_width=4),    # 33 * 8 > 256
           dict(row_threads=256, row_lines=32, col_width=3),    # 3 does not divide 64
           dict(row_threads=256, row_lines=32, col_width=64)]   # 64 * 8 > 256
    for t in bad:
        with pytest.raises(Q.QdxError) as e:
            B.set_tiling(64, 64, **t)
        assert e.value.code == -1
    with pytest.raises(Q.QdxError):
        B.set_tiling(100, 64, row_threads=256, row_lines=1, col_width=1)
    for lds in [(0, 2), (2, 3)]:
        with pytest.raises(Q.QdxError) as e:
            B.set_lds_schedule(*lds)
        assert e.value.code == -1
    for L, r in [(64, [8, 4, 4]), (64, [3, 8, 8]), (96, [8, 4, 3]), (96, [8, 12]), (64, [8] * 2 + [1])]:
        with pytest.raises(Q.QdxError) as e:
            ctx.set_radix_plan(L, r)
        assert e.value.code == -1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_installs_a_valid_schedule(ctx, precision, tmp_path):
    import qdng_amd as Q
    steps = 5
    c = Q.Context(0)
    B = Q.Batch(c, precision, 1 << 30)
    res = {}
    for n0, n1 in [(96, 64), (128, 128)]:
        r = B.autotune(n0, n1)
        note(f"autotune {n0}x{n1} {precision}: {r}")
        assert r["candidates"] >= 4
        assert 0 < r["row_us"] <= r["row_us_default"] and 0 < r["col_us"] <= r["col_us_default"]
        assert B.tiling(n0, n1) == r["tiling"]
        assert c.radix_plan(n0, precision) == r["radix0"]
        assert c.radix_plan(n1, precision) == r["radix1"]
        res[(n0, n1)] = r
    # jobs submitted afterwards run the tuned schedule within the parity bars
    rng = np.random.default_rng(54)
    jobs = [cases.driven_job([96, 64], rng, steps), cases.driven_job([128, 128], rng, steps)]
    ids = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in jobs]
    B.run(steps)
    fp32_errs = []
    for p, jid in zip(jobs, ids):
        got = B.fetch(jid)
        if precision == "fp64":
            U = O.Cheby(p.oracle_ham(), 1.0, prec=1e-10)
            psi = p.psi0.copy()
            nh = sum(U.step(psi, i) for i in range(steps))
            assert B.state(jid)[1] == nh
            assert rel(got, psi) < 1e-10
        else:
            U = Q.Cheby(p.gpu_ham(ctx), 1.0, prec=1e-10, precision="fp64")
            w = p.gpu_wf(ctx)
            U.run(w, 0, steps, observables=False)
            fp32_errs.append(rel(got, w.download()))
    # the tuning file restores the schedule in a fresh context
    path = str(tmp_path / "tuning.txt")
    B.save_tuning(path)
    c3 = Q.Context(0)
    B3 = Q.Batch(c3, precision, 1 << 24)
    B3.load_tuning(path)
    for (n0, n1), r in res.items():
        assert B3.tiling(n0, n1) == r["tiling"]
        assert c3.radix_plan(n0, precision) == r["radix0"]
        assert c3.radix_plan(n1, precision) == r["radix1"]
    assert B3.lds_schedule() == B.lds_schedule()
    bad = tmp_path / "bad.txt"
    bad.write_text("plan 96 8 4 3\n")
    with pytest.raises(Q.QdxError):
        B3.load_tuning(str(bad))
    if precision == "fp32":
        cases.bulk_bar(fp32_errs, "fp32", " (tuned schedule vs the FP64 path)")


def test_autotune_on_submit_keeps_results_mix_independent(ctx, tmp_path):
    """With autotune on, the first job of a shape triggers its tuning; jobs
    keep the plans they were submitted with, so a job's result does not
    depend on what was tuned while it ran."""
    import qdng_amd as Q
    steps = 6
    rng = np.random.default_rng(55)
    a = cases.driven_job([64, 192], rng, steps + 4)
    b = cases.driven_job([192, 192], rng, steps)
    c1 = Q.Context(0)
    A = Q.Batch(c1, "fp64", 1 << 30)
    A.set_autotune(True)
    ja = A.submit(a.gpu_ham(None), 1.0, steps + 4, a.psi0)   # tunes 64 x 192
    path = str(tmp_path / "a.txt")
    A.save_tuning(path)
    A.run(2)
    A.submit(b.gpu_ham(None), 1.0, steps, b.psi0)            # tunes 192 x 192 while a runs
    A.run(steps + 2)
    mixed = A.fetch(ja)
    # alone, with the schedule a was tuned to (tuning itself depends on timing)
    c2 = Q.Context(0)
    S = Q.Batch(c2, "fp64", 1 << 30)
    S.load_tuning(path)
    S.set_autotune(True)                                     # 64 x 192 is known: no new tuning
    js = S.submit(a.gpu_ham(None), 1.0, steps + 4, a.psi0)
    S.run(steps + 4)
    assert np.array_equal(mixed, S.fetch(js))
