"""Batched independent Chebyshev jobs (qdx_batch_*).

Each job of a batch is one QDng propagation (OCheby real-time steps of its
own Hamiltonian, ProgPropa.cpp:317-326): its result must equal the
single-propagation path's and, in FP64, the oracle's; it must not depend on
the other jobs in the batch or on the tick it started at; jobs are admitted
only while they fit into the batch's HBM budget.
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


def mixed_jobs(steps, seed=5):
    """1D lines on both row-tile block sizes (<= 2048 and 3072 / 4096 points),
    2D jobs with square and non-square grids, all laser driven."""
    rng = np.random.default_rng(seed)
    shapes = [[128], [384], [3072], [4096], [64, 64], [96, 64], [64, 192], [128, 128]]
    return [cases.driven_job(s, rng, steps) for s in shapes]


def single_path(ctx, p, dt, steps, precision):
    import qdng_amd as Q
    U = Q.Cheby(p.gpu_ham(ctx), dt, prec=1e-10, precision=precision)
    w = p.gpu_wf(ctx)
    _, _, nh, _ = U.run(w, 0, steps, observables=False)
    return w.download(), int(nh.sum())


@pytest.mark.parametrize("precision", ["fp64", "fp32", "compact"])
def test_batch_matches_single_path(ctx, precision):
    import qdng_amd as Q
    steps = 12
    jobs = mixed_jobs(steps)
    note(f"batch vs single path, {precision}, {len(jobs)} jobs")
    B = Q.Batch(ctx, precision, 1 << 30)
    ids = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in jobs]
    B.run(steps)
    # FP32: the 1D jobs run whole steps in registers (psi updated every
    # order), the single path updates psi every third order from the stored
    # FP32 terms, so the two FP32 results differ by their rounding; each is
    # held to the 1e-6 bar against the FP64 single path instead (the
    # 3072-point job's 58-term series over 12 steps misses it, ~5e-6: FP32
    # arithmetic; the compact storage meets it, test_compact_batch_vs_fp64).
    # Compact: the same FP64 arithmetic with the terms rounded to 40 bits in
    # both paths, 1D jobs exactly FP64
    tol = {"fp64": 1e-12, "compact": 1e-8}.get(precision)
    errs = []
    for p, jid in zip(jobs, ids):
        left, nh, done = B.state(jid)
        got = B.fetch(jid)
        ref, nh_ref = single_path(ctx, p, 1.0, steps, "fp64" if precision == "fp32" else precision)
        note(f"  job N={p.N} nhpsi {nh}/{nh_ref} rel {rel(got, ref):.2e}")
        assert left == 0 and done == steps
        if precision != "fp32":
            assert nh == nh_ref
            assert rel(got, ref) < tol
        errs.append(rel(got, ref))
    if precision == "fp32":
        cases.bulk_bar(errs, "fp32", f" (batch vs FP64, worst of {len(errs)} jobs)")


def test_compact_batch_vs_fp64(ctx):
    # the compact batch (FP64 arithmetic, 40-bit terms) against the FP64
    # batch on every job of the mix: within the 1e-6 relative-L2 bar
    import qdng_amd as Q
    steps = 30
    jobs = mixed_jobs(steps, seed=17)
    note("compact batch vs fp64 batch")
    res = {}
    for prec in ("fp64", "compact"):
        B = Q.Batch(ctx, prec, 1 << 30)
        ids = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in jobs]
        B.run(steps)
        res[prec] = [B.fetch(j) for j in ids]
    for p, a, b in zip(jobs, res["compact"], res["fp64"]):
        e = rel(a, b)
        note(f"  job N={p.N} compact vs fp64 {e:.2e}")
        assert e < 1e-6


def test_batch_fp64_matches_oracle(ctx):
    import qdng_amd as Q
    steps = 6
    jobs = mixed_jobs(steps, seed=11)
    pick = [jobs[1], jobs[5]]                    # 1D N=384 and 2D 96 x 64
    note("batch fp64 vs oracle")
    B = Q.Batch(ctx, "fp64", 1 << 30)
    ids = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in pick]
    B.run(steps)
    for p, jid in zip(pick, ids):
        U = O.Cheby(p.oracle_ham(), 1.0, prec=1e-10)
        psi = p.psi0.copy()
        nh = sum(U.step(psi, i) for i in range(steps))
        _, nh_b, _ = B.state(jid)
        assert nh_b == nh
        assert rel(B.fetch(jid), psi) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32", "compact"])
def test_job_result_same_alone_and_in_mix(ctx, precision):
    import qdng_amd as Q
    steps = 10
    jobs = mixed_jobs(steps + 20, seed=3)
    target = jobs[6]                              # 2D 64 x 192
    note(f"alone vs mix, {precision}")
    A = Q.Batch(ctx, precision, 1 << 30)
    a = A.submit(target.gpu_ham(None), 1.0, steps, target.psi0)
    A.run(steps)
    alone = A.fetch(a)
    _, nh_alone, _ = A.state(a)
    M = Q.Batch(ctx, precision, 1 << 30)
    others = [M.submit(p.gpu_ham(None), 1.0, steps + 20, p.psi0) for p in jobs if p is not target]
    M.run(3)                                      # the target joins a running batch
    t = M.submit(target.gpu_ham(None), 1.0, steps, target.psi0)
    M.run(4)
    M.release(others[0])                          # the mix changes while it runs
    M.run(steps)
    mixed = M.fetch(t)
    _, nh_mix, done = M.state(t)
    assert done == steps
    assert nh_mix == nh_alone
    assert np.array_equal(mixed, alone)


def test_admission_by_footprint(ctx):
    import qdng_amd as Q
    steps = 4
    jobs = mixed_jobs(steps, seed=9)
    note("admission by footprint")
    sizes = []
    probe = Q.Batch(ctx, "fp32", 1 << 30)
    for p in jobs:
        sizes.append(probe.footprint(p.gpu_ham(None), 1.0, steps, p.psi0))
    budget = sizes[4] + sizes[5] + 4096
    B = Q.Batch(ctx, "fp32", budget)
    a = B.try_submit(jobs[4].gpu_ham(None), 1.0, steps, jobs[4].psi0)
    b = B.try_submit(jobs[5].gpu_ham(None), 1.0, steps, jobs[5].psi0)
    assert a is not None and b is not None
    assert B.try_submit(jobs[7].gpu_ham(None), 1.0, steps, jobs[7].psi0) is None   # 128 x 128 does not fit
    n, used, _ = B.usage()
    assert n == 2 and used <= budget
    B.run(steps)
    B.release(a)
    B.release(b)
    c = B.try_submit(jobs[4].gpu_ham(None), 1.0, steps, jobs[4].psi0)
    assert c is not None
    B.run(steps)
    assert B.state(c)[2] == steps


def test_batch_rejects_unsupported_jobs(ctx):
    import qdng_amd as Q
    p = cases.morse_2state_2d(64)
    B = Q.Batch(ctx, "fp64", 1 << 26)
    with pytest.raises(Q.QdxError) as e:
        B.submit(p.gpu_ham(None), 1.0, 2, p.psi0)
    assert e.value.code == -2
    q = cases.laser_1d(128, nt=5)
    with pytest.raises(Q.QdxError) as e:
        B.submit(q.gpu_ham(None), 1.0, 10, q.psi0)   # laser shorter than the run
    assert e.value.code == -1


# ---------------------------------------------------------------------------
# GSPO jobs and the serve loop

def gspo_jobs(steps, seed=21):
    """two coupled states on a 2D grid, one laser-driven 1D state, one 2D state"""
    rng = np.random.default_rng(seed)
    a = cases.morse_2state_2d(64)
    b = cases.driven_job([256], rng, steps)
    c = cases.driven_job([96, 64], rng, steps)
    return [a, b, c]


def split_ops(p, ctx):
    return [p.gpu_split_op(ctx, ["kin"]), p.gpu_split_op(ctx, ["pot", "dip"])]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_gspo_matches_single_path(ctx, precision):
    import qdng_amd as Q
    steps = 9
    jobs = gspo_jobs(steps)
    note(f"batched GSPO vs single path, {precision}")
    B = Q.Batch(ctx, precision, 1 << 30)
    ids = [B.submit(dt=1.0, nsteps=steps, psi0=p.psi0, ops=split_ops(p, None)) for p in jobs]
    B.run(steps)
    tol = 1e-12 if precision == "fp64" else 1e-6
    for p, jid in zip(jobs, ids):
        G = Q.GSPO(split_ops(p, ctx), 1.0, precision=precision)
        w = p.gpu_wf(ctx)
        G.run(w, 0, steps, observables=False)
        ref = w.download()
        got = B.fetch(jid)
        left, nh, done = B.state(jid)
        note(f"  job N={p.N} S={p.S} rel {rel(got, ref):.2e}")
        assert left == 0 and done == steps and nh == steps
        assert rel(got, ref) < tol


def test_batch_gspo_fp64_matches_oracle(ctx):
    import qdng_amd as Q
    steps = 5
    p = cases.morse_2state_2d(48)
    note("batched GSPO fp64 vs oracle, coupled 2D")
    B = Q.Batch(ctx, "fp64", 1 << 28)
    jid = B.submit(dt=2.0, nsteps=steps, psi0=p.psi0, ops=split_ops(p, None))
    B.run(steps)
    Go = O.GSPO([p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])], 2.0, p.S)
    psi = p.psi0.copy()
    for i in range(steps):
        Go.step(psi, i)
    assert rel(B.fetch(jid), psi) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replay_matches_eager_ticks(ctx, precision):
    """Ticks replayed from captured hipGraphs leave the bits that ticks
    launched kernel by kernel leave, with host work between the replays
    (fetches of a running job) and across changes of the job mix."""
    import qdng_amd as Q
    steps = 8
    cheb = mixed_jobs(steps + 6, seed=41)
    gsp = gspo_jobs(steps + 6, seed=42)

    def drive(graphs):
        B = Q.Batch(ctx, precision, 1 << 30)
        B.set_graphs(graphs)
        ids = [B.submit(p.gpu_ham(None), 1.0, steps + 6, p.psi0) for p in cheb[:5]]
        ids += [B.submit(dt=1.0, nsteps=steps + 6, psi0=p.psi0, ops=split_ops(p, None)) for p in gsp]
        snaps = []
        for _ in range(4):
            B.run(1)
            snaps.append(B.fetch(ids[4]))
        late = [B.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in cheb[5:]]
        B.run(3)
        B.release(ids[0])
        B.run(steps + 6)
        out = [B.fetch(j) for j in ids[1:] + late]
        return snaps + out, B.graph_stats()

    note(f"graph replay vs eager ticks, {precision}")
    eager, g0 = drive(False)
    graph, g1 = drive(True)
    note(f"  graphs captured {g1[0]}, replays {g1[1]}")
    assert g0 == (0, 0)
    assert g1[0] >= 2 and g1[1] > g1[0]
    for a, b in zip(eager, graph):
        assert np.array_equal(a, b)


def serve_mix(steps_lo=4, steps_hi=14, seed=31):
    rng = np.random.default_rng(seed)
    shapes = [[128], [64, 64], [384], [96, 64], [4096], [64, 128], [192], [128, 128]]
    reqs = []
    for i, sh in enumerate(shapes):
        n = int(rng.integers(steps_lo, steps_hi))
        p = cases.driven_job(sh, rng, n)
        reqs.append(("cheby", p, n))
        if i % 3 == 0:
            q = cases.morse_2state_2d(64)
            reqs.append(("gspo", q, n))
    return reqs


def req_kwargs(kind, p, n):
    if kind == "cheby":
        return dict(H=p.gpu_ham(None), dt=1.0, nsteps=n, psi0=p.psi0)
    return dict(dt=1.0, nsteps=n, psi0=p.psi0, ops=split_ops(p, None))


@pytest.mark.parametrize("precision", ["fp64", "fp32", "compact"])
def test_serve_results_same_alone_and_in_mix(ctx, precision):
    import qdng_amd as Q
    reqs = serve_mix()
    note(f"serve loop, {len(reqs)} requests, {precision}")
    arrivals = np.cumsum(np.full(len(reqs), 2e-3))           # overlapping lifetimes
    S = Q.Batch(ctx, precision, 1 << 30)
    admit, finish, nh, psis = S.serve([req_kwargs(*r) for r in reqs], arrivals, want_psi=True)
    assert np.all(admit >= arrivals - 1e-9)
    assert np.all(finish >= admit - 1e-3)
    assert S.usage()[0] == 0 and S.usage()[1] == 0           # all memory returned
    for (kind, p, n), got, h in zip(reqs, psis, nh):
        A = Q.Batch(ctx, precision, 1 << 28)
        a = A.submit(**req_kwargs(kind, p, n))
        A.run(n)
        alone = A.fetch(a)
        assert A.state(a)[1] == h
        assert np.array_equal(got, alone), (kind, p.N)


def test_serve_admission_queues_on_budget(ctx):
    import qdng_amd as Q
    reqs = serve_mix(seed=32)
    probe = Q.Batch(ctx, "fp32", 1 << 28)
    sizes = [probe.footprint(**req_kwargs(*r)) for r in reqs]
    budget = max(sizes) + min(sizes) + 4096                   # at most two jobs at a time
    note(f"serve loop with a budget of {budget} bytes")
    S = Q.Batch(ctx, "fp32", budget)
    arrivals = np.zeros(len(reqs))                            # all at once: most of them wait
    admit, finish, nh = S.serve([req_kwargs(*r) for r in reqs], arrivals)
    assert np.all(np.diff(admit) >= 0)                        # admitted in arrival order
    assert np.all(nh > 0)
    order = np.argsort(finish)
    assert finish[order[-1]] > finish[order[0]]
    with pytest.raises(Q.QdxError) as e:                      # larger than the whole budget
        T = Q.Batch(ctx, "fp32", min(sizes))
        T.serve([req_kwargs(*reqs[4])], [0.0])
    assert e.value.code == -4


# ---------------------------------------------------------------------------
# per-job failure isolation: a job whose state stops being a number stops
# (Reporter.cpp:112-113 "Norm is not a number" ends a QDng propagation); its
# batch-mates' results do not change by a bit

def _poisoned_laser(p, at):
    """p with its dipole field not a number at sample `at`"""
    q = cases.Problem(p.N, p.xmin, p.xmax, p.S, p.multistate, p.unity)
    q.psi0 = p.psi0.copy()
    for t in p.terms:
        if t[0] == "dip":
            las = t[4].copy()
            las[at] = np.nan
            t = t[:4] + (las,) + t[5:]
        q.terms.append(t)
    return q


@pytest.mark.parametrize("precision", ["fp64", "fp32", "compact"])
def test_failed_job_is_isolated(ctx, precision):
    import qdng_amd as Q
    steps = 8
    jobs = mixed_jobs(steps, seed=21)      # 1D 128, 384, 3072, 4096; 2D 64^2, 96x64, 64x192, 128^2
    gs = gspo_jobs(steps)[:2]
    note(f"failure isolation, {precision}")
    # (job index, how it is poisoned, the step it stops at)
    bad_psi = {1: 0, 6: 0}                 # a NaN / an inf in psi0: the first step's start state
    bad_laser = {2: 5, 5: 3}               # NaN field at step s: the state after step s
    bad_gspo = {0: 0}                      # GSPO: the state its first step produced

    def submit(B, poison):
        ids = []
        for i, p in enumerate(jobs):
            psi0 = p.psi0.copy()
            if poison and i in bad_psi:
                psi0[len(psi0) // 3] = np.nan if i == 1 else np.inf
            q = _poisoned_laser(p, bad_laser[i]) if poison and i in bad_laser else p
            ids.append(B.submit(q.gpu_ham(None), 1.0, steps, psi0))
        for i, p in enumerate(gs):
            psi0 = p.psi0.copy()
            if poison and i in bad_gspo:
                psi0[7] = np.nan
            ids.append(B.submit(dt=1.0, nsteps=steps, psi0=psi0, ops=split_ops(p, None)))
        return ids

    A = Q.Batch(ctx, precision, 1 << 30)
    ia = submit(A, False)
    A.run(steps)
    M = Q.Batch(ctx, precision, 1 << 30)
    im = submit(M, True)
    M.run(steps)
    nj = len(jobs)
    for k, (a, m) in enumerate(zip(ia, im)):
        stop = (bad_psi.get(k, bad_laser.get(k)) if k < nj else bad_gspo.get(k - nj))
        left, nh, done = M.state(m)
        if stop is None:
            assert A.failed_step(a) == -1 and M.failed_step(m) == -1
            assert np.array_equal(M.fetch(m), A.fetch(a)), k
            assert (nh, done) == A.state(a)[1:]
            continue
        # a laser sample that is not a number spoils step s; the next
        # step starts from a state that is not a number and stops the job
        want = stop + 1 if k in bad_laser else stop
        note(f"  job {k}: failed at step {M.failed_step(m)}, {done} steps done")
        assert M.failed_step(m) == want and done == want
        assert (nh > 0) == (want > 0)
    # the serving loop reports the step as well
    S = Q.Batch(ctx, precision, 1 << 30)
    p = jobs[4]
    psi0 = p.psi0.copy()
    psi0[0] = np.nan
    reqs = [dict(H=p.gpu_ham(None), dt=1.0, nsteps=4, psi0=psi0), dict(H=p.gpu_ham(None), dt=1.0, nsteps=4,
                                                                        psi0=p.psi0)]
    S.serve(reqs, np.array([0.0, 0.0]))
    assert list(S.last_failed) == [0, -1]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32", "compact"])
def test_restart_from_snapshot_is_bit_identical(ctx, precision):
    """A job fetched after k steps and resubmitted with first_step = k (its
    laser samples from step k on) ends bit for bit where the uninterrupted
    job ends; also for a GSPO job."""
    import qdng_amd as Q
    print(f"[gpu case] restart from snapshot {precision}", flush=True)
    steps, k = 10, 4
    jobs = mixed_jobs(steps, seed=29)[:4]
    g = gspo_jobs(steps, seed=31)[:2] if precision != "compact" else []
    A = Q.Batch(ctx, precision, 1 << 30)
    ids = [A.submit(p.gpu_ham(None), 1.0, steps, p.psi0) for p in jobs]
    ids += [A.submit(dt=1.0, nsteps=steps, psi0=p.psi0, ops=split_ops(p, None)) for p in g]
    A.run(steps)
    straight = [A.fetch(j) for j in ids]
    B = Q.Batch(ctx, precision, 1 << 30)
    first = [B.submit(p.gpu_ham(None), 1.0, k, p.psi0) for p in jobs]
    first += [B.submit(dt=1.0, nsteps=k, psi0=p.psi0, ops=split_ops(p, None)) for p in g]
    B.run(k)
    snaps = [B.fetch(j) for j in first]
    # the restart in a new batch (as after the end of the process that held
    # the snapshot)
    C = Q.Batch(ctx, precision, 1 << 30)
    second = [C.submit(p.gpu_ham(None), 1.0, steps - k, s, first_step=k) for p, s in zip(jobs, snaps)]
    second += [C.submit(dt=1.0, nsteps=steps - k, psi0=s, ops=split_ops(p, None), first_step=k)
               for p, s in zip(g, snaps[len(jobs):])]
    C.run(steps - k)
    for a, j in zip(straight, second):
        assert np.array_equal(C.fetch(j), a)
    with pytest.raises(Q.QdxError):
        C.submit(jobs[0].gpu_ham(None), 1.0, steps, jobs[0].psi0, first_step=steps)   # beyond the laser
