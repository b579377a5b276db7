"""Line lengths with prime factors 5 .. 31 on the device.

QDng transforms any grid length with FFTW (src/fft/fft.cpp:95-119).  The
device takes multiples of 8 whose prime factors above 3 are at most 31: the
line transforms run those factors as generic middle stages.  The oracle's FFT
for such lengths is pinned against numpy and a brute-force DFT in
tests/test_oracle.py (lengths 40, 56, 440, 1000 among them); here the device
is compared with the oracle: H psi, Chebyshev steps on 1D lines (the
whole-step kernel at 4 and 8 points per thread), 2D two-state grids (row
kernel and column passes at 4 and 8 points per thread), GSPO, FP32 and set
radix plans.  Batched jobs keep the radix 8/4/2/3 transforms and reject
other lengths."""
import numpy as np
import pytest

import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def note(msg):
    print("[gpu case]", msg, flush=True)


@pytest.fixture(autouse=True)
def _fresh():
    O.reset()
    yield
    O.reset()


# 8 * {5, 7, 11, 13, 25, 35, 3*5*7, 31, 5*5*5, 3*11*5, 7*7*7, 5*7*11}
SIZES = [40, 56, 88, 104, 200, 280, 840, 248, 1000, 1320, 2744, 3080]


@pytest.mark.parametrize("N", SIZES)
def test_apply_1d(ctx, N):
    note(f"odd apply 1D N={N}")
    p = cases.harmonic_1d(N, L=6.0)
    rng = np.random.default_rng(N)
    psi = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    H = p.gpu_ham(ctx)
    src, dst = p.gpu_wf(ctx, psi), p.gpu_wf(ctx)
    H.apply(dst, src)
    assert rel(dst.download(), p.oracle_ham().apply(psi)) < 1e-12


def two_state(n0, n1, L=8.0):
    p = cases.Problem([n0, n1], [-L, -L], [L, L], nstates=2)
    x, y = p.axes()
    for s in (0, 1):
        p.kinetic(s, [1836.0, 3000.0])
    # Morse-like and harmonic surfaces capped at 1 au (the spectral range
    # sets the series length)
    p.potential(0, 0, np.minimum(0.2 * ((1 - np.exp(-0.5 * x)) ** 2 + 0.5 * (1 - np.exp(-0.5 * y)) ** 2), 1.0))
    p.potential(1, 1, np.minimum(0.05 * ((x - 1.0) ** 2 + (y - 1.0) ** 2) + 0.03, 1.0))
    p.potential(0, 1, 0.01 * np.exp(-(x ** 2 + y ** 2)))
    g = cases.gaussian(p, [0.4, -0.3], 0.8)
    p.psi0 = cases.normalized(p, np.concatenate([g.ravel(), 0.3 * g.ravel()]))
    return p


@pytest.mark.parametrize("dims", [[40, 56], [104, 24], [200, 88], [48, 1000]])
def test_apply_2d_two_states(ctx, dims):
    note(f"odd apply 2D {dims}")
    p = two_state(*dims)
    rng = np.random.default_rng(sum(dims))
    psi = rng.standard_normal(2 * p.size) + 1j * rng.standard_normal(2 * p.size)
    H = p.gpu_ham(ctx)
    src, dst = p.gpu_wf(ctx, psi), p.gpu_wf(ctx)
    H.apply(dst, src)
    assert rel(dst.download(), p.oracle_ham().apply(psi)) < 1e-12


def cheby_pair(ctx, p, dt, steps, precision="fp64"):
    import qdng_amd as Q
    U = Q.Cheby(p.gpu_ham(ctx), dt, prec=1e-10, precision=precision)
    Uo = O.Cheby(p.oracle_ham(), dt, prec=1e-10)
    w, psi = p.gpu_wf(ctx), p.psi0.copy()
    counts = []
    for i in range(steps):
        st = U.apply(w, i)
        counts.append((st.n_hpsi, Uo.step(psi, i)))
    return w.download(), psi, counts, U.order, Uo.order


# 1D single-state lines: whole-step kernel, 4 points per thread up to 1024
# points and 8 above (or 8 when set)
@pytest.mark.parametrize("points", [0, 8])
@pytest.mark.parametrize("N", [200, 280, 1000, 1320, 3080])
def test_cheby_1d_line(N, points):
    import qdng_amd as Q
    note(f"odd cheby 1D N={N} points {points}")
    c = Q.Context(0)
    try:
        c.set_line_points(points)
        p = cases.harmonic_1d(N, L=8.0)
        # the spectral range grows as N^2: a shorter step keeps the series
        # within OCheby's order limit on the longest line
        got, ref, counts, n_gpu, n_orc = cheby_pair(c, p, 10.0 if N <= 1320 else 1.0, 5)
        assert n_gpu == n_orc
        assert all(a == b for a, b in counts), counts
        assert rel(got, ref) < 1e-10
    finally:
        c.close()


@pytest.mark.parametrize("points", [4, 8])
@pytest.mark.parametrize("dims", [[40, 56], [280, 200]])
def test_cheby_2d_two_states(dims, points):
    import qdng_amd as Q
    note(f"odd cheby 2D {dims} points {points}")
    c = Q.Context(0)
    try:
        c.set_line_points(points)
        p = two_state(*dims)
        got, ref, counts, n_gpu, n_orc = cheby_pair(c, p, 5.0, 3)
        assert n_gpu == n_orc
        assert all(a == b for a, b in counts), counts
        assert rel(got, ref) < 1e-10
    finally:
        c.close()


def test_cheby_fp32_and_compact(ctx):
    note("odd cheby 2D fp32 / compact")
    p = two_state(56, 40)
    ref = cheby_pair(ctx, p, 5.0, 3)[0]
    for precision in ("compact", "fp32"):
        got = cheby_pair(ctx, p, 5.0, 3, precision)[0]
        cases.bulk_bar(rel(got, ref), precision, f" (56x40, {precision})")


@pytest.mark.parametrize("dims", [[280], [40, 56]])
def test_gspo(ctx, dims):
    import qdng_amd as Q
    note(f"odd gspo {dims}")
    p = cases.harmonic_1d(dims[0], L=8.0) if len(dims) == 1 else two_state(*dims)
    ops = [p.gpu_split_op(ctx, ["kin"]), p.gpu_split_op(ctx, ["pot"])]
    ops_o = [p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])]
    G, Go = Q.GSPO(ops, 2.0), O.GSPO(ops_o, 2.0, p.S)
    w, psi = p.gpu_wf(ctx), p.psi0.copy()
    for i in range(5):
        G.apply(w, i)
        Go.step(psi, i)
    assert rel(w.download(), psi) < 1e-12


def test_radix_plans_and_limits():
    import qdng_amd as Q
    note("odd radix plans")
    c = Q.Context(0)
    try:
        # a prime stage in the middle only; edges 8, 4 or 2
        c.set_radix_plan(280, [4, 7, 5, 2], "fp64")
        assert c.radix_plan(280, "fp64") == [4, 7, 5, 2]
        for bad in ([5, 7, 8], [8, 5, 7], [4, 35, 2], [4, 7, 5, 3]):
            with pytest.raises(Q.QdxError):
                c.set_radix_plan(280, bad, "fp64")
        p = cases.harmonic_1d(280, L=6.0)
        rng = np.random.default_rng(3)
        psi = rng.standard_normal(280) + 1j * rng.standard_normal(280)
        H = p.gpu_ham(c)
        src, dst = p.gpu_wf(c, psi), p.gpu_wf(c)
        H.apply(dst, src)
        assert rel(dst.download(), p.oracle_ham().apply(psi)) < 1e-12
        # prime factors above 31 and lengths that are not multiples of 8
        for n in (8 * 37, 8 * 41, 100):
            with pytest.raises(Q.QdxError):
                Q.WaveFunction(c, [n], [-1.0], [1.0], 1)
        # batched jobs keep the radix 8/4/2/3 transforms
        B = Q.Batch(c, "fp64", hbm_budget=1 << 28)
        q = cases.harmonic_1d(40)
        with pytest.raises(Q.QdxError):
            B.submit(H=q.gpu_ham(c), dt=1.0, nsteps=2, psi0=q.psi0)
        del B
    finally:
        c.close()
