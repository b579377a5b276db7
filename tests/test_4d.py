"""Four-dimensional grids on the device (GridSystem's MAX_DIMS = 4).

The strided passes run along dimensions 1 .. 3 (the k-space factor of the
full kinetic operator after forward transforms along 1 .. nd-2, inverse
ones back), the Chebyshev row kernel sums the separable filters of all four
dimensions.  Compared with the oracle, whose N-dimensional FFT is pinned to
numpy for up to three dimensions and here checked on a 4D array as well:
H psi, Chebyshev (real and imaginary time), GSPO (coupled states), <x_d> and
<p_d>."""
import numpy as np
import pytest

import cases
from oracle import oracle as O


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def note(msg):
    print("[gpu case]", msg, flush=True)


@pytest.fixture(autouse=True)
def _fresh():
    O.reset()
    yield
    O.reset()


def test_oracle_fft_4d_matches_numpy():
    dims = [8, 6, 4, 10]
    rng = np.random.default_rng(4)
    shape = dims[::-1]
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = O.fft(x.ravel(), dims, -1).reshape(shape)
    assert np.allclose(f, np.fft.fftn(x), rtol=1e-13, atol=1e-12 * np.abs(x).sum())


def problem(dims, S, coupled=True, L=5.0):
    p = cases.Problem(dims, [-L] * 4, [L] * 4, nstates=S)
    ax = p.axes()
    r2 = sum((0.8 + 0.1 * d) * a ** 2 for d, a in enumerate(ax))
    masses = [1836.0, 2500.0, 3000.0, 4000.0]
    for s in range(S):
        p.kinetic(s, masses)
        p.potential(s, s, np.minimum(0.02 * r2 + 0.01 * s + 0.003 * ax[0] * ax[3], 0.8))
    if S == 2 and coupled:
        p.potential(0, 1, 0.01 * np.exp(-r2))
    g = cases.gaussian(p, [0.3, -0.2, 0.1, 0.4], 0.9)
    p.psi0 = cases.normalized(p, np.concatenate([g.ravel()] + [0.5 * g.ravel()] * (S - 1)))
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("dims,S", [([16, 8, 8, 16], 1), ([8, 16, 24, 8], 2), ([40, 8, 16, 8], 1)])
def test_apply(ctx, dims, S):
    note(f"4D apply {dims} S={S}")
    p = problem(dims, S)
    rng = np.random.default_rng(sum(dims))
    psi = rng.standard_normal(S * p.size) + 1j * rng.standard_normal(S * p.size)
    H = p.gpu_ham(ctx)
    src, dst = p.gpu_wf(ctx, psi), p.gpu_wf(ctx)
    H.apply(dst, src)
    assert rel(dst.download(), p.oracle_ham().apply(psi)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("points", [0, 8])
@pytest.mark.parametrize("dims,S", [([16, 8, 8, 16], 1), ([8, 16, 8, 8], 2)])
def test_cheby(dims, S, points):
    import qdng_amd as Q
    note(f"4D cheby {dims} S={S} points {points}")
    c = Q.Context(0)
    try:
        c.set_line_points(points)
        p = problem(dims, S)
        U = Q.Cheby(p.gpu_ham(c), 5.0, prec=1e-10)
        Uo = O.Cheby(p.oracle_ham(), 5.0, prec=1e-10)
        assert U.order == Uo.order
        w, psi = p.gpu_wf(c), p.psi0.copy()
        for i in range(3):
            st = U.apply(w, i)
            assert st.n_hpsi == Uo.step(psi, i)
        assert rel(w.download(), psi) < 1e-10
    finally:
        c.close()


@pytest.mark.gpu
def test_imag_time(ctx):
    import qdng_amd as Q
    note("4D imaginary time")
    p = problem([8, 8, 16, 8], 1)
    U = Q.Cheby(p.gpu_ham(ctx), 10.0, prec=1e-10, imag=True)
    Uo = O.Cheby(p.oracle_ham(), 10.0, prec=1e-10, imag=True)
    w, psi = p.gpu_wf(ctx), p.psi0.copy()
    for i in range(3):
        U.apply(w, i)
        Uo.step(psi, i)
    assert rel(w.download(), psi) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("dims,S", [([16, 8, 8, 16], 1), ([8, 16, 8, 8], 2)])
def test_gspo(ctx, dims, S):
    import qdng_amd as Q
    note(f"4D gspo {dims} S={S}")
    p = problem(dims, S)
    ops = [p.gpu_split_op(ctx, ["kin"]), p.gpu_split_op(ctx, ["pot"])]
    ops_o = [p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])]
    G, Go = Q.GSPO(ops, 2.0), O.GSPO(ops_o, 2.0, p.S)
    w, psi = p.gpu_wf(ctx), p.psi0.copy()
    for i in range(4):
        G.apply(w, i)
        Go.step(psi, i)
    assert rel(w.download(), psi) < 1e-12
    # a run keeps psi transformed along dimension 0 between the steps
    w2, w0 = p.gpu_wf(ctx), p.gpu_wf(ctx)
    G.run(w2, 0, 4, w0)
    assert np.array_equal(w2.download(), w.download())


@pytest.mark.gpu
def test_expec_position_momentum(ctx):
    import qdng_amd as Q
    dims, S = [8, 16, 8, 24], 2
    note(f"4D expec x/p {dims}")
    rng = np.random.default_rng(41)
    lo = [-3.0 - 0.5 * d for d in range(4)]
    hi = [4.0 + 0.25 * d for d in range(4)]
    n = int(np.prod(dims))
    psi = rng.standard_normal(S * n) + 1j * rng.standard_normal(S * n)
    w = Q.WaveFunction(ctx, dims, lo, hi, S).upload(psi)
    g = O.Grid(dims, lo, hi)
    for d in range(4):
        x, pm = w.expec_position(d), w.expec_momentum(d)
        for s in range(S):
            part = psi[s * n:(s + 1) * n].copy()
            assert x[s] == pytest.approx(O.expec_position(g, part, d), rel=1e-12, abs=1e-12)
            assert pm[s] == pytest.approx(O.expec_momentum(g, part, d), rel=1e-11, abs=1e-11)


@pytest.mark.gpu
def test_five_dimensions_rejected(ctx):
    import qdng_amd as Q
    with pytest.raises(Q.QdxError):
        Q.WaveFunction(ctx, [8, 8, 8, 8, 8], [-1.0] * 5, [1.0] * 5, 1)
