"""Pins the FP64 CPU oracle before it is trusted as the HIP path's checker.

What pins what (QDng file:line of the behaviour in brackets):
  - QDng's own regression tests: report line and expectation values of
    TGP-Expec, TGPMS-Cheby, TGPMS-GSPO, TGPL-IR-Pulse-Cheby/GSPO,
    TGPLMS-Cheby at their printed precision   [tests/checktests.sh:129-152]
  - FFT against numpy and a brute-force DFT  [src/fft/fft.cpp:98-119]
  - Bessel J_k against scipy/AMOS vectors and the identity
    J_0^2 + 2 sum J_k^2 = 1                   [src/linalg/Bessel.cpp:22-58]
  - kinetic operator on a periodic sine      [src/tests/OGridNablaSqTest.cpp:76-137]
  - Chebyshev on a two-level system          [src/tests/TestCheby.cpp:60-125]
  - split operator full transfer             [src/tests/OGSPOTest.cpp:80-129]
  - closed forms: 2x2 Rabi propagator, harmonic-oscillator revival, norm
    conservation.
Each check is chosen so that a dropped term, a wrong sign or a wrong index in
the oracle fails it (see test_planted_mistakes).
"""
import math

import numpy as np
import pytest
import scipy.linalg

import cases
from oracle import oracle as O


@pytest.fixture(autouse=True)
def _fresh():
    O.reset()
    yield
    O.reset()


# ---------------------------------------------------------------------------
# FFT

@pytest.mark.parametrize("dims", [[8], [9], [12], [96], [128], [243], [100], [6, 4], [8, 12],
                                  [4, 6, 8], [16, 9, 5], [40], [56], [440], [1000], [3080], [40, 56]])
def test_fft_matches_numpy(dims):
    rng = np.random.default_rng(1)
    shape = dims[::-1]
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = O.fft(x.ravel(), dims, -1).reshape(shape)
    assert np.allclose(f, np.fft.fftn(x), rtol=1e-13, atol=1e-12 * np.abs(x).sum())
    b = O.fft(x.ravel(), dims, +1).reshape(shape)
    assert np.allclose(b, np.fft.ifftn(x) * x.size, rtol=1e-13, atol=1e-12 * np.abs(x).sum())


@pytest.mark.parametrize("n", [7, 24, 45, 88, 104, 248])
def test_fft_brute_force_dft(n):
    x = np.exp(1j * np.arange(n) ** 1.5) * (1 + np.arange(n) % 3)
    F = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
    assert np.allclose(O.fft(x, [n], -1), F @ x, atol=1e-12)


# ---------------------------------------------------------------------------
# Bessel

def test_bessel_against_scipy_vectors():
    d = np.load(cases.GOLDEN + "/bessel_jv.npz")
    for i, x in enumerate(d["x"]):
        ref = d["j%d" % i]
        got, nz = O.bessel_j(len(ref), float(x))
        scale = np.abs(ref).max()
        # relative where J_k is representable, absolute in the underflow tail
        big = np.abs(ref) > 1e-280
        assert np.allclose(got[big], ref[big], rtol=1e-9, atol=1e-15 * scale), x
        assert np.all(np.abs(got[~big]) < 1e-270)


@pytest.mark.parametrize("x", [0.3, 19.41, 156.6])
def test_bessel_sum_rule(x):
    j, _ = O.bessel_j(int(x) + 80, x)
    assert abs(j[0] ** 2 + 2 * np.sum(j[1:] ** 2) - 1.0) < 1e-13


# ---------------------------------------------------------------------------
# kinetic operator: OGridNablaSqTest analogue, even and odd N

@pytest.mark.parametrize("N", [8, 9, 64])
def test_kinetic_on_periodic_sine(N):
    g = O.Grid([N], [-5.0], [5.0])
    dx = 10.0 / (N - 1)
    period = 10.0 + dx
    x = -5.0 + dx * np.arange(N)
    psi = np.sin(2 * np.pi * x / period).astype(complex)
    T = O.kinetic(g, [1.0])
    Tpsi = T.apply(psi)
    expect = (2 * np.pi / period) ** 2 / 2 * psi
    assert np.allclose(Tpsi, expect, atol=1e-12)
    emin, emax = T.bounds()
    assert emin == 0 and emax == pytest.approx(np.pi ** 2 / 2 / dx ** 2, rel=1e-15)


def test_kinetic_2d_is_sum_of_dims():
    p = cases.Problem([16, 12], [-4, -3], [4, 3])
    x, y = p.axes()
    kx = 2 * np.pi / (8 + p.dx(0)) * 2
    ky = 2 * np.pi / (6 + p.dx(1)) * 1
    psi = np.exp(1j * (kx * x + ky * y)).ravel()
    p.kinetic(0, [2.0, 3.0])
    got = p.oracle_ham().apply(psi)
    assert np.allclose(got, (kx ** 2 / 4 + ky ** 2 / 6) * psi, atol=1e-12)


# ---------------------------------------------------------------------------
# Chebyshev: TestCheby analogue, H = diag(0, 0.5) on every grid point

def two_level(pot0, pot1, v01=None, N=8):
    p = cases.Problem([N], [0.0], [1.0], nstates=2)
    p.potential(0, 0, np.full(N, pot0)).potential(1, 1, np.full(N, pot1))
    if v01 is not None:
        p.potential(0, 1, np.full(N, v01))
    return p


def test_cheby_two_level_quarter_rotation():
    p = two_level(0.0, 0.5)
    U = O.Cheby(p.oracle_ham(), math.pi)
    assert U.prec == pytest.approx(1e-10)          # TestCheby.cpp:84-86
    assert U.offset == pytest.approx(0.25)          # TestCheby.cpp:87-90
    assert U.scaling == pytest.approx(0.25)         # TestCheby.cpp:91-93
    psi = np.concatenate([np.zeros(8), np.ones(8)]).astype(complex)
    U.step(psi)
    assert np.allclose(psi[8:], -1j, atol=1e-10)    # TestCheby.cpp:116-124
    assert np.allclose(psi[:8], 0, atol=1e-10)
    psi = np.concatenate([np.ones(8), np.zeros(8)]).astype(complex)
    U.step(psi)
    assert np.allclose(psi[:8], 1, atol=1e-10)      # TestCheby.cpp:107-113


def test_cheby_rabi_closed_form():
    a, b, c, dt = 0.1, 0.35, 0.07, 7.0
    p = two_level(a, b, c)
    U = O.Cheby(p.oracle_ham(), dt, prec=1e-14)
    psi = np.concatenate([np.full(8, 0.6), np.full(8, 0.8j)])
    psi0 = psi.copy()
    U.step(psi)
    M = scipy.linalg.expm(-1j * dt * np.array([[a, c], [c, b]]))
    ref = M @ np.array([psi0[0], psi0[8]])
    assert np.allclose(psi[:8], ref[0], atol=1e-12) and np.allclose(psi[8:], ref[1], atol=1e-12)


def test_gspo_two_level_full_transfer():
    # sigma_x / 2 coupling only: exp(-i pi sigma_x / 2) = -i sigma_x (OGSPOTest.cpp:80-129)
    p = cases.Problem([8], [0.0], [1.0], nstates=2)
    p.potential(0, 1, np.full(8, 0.5))
    B = p.oracle_split_op(["pot"])
    G = O.GSPO([B], math.pi, nstates=2)
    psi = np.concatenate([np.ones(8), np.zeros(8)]).astype(complex)
    G.step(psi)
    assert np.allclose(np.abs(psi[8:]) ** 2, 1, atol=1e-14)
    assert np.allclose(psi[:8], 0, atol=1e-14)


def test_harmonic_revival_and_norm():
    # Gaussian in V = k x^2 / 2 returns to itself after one period 2 pi / w
    p = cases.harmonic_1d(256, L=8.0, mass=1836.0, k=0.1)
    w = math.sqrt(0.1 / 1836.0)
    period = 2 * math.pi / w
    nsteps = 40
    U = O.Cheby(p.oracle_ham(), period / nsteps, prec=1e-14)
    g = p.oracle_grid()
    psi = p.psi0.copy()
    for i in range(nsteps):
        U.step(psi, i)
    ov = O.overlap(g, p.psi0, psi)
    # grid discretisation of the oscillator limits the revival to ~1e-7
    assert abs(abs(ov) - 1) < 1e-6
    assert abs(O.norm2(g, psi) - 1) < 1e-12


# ---------------------------------------------------------------------------
# QDng regression tests (checktests.sh:129-152)

def _digits_ok(value, golden_text, prefix):
    """Would QDng's grep pattern for `golden_text` match `value`?

    prefix=True: the printed value starts with these digits (truncation);
    otherwise the printed value has exactly these digits (rounding).  A slack
    of 5% of the last printed digit absorbs summation-order noise.
    """
    t = golden_text.lower()
    mant, ex = (t.split("e") + ["0"])[:2]
    decimals = len(mant.split(".")[1]) if "." in mant else 0
    ulp = 10.0 ** (int(ex) - decimals)
    g = float(t)
    if prefix:
        # printed value starts with the digits: truncation, or rounding up
        # into them at the printed width
        d = (value - g) if g >= 0 else (g - value)   # distance away from zero
        return -0.55 * ulp <= d < 1.05 * ulp
    return abs(value - g) <= 0.55 * ulp


def run_oracle_kat(name):
    p, kind, dt, steps, prec = cases.kat_problem(name)
    H = p.oracle_ham()
    g = p.oracle_grid()
    psi = p.psi0.copy()
    if kind == "cheby":
        U = O.Cheby(H, dt, prec=prec)
        stepper = U.step
        Hrep = H
    else:
        if p.S == 1:
            ops = [p._oracle_term(t) for t in p.terms]          # A=T, B=V, C=dipole
        else:
            ops = [p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])]
        G = O.GSPO(ops, dt, nstates=p.S)
        stepper = G.step
        Hrep = O.op_sum(*ops) if p.S > 1 else H
    rep = None
    for i in range(steps):
        if i == steps - 1:
            n = p.size
            rep = {"population": O.norm2(g, psi, p.S),
                   "pops": [O.norm2(g, psi[s * n:(s + 1) * n]) for s in range(p.S)] if p.S > 1 else [],
                   "energy": Hrep.expec(psi, i),
                   "autocorr": O.overlap(g, psi, p.psi0, p.S)}
        stepper(psi, i)
    return p, g, psi, rep


def _post_expectations(name, p, g, psi):
    """expec_post.dat values at the final time (FiltersList post filters)."""
    a, _ = cases.kat_data()
    n = p.size
    if name == "TGP-NIP-Cheby":
        # GridMomentum, GridPosition, T, V (the real Gobbler), each divided by
        # <psi|psi> (FiltersList.cpp:135-330)
        n2 = O.norm2(g, psi)
        T = O.kinetic(g, [1763.2])
        V = O.gobbler(g, lp={0: 2.0}, order=20, gain=0.3)
        return [v / n2 for v in (O.expec_momentum(g, psi), O.expec_position(g, psi), T.expec(psi), V.expec(psi))]
    if p.S == 1:
        vals = [O.expec_momentum(g, psi), O.expec_position(g, psi)]
        if name == "TGP-Expec":
            vals.append(sum(abs(O.overlap(g, a["GridPot0_ef_%d" % i], psi)) ** 2 for i in range(4)))
        else:
            vals += [abs(O.overlap(g, a["GridPot0_ef_%d" % i], psi)) ** 2 for i in range(3)]
        T = O.kinetic(g, [1763.2])
        V = O.potential(g, a["GridPot0"])
        vals += [T.expec(psi), V.expec(psi)]
        return vals
    st = [psi[s * n:(s + 1) * n] for s in range(p.S)]
    T = O.kinetic(g, [1763.2])
    V = [O.potential(g, a["GridPot0"]), O.potential(g, a["GridPot1"])]
    return ([O.expec_momentum(g, x) for x in st] + [O.expec_position(g, x) for x in st]
            + [T.expec(x) for x in st] + [V[s].expec(st[s]) for s in range(2)])


@pytest.mark.parametrize("name", cases.KAT_NAMES)
def test_qdng_regression(name):
    p, g, psi, rep = run_oracle_kat(name)
    gold = cases.kat_data()[1]["golden"][name]
    r = gold["report"]
    assert abs(rep["population"] - r["population"]) < 5e-7
    for got, want in zip(rep["pops"], r["pops"]):
        assert abs(got - want) < 5e-7
    assert abs(rep["energy"] - r["energy"]) < 6e-9, (rep["energy"], r["energy"])
    assert abs(rep["autocorr"].real - r["autocorr"][0]) < 6e-9
    assert abs(rep["autocorr"].imag - r["autocorr"][1]) < 6e-9
    vals = _post_expectations(name, p, g, psi)
    want = gold["expec_post"]["values"]
    prefix = gold["expec_post"]["prefix"]
    assert len(vals) == len(want)
    for v, w, pf in zip(vals, want, prefix):
        assert _digits_ok(v, w, pf), (name, v, w)


# ---------------------------------------------------------------------------
# planted mistakes: each variation of the oracle's inputs that mimics a
# plausible restatement error moves a KAT quantity well outside its printed
# precision, so the KATs above would catch that error.

def test_planted_mistakes():
    p, kind, dt, steps, prec = cases.kat_problem("TGP-Expec")
    g = p.oracle_grid()
    a, meta = cases.kat_data()
    gold = meta["golden"]["TGP-Expec"]["report"]

    def run(mass=1763.2, pot_scale=1.0, grid_xmax=None):
        O.reset()
        q = cases.Problem(p.N, p.xmin, [grid_xmax or p.xmax[0]])
        q.kinetic(0, mass).potential(0, 0, a["GridPot0"] * pot_scale)
        H = q.oracle_ham()
        U = O.Cheby(H, dt, prec=prec)
        psi = p.psi0.copy()
        for i in range(999):
            U.step(psi, i)
        return O.overlap(q.oracle_grid(), psi, p.psi0)

    good = run()
    want = complex(*gold["autocorr"])
    assert abs(good - want) < 1e-8
    # a mass factor of 2 (e.g. -1/m instead of -1/(2m)), a 1% potential scale
    # and a grid spacing (xmax-xmin)/N instead of /(N-1) each move the
    # autocorrelation far beyond its 1e-8 printed precision
    assert abs(run(mass=2 * 1763.2) - want) > 1e-4
    assert abs(run(pot_scale=1.01) - want) > 1e-4
    dx_wrong_xmax = p.xmin[0] + (p.xmax[0] - p.xmin[0]) * p.N[0] / (p.N[0] - 1)
    assert abs(run(grid_xmax=dx_wrong_xmax) - want) > 1e-4


# ---------------------------------------------------------------------------
# coupled-state GSPO split (ODMultistate.cpp:109-159, TransformMSD.cpp:31-101)
# against the dense Strang product of the full 2N x 2N matrices

def _dense_kinetic(p, mass):
    # F^-1 diag(k^2 / 2m) F on a periodic box of N dx (numpy's DFT matrices)
    N = p.N[0]
    dx = p.dx(0)
    k = 2 * np.pi * np.fft.fftfreq(N, d=dx)
    F = np.fft.fft(np.eye(N), axis=0)
    return np.linalg.inv(F) @ np.diag(k ** 2 / (2 * mass)) @ F


def _coupled_problem(N=24):
    p = cases.Problem([N], [-3.0], [3.0], nstates=2)
    x = p.axes()[0]
    p.kinetic(0, 700.0).kinetic(1, 700.0)
    p.potential(0, 0, 0.05 * x ** 2)
    p.potential(1, 1, 0.1 + 0.03 * (x - 0.5) ** 2)
    p.potential(0, 1, 0.02 * np.exp(-x ** 2) + 0.01)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal(2 * N) + 1j * rng.standard_normal(2 * N)
    return p, x, psi / np.linalg.norm(psi)


def test_gspo_coupled_split_vs_dense_strang():
    N, dt = 24, 3.0
    p, x, psi0 = _coupled_problem(N)
    K = _dense_kinetic(p, 700.0)
    T = np.block([[K, np.zeros((N, N))], [np.zeros((N, N)), K]])
    V = np.block([[np.diag(0.05 * x ** 2), np.diag(0.02 * np.exp(-x ** 2) + 0.01)],
                  [np.diag(0.02 * np.exp(-x ** 2) + 0.01), np.diag(0.1 + 0.03 * (x - 0.5) ** 2)]])
    half = scipy.linalg.expm(-0.5j * dt * V)
    ref = half @ scipy.linalg.expm(-1j * dt * T) @ half @ psi0
    # the Strang product differs from the exact propagator at O(dt^3):
    # the check below separates it from exp(-i H dt) as well
    exact = scipy.linalg.expm(-1j * dt * (T + V)) @ psi0
    assert np.linalg.norm(exact - ref) > 1e-6

    def run(plant):
        O.plant_mistake(plant)
        try:
            A = p.oracle_split_op(["kin"])
            B = p.oracle_split_op(["pot"])
            G = O.GSPO([A, B], dt, nstates=2)
            psi = psi0.copy()
            G.step(psi)
            return psi
        finally:
            O.plant_mistake(0)

    got = run(0)
    assert np.linalg.norm(got - ref) < 1e-12, np.linalg.norm(got - ref)
    # the planted mistake (X^T e^{c Lambda} X instead of X e^{c Lambda} X^T)
    # fails the same check: with distinct diagonals X is not symmetric
    bad = run(1)
    assert np.linalg.norm(bad - ref) > 1e-3
    # and in imaginary time (exponent -dt): exp(-V dt/2) exp(-T dt) exp(-V dt/2)
    G = O.GSPO([p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])], dt, nstates=2, imag=True)
    psi = psi0.copy()
    G.step(psi)
    h2 = scipy.linalg.expm(-0.5 * dt * V)
    ref_i = h2 @ scipy.linalg.expm(-dt * T) @ h2 @ psi0
    assert np.linalg.norm(psi - ref_i) < 1e-12 * np.linalg.norm(ref_i)


# ---------------------------------------------------------------------------
# backward real time (OPropagator::Backward, OPropagator.h:82-100): the
# exponent +i dt, so one step is exp(+i H dt); checked against the matrix
# exponential of the discretised H and by time reversal

def _dense_h(H, n):
    M = np.zeros((n, n), dtype=complex)
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1
        M[:, j] = H.apply(e)
    return M


@pytest.mark.parametrize("kind", ["cheby", "sil", "arnoldi", "gspo"])
def test_backward_real_time_is_exp_plus_i_h_dt(kind):
    p = cases.harmonic_1d(48, L=6.0, mass=900.0)
    n = p.size
    dt = 3.0
    H = p.oracle_ham()
    M = _dense_h(H, n)
    if kind == "cheby":
        U = O.Cheby(H, dt, prec=1e-14)
    elif kind == "sil":
        U = O.SIL(H, dt, order=30, err=1e-14)
    elif kind == "arnoldi":
        U = O.Arnoldi(H, dt, order=30, conv=1e-14)
    else:
        U = O.GSPO([p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])], dt, 1)
    psi0 = p.psi0.copy()
    fwd = psi0.copy()
    U.step(fwd)
    U.backward(True)
    bwd = psi0.copy()
    U.step(bwd)
    if kind == "gspo":
        K = _dense_h(p.oracle_split_op(["kin"]), n)
        V = _dense_h(p.oracle_split_op(["pot"]), n)
        half = scipy.linalg.expm(0.5j * dt * V)
        ref = half @ scipy.linalg.expm(1j * dt * K) @ half @ psi0
        tol = 1e-12
    else:
        ref = scipy.linalg.expm(1j * dt * M) @ psi0
        tol = 1e-10
    assert np.linalg.norm(bwd - ref) < tol * np.linalg.norm(ref)
    # the forward step is the other sign, far from it
    assert np.linalg.norm(fwd - ref) > 1e-3
    # time reversal: forward then backward returns psi0
    back = fwd.copy()
    U.step(back)
    assert np.linalg.norm(back - psi0) < 10 * tol
    U.backward(False)
    again = psi0.copy()
    U.step(again)
    assert np.array_equal(again, fwd)


# ---------------------------------------------------------------------------
# Gobbler / negative imaginary potential (OGobbler.cpp): pinned by the
# TGP-NIP-Cheby regression case above; the checks below pin the pieces.

def _nip_grid():
    a, meta = cases.kat_data()
    g = meta["grids"]["GridPot0"]
    return O.Grid(g["N"], g["xmin"], g["xmax"]), a


def test_gobbler_bounds_and_nip_apply():
    g, a = _nip_grid()
    psi = cases.normalized(cases.Problem(g.N, g.xmin, g.xmax), a["GridPot0_ef_0"])
    nip = O.gobbler(g, rp={0: 5.5}, order=5, gain=0.02, nip=True)
    real = O.gobbler(g, rp={0: 5.5}, order=5, gain=0.02, nip=False)
    # OGobbler::Emax / Emin: gain and -i gain (nip), gain and 0 (otherwise)
    assert nip.bounds_c() == (-0.02j, 0.02 + 0j)
    assert real.bounds_c() == (0j, 0.02 + 0j)
    # one high-pass edge: 0 + f (nip) and 1 * f (real) are the same filter,
    # so the NIP applies -i times the real Gobbler
    assert np.allclose(nip.apply(psi), -1j * real.apply(psi), rtol=0, atol=1e-15)
    # filter shape: rising from 0 at the first point through gain / 2 at the
    # edge (x in grid points from xmin) towards gain
    x = g.axis(0)
    v = real.apply(np.ones(g.size, complex)).real
    assert v[0] == 0 and np.all(np.diff(v) > 0) and v[-1] < 0.02
    assert v[x < 4.0].max() < 0.01 * 0.02 and v[-1] > 0.6 * 0.02
    xe = (5.5 - g.xmin[0]) / g.dx(0)
    lo = int(np.floor(xe))
    assert v[lo] < 0.01 < v[lo + 1]
    # an edge outside the grid is refused as in OGobbler::Init
    with pytest.raises(O.OracleError):
        O.gobbler(g, rp={0: 100.0})


def test_nip_planted_mistakes():
    p, kind, dt, steps, prec = cases.kat_problem("TGP-NIP-Cheby")
    gold = cases.kat_data()[1]["golden"]["TGP-NIP-Cheby"]["report"]

    def run(plant):
        O.plant_mistake(plant)
        try:
            O.reset()
            if hasattr(p, "_og"):
                del p._og
            U = O.Cheby(p.oracle_ham(), dt, prec=prec)
            psi = p.psi0.copy()
            for i in range(steps - 1):
                U.step(psi, i)
            return O.norm2(p.oracle_grid(), psi), O.overlap(p.oracle_grid(), psi, p.psi0)
        finally:
            O.plant_mistake(0)

    n, ac = run(0)
    assert abs(n - gold["population"]) < 5e-7 and abs(ac - complex(*gold["autocorr"])) < 1e-8
    # +i v (growth instead of absorption), the coefficient phase of the real
    # offset only, and a multiplied NIP high pass (v = 0) each miss the
    # golden line
    for plant in (2, 3, 4):
        n, ac = run(plant)
        assert abs(n - gold["population"]) > 1e-3 or abs(ac - complex(*gold["autocorr"])) > 1e-4, plant


def test_nip_refuses_backward():
    p, kind, dt, steps, prec = cases.kat_problem("TGP-NIP-Cheby")
    U = O.Cheby(p.oracle_ham(), dt, prec=prec)
    with pytest.raises(O.OracleError):
        U.backward()
