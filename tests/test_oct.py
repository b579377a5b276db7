"""Optimal control (QDng ProgOCT, src/ProgOCT.cpp): the oracle's restatement
pinned by the mathematics of the method, and the device driver
(qdx_oct_run) against it.

Pins of the oracle (CPU):
- gradient: the first Krotov correction with a very large alpha is
  -shape/alpha * Im(<psi(s)|psi_t(s)> <psi_t(s)|mu psi(s)>), and
  dJ/dE(s) = -2 dt Im(...) for J = |<psi_t|psi(T)>|^2 and H = H0 - mu E(t);
  so 2 alpha dt * (E_1 - E_0)(s) must equal the finite-difference derivative
  of J with respect to the field sample E(s).  A wrong sign, a missing
  conjugate or the wrong operator in CalcCorr fails it.
- monotonic convergence: Krotov's yield increases from iteration to
  iteration (a property of the method for a small enough step), for an
  overlap and a projection target.
- the projection operator against an explicit projector matrix.
Device (GPU): every report row and the optimised field against the oracle
for Krotov, Rabitz and Rabitz feedback, overlap and projection targets, one
and two targets, phase-sensitive objectives, a two-state Hamiltonian and a
gobbler.
"""
import numpy as np
import pytest

import cases
from oracle import oracle as O

DT = 0.2


def two_level_1d(nsteps=150, E0=0.02, w=1.0, N=128, dt=DT):
    """1D harmonic oscillator (omega 1, mass 1) on [-8, 8], dipole mu = x,
    guess field E0 sin(w t); psi_i the ground state, psi_t the first excited
    state (Gaussian and x * Gaussian, normalised on the grid)."""
    p = cases.Problem([N], [-8.0], [8.0])
    (x,) = p.axes()
    t = np.arange(nsteps + 5) * dt
    p.kinetic(0, [1.0]).potential(0, 0, 0.5 * x ** 2).dipole(0, 0, x, E0 * np.sin(w * t))
    g0 = np.exp(-0.5 * x ** 2).astype(np.complex128)
    g1 = (x * np.exp(-0.5 * x ** 2)).astype(np.complex128)
    p.psi0 = cases.normalized(p, g0)
    p.psi_t = cases.normalized(p, g1)
    p.shape = np.sin(np.pi * t / (nsteps * dt)) ** 2
    return p


def two_state_1d(nsteps=150, N=128):
    """Two displaced harmonic surfaces coupled by a constant transition dipole
    (off-diagonal GridDipole, mirrored), the ground state on surface 0."""
    p = cases.Problem([N], [-8.0], [8.0], nstates=2)
    (x,) = p.axes()
    t = np.arange(nsteps + 5) * DT
    p.kinetic(0, [1.0]).kinetic(1, [1.0])
    p.potential(0, 0, 0.5 * x ** 2).potential(1, 1, 0.5 * (x - 0.5) ** 2 + 1.0)
    p.dipole(1, 0, np.ones_like(x), 0.05 * np.sin(1.0 * t))
    g0 = np.exp(-0.5 * x ** 2).astype(np.complex128)
    psi = np.zeros(2 * N, np.complex128)
    psi[:N] = g0
    p.psi0 = cases.normalized(p, psi)
    p.shape = np.sin(np.pi * t / (nsteps * DT)) ** 2
    return p


def reset(p):
    """Free the oracle's objects (and the problem's cached grid)."""
    O.reset()
    p.__dict__.pop("_og", None)


def oracle_coupling(p):
    """ProgOCT's coupling operator: the dipole arrays without the field
    (single state: their sum; multistate: the off-diagonal ones, unity off)."""
    g = p.oracle_grid()
    dips = [t for t in p.terms if t[0] == "dip"]
    if p.S == 1:
        return O.op_sum(*[O.potential(g, t[3]) for t in dips])
    return O.multistate(p.S, {(t[1], t[2]): O.potential(g, t[3]) for t in dips if t[1] != t[2]}, unity=False)


def oracle_oct(p, nsteps, dt=DT, **kw):
    H = p.oracle_ham()
    U = O.Cheby(H, dt, prec=1e-12)
    laser0 = [t for t in p.terms if t[0] == "dip"][0][4]
    return O.oct_run(U, p.oracle_dipoles, laser0, oracle_coupling(p), p.oracle_grid(), p.S, p.shape,
                     [p.psi0.copy()], nsteps, dt, **kw)


def yield_of(p, field, nsteps, dt):
    """J = |<psi_t|psi(T)>|^2 propagating psi_i with the given field."""
    H = p.oracle_ham()
    for d in p.oracle_dipoles:
        O.set_laser(d, field)
    U = O.Cheby(H, dt, prec=1e-12)
    psi = p.psi0.copy()
    for s in range(nsteps):
        U.step(psi, s)
    return abs(O.overlap(p.oracle_grid(), p.psi_t, psi, p.S)) ** 2


def gradient_mismatch(dt, nsteps):
    """Largest |K - FD| / |FD| over three field samples, K from the first
    Krotov correction at alpha 1e6 (2 alpha dt dE / shape), FD the central
    finite difference of the yield."""
    p = two_level_1d(nsteps, dt=dt)
    reset(p)
    alpha = 1e6
    rows, field, _ = oracle_oct(p, nsteps, dt=dt, psi_t=[p.psi_t.copy()], method="krotov", iterations=1,
                                alpha=alpha, conv=0.0)
    e0 = [t for t in p.terms if t[0] == "dip"][0][4].copy()
    worst = 0.0
    for s in (nsteps // 7, nsteps // 2, 6 * nsteps // 7):
        k = 2 * alpha * dt * (field[s] - e0[s]) / p.shape[s]
        eps = 1e-4
        ep, em = e0.copy(), e0.copy()
        ep[s] += eps
        em[s] -= eps
        reset(p)
        fd = (yield_of(p, ep, nsteps, dt) - yield_of(p, em, nsteps, dt)) / (2 * eps)
        assert abs(fd) > 1e-4
        assert np.sign(k) == np.sign(fd), (s, k, fd)
        worst = max(worst, abs(k - fd) / abs(fd))
    return worst


def test_oracle_krotov_correction_is_the_gradient():
    # CalcCorr takes psi and psi_t at the start of step s, the derivative of
    # that step's exponential weighs the whole step: they agree to O(dt)
    coarse = gradient_mismatch(0.2, 150)
    fine = gradient_mismatch(0.05, 600)
    assert fine < 0.06, fine
    assert fine < 0.5 * coarse, (fine, coarse)


@pytest.mark.parametrize("target", ["overlap", "projection"])
def test_oracle_krotov_yield_increases(target):
    nsteps = 150
    p = two_level_1d(nsteps)
    reset(p)
    kw = dict(method="krotov", iterations=5, alpha=2.0, conv=0.0)
    if target == "overlap":
        rows, _, _ = oracle_oct(p, nsteps, psi_t=[p.psi_t.copy()], **kw)
        y = rows[:, 5 + 2]
    else:
        rows, _, _ = oracle_oct(p, nsteps, targets=[([p.psi_t.copy()], True, False)], **kw)
        y = rows[:, 5 + 4]
    assert np.all(np.diff(y) > 0), y
    assert y[-1] > 2 * y[0]


def test_oracle_projection_matches_matrix():
    p = two_level_1d(10)
    g = p.oracle_grid()
    rng = np.random.default_rng(3)
    vecs = [cases.normalized(p, rng.standard_normal(p.size) + 1j * rng.standard_normal(p.size)) for _ in range(3)]
    psi = rng.standard_normal(p.size) + 1j * rng.standard_normal(p.size)
    w = p.weight()
    M = sum(np.outer(v, v.conj()) * w for v in vecs)
    assert np.allclose(O._proj((vecs, True, False), psi, g, 1), M @ psi, atol=1e-12)
    assert np.allclose(O._proj((vecs, False, False), psi, g, 1), -(M @ psi), atol=1e-12)
    assert np.allclose(O._proj((vecs, True, True), psi, g, 1), psi - M @ psi, atol=1e-12)


# ---------------------------------------------------------------------------
# device against the oracle

def _device(ctx, p, nsteps, targets_kind="overlap", ntargets=1, gobbler=None, **kw):
    import qdng_amd as Q
    H = p.gpu_ham(ctx)
    U = Q.Cheby(H, DT, prec=1e-12)
    psi_i = [p.gpu_wf(ctx) for _ in range(ntargets)]
    extra = {}
    if targets_kind == "overlap":
        extra["psi_t"] = [p.gpu_wf(ctx, p.psi_t) for _ in range(ntargets)]
    else:
        extra["targets"] = [Q.Projection([p.gpu_wf(ctx, v) for v in p.proj_vecs]) for _ in range(ntargets)]
    G = None
    if gobbler is not None:
        G = Q.Hamiltonian(ctx, p.N, p.xmin, p.xmax, p.S, p.S > 1, True)
        G.add_gobbler(0, 0, **gobbler)
    rows, conv = Q.oct_run(U, p.shape, psi_i, nsteps, gobbler=G, **extra, **kw)
    return rows, H.laser(), conv


def _oracle(p, nsteps, targets_kind="overlap", ntargets=1, gobbler=None, **kw):
    reset(p)
    H = p.oracle_ham()
    U = O.Cheby(H, DT, prec=1e-12)
    laser0 = [t for t in p.terms if t[0] == "dip"][0][4]
    extra = {}
    if targets_kind == "overlap":
        extra["psi_t"] = [p.psi_t.copy() for _ in range(ntargets)]
    else:
        extra["targets"] = [([v.copy() for v in p.proj_vecs], True, False) for _ in range(ntargets)]
    G = None
    if gobbler is not None:
        G = O.gobbler(p.oracle_grid(), **gobbler)
    return O.oct_run(U, p.oracle_dipoles, laser0, oracle_coupling(p), p.oracle_grid(), p.S, p.shape,
                     [p.psi0.copy() for _ in range(ntargets)], nsteps, DT, gobbler=G, **extra, **kw)


CASES = [
    ("krotov", "overlap", 1, False, "two_level", None),
    ("rabitz", "overlap", 1, False, "two_level", None),
    ("rabitzfb", "overlap", 1, False, "two_level", None),
    ("krotov", "overlap", 1, True, "two_level", None),
    ("krotov", "overlap", 2, False, "two_level", None),
    ("krotov", "projection", 1, False, "two_level", None),
    ("rabitzfb", "projection", 1, False, "two_level", None),
    ("krotov", "projection", 1, False, "two_state", None),
    ("krotov", "overlap", 1, False, "two_level", {"rp": {0: 6.0}, "order": 8, "gain": 1.0}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("method,target,ntargets,phase,system,gob", CASES)
def test_device_oct_matches_oracle(ctx, method, target, ntargets, phase, system, gob):
    print(f"[gpu case] oct {method} {target} x{ntargets} phase={phase} {system} gobbler={gob is not None}",
          flush=True)
    nsteps = 120
    p = two_level_1d(nsteps) if system == "two_level" else two_state_1d(nsteps)
    if system == "two_state":
        (x,) = p.axes()
        N = p.N[0]
        ex = np.zeros(2 * N, np.complex128)
        ex[N:] = np.exp(-0.5 * (x - 0.5) ** 2)
        p.proj_vecs = [cases.normalized(p, ex)]
    else:
        p.proj_vecs = [p.psi_t.copy()]
    kw = dict(method=method, phase=phase, iterations=3, alpha=4.0 if method != "krotov" else 2.0, conv=0.0)
    rows_o, field_o, _ = _oracle(p, nsteps, target, ntargets, gob, **kw)
    rows_d, field_d, _ = _device(ctx, p, nsteps, target, ntargets, gob, **kw)
    assert rows_d.shape == rows_o.shape
    assert np.array_equal(np.isnan(rows_d), np.isnan(rows_o))
    m = ~np.isnan(rows_o)
    # relative 1e-8; the phase / pi columns absolute 1e-8 (a phase near zero
    # after several phase-sensitive iterations)
    tol = 1e-8 * np.abs(rows_o) + 1e-13
    for t in range(ntargets):
        tol[:, 5 + 5 * t + 3] = 1e-8
    diff = np.abs(rows_d - rows_o)
    print(f"  rows max rel {np.max(diff[m] / (np.abs(rows_o[m]) + 1e-300)):.2e}, yields {rows_o[:, 4]}, "
          f"phases {rows_o[:, 8]}", flush=True)
    assert np.all(diff[m] < tol[m]), np.argwhere((diff > tol) & m)
    n = field_o.size
    assert np.max(np.abs(field_d[:n] - field_o)) < 1e-9 * max(1.0, np.max(np.abs(field_o)))
    # the optimisation did something
    assert rows_o[-1, 4] != rows_o[0, 4]
