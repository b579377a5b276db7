"""Imaginary-time Chebyshev propagation (OCheby::imag_time_step,
OCheby.cpp:200-297, set up by ProgEigen as ImaginaryTime() + Backward(),
exponent -dt) and eigenstates by imaginary-time relaxation
(ProgEigen::ImagTimeEF + DiagBasis, ProgEigen.cpp:93-412).

What pins it: the scaled modified Bessel values against mpmath at 40 digits
and the sum rule; one oracle step against exp(-H dt) from an eigendecomposition of
the discretised H; QDng's TGE-Cheby regression energies
(tests/checktests.sh:4-28) and initial energy (checktests.sh:56) from the
oracle driven through ImagTimeEF; the device against the oracle step by step
and against the same regression energies."""
import os

import numpy as np
import pytest
import mpmath

import cases
from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "qdng_files")

# TGE-Cheby, method "imag": energy.dat lines "i<TAB>E" must start with these
# strings (checktests.sh:4-28); 10 eigenfunctions, dt 5, at most 1000 steps
# each (TGE-Cheby.qdi.in), convergence 1e-5 (ProgEigen.h:17, diag default on)
TGE_ENERGIES = ["0.0037100", "0.0110668", "0.0181411", "0.0249017", "0.0313281", "0.0374320",
                "0.0431786", "0.0485", "0.0535", "0.058"]
TGE_E0 = "1.205115e-02"   # "Initial Norm & energy: 1.000000e+00 ... 1.205115e-02" (checktests.sh:56)


def tge_problem():
    a, meta = cases.kat_data()
    g = meta["grids"]["GridPot0"]
    p = cases.Problem(g["N"], g["xmin"], g["xmax"])
    p.kinetic(0, 1763.2).potential(0, 0, a["GridPot0"])
    gauss = np.fromfile(os.path.join(FIX, "GridGauss.wf"), dtype="<c16")
    p.psi0 = cases.normalized(p, gauss)
    return p


# ---------------------------------------------------------------------------
# Bessel I_k(x) exp(-x)

def ive_mp(n, x):
    """I_k(x) exp(-x), k < n, at 40 significant digits (mpmath)"""
    with mpmath.workdps(40):
        return np.array([float(mpmath.besseli(k, x) * mpmath.exp(-x)) for k in range(n)])


@pytest.mark.parametrize("x", [0.3, 2.0, 14.6, 156.6])
def test_bessel_i_scaled_against_mpmath(x):
    # the oracle takes libstdc++'s cyl_bessel_i: good to ~1e-11 in far tails
    n = int(2 * x) + 60
    got, nz = O.bessel_i_scaled(n, x)
    ref = ive_mp(n, x)
    big = ref > 1e-280
    assert np.allclose(got[big], ref[big], rtol=1e-10, atol=0), x
    assert np.all(got[~big] < 1e-270)
    # e^x = I_0 + 2 sum_k I_k  (generating function at t = 1)
    assert abs(got[0] + 2 * got[1:].sum() - 1.0) < 1e-13


# ---------------------------------------------------------------------------
# one step against exp(-H dt)

def dense_h(p):
    H = p.oracle_ham()
    n = p.size
    M = np.zeros((n, n), dtype=complex)
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1
        M[:, j] = H.apply(e)
    return H, M


# QDng's series sums sum_k (2 - d_k0) I_k(z) e^{-z} (-1)^k T_k(x), x = (H -
# offset)/R in [0, 1), z = R dt, which is exp(-z (1 + x)): with the phase
# exp(-dt offset) one step gives exp(-R dt) exp(-H dt) psi.  The constant
# factor cancels in ProgEigen's normalisation; the sum cancels down to it, so
# rounding grows like exp(R dt) (R dt = 5.8 and 14.5 here).
@pytest.mark.parametrize("dt,tol", [(2.0, 1e-12), (5.0, 2e-9)])
def test_oracle_imag_step_is_exp_minus_h_dt(dt, tol):
    p = tge_problem()
    H, M = dense_h(p)
    assert np.allclose(M, M.conj().T, atol=1e-12 * np.abs(M).max())
    w, V = np.linalg.eigh(0.5 * (M + M.conj().T))
    U = O.Cheby(H, dt, prec=1e-14, imag=True)
    ref = np.exp(-U.scaling * dt) * (V @ (np.exp(-w * dt) * (V.conj().T @ p.psi0)))
    # QDng's imaginary-time parameters (OCheby.cpp:406-416)
    emin, emax = H.bounds()
    assert U.offset == pytest.approx(emin - abs(emin) * 0.05, rel=1e-15)
    assert U.scaling == pytest.approx(abs(emax - emin) * 1.05, rel=1e-15)
    c = U.coeffs()
    b, _ = O.bessel_i_scaled(len(c), U.scaling * dt)
    assert c[0] == b[0] and np.array_equal(c[1:], 2 * b[1:])
    psi = p.psi0.copy()
    U.step(psi, 0)
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < tol
    # the bar discriminates: exp(+H dt) (a sign mistake) differs in shape
    # from the result by far more than the tolerance
    bad = V @ (np.exp(+w * dt) * (V.conj().T @ p.psi0))
    unit = lambda v: v / np.linalg.norm(v)
    assert np.linalg.norm(unit(bad) - unit(ref)) > 1e3 * tol


# ---------------------------------------------------------------------------
# eigenstates by imaginary-time relaxation

def imag_time_ef(step, project_out, normalize, overlap, expec, matrix_element, psi0, nef, max_steps, conv):
    """ProgEigen::ImagTimeEF (ProgEigen.cpp:318-412) + DiagBasis (93-138) over
    callables, so that the oracle and the device run the same driver."""
    efs, e_raw, steps = [], [], []
    for i in range(nef):
        psi = psi0()
        s, diff = 0, 1.0
        while s < max_steps and diff > conv:
            old = psi.copy() if isinstance(psi, np.ndarray) else psi.clone()
            step(psi, s)
            if i > 0:
                project_out(psi, efs)
            normalize(psi)
            diff = abs(1 - overlap(old, psi))
            s += 1
        efs.append(psi)
        e_raw.append(expec(psi))
        steps.append(s)
    S = np.diag(e_raw)
    for i in range(nef):
        for j in range(i):
            S[j, i] = S[i, j] = matrix_element(efs[i], efs[j]).real
    e_diag = np.linalg.eigvalsh(S, UPLO="U")
    return np.array(e_raw), e_diag, steps


def assert_tge_energies(e):
    got = ["%.8f" % v for v in e]
    for g, want in zip(got, TGE_ENERGIES):
        assert g.startswith(want), (got, TGE_ENERGIES)


def test_oracle_eigenstates_match_qdng_tge():
    p = tge_problem()
    H = p.oracle_ham()
    g = p.oracle_grid()
    assert "%.6e" % H.expec(p.psi0) == TGE_E0
    U = O.Cheby(H, 5.0, imag=True)

    def step(psi, s):
        U.step(psi, s)

    def project_out(psi, efs):
        buf = np.zeros_like(psi)
        for ef in efs:
            buf += ef * O.overlap(g, ef, psi)
        psi -= buf

    def normalize(psi):
        psi /= np.sqrt(O.norm2(g, psi))

    e_raw, e_diag, steps = imag_time_ef(
        step, project_out, normalize, lambda a, b: O.overlap(g, a, b), lambda w: H.expec(w),
        lambda bra, ket: O.overlap(g, bra, H.apply(ket)), lambda: p.psi0.copy(), 10, 1000, 1e-5)
    assert_tge_energies(e_diag)
    assert max(steps) < 1000


# ---------------------------------------------------------------------------
# the library's own Bessel I (host, no device)

@pytest.mark.parametrize("x", [0.3, 2.0, 14.6, 156.6, 640.0])
def test_library_bessel_i0_against_mpmath(x):
    import qdng_amd as Q
    n = int(2 * x) + 60
    got, nz = Q.bessel_i0(n, x)
    ref = ive_mp(n, x)
    big = ref > 1e-280
    assert np.allclose(got[big], ref[big], rtol=1e-15, atol=0), x
    assert np.all(got[ref < 1e-300] < 1e-270)
    assert abs(got[0] + 2 * got[1:].sum() - 1.0) < 1e-14


# ---------------------------------------------------------------------------
# device against the oracle

def device_pair(ctx, p, dt, steps, prec=1e-10):
    import qdng_amd as Q
    U = Q.Cheby(p.gpu_ham(ctx), dt, prec=prec, imag=True)
    Uo = O.Cheby(p.oracle_ham(), dt, prec=prec, imag=True)
    assert U.order == Uo.order
    assert U.info.offset == pytest.approx(Uo.offset, rel=1e-14)
    assert U.info.scaling == pytest.approx(Uo.scaling, rel=1e-14)
    # (the oracle's libstdc++ values are good to ~1e-11 in the tail, the
    # library's to an ulp: both checked against mpmath above)
    assert np.allclose(U.coeffs(), Uo.coeffs(), rtol=1e-10, atol=1e-300)
    assert U.info.scaling * dt <= 15.0
    w = p.gpu_wf(ctx)
    psi = p.psi0.copy()
    counts = []
    for s in range(steps):
        st = U.apply(w, s)
        n = Uo.step(psi, s)
        counts.append((st.n_hpsi, n))
        # normalise both, as ProgEigen does (the constant exp(-R dt) would
        # underflow over many steps)
        w.normalize()
        psi /= np.sqrt(O.norm2(p.oracle_grid(), psi, p.S))
    return w.download(), psi, counts


# The series cancels by exp(-R dt) (above), so any two implementations agree
# only to ~1e-16 exp(R dt): the cases keep R dt <= 15 (QDng's TGE-Cheby runs
# at R dt = 14.5); at R dt = 29 (anharmonic_3d(16), dt 3.5) device and
# oracle differ by 5e-5 relative, both being rounding at that size.
@pytest.mark.gpu
@pytest.mark.parametrize("name,builder,dt,steps", [
    ("tge_1d", tge_problem, 2.0, 10),
    ("tge_1d_dt5", tge_problem, 5.0, 4),
    ("morse_2state_2d", lambda: cases.morse_2state_2d(64), 5.0, 3),
    ("anharmonic_3d", lambda: cases.anharmonic_3d(16), 1.0, 3),
])
def test_device_imag_step_matches_oracle(ctx, name, builder, dt, steps):
    p = builder()
    got, ref, counts = device_pair(ctx, p, dt, steps)
    assert all(a == b for a, b in counts), counts
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-10


@pytest.mark.gpu
def test_device_imag_rejects_fp32(ctx):
    import qdng_amd as Q
    p = tge_problem()
    with pytest.raises(Q.QdxError):
        Q.Cheby(p.gpu_ham(ctx), 5.0, imag=True, precision="fp32")


@pytest.mark.gpu
def test_device_eigenstates_match_qdng_tge(ctx):
    import qdng_amd as Q
    p = tge_problem()
    H = p.gpu_ham(ctx)
    psi0 = p.gpu_wf(ctx)
    assert "%.6e" % H.expec(psi0) == TGE_E0
    U = Q.Cheby(H, 5.0, imag=True)
    efs, e, e_raw, steps = Q.eigen_imag(U, psi0, 10, max_steps=1000, conv=1e-5, diag=True)
    assert_tge_energies(e)
    assert max(steps) < 1000
    # orthonormal eigenstates of the diagonalised basis
    g = p.oracle_grid()
    vecs = [w.download() for w in efs]
    ov = np.array([[O.overlap(g, a, b) for b in vecs] for a in vecs])
    assert np.allclose(ov, np.eye(10), atol=1e-9)
    Ho = p.oracle_ham()
    assert np.allclose([Ho.expec(v) for v in vecs], e, rtol=1e-8)
    # raw relaxation without the basis diagonalisation (QDng's conv 1e-8)
    # against the same driver on the oracle: energies and steps per state
    efs2, e2, e2_raw, steps2 = Q.eigen_imag(U, psi0, 3, max_steps=1000, diag=False)
    assert np.array_equal(e2, e2_raw)
    Uo = O.Cheby(Ho, 5.0, imag=True)

    def step(psi, s):
        Uo.step(psi, s)

    def project_out(psi, efs):
        buf = np.zeros_like(psi)
        for ef in efs:
            buf += ef * O.overlap(g, ef, psi)
        psi -= buf

    def normalize(psi):
        psi /= np.sqrt(O.norm2(g, psi))

    e_o, _, steps_o = imag_time_ef(
        step, project_out, normalize, lambda a, b: O.overlap(g, a, b), lambda w: Ho.expec(w),
        lambda bra, ket: O.overlap(g, bra, Ho.apply(ket)), lambda: p.psi0.copy(), 3, 1000, 1e-8)
    assert list(steps2) == list(steps_o)
    assert np.allclose(e2, e_o, rtol=1e-9, atol=0)


# ---------------------------------------------------------------------------
# short iterative Lanczos (OSIL.cpp:79-190), real and imaginary time

@pytest.mark.parametrize("imag,dt,order,tol", [(False, 0.5, 20, 1e-12), (False, 5.0, 25, 1e-9),
                                               (True, 5.0, 25, 1e-10)])
def test_oracle_sil_step_is_matrix_exponential(imag, dt, order, tol):
    p = tge_problem()
    H, M = dense_h(p)
    w, V = np.linalg.eigh(0.5 * (M + M.conj().T))
    f = np.exp(-w * dt) if imag else np.exp(-1j * w * dt)
    ref = V @ (f * (V.conj().T @ p.psi0))
    U = O.SIL(H, dt, order=order, imag=imag)
    psi = p.psi0.copy()
    n = U.step(psi, 0)
    assert 3 <= n <= order
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < tol
    # the wrong exponent sign is far outside the bar
    fbad = np.exp(+w * dt) if imag else np.conj(f)
    bad = V @ (fbad * (V.conj().T @ p.psi0))
    assert np.linalg.norm(bad / np.linalg.norm(bad) - ref / np.linalg.norm(ref)) > 1e3 * tol


def test_oracle_sil_eigenstates_match_qdng_tge():
    # TGE-SIL (tests/TGE-SIL.qdi.in: SIL order 25, dt 5) checks the same
    # energies as TGE-Cheby (checktests.sh:124-127)
    p = tge_problem()
    H = p.oracle_ham()
    g = p.oracle_grid()
    U = O.SIL(H, 5.0, order=25, imag=True)

    def project_out(psi, efs):
        buf = np.zeros_like(psi)
        for ef in efs:
            buf += ef * O.overlap(g, ef, psi)
        psi -= buf

    def normalize(psi):
        psi /= np.sqrt(O.norm2(g, psi))

    e_raw, e_diag, steps = imag_time_ef(
        lambda psi, s: U.step(psi, s), project_out, normalize, lambda a, b: O.overlap(g, a, b),
        lambda w: H.expec(w), lambda bra, ket: O.overlap(g, bra, H.apply(ket)), lambda: p.psi0.copy(),
        10, 1000, 1e-5)
    assert_tge_energies(e_diag)


@pytest.mark.gpu
@pytest.mark.parametrize("name,builder,dt,order,imag,steps", [
    ("tge_1d_real", tge_problem, 0.5, 20, False, 5),
    ("tge_1d_real_dt5", tge_problem, 5.0, 25, False, 4),
    ("tge_1d_imag", tge_problem, 5.0, 25, True, 4),
    ("morse_2state_2d_real", lambda: cases.morse_2state_2d(64), 1.0, 20, False, 3),
    ("anharmonic_3d_imag", lambda: cases.anharmonic_3d(16), 1.0, 20, True, 2),
])
def test_device_sil_matches_oracle(ctx, name, builder, dt, order, imag, steps):
    import qdng_amd as Q
    p = builder()
    U = Q.SIL(p.gpu_ham(ctx), dt, order=order, imag=imag)
    Uo = O.SIL(p.oracle_ham(), dt, order=order, imag=imag)
    w = p.gpu_wf(ctx)
    psi = p.psi0.copy()
    counts = []
    for s in range(steps):
        st = U.apply(w, s)
        n = Uo.step(psi, s)
        counts.append(n)
        assert st.n_hpsi == n, (s, st.n_hpsi, n)
    got = w.download()
    assert np.linalg.norm(got - psi) / np.linalg.norm(psi) < 1e-10
    # the run loop takes the same steps and reports the same H counts
    w2 = p.gpu_wf(ctx)
    _, _, nh, _ = U.run(w2, 0, steps, observables=False)
    assert list(nh) == counts
    assert np.array_equal(w2.download(), got)


@pytest.mark.gpu
def test_device_sil_eigenstates_match_qdng_tge(ctx):
    import qdng_amd as Q
    p = tge_problem()
    U = Q.SIL(p.gpu_ham(ctx), 5.0, order=25, imag=True)
    efs, e, e_raw, steps = Q.eigen_imag(U, p.gpu_wf(ctx), 10, max_steps=1000, conv=1e-5, diag=True)
    assert_tge_energies(e)
    with pytest.raises(Q.QdxError):
        Q.eigen_imag(Q.SIL(p.gpu_ham(ctx), 5.0, order=25), p.gpu_wf(ctx), 2)   # real time: rejected


# ---------------------------------------------------------------------------
# GSPO in imaginary time (exponent -dt): every split operator a real decay

def test_oracle_gspo_imag_is_strang_product():
    # ops (kin, pot): exp(-V dt/2) exp(-T dt) exp(-V dt/2) (OGSPO.cpp:114-139)
    p = tge_problem()
    kin = p.oracle_split_op(["kin"])
    pot = p.oracle_split_op(["pot"])
    n = p.size
    Tm = np.zeros((n, n), dtype=complex)
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1
        Tm[:, j] = kin.apply(e)
    v = pot.apply(np.ones(n, dtype=complex)).real
    wt, Vt = np.linalg.eigh(0.5 * (Tm + Tm.conj().T))
    dt = 5.0
    half = np.exp(-0.5 * dt * v)
    ref = half * (Vt @ (np.exp(-dt * wt) * (Vt.conj().T @ (half * p.psi0))))
    G = O.GSPO([kin, pot], dt, 1, imag=True)
    psi = p.psi0.copy()
    G.step(psi, 0)
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < 1e-12
    # real time differs completely
    Gr = O.GSPO([kin, pot], dt, 1)
    psr = p.psi0.copy()
    Gr.step(psr, 0)
    assert np.linalg.norm(psr - ref) > 0.1 * np.linalg.norm(ref)


def test_oracle_gspo_eigenstates_match_qdng_tge():
    # TGE-GSPO (tests/TGE-GSPO.qdi.in: A kinetic, B potential, dt 5) checks
    # the same energies as TGE-Cheby (checktests.sh:124-127)
    p = tge_problem()
    H = p.oracle_ham()
    g = p.oracle_grid()
    G = O.GSPO([p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])], 5.0, 1, imag=True)

    def project_out(psi, efs):
        buf = np.zeros_like(psi)
        for ef in efs:
            buf += ef * O.overlap(g, ef, psi)
        psi -= buf

    def normalize(psi):
        psi /= np.sqrt(O.norm2(g, psi))

    e_raw, e_diag, steps = imag_time_ef(
        lambda psi, s: G.step(psi, s), project_out, normalize, lambda a, b: O.overlap(g, a, b),
        lambda w: H.expec(w), lambda bra, ket: O.overlap(g, bra, H.apply(ket)), lambda: p.psi0.copy(),
        10, 1000, 1e-5)
    assert_tge_energies(e_diag)


@pytest.mark.gpu
@pytest.mark.parametrize("name,builder,dt,steps", [
    ("tge_1d", tge_problem, 5.0, 4),
    ("morse_2state_2d_coupled", lambda: cases.morse_2state_2d(64), 5.0, 3),
    ("anharmonic_3d", lambda: cases.anharmonic_3d(16), 3.5, 2),
])
def test_device_gspo_imag_matches_oracle(ctx, name, builder, dt, steps):
    import qdng_amd as Q
    p = builder()
    G = Q.GSPO([p.gpu_split_op(ctx, ["kin"]), p.gpu_split_op(ctx, ["pot"])], dt, imag=True)
    Go = O.GSPO([p.oracle_split_op(["kin"]), p.oracle_split_op(["pot"])], dt, p.S, imag=True)
    w = p.gpu_wf(ctx)
    psi = p.psi0.copy()
    for s in range(steps):
        G.apply(w, s)
        Go.step(psi, s)
    assert np.linalg.norm(w.download() - psi) / np.linalg.norm(psi) < 1e-12


@pytest.mark.gpu
def test_device_gspo_eigenstates_match_qdng_tge(ctx):
    import qdng_amd as Q
    p = tge_problem()
    G = Q.GSPO([p.gpu_split_op(ctx, ["kin"]), p.gpu_split_op(ctx, ["pot"])], 5.0, imag=True)
    with pytest.raises(Q.QdxError):   # split operators form no Hamiltonian: H is required
        Q.eigen_imag(G, p.gpu_wf(ctx), 2)
    efs, e, e_raw, steps = Q.eigen_imag(G, p.gpu_wf(ctx), 10, max_steps=1000, conv=1e-5, diag=True,
                                        H=p.gpu_ham(ctx))
    assert_tge_energies(e)


# ---------------------------------------------------------------------------
# Arnoldi (OArnoldi.cpp:64-190), real and imaginary time

@pytest.mark.parametrize("imag,dt,order,tol", [(False, 0.5, 20, 1e-12), (False, 5.0, 20, 1e-8),
                                               (True, 5.0, 15, 1e-9)])
def test_oracle_arnoldi_step_is_matrix_exponential(imag, dt, order, tol):
    p = tge_problem()
    H, M = dense_h(p)
    w, V = np.linalg.eigh(0.5 * (M + M.conj().T))
    f = np.exp(-w * dt) if imag else np.exp(-1j * w * dt)
    ref = V @ (f * (V.conj().T @ p.psi0))
    U = O.Arnoldi(H, dt, order=order, imag=imag)
    psi = p.psi0.copy()
    n = U.step(psi, 0)
    assert 2 <= n <= order
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < tol
    fbad = np.exp(+w * dt) if imag else np.conj(f)
    bad = V @ (fbad * (V.conj().T @ p.psi0))
    assert np.linalg.norm(bad / np.linalg.norm(bad) - ref / np.linalg.norm(ref)) > 1e3 * tol


def test_oracle_arnoldi_eigenstates_match_qdng_tge():
    # TGE-Arnoldi (tests/TGE-Arnoldi.qdi.in: order 15, dt 5) checks the same
    # energies as TGE-Cheby (checktests.sh:124-127)
    p = tge_problem()
    H = p.oracle_ham()
    g = p.oracle_grid()
    U = O.Arnoldi(H, 5.0, order=15, imag=True)

    def project_out(psi, efs):
        buf = np.zeros_like(psi)
        for ef in efs:
            buf += ef * O.overlap(g, ef, psi)
        psi -= buf

    def normalize(psi):
        psi /= np.sqrt(O.norm2(g, psi))

    e_raw, e_diag, steps = imag_time_ef(
        lambda psi, s: U.step(psi, s), project_out, normalize, lambda a, b: O.overlap(g, a, b),
        lambda w: H.expec(w), lambda bra, ket: O.overlap(g, bra, H.apply(ket)), lambda: p.psi0.copy(),
        10, 1000, 1e-5)
    assert_tge_energies(e_diag)


@pytest.mark.gpu
@pytest.mark.parametrize("name,builder,dt,order,imag,steps", [
    ("tge_1d_real", tge_problem, 0.5, 20, False, 4),
    ("tge_1d_imag", tge_problem, 5.0, 15, True, 4),
    ("morse_2state_2d_real", lambda: cases.morse_2state_2d(64), 1.0, 20, False, 3),
    ("anharmonic_3d_imag", lambda: cases.anharmonic_3d(16), 1.0, 20, True, 2),
])
def test_device_arnoldi_matches_oracle(ctx, name, builder, dt, order, imag, steps):
    import qdng_amd as Q
    p = builder()
    U = Q.Arnoldi(p.gpu_ham(ctx), dt, order=order, imag=imag)
    Uo = O.Arnoldi(p.oracle_ham(), dt, order=order, imag=imag)
    w = p.gpu_wf(ctx)
    psi = p.psi0.copy()
    for s in range(steps):
        st = U.apply(w, s)
        n = Uo.step(psi, s)
        assert st.n_hpsi == n, (s, st.n_hpsi, n)
    assert np.linalg.norm(w.download() - psi) / np.linalg.norm(psi) < 1e-10


@pytest.mark.gpu
def test_device_arnoldi_eigenstates_match_qdng_tge(ctx):
    import qdng_amd as Q
    p = tge_problem()
    U = Q.Arnoldi(p.gpu_ham(ctx), 5.0, order=15, imag=True)
    efs, e, e_raw, steps = Q.eigen_imag(U, p.gpu_wf(ctx), 10, max_steps=1000, conv=1e-5, diag=True)
    assert_tge_energies(e)
