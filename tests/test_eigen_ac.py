"""Eigenstates from the autocorrelation (ProgEigen method "ac",
ProgEigen::AutoCorrEF, src/ProgEigen.cpp:414-600; PeakFinder,
src/tools/PeakFinder.cpp).

What pins it: the peak finder on hand-built spectra whose peaks are known;
the oracle chain (real-time Chebyshev, Hann window, FFT, peaks, integration)
on QDng's TGE input (the LC of the six lowest eigenfunctions of the
discretised H) against the energies CheckTGE-ac-Energy greps for
(tests/checktests.sh:31-45); the device driver (qdx_eigen_ac) against the
oracle on the same input: the same peaks and autocorrelation and energies to
1e-10."""
import numpy as np
import pytest

from oracle import oracle as O
import test_imag_time as T

TGE_AC_ENERGIES = ["0.003714", "0.011074", "0.018152", "0.024892", "0.031319", "0.037427"]


def test_peak_regions():
    # two bumps above mean + sigma: one maximum each
    d = [0, 1, 5, 1, 0, 0, 3, 0, 0, 0]
    assert O.peaks_regions(d) == [2, 6]
    # a run open at the end counts only when its last value is not the maximum
    assert O.peaks_regions([0, 0, 0, 0, 0, 0, 4, 6, 5]) == [7]
    assert O.peaks_regions([0, 0, 0, 0, 0, 0, 4, 5, 6]) == []
    # a higher tolerance drops the small bump
    assert O.peaks_regions(d, tol=2.0) == [2]


def test_peak_window():
    # a small local maximum beside a large one: the region search merges
    # them into one run, the windowed search keeps both
    d = np.array([0, 0, 0, 1, 10, 4, 6, 1, 0, 0, 0, 0, 0, 0, 0, 0], dtype=float)
    assert O.peaks_regions(d, tol=0.2) == [4]
    assert O.peaks_window(d, 0.2, 2) == [4, 6]
    # plateaus and the first/last points are no strict maxima
    assert O.peaks_window([3, 1, 2, 2, 1, 0, 5], 0.0, 3) == []


def tge_ac_input():
    p = T.tge_problem()
    g = p.oracle_grid()
    _, M = T.dense_h(p)
    w, V = np.linalg.eigh(0.5 * (M + M.conj().T))
    efs = []
    for k in range(6):
        v = V[:, k].astype(complex)
        efs.append(v / np.sqrt(O.norm2(g, v)))
    psi0 = 2 * sum(efs)
    return p, psi0 / np.sqrt(O.norm2(g, psi0))


def oracle_ac(p, psi0, nsteps=2048, dt=5.0):
    H, g = p.oracle_ham(), p.oracle_grid()
    U = O.Cheby(H, dt)
    return O.autocorr_ef(lambda ps, k: U.step(ps, k), lambda a, b: O.overlap(g, a, b), lambda f: H.expec(f),
                         psi0, nsteps, dt)


def test_oracle_ac_matches_qdng_tge():
    p, psi0 = tge_ac_input()
    e, pk, ac, efs = oracle_ac(p, psi0)
    assert len(pk) == 6
    got = ["%.8f" % v for v in e]
    for g_, want in zip(got, TGE_AC_ENERGIES):
        assert g_.startswith(want), (got, TGE_AC_ENERGIES)


@pytest.mark.gpu
def test_device_ac_matches_oracle(ctx):
    import qdng_amd as Q
    print("[gpu case] eigen ac TGE", flush=True)
    p, psi0 = tge_ac_input()
    e_o, pk, ac_o, efs_o = oracle_ac(p, psi0)
    H = p.gpu_ham(ctx)
    U = Q.Cheby(H, 5.0)
    w0 = p.gpu_wf(ctx, psi0)
    efs, e, er, ac = Q.eigen_ac(U, w0, 2048)
    assert len(efs) == len(pk) == 6
    assert np.allclose(ac, ac_o, rtol=0, atol=1e-10)
    assert np.allclose(er, e_o, rtol=1e-10, atol=0)
    assert np.array_equal(e, er)           # diag off: the raw energies
    g = p.oracle_grid()
    for f, fo in zip(efs, efs_o):
        x = f.download()
        assert abs(abs(O.overlap(g, fo, x)) - 1) < 1e-10
    # diag on: DiagBasis over the six functions
    efs2, e2, er2, _ = Q.eigen_ac(U, w0, 2048, diag=True)
    assert np.allclose(er2, er, rtol=1e-10, atol=0)
    Ho = p.oracle_ham()
    S = np.diag(e_o)
    for i in range(6):
        for j in range(i):
            S[i, j] = S[j, i] = O.overlap(g, efs_o[i], Ho.apply(efs_o[j])).real
    assert np.allclose(e2, np.linalg.eigvalsh(S), rtol=1e-9, atol=0)
    # errors: imaginary-time propagator, too few steps, too many peaks
    Ui = Q.Cheby(H, 5.0, imag=True)
    with pytest.raises(Q.QdxError):
        Q.eigen_ac(Ui, w0, 64)
    with pytest.raises(Q.QdxError):
        Q.eigen_ac(U, w0, 1)
    with pytest.raises(Q.QdxError):
        Q.eigen_ac(U, w0, 2048, max_ef=3)


# the other branches of the device driver: a record length that is not a
# power of two (direct transform), no window, the windowed peak search
@pytest.mark.gpu
@pytest.mark.parametrize("nsteps,win,width,tol", [(1500, True, -1, 1.0), (1024, False, -1, 1.0),
                                                  (2048, True, 6, 2.0), (1536, False, 10, 1.5)])
def test_device_ac_variants_match_oracle(ctx, nsteps, win, width, tol):
    import qdng_amd as Q
    print("[gpu case] eigen ac", nsteps, win, width, tol, flush=True)
    p, psi0 = tge_ac_input()
    H, g = p.oracle_ham(), p.oracle_grid()
    U = O.Cheby(H, 5.0)
    e_o, pk, ac_o, _ = O.autocorr_ef(lambda ps, k: U.step(ps, k), lambda a, b: O.overlap(g, a, b),
                                     lambda f: H.expec(f), psi0, nsteps, 5.0, win=win, tol=tol, width=width)
    assert len(pk) >= 1
    Ud = Q.Cheby(p.gpu_ham(ctx), 5.0)
    efs, e, er, ac = Q.eigen_ac(Ud, p.gpu_wf(ctx, psi0), nsteps, win=win, tol=tol, width=width)
    assert len(efs) == len(pk), (len(efs), pk)
    assert np.allclose(ac, ac_o, rtol=0, atol=1e-10)
    assert np.allclose(er, e_o, rtol=1e-10, atol=0)
