"""CPU checks of libqdx.so: it loads, exports every entry point of
include/qdx.h and include/qdx_io.h, and its host-side rules (QDng bounds, Chebyshev order and
coefficients, Bessel values, argument checks) agree with the oracle.  No
device call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cases
import qdng_amd as Q
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for h in ("qdx.h", "qdx_io.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        syms |= set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(qdx_\w+)\s*\(", text, re.M))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 30
    L = C.CDLL(Q.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(Q.EXPORTS) <= set(syms)


def test_no_device_context_without_gpu():
    # on a machine without a HIP device the library must fail loudly
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(Q.QdxError):
        Q.Context(0)


def host_ham(p):
    H = Q.Hamiltonian(None, p.N, p.xmin, p.xmax, p.S, p.multistate, p.unity)
    for t in p.terms:
        p._gpu_add(H, t)
    return H


def problems():
    out = {}
    for name in cases.KAT_NAMES:
        p, kind, dt, steps, prec = cases.kat_problem(name)
        if kind == "cheby" and name != "TGP-NIP-Cheby":   # complex bounds: its own test below
            out[name] = (p, dt, prec)
    out["harmonic_1d"] = (cases.harmonic_1d(256), 10.0, 1e-12)
    out["morse_2state_2d"] = (cases.morse_2state_2d(64), 5.0, 1e-10)
    out["anharmonic_3d"] = (cases.anharmonic_3d(16), 3.5, 1e-10)
    out["anharmonic_3d_long"] = (cases.anharmonic_3d(16), 37.0, 1e-10)
    out["laser_1d"] = (cases.laser_1d(128), 1.0, 1e-12)
    return out


@pytest.mark.parametrize("name", list(problems().keys()))
def test_bounds_and_chebyshev_parameters_match_oracle(name):
    O.reset()
    p, dt, prec = problems()[name]
    H = host_ham(p)
    Ho = p.oracle_ham()
    emin, emax = H.bounds()
    omin, omax = Ho.bounds()
    assert emin == pytest.approx(omin, rel=1e-14, abs=1e-300)
    assert emax == pytest.approx(omax, rel=1e-14)
    info, coeff = H.cheby_params(dt, prec=prec)
    U = O.Cheby(Ho, dt, prec=prec)
    assert info.order == U.order
    assert info.offset == pytest.approx(U.offset, rel=1e-14)
    assert info.scaling == pytest.approx(U.scaling, rel=1e-14)
    assert info.prec == U.prec
    ref = U.coeffs()
    # libstdc++'s cyl_bessel_j (oracle) is good to ~1e-11 near zeros of J_k at
    # large arguments; the library's Miller recurrence agrees with scipy to 1e-12
    assert np.allclose(coeff, ref, rtol=1e-10, atol=1e-13 * np.abs(ref).max())


def test_multistate_bounds_use_absolute_values():
    # OMultistate.cpp:183-204 diagonalises |Emin_ij|: a negative potential
    # minimum enters as a positive number
    p = cases.Problem([16], [0.0], [1.0], nstates=2)
    p.potential(0, 0, np.linspace(-0.3, 0.2, 16)).potential(1, 1, np.linspace(0.1, 0.4, 16))
    p.potential(0, 1, np.full(16, 0.05))
    O.reset()
    emin, emax = host_ham(p).bounds()
    omin, omax = p.oracle_ham().bounds()
    assert emin == pytest.approx(omin) and emax == pytest.approx(omax)
    assert emin > 0


def test_forced_order_and_scaling():
    O.reset()
    p = cases.harmonic_1d(128)
    info, coeff = host_ham(p).cheby_params(10.0, order=90, scaling=2.5)
    U = O.Cheby(p.oracle_ham(), 10.0, order=90, scaling=2.5)
    assert info.order == U.order == 90 - info.nz
    assert info.scaling == 2.5
    assert np.allclose(coeff, U.coeffs(), rtol=1e-11, atol=1e-300)


def test_order_too_small_is_rejected():
    p = cases.harmonic_1d(128)
    with pytest.raises(Q.QdxError):
        host_ham(p).cheby_params(10.0, order=5)   # order <= R*dt (OCheby.cpp:467-468)


def test_bessel_against_scipy_vectors():
    d = np.load(cases.GOLDEN + "/bessel_jv.npz")
    for i, x in enumerate(d["x"]):
        ref = d["j%d" % i]
        got, nz = Q.bessel_j0(len(ref), float(x))
        big = np.abs(ref) > 1e-280
        assert np.allclose(got[big], ref[big], rtol=1e-9, atol=1e-15 * np.abs(ref).max()), x
        assert np.all(np.abs(got[~big]) < 1e-270)
        assert nz == np.sum(np.cumprod((np.abs(got) == 0)[::-1]))


@pytest.mark.parametrize("N", [[6], [12], [100], [8192], [1], [64, 10]])
def test_invalid_grids_are_rejected(N):
    with pytest.raises(Q.QdxError):
        Q.Hamiltonian(None, N, [0.0] * len(N), [1.0] * len(N))


def test_argument_errors():
    with pytest.raises(Q.QdxError):
        Q.Hamiltonian(None, [16], [0.0], [1.0], nstates=5, multistate=True)
    with pytest.raises(Q.QdxError):
        Q.Hamiltonian(None, [16], [0.0], [1.0], nstates=2, multistate=False)
    H = Q.Hamiltonian(None, [16], [0.0], [1.0])
    with pytest.raises(Q.QdxError):
        H.add_kinetic(0, [0.0])                  # Zero mass defined
    with pytest.raises(Q.QdxError):
        H.add_potential(1, 0, np.zeros(16))       # index out of range
    with pytest.raises(Q.QdxError):
        H.bounds()                                 # empty operator
    H.add_potential(0, 0, np.zeros(16))
    with pytest.raises(Q.QdxError):
        H.cheby_params(0.0)                        # No time step defined


def test_gobbler_chebyshev_parameters_match_oracle():
    # a negative imaginary potential makes Emin complex (OGobbler.cpp:183-192):
    # offset (Re(Emax + Emin) / 2, Im(Emax + Emin)), R = 1.05 |Emax - Emin|
    # and the whole series (no break test, OCheby.cpp:181, 406-416)
    O.reset()
    p, kind, dt, steps, prec = cases.kat_problem("TGP-NIP-Cheby")
    H = host_ham(p)
    Ho = p.oracle_ham()
    omin, omax = Ho.bounds_c()
    assert omin.imag == -0.02 and omax.imag == 0
    emin, emax = H.bounds()
    assert emin == omin.real and emax == pytest.approx(omax.real, rel=1e-15)
    info, coeff = H.cheby_params(dt, prec=prec)
    U = O.Cheby(Ho, dt, prec=prec)
    assert info.order == U.order
    assert info.offset == pytest.approx(U.offset, rel=1e-15)
    assert info.offset_i == -0.02
    assert info.scaling == pytest.approx(U.scaling, rel=1e-15)
    assert info.scaling == pytest.approx(1.05 * abs(omax - omin), rel=1e-15)
    assert info.prec == 0
    assert np.allclose(coeff, U.coeffs(), rtol=1e-10, atol=1e-13 * np.abs(U.coeffs()).max())


def test_gobbler_argument_errors():
    p = cases.Problem([64], [0.0], [4.0])
    H = Q.Hamiltonian(None, p.N, p.xmin, p.xmax)
    with pytest.raises(Q.QdxError) as e:
        H.add_gobbler(0, 0, rp={0: 9.0})            # edge beyond the grid (OGobbler.cpp:63-64)
    assert e.value.code == -1
    with pytest.raises(Q.QdxError):
        H.add_gobbler(0, 0, lp={0: 2.0}, order=0)  # "Filter order to small"
    M = Q.Hamiltonian(None, p.N, p.xmin, p.xmax, 2, True)
    with pytest.raises(Q.QdxError) as e:
        M.add_gobbler(0, 0, rp={0: 3.0}, nip=True)  # NIP: single-state Sum only
    assert e.value.code == -2
    # a real Gobbler is a potential with the bounds gain and 0 anywhere
    M.add_gobbler(1, 1, rp={0: 3.0}, gain=0.5).add_kinetic(0, [1.0]).add_kinetic(1, [1.0])
    O.reset()
    po = cases.Problem(p.N, p.xmin, p.xmax, nstates=2)
    po.gobbler(1, 1, rp={0: 3.0}, gain=0.5).kinetic(0, [1.0]).kinetic(1, [1.0])
    assert M.bounds() == pytest.approx(po.oracle_ham().bounds(), rel=1e-14)
