"""qdx_ham_expec_noalloc (the allocation-free <psi|H|psi>) against
qdx_ham_expec, both called through the C ABI: two calls in a row, the
second on a smaller grid, then a third on a two-state 2D grid; the size
limit and the set-up rule."""
import ctypes as C

import numpy as np
import pytest

import cases
import qdng_amd as Q

pytestmark = pytest.mark.gpu


def _noalloc(H, w, step):
    out = C.c_double(np.nan)
    rc = Q.lib().qdx_ham_expec_noalloc(H.h, w.h, int(step), C.byref(out))
    return rc, out.value


def _expec(H, w, step):
    out = C.c_double(np.nan)
    rc = Q.lib().qdx_ham_expec(H.h, w.h, int(step), C.byref(out))
    assert rc == 0, Q.lib().qdx_last_error().decode()
    return out.value


def test_noalloc_expec_matches_expec(ctx):
    print("[gpu case] qdx_ham_expec_noalloc: 512^2 x2, 1D 512 laser, 2D 64 x2", flush=True)
    runs = [(cases.morse_2state_2d(512), 0),          # S * N = 524288, the limit
            (cases.laser_1d(512), 7),                 # smaller, time-dependent (dipole)
            (cases.morse_2state_2d(64), 0)]
    for p, step in runs:
        H = p.gpu_ham(ctx)
        w = p.gpu_wf(ctx)
        # before set-up (the first apply uploads the operator arrays)
        rc, _ = _noalloc(H, w, step)
        assert rc == -6                   # QDX_ERR_STATE
        ref = _expec(H, w, step)
        for _ in range(2):
            rc, got = _noalloc(H, w, step)
            assert rc == 0, Q.lib().qdx_last_error().decode()
            # the same kernels in the same order: equal to the last bit
            # expected; FP64 bar 1e-14 relative
            assert abs(got - ref) <= 1e-14 * abs(ref), (p.N, got, ref)
        assert np.isfinite(ref) and ref != 0.0


def test_noalloc_expec_rejects_oversize(ctx):
    print("[gpu case] qdx_ham_expec_noalloc: 3D 128^3 over the limit", flush=True)
    p = cases.anharmonic_3d(128)          # 2097152 points > 524288
    H = p.gpu_ham(ctx)
    w = p.gpu_wf(ctx)
    rc, _ = _noalloc(H, w, 0)
    assert rc == -4                       # QDX_ERR_MEMORY
    assert "exceed" in Q.lib().qdx_last_error().decode()
