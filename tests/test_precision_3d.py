"""Precision variant: 3D 128^3 single-state anharmonic oscillator on [-6, 6]^3,
Chebyshev at order ~64 (dt 3.5) and ~256 (dt 37), 20 steps, the bulk paths
against the plain FP64 path on the same complex64 initial state.

- "compact": FP64 arithmetic with the recurrence terms and the kinetic pass
  buffer stored in 40 bits per component (relative rounding 2^-29); it meets
  the 1e-6 relative-L2 bar at both orders.
- "fp32": FP32 transforms and terms.  Its error is dominated by the FP32
  arithmetic of the transforms (a CPU model with FP64 transforms and FP32
  terms gives a fifth of it) and grows like the square root of the number of
  steps; at order ~256 it is above the 1e-6 bar (measured 4.5e-6 on MI355X):
  that case is an expected failure that reports its measured error.

The FP64 device path is the reference at this size; it is itself tied to the
FP64 oracle (<= 1e-10) on the same Hamiltonian at 24^3 and 32^3 in
test_gpu_parity.py (anharmonic_3d cases at both time steps).

Reported per case: relative L2 difference of psi after 20 steps, max norm
drift |<psi(t)|psi(t)> - <psi(0)|psi(0)>| over the steps for both paths, the
largest autocorrelation difference, and the series lengths.
"""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

BAR = cases.BULK_BAR   # relative L2 of the bulk paths against FP64


def _run(ctx, p, dt, precision, steps):
    import qdng_amd as Q
    U = Q.Cheby(p.gpu_ham(ctx), dt, prec=1e-10, precision=precision)
    w, w0 = p.gpu_wf(ctx), p.gpu_wf(ctx)
    n2, ac, nh, st = U.run(w, 0, steps, w0)
    # norm of the final state (run() records the norm before each step)
    n_end = w.norm()[0] ** 2
    return w.download(), np.append(n2, n_end), ac, nh, U.order


_FP64 = {}


def _fp64(ctx, p, dt, steps):
    if dt not in _FP64:
        _FP64[dt] = _run(ctx, p, dt, "fp64", steps)
    return _FP64[dt]


@pytest.mark.parametrize("precision", ["compact", "fp32"])
@pytest.mark.parametrize("dt,order_lo,order_hi", [(3.5, 48, 80), (37.0, 224, 288)])
def test_bulk_vs_fp64_3d_128(ctx, dt, order_lo, order_hi, precision):
    print(f"[gpu case] precision 3D 128^3 dt={dt} {precision}", flush=True)
    p = cases.anharmonic_3d(128)
    p.psi0 = p.psi0.astype(np.complex64).astype(np.complex128)   # complex64 initial state
    steps = 20
    g64, n64, ac64, nh64, order64 = _fp64(ctx, p, dt, steps)
    gb, nb, acb, nhb, orderb = _run(ctx, p, dt, precision, steps)
    assert order_lo <= order64 <= order_hi
    err = np.linalg.norm(gb - g64) / np.linalg.norm(g64)
    drift64 = np.max(np.abs(n64 - n64[0]))
    driftb = np.max(np.abs(nb - nb[0]))
    dac = np.max(np.abs(acb - ac64))
    print(f"[precision] dt={dt} {precision} order fp64 {order64} bulk {orderb} "
          f"hpsi/step fp64 {nh64.min()}-{nh64.max()} bulk {nhb.min()}-{nhb.max()} "
          f"relL2 {err:.3e} drift64 {drift64:.3e} drift {driftb:.3e} autocorr {dac:.3e}", flush=True)
    sys.stdout.flush()
    assert drift64 < 1e-8
    if precision == "compact":
        # the difference is mostly the FP64 series' truncation: it stops at
        # QDng's norm test (prec 1e-10), the compact series runs its fixed
        # length (a few more terms); the 40-bit storage adds ~1e-8
        assert err < BAR
        assert driftb < 1e-8
        assert dac < 1e-7
        return
    assert driftb < 1e-6
    assert dac < 1e-6
    # FP32 arithmetic at order ~256 misses the bar (measured 4.5e-6 after
    # 20 steps: ~1e-6 per sqrt(step) at this series length), reported as an
    # expected failure with its value
    cases.bulk_bar(err, precision, f" (128^3, dt {dt}, order {order64}, {steps} steps)")
