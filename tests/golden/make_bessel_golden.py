"""Golden Bessel vectors J_k(x) from scipy.special.jv (AMOS zbesj, the
algorithm family QDng calls through SLATEC ZBESJ in linalg/Bessel.cpp:22-58).

Writes tests/golden/bessel_jv.npz with, for each argument x, J_k(x) for
k = 0..K-1.  The arguments are the R*dt values of the benchmark
configurations (1D harmonic dt=10, 2D two-state dt=5, 3D dt=3.5 and dt=37,
the QDng regression tests) plus small and large cases.
"""
import os

import numpy as np
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
XS = [0.5, 1.0, 7.3, 13.7766, 14.8, 19.41, 27.55, 37.0, 74.6, 156.6, 300.0]
K = 512


def main():
    ks = np.arange(K)
    out = {"x": np.array(XS), "k": ks}
    for i, x in enumerate(XS):
        out["j%d" % i] = sp.jv(ks, x)
    np.savez_compressed(os.path.join(HERE, "bessel_jv.npz"), **out)
    print("wrote", len(XS), "vectors; scipy", sp.__name__)


if __name__ == "__main__":
    main()
