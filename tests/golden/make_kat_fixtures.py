"""Extract QDng's regression-test inputs and golden lines into small fixtures.

Reads the QDng source tree (default /root/reference) and writes
  tests/golden/qdng_kat.npz   the binary inputs the regression tests use:
                              potentials, dipoles, laser fields, eigenfunctions
                              (raw little-endian float64 / complex128, sized by
                              their .meta files; FileSingle.h:1237-1286)
  tests/golden/qdng_kat.json  grid geometry from the .meta files and the golden
                              lines of tests/checktests.sh:129-152 parsed into
                              numbers

Only data is copied.  Run once in a container where the QDng tree exists:
    python tests/golden/make_kat_fixtures.py [/path/to/QDng]
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# regression tests that exercise only the grid Chebyshev / GSPO path
IN_SCOPE = ["TGP-Expec", "TGPMS-Cheby", "TGPMS-GSPO", "TGPL-IR-Pulse-Cheby",
            "TGPL-IR-Pulse-GSPO", "TGPLMS-Cheby", "TGP-NIP-Cheby"]


def read_meta(path):
    meta = {}
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#") or "=" not in line:
            continue
        k, v = line.split("=", 1)
        meta[k.strip()] = v.strip()
    return meta


def main(root):
    tdir = os.path.join(root, "tests")
    arrays = {}
    grids = {}
    # grid operators: 128 real values each
    for name in ["GridPot0", "GridPot1", "GridDip0", "GridDip01"]:
        meta = read_meta(os.path.join(tdir, name + ".meta"))
        n = int(meta["N0"])
        arrays[name] = np.fromfile(os.path.join(tdir, name + ".op"), dtype="<f8", count=n)
        grids[name] = {"N": [n], "xmin": [float(meta["xmin0"])], "xmax": [float(meta["xmax0"])]}
    meta = read_meta(os.path.join(tdir, "GridPot0_ef.meta"))
    n = int(meta["N0"])
    for i in range(4):
        arrays["GridPot0_ef_%d" % i] = np.fromfile(os.path.join(tdir, "GridPot0_ef_%d.wf" % i),
                                                   dtype="<c16", count=n)
    for name in ["GridLaser_w0.00713", "GridLaser_w0.176"]:
        meta = read_meta(os.path.join(tdir, name + ".meta"))
        nt = int(meta["Nt"])
        arrays[name] = np.fromfile(os.path.join(tdir, name + ".op"), dtype="<f8", count=nt)
        grids[name] = {"Nt": nt, "dt": float(meta["dt"])}
    for k, v in arrays.items():
        assert np.all(np.isfinite(v)), k

    # golden report lines: step time population [pops...] energy re,imi
    golden = {}
    text = open(os.path.join(tdir, "checktests.sh")).read()
    for m in re.finditer(r'^CheckT\s+(\S+)\s+"([^"]+)"', text, re.M):
        name, line = m.group(1), m.group(2)
        if name not in IN_SCOPE:
            continue
        toks = line.split()
        re_im = toks[-1].rstrip("i").split(",")
        golden.setdefault(name, {})["report"] = {
            "step": int(toks[0]), "time": float(toks[1]),
            "population": float(toks[2]),
            "pops": [float(t) for t in toks[3:-2]],
            "energy": float(toks[-2]),
            "autocorr": [float(re_im[0]), float(re_im[1])],
            "energy_digits": len(toks[-2].split(".")[1]),
            "source": "tests/checktests.sh",
        }
    for m in re.finditer(r'^CheckT-Expec\s+(\S+)\s+"([^"]+)"', text, re.M):
        name, line = m.group(1), m.group(2)
        if name not in golden:
            continue
        toks = line.split()
        # a trailing [0-9]* in the grep pattern means "printed value starts
        # with these digits" (truncation); otherwise the digits are complete
        vals = [t.replace("[0-9]*", "").replace("[0-9]", "") for t in toks]
        prefix = [("[0-9]" in t) for t in toks]
        golden[name]["expec_post"] = {"time": float(vals[0]),
                                      "values": [v for v in vals[1:]],
                                      "prefix": prefix[1:]}

    np.savez_compressed(os.path.join(HERE, "qdng_kat.npz"), **arrays)
    with open(os.path.join(HERE, "qdng_kat.json"), "w") as f:
        json.dump({"grids": grids, "golden": golden}, f, indent=1, sort_keys=True)
    print("wrote", sorted(arrays), sorted(golden))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
