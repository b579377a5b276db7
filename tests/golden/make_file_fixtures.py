"""Copy a few of QDng's regression-test data files, unchanged, as fixtures of
the file-format tests (tests/test_io.py) and of the program-input tests
(tests/test_qdi.py): raw .op / .wf data and their .meta files (potential,
dipole, laser, initial wave function, eigenfunctions) and the XML program
inputs of the grid regression tests.

Only data is copied.  Run once in a container where the QDng tree exists:
    python tests/golden/make_file_fixtures.py [/path/to/QDng]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["GridPot0.meta", "GridPot0.op", "GridPot1.meta", "GridPot1.op", "GridDip0.meta", "GridDip0.op",
         "GridGauss.meta", "GridGauss.wf", "GridPot0_ef.meta", "GridPot0_ef_0.wf", "GridPot0_ef_1.wf",
         "GridLaser_w0.176.meta", "GridLaser_w0.176.op",
         "GridPot0_ef_2.wf", "GridPot0_ef_3.wf", "GridDip01.meta", "GridDip01.op",
         "GridLaser_w0.00713.meta", "GridLaser_w0.00713.op"]
# the program inputs of the grid regression tests (XML, @srcdir@ names the
# folder of the data files), run end to end by tests/test_qdi.py
INPUTS = ["TGP-Expec", "TGPMS-Cheby", "TGPMS-GSPO", "TGPL-IR-Pulse-Cheby", "TGPL-IR-Pulse-GSPO",
          "TGPLMS-Cheby", "TGPLMS-SIL", "TGP-Kick", "TGE-Cheby", "TGE-GSPO", "TGE-SIL", "TGE-Arnoldi"]


def main(root):
    out = os.path.join(HERE, "qdng_files")
    os.makedirs(out, exist_ok=True)
    for f in FILES + [n + ".qdi.in" for n in INPUTS]:
        shutil.copyfile(os.path.join(root, "tests", f), os.path.join(out, f))
    print("copied", len(FILES) + len(INPUTS), "files to", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
