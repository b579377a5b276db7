"""Copy a few of QDng's regression-test data files, unchanged, as fixtures of
the file-format tests (tests/test_io.py): raw .op / .wf data and their .meta
files (potential, dipole, laser, initial wave function, eigenfunctions).

Only data is copied.  Run once in a container where the QDng tree exists:
    python tests/golden/make_file_fixtures.py [/path/to/QDng]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["GridPot0.meta", "GridPot0.op", "GridPot1.meta", "GridPot1.op", "GridDip0.meta", "GridDip0.op",
         "GridGauss.meta", "GridGauss.wf", "GridPot0_ef.meta", "GridPot0_ef_0.wf", "GridPot0_ef_1.wf",
         "GridLaser_w0.176.meta", "GridLaser_w0.176.op"]


def main(root):
    out = os.path.join(HERE, "qdng_files")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(root, "tests", f), os.path.join(out, f))
    print("copied", len(FILES), "files to", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
