"""The QDng-style C++ classes (include/qdx_qdlib.hpp) over the C ABI.

tests/cpp/test_qdlib.cpp follows QDng's CppUnit tests (TestCheby, OGSPOTest,
OGridNablaSqTest) with grid operators; it is built by qdng_amd/build.py
(build_cpp_tests) and run here on the GPU.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_qdlib")


def test_header_compiles_standalone():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-x", "c++", "-"], input='#include "qdx_qdlib.hpp"\nint main() { return 0; }\n',
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_qdlib_classes_on_device():
    assert os.path.exists(EXE), "build the C++ tests first: python qdng_amd/build.py"
    print("[gpu case] C++ qdlib-style classes", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=180)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
