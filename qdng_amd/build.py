"""Build libqdx.so (HIP kernels for gfx950 + host runtime) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqdx.so")
SOURCES = ["qdx_cheby64.hip", "qdx_cheby32.hip", "qdx_lines.hip", "qdx_gspo.hip", "qdx_kernels.hip", "qdx_batch.hip", "qdx_host.cpp",
           "qdx_batch.cpp", "qdx_io.cpp", "qdx_oct.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are switched off for all device code.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    objs, procs = [], []
    for src in SOURCES:   # translation units compile in parallel
        obj = os.path.join(CSRC, src.replace(".", "_") + ".o")
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, pr in procs if pr.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-shared", "-o", LIB + ".tmp"] + objs + ["-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


REPO = os.path.dirname(HERE)
CPP_TESTS = [("tests/cpp/test_qdlib.cpp", "tests/cpp/test_qdlib")]


def build_cpp_tests(verbose: bool = False) -> None:
    """Host-side C++ test programs of the QDng-style classes
    (include/qdx_qdlib.hpp), linked against libqdx.so."""
    for src, exe in CPP_TESTS:
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(REPO, "include"), os.path.join(REPO, src),
               "-L" + HERE, "-lqdx", "-Wl,-rpath,$ORIGIN/../../qdng_amd", "-o", os.path.join(REPO, exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    build_cpp_tests(verbose="-v" in sys.argv)
