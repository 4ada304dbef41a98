"""Build the gfx950 shared library libpisco.so in-tree with hipcc.

    python -m pisco4cmb_amd.build            # library (+ the tests' C++ driver)
    python -m pisco4cmb_amd.build --oracle   # also the CPU oracle (make -C oracle)

The library is the product: hand-written HIP kernels for gfx950 and the
extern "C" boundary declared in include/pisco.h.  Flags that matter for
correctness: -ffp-contract=off (the index path must evaluate exactly the
IEEE operations it is written with, so that pixel indices match the CPU
oracle bit for bit) and the packed-FP32 feature switched off.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpisco.so")
SOURCES = ["kernels.hip", "skymap.hip", "focalplane.hip", "textio.hip", "capi.hip",
           "noalloc.hip"]
ARCH = os.environ.get("PISCO_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build")


def lib_flags():
    return [
        "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I" + os.path.join(ROOT, "include"),
    ]


def build_library(verbose=False):
    cmd = [hipcc()] + lib_flags() + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libpisco.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "drop_in_driver")


def build_cpp_driver(verbose=False):
    """The C++ program of the test suite that uses include/pisco.hpp the way a
    reference program uses PISCO's classes (plain g++, links libpisco)."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC,
           "-L" + HERE, "-lpisco", "-Wl,-rpath,$ORIGIN/../../pisco4cmb_amd", "-o", DRIVER]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the C++ drop-in driver failed")
    return DRIVER


def build_oracle(verbose=False):
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"],
                       capture_output=True, text=True)
    if verbose:
        sys.stdout.write(r.stdout)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the CPU oracle failed")


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_library(verbose=True)
    build_cpp_driver(verbose=True)
    if "--oracle" in argv:
        build_oracle(verbose=True)


if __name__ == "__main__":
    main()
