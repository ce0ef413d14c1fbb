"""Build gshtrans_amd/libgshtrans_hip.so with hipcc for gfx950 (in-tree).

    python -m gshtrans_amd.build        # or gshtrans_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgshtrans_hip.so")

SOURCES = ["gsh_capi.cpp", "gsh_legendre.hip", "gsh_fft.hip", "gsh_fft_wave.hip", "gsh_mixed.hip",
           "gsh_batch.hip", "gsh_wigner.hip",
           "gsh_grid_ops.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # Keep hipcc from emitting packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


ROOT = os.path.dirname(HERE)
FACADE_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_facade.cpp")
FACADE_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_facade")


def cxx23():
    """A C++23 host compiler (multi-argument operator[], as the reference's
    headers use): ROCm's clang++."""
    for c in (os.environ.get("GSH_CXX23"), "/opt/rocm/llvm/bin/clang++", "clang++"):
        if c and (not os.path.isabs(c) or os.path.exists(c)):
            return c
    return "clang++"


def build_facade_test(verbose: bool = False) -> str:
    """Compile the C++ facade test (include/GSHTrans) against the library;
    the binary finds the library through an $ORIGIN-relative rpath."""
    cmd = [cxx23(), "-std=c++23", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
           FACADE_TEST_SRC, "-o", FACADE_TEST_BIN + ".tmp", "-L", HERE, "-lgshtrans_hip",
           "-Wl,-rpath,$ORIGIN/../../gshtrans_amd", "-Wl,--allow-shlib-undefined"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stderr)
    os.replace(FACADE_TEST_BIN + ".tmp", FACADE_TEST_BIN)
    return FACADE_TEST_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_facade_test(verbose=True))
