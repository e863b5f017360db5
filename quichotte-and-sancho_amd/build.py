"""Build the gfx950 shared library lib/libqs_hip.so (and the host driver) in-tree.

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
switched off for all device code (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
SO = os.path.join(LIB, "libqs_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["qs_engine.cpp", "qs_dyson.hip", "qs_dyson_wave.hip", "qs_dyson_big.hip", "qs_sigma2.hip", "qs_vector.hip",
           "qs_sancho.hip"]
# the large-Norb Dyson path calls rocSOLVER (batched zgetrf / zgetri) and rocBLAS (batched zgemm)
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = SO + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-shared", "-std=c++17", *DEVICE_FLAGS,
           "-I" + os.path.join(REPO, "include"), "-I" + CSRC, *srcs, *LIBS, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libqs_hip.so")
    os.replace(tmp, SO)
    # the drop-in driver: plain C++ on the C ABI
    exe = os.path.join(LIB, "quichotte_hip")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(REPO, "include"),
           os.path.join(HERE, "host", "quichotte_hip.cpp"), "-L" + LIB, "-lqs_hip", "-Wl,-rpath,$ORIGIN",
           "-o", exe + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for quichotte_hip")
    os.replace(exe + ".tmp", exe)
    return SO


if __name__ == "__main__":
    print(build(verbose=True))
