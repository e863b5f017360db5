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
OBJ = os.path.join(LIB, "_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["qs_engine.cpp", "qs_dyson.hip", "qs_dyson_wave.hip", "qs_dyson_big.hip", "qs_gj_mfma.hip", "qs_sigma2.hip",
           "qs_vector.hip", "qs_control.hip", "qs_sancho.hip"]
LIBS = []


def _compile(src, obj, verbose):
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", *DEVICE_FLAGS,
           "-I" + os.path.join(REPO, "include"), "-I" + CSRC, "-c", src, "-o", obj + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return "hipcc failed for %s:\n%s%s" % (os.path.basename(src), r.stdout, r.stderr)
    os.replace(obj + ".tmp", obj)
    return None


def build(verbose=False, incremental=False):
    """Compile every source (in parallel, one hipcc per file) and link
    lib/libqs_hip.so.  incremental=True skips a source whose object is newer
    than it and than every header of csrc/ and include/."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers += [os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include"))]
    newest_h = max(os.path.getmtime(f) for f in headers)
    jobs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, s + ".o")
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_h):
            continue
        jobs.append((src, obj))
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4, 8) or 1) as ex:
        errs = [e for e in ex.map(lambda j: _compile(j[0], j[1], verbose), jobs) if e]
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed for libqs_hip.so")
    tmp = SO + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, *LIBS, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed for libqs_hip.so")
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
    print(build(verbose=True, incremental="--incremental" in sys.argv))
