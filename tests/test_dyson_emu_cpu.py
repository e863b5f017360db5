"""CPU builds of the register-resident Dyson/Keldysh kernels under the host
sanitizers (no GPU).

`tests/emu/hip/hip_runtime.h` runs every GPU thread of a workgroup as a host
thread, with __syncthreads and the wavefront exchanges as barriers, so the
kernel sources of quichotte-and-sancho_amd/csrc compile with g++ unchanged
except for their dynamic-LDS declaration.  AddressSanitizer then checks every
LDS and global access against the launch's exact LDS size and the buffers'
sizes, ThreadSanitizer checks that every pair of LDS accesses of different
threads is ordered by a barrier, and LDS starts as NaN bytes so a read of
anything unwritten shows in the outputs.  The outputs are compared with numpy
(init2: G^r and G^< of one energy, src/scf.f90:315-355) and between launch
shapes, which compute the same products in the same order.

The large-Norb path (chunked build, blocked Gauss-Jordan and products on MFMA,
packing) is built with clang++ and checked against numpy and the register kernel.

k_dyson<16, 5, 5> (256 threads at 65..80 orbitals) is the shape switched off
in the library after a wrong result and a memory fault on the GPU; these
builds are how its source was checked for a race or an out-of-range access.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quichotte-and-sancho_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "emu")
LDS_DECL = "extern __shared__ __align__(16) char lds_raw[];"

M_INIT2, M_SCF2, M_FINAL = 1, 2, 4
KB = 8.617333262e-5

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed for the CPU builds")


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    """{name: executable}: k_dyson and k_dyson_w, each under ASan and TSan."""
    d = tmp_path_factory.mktemp("emu")
    srcs = {}
    for name, fn in (("k_dyson", "qs_dyson.hip"), ("k_dyson_w", "qs_dyson_wave.hip")):
        text = open(os.path.join(CSRC, fn)).read()
        assert text.count(LDS_DECL) >= 1
        p = d / ("%s_cpu.hip" % name)
        p.write_text(text.replace(LDS_DECL, "char *lds_raw = emu::lds();"))
        srcs[name] = str(p)
    procs, exes = [], {}
    for name, src in srcs.items():
        for san in ("address", "thread"):
            exe = str(d / ("%s_%s" % (name, san)))
            cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=" + san, "-fno-omit-frame-pointer",
                   "-I", EMU, "-I", CSRC, '-DQS_DYSON_SOURCE="%s"' % src, os.path.join(EMU, "dyson_emu.cpp"),
                   "-o", exe, "-pthread"]
            if name == "k_dyson_w":
                cmd.insert(1, "-DQS_EMU_WAVE")
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            exes[(name, san)] = exe
    for cmd, p in procs:
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, "CPU build failed: %s\n%s" % (" ".join(cmd), out[-4000:])
    return exes, d


def make_problem(path, n, count, mode, use_sigma, seed=1, V=0.5):
    """A dense random system with every orbital active and on the support."""
    rng = np.random.default_rng(seed)
    H = rng.uniform(-1, 1, (n, n)) + 0.1j * rng.uniform(-1, 1, (n, n))
    H = (H + H.conj().T) / 2
    GL = np.zeros((n, n), complex)
    GR = np.zeros((n, n), complex)
    GL[0, 0] = GL[1, 1] = 0.1
    GL[0, 1] = GL[1, 0] = 0.03
    GR[n - 2, n - 2] = GR[n - 1, n - 1] = 0.1
    Hemb = -H + 0.5j * (GL + GR)
    ns = n // 2
    sig_ld = (12 * (ns * (ns + 1) // 2) + 7) // 8 * 8
    P = dict(n=n, Hemb=Hemb, GL=GL, GR=GR, w_init=-2.0, delta=0.37, mu=0.0, beta=1.0 / (KB * 300.0), V=V)
    with open(path, "wb") as f:
        np.array([n, count, mode, use_sigma, 2], np.int32).tofile(f)
        np.array([P["w_init"], P["delta"], P["mu"], P["beta"], V]).tofile(f)
        for a in (Hemb, GL, GR):
            np.ascontiguousarray(a, dtype=np.complex128).tofile(f)
        np.ascontiguousarray(GL[:2, :2], dtype=np.complex128).tofile(f)
        np.array([0, 1], np.int32).tofile(f)
        (0.1 * (rng.uniform(-1, 1, 4 * ns) + 1j * rng.uniform(-1, 1, 4 * ns))).tofile(f)
        if use_sigma:
            sig = 0.05 * (rng.standard_normal((count, sig_ld)) + 1j * rng.standard_normal((count, sig_ld)))
            # Sigma^< is anti-Hermitian; its site-diagonal 2x2 blocks (the only lesser
            # blocks stored whole) are made so, as the physics has them
            for I in range(ns):
                b = 12 * (I * ns - I * (I - 1) // 2) + 4
                blk = sig[:, b:b + 4].reshape(count, 2, 2)
                blk = (blk - np.conj(np.transpose(blk, (0, 2, 1)))) / 2
                sig[:, b:b + 4] = blk.reshape(count, 4)
            sig.tofile(f)
        if mode == M_SCF2:
            rng.standard_normal(count * 3 * n * n).tofile(f)
            rng.standard_normal(count * 3 * n * n).tofile(f)
    return P


def read_out(path, n, count):
    raw = np.fromfile(path, np.uint8)
    out, off = {}, 0
    for name, dt, k in (("x", np.float64, count * 3 * n * n), ("partials", np.float64, count * 4),
                        ("diagR", np.float64, count * n), ("diagL", np.float64, count * n),
                        ("onsL", np.complex128, count * 4 * (n // 2)), ("dosd", np.float64, count * n),
                        ("trUD", np.float64, count * 2), ("status", np.int32, 1)):
        nb = np.dtype(dt).itemsize * k
        out[name] = raw[off:off + nb].view(dt)
        off += nb
    assert off == raw.size
    return out


def run(exe, prob, outp, shape):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, str(prob), str(outp)] + list(shape), capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "Sanitizer" not in r.stderr, "%s %s: rc %d\n%s" % (
        os.path.basename(exe), shape, r.returncode, r.stderr[-6000:])


def init2_reference(P, count):
    """Packed [Re G^r | Im G^r | Re G^<(i<j) | Im G^<(i<=j)] per energy (common/linear_algebra.f90:142-191)."""
    n = P["n"]
    xs = []
    for iw in range(count):
        w = P["w_init"] + (iw + 1) * P["delta"]
        f = lambda V: 1.0 / (np.exp(min((w - P["mu"] + V) * P["beta"], 700.0)) + 1.0)
        G = np.linalg.inv(w * np.eye(n) + P["Hemb"])
        L = G @ (1j * (f(P["V"]) * P["GL"] + f(0.0) * P["GR"])) @ G.conj().T
        xs.append(np.concatenate([G.real.ravel(), G.imag.ravel(), L.real[np.triu_indices(n, 1)],
                                  L.imag[np.triu_indices(n)]]))
    return np.concatenate(xs)


def written(o, mode):
    """the outputs each mode writes (the rest keep their NaN fill)"""
    if mode == M_INIT2:
        return ["x"]
    if mode == M_SCF2:
        return ["x", "partials", "diagR", "diagL"]
    return ["onsL", "dosd", "trUD"]


def compare(a, b, names, rtol):
    for k in names:
        x, y = a[k], b[k]
        assert not np.isnan(x).any() and not np.isnan(y).any(), k
        scale = max(np.max(np.abs(x)), 1e-300)
        assert np.max(np.abs(x - y)) <= rtol * scale, (k, np.max(np.abs(x - y)) / scale)


@pytest.mark.parametrize("n,shapes", [(80, [("16", "5", "5"), ("32", "3", "5")]),
                                      (88, [("16", "6", "6"), ("32", "3", "6")]),
                                      (20, [("16", "2", "2")])])
@pytest.mark.parametrize("mode", [M_INIT2, M_SCF2, M_FINAL])
def test_k_dyson_shapes_address_sanitizer(builds, tmp_path, n, shapes, mode):
    exes, _ = builds
    count = 2
    P = make_problem(tmp_path / "p.bin", n, count, mode, use_sigma=int(mode != M_INIT2))
    outs = []
    for sh in shapes:
        run(exes[("k_dyson", "address")], tmp_path / "p.bin", tmp_path / "o.bin", sh)
        outs.append(read_out(tmp_path / "o.bin", n, count))
    if mode == M_INIT2:
        ref = init2_reference(P, count)
        for o in outs:
            assert np.max(np.abs(o["x"] - ref)) <= 1e-12 * np.max(np.abs(ref))
    # the same products in the same order: equal bits, except the block
    # reductions of |F|^2 and |F - Fold|^2, whose order follows the thread count
    for o in outs[1:]:
        names = written(o, mode)
        compare(outs[0], o, [k for k in names if k != "partials"], 0.0)
        if "partials" in names:
            compare(outs[0], o, ["partials"], 1e-13)


@pytest.mark.parametrize("n", [16, 44, 48, 80, 94])
@pytest.mark.parametrize("mode", [M_INIT2, M_SCF2, M_FINAL])
def test_small_lds_kernel_address_sanitizer(builds, tmp_path, n, mode):
    """k_dyson_w (8x8 tile at 16 orbitals, 16x16 above) against numpy and k_dyson, and
    its split form (Gauss-Jordan kernel, then Keldysh kernel) against itself."""
    exes, _ = builds
    count = 2
    P = make_problem(tmp_path / "p.bin", n, count, mode, use_sigma=int(mode != M_INIT2), seed=n)
    run(exes[("k_dyson_w", "address")], tmp_path / "p.bin", tmp_path / "w.bin", ("wave", "0", "0"))
    w = read_out(tmp_path / "w.bin", n, count)
    for k in written(w, mode):
        assert not np.isnan(w[k]).any(), k
    if n > 16:
        # split into the Gauss-Jordan and Keldysh kernels: the same operations, the same bits
        run(exes[("k_dyson_w", "address")], tmp_path / "p.bin", tmp_path / "s.bin", ("split", "0", "0"))
        compare(w, read_out(tmp_path / "s.bin", n, count), written(w, mode), 0.0)
        # the Keldysh products on MFMA (emulated): the same values to rounding
        run(exes[("k_dyson_w", "address")], tmp_path / "p.bin", tmp_path / "m.bin", ("mfsplit", "0", "0"))
        compare(w, read_out(tmp_path / "m.bin", n, count), written(w, mode), 1e-12)
    if n == 80:
        run(exes[("k_dyson", "address")], tmp_path / "p.bin", tmp_path / "v.bin", ("32", "3", "5"))
        compare(w, read_out(tmp_path / "v.bin", n, count), written(w, mode), 1e-12)
    if mode == M_INIT2:
        ref = init2_reference(P, count)
        assert np.max(np.abs(w["x"] - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("kernel,n,shape", [("k_dyson", 80, ("16", "5", "5")), ("k_dyson", 80, ("32", "3", "5")),
                                            ("k_dyson_w", 80, ("wave", "0", "0")),
                                            ("k_dyson_w", 80, ("split", "0", "0")),
                                            ("k_dyson_w", 80, ("mfsplit", "0", "0")),
                                            ("k_dyson_w", 16, ("wave", "0", "0")),
                                            ("k_dyson_w", 48, ("wave", "0", "0")),
                                            ("k_dyson_w", 48, ("split", "0", "0"))])
@pytest.mark.parametrize("mode", [M_SCF2, M_FINAL])
def test_no_unordered_lds_access_thread_sanitizer(builds, tmp_path, kernel, n, shape, mode):
    exes, _ = builds
    make_problem(tmp_path / "p.bin", n, 1, mode, use_sigma=1)
    run(exes[(kernel, "thread")], tmp_path / "p.bin", tmp_path / "o.bin", shape)
    o = read_out(tmp_path / "o.bin", n, 1)
    for k in written(o, mode):
        assert not np.isnan(o[k]).any(), k


@pytest.mark.parametrize("sizes,kind", [((80, 72), "group"), ((80, 66), "gsplit"), ((16, 10), "group"),
                                        ((44, 40), "gsplit"), ((80, 70), "gmfsplit"), ((30, 20), "gmfsplit")])
@pytest.mark.parametrize("mode", [M_INIT2, M_SCF2, M_FINAL])
def test_grouped_launch_gives_each_member_its_own_bits(builds, tmp_path, sizes, kind, mode):
    """Two solves of different orbital counts and grid lengths, one tile class, in one
    grouped launch (fused or split): each member's outputs equal its own launch bit
    for bit, and the emulated LDS and buffers stay in bounds (AddressSanitizer)."""
    exes, _ = builds
    exe = exes[("k_dyson_w", "address")]
    solo = {"group": "wave", "gsplit": "split", "gmfsplit": "mfsplit"}[kind]
    counts = (2, 4) if kind == "gmfsplit" else (2, 3)   # the MFMA Keldysh part pairs energies
    for k, (n, c) in enumerate(zip(sizes, counts)):
        make_problem(tmp_path / ("p%d.bin" % k), n, c, mode, use_sigma=int(mode != M_INIT2), seed=n + k,
                     V=0.3 + 0.2 * k)
        run(exe, tmp_path / ("p%d.bin" % k), tmp_path / ("s%d.bin" % k), (solo, "0", "0"))
    run(exe, tmp_path / "p0.bin", tmp_path / "g0.bin", (kind, str(tmp_path / "p1.bin"), str(tmp_path / "g1.bin")))
    for k, (n, c) in enumerate(zip(sizes, counts)):
        s = read_out(tmp_path / ("s%d.bin" % k), n, c)
        g = read_out(tmp_path / ("g%d.bin" % k), n, c)
        compare(s, g, written(s, mode), 0.0)


@pytest.mark.parametrize("n,count", [(80, 3), (36, 1)])
def test_mfma_keldysh_odd_energy_count(builds, tmp_path, n, count):
    """The MFMA Keldysh part pairs energies per workgroup; an odd count's last energy
    goes through a one-energy launch: every energy against the fused kernel."""
    exes, _ = builds
    exe = exes[("k_dyson_w", "address")]
    make_problem(tmp_path / "p.bin", n, count, M_SCF2, use_sigma=1, seed=7 * n)
    run(exe, tmp_path / "p.bin", tmp_path / "w.bin", ("wave", "0", "0"))
    run(exe, tmp_path / "p.bin", tmp_path / "m.bin", ("mfsplit", "0", "0"))
    w, m = read_out(tmp_path / "w.bin", n, count), read_out(tmp_path / "m.bin", n, count)
    compare(w, m, written(w, M_SCF2), 1e-12)


# ---- the large-Norb path (qs_dyson_big.hip + qs_gj_mfma.hip) ------------------------
# Its kernels use clang's ext_vector_type, so these builds use clang++ (ROCm's, on
# the host); the MFMA instructions (FP64 and FP32 16 x 16 x 4) are emulated lane by
# lane.  The G^< product forms only the 64 x 64 tiles on and above the diagonal and
# the packing reads a tile below it through its mirror: at 100 and 130 orbitals
# (padded to 112 and 144: 2 and 3 tile rows) some outputs come from mirrored tiles.
CLANG = "/opt/rocm/llvm/bin/clang++" if os.path.exists("/opt/rocm/llvm/bin/clang++") else shutil.which("clang++")
GJ_DECL = "extern __shared__ __align__(16) char smem[];"


@pytest.fixture(scope="module")
def big_builds(tmp_path_factory):
    if CLANG is None:
        pytest.skip("clang++ is needed for the large-Norb CPU builds")
    d = tmp_path_factory.mktemp("emu_big")
    text = open(os.path.join(CSRC, "qs_gj_mfma.hip")).read()
    assert text.count(GJ_DECL) == 1
    gj = d / "qs_gj_mfma_cpu.hip"
    gj.write_text(text.replace(GJ_DECL, "char *smem = emu::lds();"))
    exes, procs = {}, []
    for san in ("address", "thread"):
        base = [CLANG, "-x", "c++", "-std=c++17", "-O1", "-g", "-fsanitize=" + san, "-fno-omit-frame-pointer",
                "-I", EMU, "-I", CSRC, "-c"]
        o1, o2 = str(d / ("gj_%s.o" % san)), str(d / ("drv_%s.o" % san))
        procs.append((san, [base + [str(gj), "-o", o1],
                            base + ["-DQS_EMU_BIG", '-DQS_DYSON_SOURCE="%s"' % os.path.join(CSRC, "qs_dyson_big.hip"),
                                    os.path.join(EMU, "dyson_emu.cpp"), "-o", o2]],
                      [CLANG, "-fsanitize=" + san, o1, o2, "-o", str(d / ("big_%s" % san)), "-pthread"]))
        exes[san] = str(d / ("big_%s" % san))
    running = [(cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
               for _, comp, _ in procs for cmd in comp]
    for cmd, p in running:
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, "CPU build failed: %s\n%s" % (" ".join(cmd), out[-4000:])
    for _, _, link in procs:
        r = subprocess.run(link, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, "link failed: %s\n%s" % (" ".join(link), r.stderr[-4000:])
    return exes


@pytest.mark.parametrize("n", [44, 100, 130])
def test_large_norb_path_init2_against_numpy(big_builds, tmp_path, n):
    """G^r and G^< of two energies, FP64, to 1e-12 of numpy; the complex-FP32 form to 1e-4."""
    P = make_problem(tmp_path / "p.bin", n, 2, M_INIT2, use_sigma=0, seed=n)
    run(big_builds["address"], tmp_path / "p.bin", tmp_path / "b.bin", ("big", "0", "0"))
    run(big_builds["address"], tmp_path / "p.bin", tmp_path / "f.bin", ("big32", "0", "0"))
    ref = init2_reference(P, 2)
    b, f = read_out(tmp_path / "b.bin", n, 2), read_out(tmp_path / "f.bin", n, 2)
    assert np.max(np.abs(b["x"] - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(f["x"] - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("mode", [M_SCF2, M_FINAL])
def test_large_norb_path_against_the_register_kernel(builds, big_builds, tmp_path, mode):
    """At 44 orbitals (with Sigma2) the chunked path gives the register kernel's
    outputs to rounding in every written output."""
    n = 44
    make_problem(tmp_path / "p.bin", n, 2, mode, use_sigma=1, seed=3)
    exes, _ = builds
    run(exes[("k_dyson_w", "address")], tmp_path / "p.bin", tmp_path / "w.bin", ("wave", "0", "0"))
    run(big_builds["address"], tmp_path / "p.bin", tmp_path / "b.bin", ("big", "0", "0"))
    w, b = read_out(tmp_path / "w.bin", n, 2), read_out(tmp_path / "b.bin", n, 2)
    compare(w, b, written(w, mode), 1e-12)


def test_large_norb_path_thread_sanitizer(big_builds, tmp_path):
    n = 100
    make_problem(tmp_path / "p.bin", n, 1, M_SCF2, use_sigma=1, seed=5)
    run(big_builds["thread"], tmp_path / "p.bin", tmp_path / "t.bin", ("big", "0", "0"))
    o = read_out(tmp_path / "t.bin", n, 1)
    for k in written(o, M_SCF2):
        assert not np.isnan(o[k]).any(), k
