"""CPU builds of the self-energy, gather/update, streaming and control kernels
under the host sanitizers (no GPU).

As for the Dyson kernels (test_dyson_emu_cpu.py), `tests/emu/hip/hip_runtime.h`
runs every GPU thread of a workgroup as a host thread with the kernels' barriers
as barriers, AddressSanitizer checks every access against the buffers and the
launch's exact LDS size, and ThreadSanitizer checks that every pair of accesses
of different threads to one LDS array is ordered by a barrier.  The drivers
(tests/emu/sigma2_emu.cpp, tests/emu/vector_emu.cpp) run the launches of one SCF
step on random inputs and check what has a plain host form against it: the two
forms of the gather agree bit for bit, the fused Broyden update gives the plain
update's x_new, its dx and the gather of x_new to the bit, the history pass
gives the same bits from dx as from x and xold, and the column sums, Gram and
projection sums, equal-time G^< and Sigma1 match host loops.  The self-energy
passes run at grid sizes whose FFT splits use each LDS layout of the kernels
(N2 = 8, 16, 128 and 256; N1 = 16, 128 and 256).

SANCHO's self-energy kernels (qs_sancho.hip, tests/emu/sancho_emu.cpp) run through
the library's own entry points on physical random inputs (G^a = G^r^dagger, G^<
anti-Hermitian) and are compared with the oracle (oracle/sancho_oracle.c,
SANCHO/src/self_energies.f90:38-207) to 1e-12; the entry point that allocates
nothing must give the handle's outputs to the bit, a smaller second call must
match the oracle too, and an oversized call must be refused.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from qs_testlib import sancho_inputs, sancho_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quichotte-and-sancho_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "emu")
LDS_DECL = "extern __shared__ __align__(16) char lds_raw[];"

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed for the CPU builds")


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    d = tmp_path_factory.mktemp("emu_sv")
    text = open(os.path.join(CSRC, "qs_sigma2.hip")).read()
    assert text.count(LDS_DECL) >= 4
    s2 = d / "qs_sigma2_cpu.hip"
    s2.write_text(text.replace(LDS_DECL, "char *lds_raw = emu::lds();"))
    procs, exes = [], {}
    for san in ("address", "thread"):
        base = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=" + san, "-fno-omit-frame-pointer", "-I", EMU, "-I", CSRC,
                "-pthread"]
        exes[("sigma2", san)] = str(d / ("sigma2_%s" % san))
        exes[("vector", san)] = str(d / ("vector_%s" % san))
        cmds = [base + ['-DQS_SIGMA2_SOURCE="%s"' % s2, os.path.join(EMU, "sigma2_emu.cpp"), "-o",
                        exes[("sigma2", san)]],
                base + ['-DQS_VECTOR_SOURCE="%s"' % os.path.join(CSRC, "qs_vector.hip"),
                        '-DQS_CONTROL_SOURCE="%s"' % os.path.join(CSRC, "qs_control.hip"),
                        os.path.join(EMU, "vector_emu.cpp"), "-o", exes[("vector", san)]]]
        exes[("sancho", san)] = str(d / ("sancho_%s" % san))
        cmds.append(base + ["-I", os.path.join(ROOT, "include"),
                            '-DQS_SANCHO_SOURCE="%s"' % os.path.join(CSRC, "qs_sancho.hip"),
                            os.path.join(EMU, "sancho_emu.cpp"), "-o", exes[("sancho", san)]])
        for cmd in cmds:
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for cmd, p in procs:
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, "CPU build failed: %s\n%s" % (" ".join(cmd), out[-4000:])
    return exes


ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TSAN_OPTIONS="halt_on_error=1 report_signal_unsafe=0")

# every run of this file: (name, sanitizer, args); they go out at once (each is a
# single process of host threads), and the tests read their results
SIGMA2_ASAN = [(8, 64, False), (6, 100, False), (12, 100, False), (4, 8192, True), (4, 32768, True)]
SIGMA2_TSAN = [(8, 64, False), (6, 100, False), (4, 8192, True)]
# (sites, energies, order)
SANCHO_CASES = [(3, 64, 2), (4, 96, 3)]


@pytest.fixture(scope="module")
def runs(builds, tmp_path_factory):
    d = tmp_path_factory.mktemp("emu_sv_runs")
    jobs = {}
    for nA, nw, fft in SIGMA2_ASAN:
        jobs[("sigma2", "address", nA, nw)] = [builds[("sigma2", "address")], nA, nw, 7, d / ("a_%d_%d.bin" % (nA, nw))] + (
            ["fft"] if fft else [])
    for nA, nw, fft in SIGMA2_TSAN:
        jobs[("sigma2", "thread", nA, nw)] = [builds[("sigma2", "thread")], nA, nw, 7, d / ("t_%d_%d.bin" % (nA, nw))] + (
            ["fft"] if fft else [])
    for san in ("address", "thread"):
        jobs[("vector", san)] = [builds[("vector", san)], 4, 24, 3, 11]
    for na, nw, order in SANCHO_CASES:
        GR, GL, U, delta = sancho_inputs(na, nw, seed=na)
        path = d / ("sn_in_%d_%d_%d.bin" % (na, nw, order))
        with open(path, "wb") as f:
            np.array([na, nw, order], np.int32).tofile(f)
            np.array([delta]).tofile(f)
            np.asarray(U, np.float64).tofile(f)
            GR.ravel(order="F").astype(np.complex128).tofile(f)
            GL.ravel(order="F").astype(np.complex128).tofile(f)
        for san in ("address", "thread"):
            out = d / ("sn_out_%s_%d_%d_%d.bin" % (san, na, nw, order))
            jobs[("sancho", san, na, nw, order)] = [builds[("sancho", san)], path, out]
    procs = {k: subprocess.Popen([str(a) for a in v], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                                 env=ENV) for k, v in jobs.items()}
    out = {}
    for k, p in procs.items():
        text, _ = p.communicate(timeout=1200)
        out[k] = (p.returncode, text, jobs[k])
    return out


def check(result):
    rc, text, args = result
    name = " ".join(str(a) for a in args[1:])
    assert rc == 0, "%s %s: exit %d\n%s" % (os.path.basename(str(args[0])), name, rc, text[-6000:])
    assert "WARNING: ThreadSanitizer" not in text, text[-6000:]


def sig_out(path, nA, nw):
    nAs = nA // 2
    ncol = (12 * (nAs * (nAs + 1) // 2) + 7) // 8 * 8
    raw = np.fromfile(path, dtype=np.uint8)
    nsig = nw * ncol * 16
    return raw[:nsig], raw[nsig:]


# (nA, nw): Nf = 128 (N1 16, N2 8), 256 (16 x 16), 2^14 (128 x 128), 2^16 (256 x 256);
# the long grids run the self-energy passes only ("fft"), on the host's gather
@pytest.mark.parametrize("nA,nw,fft", SIGMA2_ASAN)
def test_sigma2_step_address_sanitizer(runs, nA, nw, fft):
    check(runs[("sigma2", "address", nA, nw)])


@pytest.mark.parametrize("nA,nw,fft", SIGMA2_TSAN)
def test_sigma2_step_thread_sanitizer_and_same_bits(runs, nA, nw, fft):
    """No unordered LDS access; the TSan and ASan builds give the same bits."""
    check(runs[("sigma2", "thread", nA, nw)])
    check(runs[("sigma2", "address", nA, nw)])
    ts, tx = sig_out(runs[("sigma2", "thread", nA, nw)][2][4], nA, nw)
    as_, ax = sig_out(runs[("sigma2", "address", nA, nw)][2][4], nA, nw)
    assert np.array_equal(ts, as_) and np.array_equal(tx, ax)


@pytest.mark.parametrize("san", ["address", "thread"])
def test_vector_and_control_kernels(runs, san):
    """Several device-controlled steps (the history ring of 3 wraps) at 4 active
    spin-orbitals on 24 energies."""
    check(runs[("vector", san)])


@pytest.mark.parametrize("san", ["address", "thread"])
@pytest.mark.parametrize("na,nw,order", SANCHO_CASES)
def test_sancho_self_energies_against_the_oracle(runs, san, na, nw, order):
    check(runs[("sancho", san, na, nw, order)])
    out = runs[("sancho", san, na, nw, order)][2][2]
    GR, GL, U, delta = sancho_inputs(na, nw, seed=na)
    nout = 6 if order == 3 else 3
    raw = np.fromfile(out, np.complex128)
    nm, nw2 = na * na * nw, nw // 2
    nm2 = na * na * nw2
    ref = sancho_oracle(GR, GL, U, delta, order)
    ref2 = sancho_oracle(np.asfortranarray(GR[:, :, :nw2]), np.asfortranarray(GL[:, :, :nw2]), U, delta, order)
    for i in range(nout):
        got = raw[i * nm:(i + 1) * nm].reshape((na, na, nw), order="F")
        assert np.max(np.abs(got - ref[i])) <= 1e-12 * np.max(np.abs(ref[i])), i
        got2 = raw[nout * nm + i * nm2:nout * nm + (i + 1) * nm2].reshape((na, na, nw2), order="F")
        assert np.max(np.abs(got2 - ref2[i])) <= 1e-12 * np.max(np.abs(ref2[i])), i
