"""The C ABI called from Fortran through iso_c_binding
(quichotte-and-sancho_amd/fortran/qs_hip_binding.f90), built with amdflang.
The demo program is a 3-voltage order-2 loop (qs_fortran_demo.f90)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from qs_testlib import KB, Oracle, Problem, REPO

FDIR = os.path.join(REPO, "quichotte-and-sancho_amd", "fortran")
LIB = os.path.join(REPO, "quichotte-and-sancho_amd", "lib")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def build(tmp):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    subprocess.run([FC, "-O2", "-c", os.path.join(FDIR, "qs_hip_binding.f90"), "-o", str(tmp / "b.o"),
                    "-J" + str(tmp)], check=True)
    exe = str(tmp / "demo")
    subprocess.run([FC, "-O2", "-J" + str(tmp), os.path.join(FDIR, "qs_fortran_demo.f90"), str(tmp / "b.o"),
                    "-L" + LIB, "-lqs_hip", "-Wl,-rpath," + LIB, "-o", exe], check=True)
    return exe


def demo_problem():
    n, nsite = 8, 4
    H = np.zeros((n, n), complex)
    for i in range(nsite):
        H[2 * i, 2 * i] = H[2 * i + 1, 2 * i + 1] = 2.0
        if i < nsite - 1:
            for s in (0, 1):
                H[2 * i + s, 2 * i + 2 + s] = H[2 * i + 2 + s, 2 * i + s] = -0.4
    GL = np.zeros((n, n), complex)
    GR = np.zeros((n, n), complex)
    GL[0, 0], GL[1, 1], GR[n - 2, n - 2], GR[n - 1, n - 1] = 1.5, 1.0, 1.0, 1.0
    U = np.zeros(n)
    U[0::2] = 0.8
    return Problem(H, GL, GR, U, 2, 300.0, 2.0, 0, 0, 4.0 / 256, [0.0, 0.25, 0.5], n_omega=256, w_init=0.0,
                   nhist=8, maxscf=200, tolf=1e-10)


def test_fortran_binding_builds_and_fails_cleanly_without_device(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        assert "qs_create failed" in r.stdout
    else:
        assert len(r.stdout.split("\n")) >= 3


@pytest.mark.gpu
def test_fortran_loop_matches_oracle(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [list(map(float, l.split())) for l in r.stdout.strip().split("\n") if l.strip()]
    prob = demo_problem()
    assert abs(prob.beta - 1.0 / (KB * 300.0)) < 1e-9
    o = Oracle(prob)
    fresh = True
    Jscale = None
    for row, V in zip(rows, prob.voltages):
        ro = o.solve(V, fresh)
        fresh = not ro["converged"]
        ju, jd = o.current()
        assert abs(row[0] - V) < 1e-12
        # J(V = 0) is round-off of order tolf: absolute floor 1e-11
        assert abs(row[1] - ju) <= 1e-10 * abs(ju) + 1e-11
        assert abs(row[2] - jd) <= 1e-10 * abs(jd) + 1e-11
        assert int(row[3]) == 1 and int(row[4]) == ro["n_iter"]


def build_sancho(tmp):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    subprocess.run([FC, "-O2", "-c", os.path.join(FDIR, "qs_hip_binding.f90"), "-o", str(tmp / "b.o"),
                    "-J" + str(tmp)], check=True)
    exe = str(tmp / "sancho_demo")
    subprocess.run([FC, "-O2", "-J" + str(tmp), os.path.join(FDIR, "qs_sancho_demo.f90"), str(tmp / "b.o"),
                    "-L" + LIB, "-lqs_hip", "-Wl,-rpath," + LIB, "-o", exe], check=True)
    return exe


def test_fortran_sancho_binding_builds(tmp_path):
    assert os.path.exists(build_sancho(tmp_path))


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3])
def test_fortran_sancho_self_energies_match_oracle(tmp_path, order):
    from qs_testlib import sancho_inputs, sancho_oracle
    exe = build_sancho(tmp_path)
    na, nw = 4, 200
    print("fortran sancho order=%d na=%d nw=%d" % (order, na, nw), flush=True)
    GR, GL, U, delta = sancho_inputs(na, nw, 31)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([na, nw, order], np.int32).tobytes())
        f.write(np.array([delta], np.float64).tobytes())
        f.write(U.astype(np.float64).tobytes())
        f.write(np.asfortranarray(GR).tobytes(order="F"))
        f.write(np.asfortranarray(GL).tobytes(order="F"))
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nout = 6 if order == 3 else 3
    got = np.frombuffer(open(fout, "rb").read(), np.complex128).reshape((na, na, nw, nout), order="F")
    ref = sancho_oracle(GR, GL, U, delta, order)
    for q in range(nout):
        assert np.max(np.abs(got[:, :, :, q] - ref[q])) <= 1e-12 * np.max(np.abs(ref[q]))
