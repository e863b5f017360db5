"""The drop-in driver's input handling and error exits (no GPU needed: every
case here stops before the device is touched).

Expected exit codes from the reference's test list (tests/cases.txt):
error cases exit 1 (src/input_output.f90:396-430, common/input_parser.f90:237-249).
"""
import os
import shutil
import subprocess

import pytest

from qs_testlib import GOLDEN, REPO

EXE = os.path.join(REPO, "quichotte-and-sancho_amd", "lib", "quichotte_hip")


def run_in(tmp_path, files):
    for name, text in files.items():
        (tmp_path / name).write_text(text)
    return subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=60)


BASE = """T = 300.d0
V = -0.5d0
Vf = 0.5d0
delv = 0.25
mu = 2
N_ions = 4
N_turns = 1
hel_length = 5.d0
hel_radius = 1.d0
hand = 1
order = 2
dw = 1
up = 1
delta = 0.01
E_CC = 2.d0
t_hop = -0.4d0
lamb = 0.05d0
U_int = 0.8d0
GammaL_up = 1.5d0
GammaL_down = 1.d0
GammaR = 1.d0
"""

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")


@pytest.mark.parametrize("case", ["error_N_sites_mismatch", "error_unknown_hamiltonian", "error_dftb_lead_overlap",
                                  "error_dftb_hs_no_lmax"])
def test_reference_error_cases_exit_1(tmp_path, case):
    src = os.path.join(GOLDEN, "cases", case)
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if os.path.isfile(p):
            shutil.copy(p, tmp_path / f)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, r.stdout + r.stderr


def test_missing_input_file(tmp_path):
    r = run_in(tmp_path, {})
    assert r.returncode == 1 and "no input file found" in r.stderr


def test_missing_required_key(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE.replace("delta = 0.01\n", "")})
    assert r.returncode == 1 and 'required key "delta"' in r.stderr


def test_duplicate_key(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE + "mu = 3\n"})
    assert r.returncode == 1 and "given twice" in r.stderr


def test_bad_value_reports_line(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE.replace("order = 2", "order = two")})
    assert r.returncode == 1 and 'cannot read the value of "order"' in r.stderr


def test_non_integer_int(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE.replace("order = 2", "order = 2.5")})
    assert r.returncode == 1


def test_check_input_order_and_steps(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE.replace("order = 2", "order = 3")})
    assert r.returncode == 1 and "order must be 0, 1 or 2" in r.stderr
    r = run_in(tmp_path, {"input.dat": BASE.replace("delv = 0.25", "delv = 0")})
    assert r.returncode == 1 and "delv" in r.stderr


def test_lead_overlap_is_rejected(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE + "lead_left_atoms = 1 2\nlead_right_atoms = 2:4\n"})
    assert r.returncode == 1 and "both lead_left_atoms and lead_right_atoms" in r.stderr


def test_hubbard_file_too_short(tmp_path):
    r = run_in(tmp_path, {"input.dat": BASE + "hubbard_file = U.dat\n", "U.dat": "0.8\n0.6\n"})
    assert r.returncode == 1 and "one U per site" in r.stderr


def test_dftb_hs_hamiltonian_matches_reference(tmp_path):
    """H = S^-1/2 (H0 + SOC) S^-1/2 built by the driver before the device is
    used, against the reference's hamiltonian_matrix.dat fixture."""
    from qs_testlib import assert_close_table, read_table
    case = "order2_dftb_hs"
    src = os.path.join(GOLDEN, "cases", case)
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if os.path.isfile(p):
            shutil.copy(p, tmp_path / f)
    subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    new = read_table(str(tmp_path / "hamiltonian_matrix.dat"))
    ref = read_table(os.path.join(src, "reference", "hamiltonian_matrix.dat"))
    assert_close_table(new, ref)
    # integer index columns identical
    assert [r[:2] for r in new] == [r[:2] for r in ref]
