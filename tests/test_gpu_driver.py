"""The drop-in driver on the reference's regression cases (tests/cases.txt),
compared file by file with the reference outputs at the reference runner's
tolerance (1e-6 abs + 1e-5 rel, tests/run_tests.f90:14-19), and exit codes.
"""
import os
import shutil
import subprocess

import pytest

from qs_testlib import GOLDEN, REPO, assert_close_table, read_table

pytestmark = pytest.mark.gpu
EXE = os.path.join(REPO, "quichotte-and-sancho_amd", "lib", "quichotte_hip")

CASES = [
    ("order0_noninteracting", 0, ["Volt_Current_0.dat", "populations.dat"]),
    ("order1_hartree_fock", 0, ["Volt_Current_1.dat", "populations.dat"]),
    ("order2_second_order", 0, ["Volt_Current_2.dat", "populations.dat"]),
    ("order2_lefthanded_siteU", 0, ["Volt_Current_2.dat", "populations.dat"]),
    ("legacy_inputSOC", 0, ["Volt_Current_1.dat"]),
    ("order2_from_xyz", 0, ["Volt_Current_2.dat", "populations.dat"]),
    ("order1_custom_template", 0, ["Volt_Current_1.dat", "populations.dat"]),
    ("order0_dftb_blocks", 0, ["Volt_Current_0.dat", "populations.dat", "populations_atoms.dat"]),
    ("order2_dftb_blocks", 0, ["Volt_Current_2.dat", "populations.dat", "populations_atoms.dat"]),
    ("order2_dftb_hs", 0, ["Volt_Current_2.dat", "populations.dat", "hamiltonian_matrix.dat",
                           "populations_atoms.dat"]),
]


@pytest.mark.parametrize("case,rc,files", CASES)
def test_driver_reproduces_reference_case(tmp_path, case, rc, files):
    src = os.path.join(GOLDEN, "cases", case)
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if os.path.isfile(p):
            shutil.copy(p, tmp_path / f)
    print("driver case", case, flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=110)
    assert r.returncode == rc, r.stdout[-2000:] + r.stderr[-2000:]
    for f in files:
        new = read_table(str(tmp_path / f))
        ref = read_table(os.path.join(src, "reference", f))
        assert_close_table(new, ref)
    # the other outputs exist and are readable
    for f in ["spectral_function.dat", "info_Hamiltonian.dat", "runtime_datasheet.dat"]:
        assert (tmp_path / f).exists()
    logs = [p for p in os.listdir(tmp_path) if p.startswith("scf_V_")]
    assert len(logs) >= 3
