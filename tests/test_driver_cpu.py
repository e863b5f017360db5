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


def _spin_block_report(H):
    """max |Im h0|, max |Re h|, max |Im h| over the 2x2 spin blocks h0*1 + h.sigma of H."""
    n = H.shape[0] // 2
    B = H.reshape(n, 2, n, 2).transpose(0, 2, 1, 3)
    h0 = 0.5 * (B[..., 0, 0] + B[..., 1, 1])
    hz = 0.5 * (B[..., 0, 0] - B[..., 1, 1])
    hx = 0.5 * (B[..., 0, 1] + B[..., 1, 0])
    hy = 0.5j * (B[..., 0, 1] - B[..., 1, 0])
    import numpy as np
    return (np.max(np.abs(h0.imag)), max(np.max(np.abs(h.real)) for h in (hx, hy, hz)),
            max(np.max(np.abs(h.imag)) for h in (hx, hy, hz)))


def _report_values(stdout):
    vals = {}
    for line in stdout.splitlines():
        for key in ("max |Im h0|", "max |Re h|", "max |Im h|"):
            if line.strip().startswith(key):
                vals[key] = float(line.split("=")[1].split()[0])
    return vals


def test_dftb_hs_time_reversal_report(tmp_path):
    """The time-reversal report of src/hamiltonian.f90:477-504 after the dftb_hs
    build: the three maxima against the spin blocks of the reference's own
    hamiltonian_matrix.dat, printed to two decimals (es10.2)."""
    import numpy as np
    from qs_testlib import read_table
    src = os.path.join(GOLDEN, "cases", "order2_dftb_hs")
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if os.path.isfile(p):
            shutil.copy(p, tmp_path / f)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert " Time reversal (all spin blocks h0*1 + h.sigma):" in r.stdout
    assert "   element C : l_max = 1, xi(p) =   8.4900E-03 eV" in r.stdout
    ref = read_table(os.path.join(src, "reference", "hamiltonian_matrix.dat"))
    n = int(max(max(row[0], row[1]) for row in ref))
    H = np.zeros((n, n), complex)
    for i, j, re, im in ref:
        H[int(i) - 1, int(j) - 1] = re + 1j * im
    low = np.tril(H, -1)
    H = np.diag(np.diag(H)) + low + low.conj().T if not np.triu(H, 1).any() else H
    odd0, oddh, soc = _spin_block_report(H)
    got = _report_values(r.stdout)
    assert soc > 1e-4 and abs(got["max |Im h|"] - soc) <= 0.006 * soc
    # a non-magnetic H: both odd parts at round-off, and no warning
    assert got["max |Im h0|"] < 1e-12 and got["max |Re h|"] < 1e-12 and odd0 < 1e-12 and oddh < 1e-12
    assert "WARNING: H breaks time reversal" not in r.stdout


def test_peptide_report(tmp_path):
    """Peptide builder's summary (src/hamiltonian.f90:260-271): the hopping table's
    bond lengths from the helix geometry, and the time-reversal report."""
    import numpy as np
    text = BASE.replace("N_ions = 4\n", "").replace("N_turns = 1\n", "") + (
        "hamiltonian = peptide\nN_sites = 12\npep_twist = 100\npep_rise = 1.5\nhop_range = 3\nhop_decay = 0.7\n"
        "soc_model = radial\ndipole_ramp = 0.2\n")
    r = run_in(tmp_path, {"input.dat": text})
    assert " Peptide helix: twist 100.00 deg, rise 1.500 A, radius 1.000 A, hand  1.0, SOC radial" in r.stdout, \
        r.stdout + r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if len(l.split()) == 3 and l.split()[0].isdigit()]
    assert [int(x[0]) for x in rows] == [1, 2, 3]
    phi = np.arange(1, 13) * np.radians(100.0)
    rs = np.stack([np.cos(phi), np.sin(phi), 1.5 * np.arange(12)], 1)
    l = np.linalg.norm(rs[1:4] - rs[0], axis=1)
    assert np.allclose([float(x[1]) for x in rows], l, atol=6e-5)
    assert np.allclose([float(x[2]) for x in rows], np.exp(-(l - l[0]) / 0.7), rtol=6e-5)
    got = _report_values(r.stdout)
    assert got["max |Im h0|"] == 0.0 and got["max |Re h|"] == 0.0 and got["max |Im h|"] > 0.0
