"""The SANCHO self-energy oracle end to end against the reference's own SANCHO
regression fixtures (SANCHO/tests/cases/, copied to tests/golden/sancho/):
the voltage sweep of tests/sancho_scf.py with the oracle's order-2/3
self-energies reproduces Volt_Current_<order>.dat and populations.dat within
the reference's tolerance (1e-6 abs + 1e-5 rel, SANCHO/tests/run_tests.f90)."""
import os

import pytest

from qs_testlib import assert_close_table, read_table, sancho_oracle
from sancho_scf import SanchoSweep

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sancho")


def cases():
    out = []
    for line in open(os.path.join(GOLD, "cases.txt")):
        t = line.split()
        if not t or t[0].startswith("#") or t[1] != "0":
            continue
        out.append((t[0], t[2:]))
    return out


def oracle_sigma(GR, GL, U, delta, order):
    return sancho_oracle(GR, GL, U, delta, order)


@pytest.mark.parametrize("case,files", cases(), ids=[c for c, _ in cases()])
def test_sancho_case_with_oracle_self_energies(case, files):
    sw = SanchoSweep(os.path.join(GOLD, case, "input.dat"), oracle_sigma)
    iv, pops = sw.run()
    for f in files:
        ref = read_table(os.path.join(GOLD, case, "reference", f))
        new = iv if f.startswith("Volt_Current") else pops
        assert_close_table(new, ref)


@pytest.mark.parametrize("plant", ["sigma3_scaled", "sigma3_lesser_dropped", "sigma2_lesser_conj"])
def test_fixtures_catch_a_planted_self_energy_mistake(plant):
    """the order-3 fixture is sensitive to the self-energies: a 10 % error in
    Sigma3, a dropped lesser Sigma3, or an elementwise-conjugated lesser Sigma2
    each fails it (the mistake is planted in the oracle's output, on the CPU)"""
    def bad(GR, GL, U, delta, order):
        outs = sancho_oracle(GR, GL, U, delta, order)
        if plant == "sigma3_scaled":
            outs[3], outs[4] = 0.9 * outs[3], 0.9 * outs[4]
        elif plant == "sigma3_lesser_dropped":
            outs[4] = 0 * outs[4]
        else:
            outs[1] = -outs[1].conj()
        return outs

    case = "order3_third_order"
    sw = SanchoSweep(os.path.join(GOLD, case, "input.dat"), bad)
    iv, pops = sw.run()
    worst = []
    for f, new in (("Volt_Current_3.dat", iv), ("populations.dat", pops)):
        ref = read_table(os.path.join(GOLD, case, "reference", f))
        try:
            assert_close_table(new, ref)
            worst.append(False)
        except AssertionError:
            worst.append(True)
    assert any(worst), "the fixtures did not see the planted mistake"
