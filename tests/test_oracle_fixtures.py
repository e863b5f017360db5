"""The CPU oracle against the reference's committed regression fixtures.

Fixture data: tests/golden/cases/<case>/ (input.dat and reference/*.dat, copied
from the reference's tests/cases/).  Tolerance as tests/run_tests.f90:14-19:
|new - ref| <= 1e-6 + 1e-5 |ref| per numeric field.
"""
import os

import pytest

from qs_testlib import GOLDEN, Oracle, Problem, assert_close_table, read_table, spin_texture_rows

CASES = [
    "order0_noninteracting",
    "order1_hartree_fock",
    "order2_second_order",
    "order2_lefthanded_siteU",
    "legacy_inputSOC",
    "order1_custom_template",
]


def run_case(case):
    cd = os.path.join(GOLDEN, "cases", case)
    prob = Problem.from_case(cd)
    o = Oracle(prob)
    iv, pops = [], []
    fresh = True
    for V in prob.voltages:
        r = o.solve(V, fresh)
        fresh = (not prob.continuation) or (not r["converged"])
        ju, jd = o.current()
        iv.append([V, ju, jd, ju + jd, 1 if r["converged"] else 0])
        pops += spin_texture_rows(V, o.g_less_obs())
    return cd, prob, iv, pops


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference_fixture(case):
    cd, prob, iv, pops = run_case(case)
    ref = read_table(os.path.join(cd, "reference", "Volt_Current_%d.dat" % prob.order))
    assert_close_table(iv, ref)
    pf = os.path.join(cd, "reference", "populations.dat")
    if os.path.exists(pf):
        assert_close_table(pops, read_table(pf))
