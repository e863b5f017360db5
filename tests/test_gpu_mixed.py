"""The mixed-precision form: complex-FP32 MFMA Dyson/Keldysh sweeps while
max|F| is above the switch level, FP64 steps to tolf / 10 after that, FP64
final sweep.  The converged I(V) must agree with the FP64 fixed point (the
oracle converged to 1e-11) to 1e-8 relative (the precision target of the form;
an absolute floor of 1e-11 covers J(V=0), which is round-off), and the error of
every current is printed.  The FP32
sweeps themselves are FP32-accurate, not FP64: checked on one sweep."""
import numpy as np
import pytest

from qs_testlib import Oracle, gpu_engine, helix_problem, random_problem, spin_texture_rows

pytestmark = pytest.mark.gpu


def helix16(voltages):
    return helix_problem(4, 2, 1.0, order=2, n_omega=1 << 14, wlo=-10.0, whi=10.0, E_CC=0.0, t_hop=-1.0,
                         lamb=0.01, gamma=0.1, mu=0.0, tolf=1e-8, nhist=16, voltages=voltages)


def compare_sweep(prob, label):
    """The oracle solves to tolf = 1e-11, so that its solution is the FP64 fixed
    point to well below the 1e-8 comparison level (at the GPU's tolf = 1e-8 an
    FP64 solution is itself only fixed to O(tolf), ~1e-8 relative in J)."""
    import copy
    po = copy.copy(prob)
    po.tolf = 1e-11
    o = Oracle(po)
    g = gpu_engine(prob)
    g.set_precision(1, 1e-5)
    fresh = True
    worst = 0.0
    for V in prob.voltages:
        ro = o.solve(V, fresh)
        rg = g.solve_scf(V, fresh)
        n32, n64 = g.precision_steps()
        fresh = not rg.converged
        assert ro["converged"] and rg.converged
        assert n32 > 0 and n64 > 0 and n32 + n64 == rg.n_iter
        Jo, Jg = np.array(o.current()), np.array(g.current())
        err = np.abs(Jg - Jo)
        rel = err / np.abs(Jo)
        print("%s V=%+.2f: %d FP32 + %d FP64 steps (oracle %d); J = %s, |dJ|/|J| = %s"
              % (label, V, n32, n64, ro["n_iter"], Jg, rel))
        assert np.all(err <= 1e-8 * np.abs(Jo) + 1e-11), (V, Jg, Jo)
        rows_o = np.array(spin_texture_rows(V, o.g_less_obs()))[:, 2:]
        rows_g = np.array(spin_texture_rows(V, g.g_less_obs()))[:, 2:]
        assert np.abs(rows_g - rows_o).max() <= 1e-8 * np.abs(rows_o).max()
        worst = max(worst, float(np.max(np.where(np.abs(Jo) > 1e-9, rel, 0.0))))
    g.close()
    return worst


def test_mixed_helix16_matches_fp64_oracle():
    """The 16-spin-orbital benchmark system, 2^14 energies, three bias points
    (V = 0 fresh, then 0.5 and 1.0 continued)."""
    compare_sweep(helix16([0.0, 0.5, 1.0]), "helix16 mixed")


def test_mixed_helix80_matches_fp64_oracle_at_2e11_energies():
    """The 80-spin-orbital benchmark system on 2^11 energies of the same window
    (a size the oracle solves in minutes), from a fresh start at V = -1."""
    prob = helix_problem(8, 5, 2.0, order=2, n_omega=1 << 11, wlo=-12.0, whi=12.0, E_CC=0.0, t_hop=-1.0, lamb=0.01,
                         gamma=0.1, mu=0.0, tolf=1e-8, nhist=16, voltages=[-1.0])
    compare_sweep(prob, "helix80/2^11 mixed")


@pytest.mark.parametrize("ions,turns", [(4, 2), (8, 5), (8, 10)])
def test_fp32_sweeps_stall_at_single_precision_and_fp64_steps_finish(ions, turns):
    """With the switch level out of reach the FP32 phase runs until max|F| stops
    falling: it stalls at the single-precision floor of the residual (well above
    an FP64 sweep's), and the FP64 steps then take max|F| below tolf / 10; the
    fixed point is the FP64 oracle's (helix with 1 eV leads on 256 energies)."""
    prob = helix_problem(ions, turns, 0.8, order=2, n_omega=256, wlo=0.0, whi=4.0, E_CC=2.0, t_hop=-0.4, lamb=0.05,
                         gamma=1.0, mu=2.0, tolf=1e-9, nhist=8, voltages=[0.3])
    o = Oracle(prob)
    g = gpu_engine(prob)
    g.set_precision(1, 1e-30)
    r = g.solve_scf(0.3, True)
    n32, n64 = g.precision_steps()
    ro = o.solve(0.3, True)
    assert r.converged and ro["converged"] and n32 >= 6 and n64 >= 1
    floor32 = r.log_fmax[:n32].min()
    print("helix %dx%d: %d FP32 steps (lowest max|F| %.2e), %d FP64 steps (last %.2e); oracle %d steps"
          % (ions, turns, n32, floor32, n64, r.log_fmax[-1], ro["n_iter"]))
    assert floor32 > 1e-11 and r.log_fmax[-1] < 1e-10
    Jo, Jg = np.array(o.current()), np.array(g.current())
    assert np.all(np.abs(Jg - Jo) <= 1e-8 * np.abs(Jo) + 1e-11), (Jg, Jo)
    g.close()


def test_precision_argument_errors():
    import importlib
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    prob = random_problem(8, 16, seed=1)
    g = gpu_engine(prob)
    for mode, sw in [(2, 1e-5), (1, 0.0), (1, -1.0)]:
        with pytest.raises(pkg.QsError):
            g.set_precision(mode, sw)
    g.set_precision(0, 1e-5)
    g.close()
