"""Order-2 SCF steps with their scalar logic on the device (QsCtrl) and replayed
from captured graphs, against the host-driven loop.

* device-controlled launches (mode 1) and graph replays (mode 2) run the same
  kernels in the same order: identical step counts and currents, bit for bit,
  over every replay of a solve (the host reads each step's result in between);
* against the host loop (mode 0) the only difference is where the small Gram
  solve is evaluated (host or device rounding): step counts within one, and
  currents equal to O(tolf) (1e-8 relative), where two FP64 paths to
  max|F| < tolf part;
* the end of a solve at maxscf and a runaway abort behave as in the host loop.
"""
import numpy as np
import pytest

from qs_testlib import Oracle, gpu_engine, helix_problem, random_problem

pytestmark = pytest.mark.gpu


def helix16(voltages, **kw):
    return helix_problem(4, 2, 1.0, order=2, n_omega=1 << 14, wlo=-10.0, whi=10.0, E_CC=0.0, t_hop=-1.0,
                         lamb=0.01, gamma=0.1, mu=0.0, tolf=1e-8, nhist=16, voltages=voltages, **kw)


def sweep(prob, mode):
    g = gpu_engine(prob)
    g.set_step_mode(mode)
    out, fresh = [], True
    for V in prob.voltages:
        r = g.solve_scf(V, fresh)
        fresh = not r.converged
        active, nrep = g.step_info()
        out.append((r.n_iter, r.converged, r.runaway, np.array(g.current()), g.g_less_obs().copy(), active, nrep,
                    r.log_fmax.copy()))
    g.close()
    return out


def test_graph_replays_equal_device_launches_and_host_loop():
    prob = helix16([0.0, 0.3, 0.6])
    host, dev, gra = sweep(prob, 0), sweep(prob, 1), sweep(prob, 2)
    for h, d, g in zip(host, dev, gra):
        assert not h[5] and d[5] and g[5]          # modes 1 and 2 take the device path here
        assert d[6] == 0 and g[6] == g[0] - 2      # every step after the second is a graph replay
        assert d[0] == g[0] and np.array_equal(d[3], g[3]) and np.array_equal(d[4], g[4])
        assert np.array_equal(d[7], g[7])          # the same max|F| at every step
        # the Gram solve rounds differently on the host and the device, which can move the
        # first step with max|F| < tolf by one and the result by O(tolf): 1e-8 relative,
        # a floor for J(V = 0)
        assert h[1] and g[1] and abs(h[0] - g[0]) <= 1, (h[0], g[0])
        assert np.all(np.abs(g[3] - h[3]) <= 1e-8 * np.abs(h[3]) + 1e-13), (g[3], h[3])
        assert np.abs(g[4] - h[4]).max() <= 1e-8 * np.abs(h[4]).max()


def test_graph_steps_match_oracle():
    prob = helix16([0.5])
    o = Oracle(prob, nthreads=16)
    ro = o.solve(0.5, True)
    g = sweep(prob, 2)[0]
    # at tolf 1e-8 the step count moves with rounding (the oracle's own with its reduction order)
    assert g[1] and ro["converged"] and abs(g[0] - ro["n_iter"]) <= 3
    Jo = np.array(o.current())
    assert np.all(np.abs(g[3] - Jo) <= 1e-8 * np.abs(Jo)), (g[3], Jo)


def test_graph_solve_stopped_at_maxscf():
    """The step that can end the loop at maxscf keeps Sigma2 for the final sweep."""
    prob = helix_problem(4, 2, 0.8, order=2, n_omega=400, wlo=0.0, whi=4.0, E_CC=2.0, t_hop=-0.4, lamb=0.05,
                         gamma=1.0, mu=2.0, tolf=1e-14, nhist=8, voltages=[0.3])
    prob.maxscf = 5
    h, g = sweep(prob, 0)[0], sweep(prob, 2)[0]
    assert h[0] == g[0] == 5 and not h[1] and not g[1]
    assert np.all(np.abs(g[3] - h[3]) <= 1e-12 * np.abs(h[3])), (g[3], h[3])
    o = Oracle(prob)
    o.solve(0.3, True)
    Jo = np.array(o.current())
    assert np.all(np.abs(g[3] - Jo) <= 1e-10 * np.abs(Jo)), (g[3], Jo)


def test_graph_runaway_abort_matches_host_loop():
    """A system whose SCF runs away within a few steps: the device sanity check
    aborts at the same step as the host's, with the same flags."""
    prob = random_problem(80, 64, seed=120, U=0.5, gamma=1.0)
    h, g = sweep(prob, 0)[0], sweep(prob, 2)[0]
    assert h[2] and g[2] and not h[1] and not g[1]
    assert h[0] == g[0]


@pytest.mark.parametrize("mode", [1, 2])
def test_device_steps_without_the_dx_buffer_give_the_same_bits(mode, monkeypatch):
    """The optional x - xold buffer (left out with QS_DX_BUFFER=0, as a failed
    allocation leaves it out) only saves the history pass reading x and xold: a
    handle without it still takes device-controlled steps, and its sweep is the
    same bit for bit."""
    prob = helix16([0.0, 0.1, 0.2])
    g = gpu_engine(prob)
    full = g.device_bytes()
    g.close()
    with_dx = sweep(prob, mode)
    monkeypatch.setenv("QS_DX_BUFFER", "0")
    g = gpu_engine(prob)
    assert g.device_bytes() < full
    g.close()
    without = sweep(prob, mode)
    for a, b in zip(with_dx, without):
        assert a[5] and b[5]                      # device-controlled steps in both
        assert a[:3] == b[:3]                     # steps, converged, runaway
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
        assert np.array_equal(a[7], b[7])         # every step's max|F|
