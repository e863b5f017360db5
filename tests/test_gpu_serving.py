"""Serving loop: requests of different sizes admitted by footprint and run
concurrently give the same result as alone (bit-for-bit: each request's
kernels run with the same launch shapes in its own handle), and one of them
matches the oracle."""
import importlib

import numpy as np
import pytest

from qs_testlib import Oracle, helix_problem

pytestmark = pytest.mark.gpu


def make_requests():
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    serving = importlib.import_module("quichotte-and-sancho_amd.serving")
    rng = np.random.default_rng(11)
    reqs, probs = [], []
    shapes = [(2, 2, 256), (4, 2, 512), (4, 1, 1024), (6, 4, 256), (4, 2, 256), (2, 3, 512), (4, 3, 384), (2, 2, 1024)]
    for rid, (ni, nt, nw) in enumerate(shapes):
        U = float(rng.uniform(0.3, 1.0))
        V = float(rng.uniform(-1.0, 1.0))
        p = helix_problem(ni, nt, U, order=2, n_omega=nw, wlo=0.0, whi=4.0, E_CC=2.0, t_hop=-0.4, lamb=0.05,
                          gamma=1.0, mu=2.0, tolf=1e-10, nhist=8, voltages=[V])
        probs.append(p)
        sp = pkg.ScfProblem(H=p.H, gammaL=p.GL, gammaR=p.GR, U_site=p.U_site, order=2, n_omega=p.n_omega,
                            w_init=p.w_init, delta=p.delta, beta=p.beta, mu=p.mu, nhist=8, tolf=1e-10)
        reqs.append(serving.Request(rid, sp, V, arrival=0.002 * rid))
    return pkg, serving, reqs, probs


def test_mix_equals_alone_and_oracle():
    pkg, serving, reqs, probs = make_requests()
    alone = []
    for r in reqs:
        e = pkg.ScfEngine(r.problem)
        res = e.solve_scf(r.V, True)
        alone.append((e.current(), res.n_iter))
        e.close()
    srv = serving.Server(max_running=4)
    out = srv.run(reqs)
    srv.shutdown()
    assert len(out) == len(reqs)
    for r, (J, nit) in zip(out, alone):
        assert r.error == ""
        assert r.converged and r.n_iter == nit
        assert r.J_up == J[0] and r.J_down == J[1]
    o = Oracle(probs[1])
    ro = o.solve(reqs[1].V, True)
    Jo = np.array(o.current())
    assert ro["n_iter"] == out[1].n_iter
    assert np.all(np.abs(np.array([out[1].J_up, out[1].J_down]) - Jo) <= 1e-10 * np.abs(Jo) + 1e-13)


def test_admission_respects_budget():
    pkg, serving, reqs, probs = make_requests()
    srv = serving.Server(max_running=8)
    fp = [pkg.footprint_bytes(r.problem) for r in reqs]
    srv.budget = max(fp) + min(fp) // 2      # at most ~one large request at a time
    out = srv.run(reqs[:4])
    assert all(r.error == "" and r.converged for r in out)
    # intervals of requests whose footprints cannot coexist do not overlap
    for a in out:
        for b in out:
            if a.rid < b.rid and a.footprint + b.footprint > srv.budget:
                assert a.done <= b.start + 1e-6 or b.done <= a.start + 1e-6


def test_pooled_handle_reset_equals_fresh_handle():
    """A handle reset to another system of the same sizes (qs_reset_problem)
    solves it to the same bits as a handle made for it; a system with another
    active subspace is refused and leaves the handle as it was."""
    pkg = importlib.import_module("quichotte-and-sancho_amd")

    def sp(U, mu, T, V_, U0=None):
        p = helix_problem(4, 2, U, order=2, n_omega=384, wlo=0.0, whi=4.0, E_CC=2.0, t_hop=-0.4, lamb=0.05,
                          gamma=1.0, mu=mu, T=T, tolf=1e-10, nhist=8, voltages=[V_])
        Us = p.U_site.copy()
        if U0 is not None:
            Us[U0] = 0.0
        return pkg.ScfProblem(H=p.H, gammaL=p.GL, gammaR=p.GR, U_site=Us, order=2, n_omega=p.n_omega,
                              w_init=p.w_init, delta=p.delta, beta=p.beta, mu=p.mu, nhist=8, tolf=1e-10)

    a, b = sp(0.6, 2.0, 300.0, 0.2), sp(0.9, 1.9, 150.0, -0.3)
    e = pkg.ScfEngine(a)
    ra = e.solve_scf(0.2, True)
    Ja = e.current()
    e.reset(b)
    rb = e.solve_scf(-0.3, True)
    Jb = e.current()
    f = pkg.ScfEngine(b)
    rf = f.solve_scf(-0.3, True)
    assert rb.n_iter == rf.n_iter and Jb == f.current()
    assert np.array_equal(e.g_less_obs(), f.g_less_obs())
    f.close()
    # U = 0 on a non-contact site: a smaller active subspace than the handle's
    with pytest.raises(pkg.QsError) as err:
        e.reset(sp(0.6, 2.0, 300.0, 0.2, U0=4))
    assert err.value.code == 6
    e.reset(a)
    ra2 = e.solve_scf(0.2, True)
    assert ra2.n_iter == ra.n_iter and e.current() == Ja
    e.close()


def test_pool_reuses_handles_of_equal_sizes():
    pkg, serving, reqs, probs = make_requests()
    # the same request list twice, the second round arriving after the first has finished
    again = [serving.Request(r.rid + 100, r.problem, r.V, arrival=1.0 + r.arrival) for r in reqs]
    srv = serving.Server(max_running=4)
    out = srv.run(reqs + again)
    srv.shutdown()
    assert srv.reused >= len(reqs) // 2
    first = {r.rid: r for r in out if r.rid < 100}
    for r in out:
        if r.rid >= 100:
            a = first[r.rid - 100]
            assert r.error == "" and r.n_iter == a.n_iter and r.J_up == a.J_up and r.J_down == a.J_down


@pytest.mark.parametrize("ions,turns,nw,variant", [(8, 5, 2048, "4"), (4, 2, 512, "2"), (6, 4, 1024, "1"),
                                                   (10, 5, 512, "3")])
def test_reused_handle_equals_a_fresh_one(ions, turns, nw, variant, monkeypatch):
    """qs_reset_problem: a handle that solved one system (other U, V and lead
    couplings) gives a second system the same bits as a handle made for it, in the
    register, split and large-Norb shapes (a pool's handle carries nothing over)."""
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    monkeypatch.setenv("QS_DYSON_VARIANT", variant)

    def prob(U, gam):
        p = helix_problem(ions, turns, U, order=2, n_omega=nw, wlo=-12.0, whi=12.0, E_CC=0.0, t_hop=-1.0,
                          lamb=0.01, gamma=gam, mu=0.0, tolf=1e-8, nhist=16, voltages=[-1.0])
        return pkg.ScfProblem(H=p.H, gammaL=p.GL, gammaR=p.GR, U_site=p.U_site, order=2, n_omega=p.n_omega,
                              w_init=p.w_init, delta=p.delta, beta=p.beta, mu=p.mu, nhist=16, tolf=1e-8, maxscf=60)

    A, B = prob(2.5, 0.15), prob(1.2, 0.1)
    e = pkg.ScfEngine(A)
    e.solve_scf(-0.8, True)
    e.reset(B)
    rb = e.solve_scf(-1.0, True)
    got = (rb.n_iter, rb.converged, rb.fmax, e.current(), e.g_less_obs().copy(), rb.log_fmax.copy())
    e.close()
    f = pkg.ScfEngine(B)
    rf = f.solve_scf(-1.0, True)
    want = (rf.n_iter, rf.converged, rf.fmax, f.current(), f.g_less_obs().copy(), rf.log_fmax.copy())
    f.close()
    print("reused: %d steps, fresh: %d steps" % (got[0], want[0]))
    assert got[:4] == want[:4]
    assert np.array_equal(got[4], want[4]) and np.array_equal(got[5], want[5])


def test_fused_group_serving_equals_alone():
    """fuse=True: every request of the stream (all of at most 82 active
    spin-orbitals) is stepped in the server's group, joining as it is admitted;
    each result is its solve alone, bit for bit."""
    pkg, serving, reqs, probs = make_requests()
    alone = []
    for r in reqs:
        e = pkg.ScfEngine(r.problem)
        res = e.solve_scf(r.V, True)
        alone.append((e.current(), res.n_iter))
        e.close()
    srv = serving.Server(max_running=2, fuse=True, max_grouped=6)
    out = srv.run(reqs)
    srv.shutdown()
    assert len(out) == len(reqs)
    for r, (J, nit) in zip(out, alone):
        assert r.error == ""
        assert r.grouped
        assert r.converged and r.n_iter == nit
        assert r.J_up == J[0] and r.J_down == J[1]
