"""Group steps (qs_group_*): several order-2 solves stepped together, with
the Dyson/Keldysh sweeps of the members that share a kernel instance in one
launch.  Each member's results equal qs_solve_scf on its own handle bit for
bit (the same per-energy arithmetic and the same Broyden steps), whatever joins
or leaves the group around it; one member is also compared with the oracle."""
import numpy as np
import pytest

from qs_testlib import Oracle, gpu_engine, helix_problem

pytestmark = pytest.mark.gpu

QS_ERR_UNSUPPORTED, QS_ERR_STATE = 5, 6


def chain(ni, nt, nw, U, V, order=2, **kw):
    """A helix of ni x nt sites (2 ni nt spin-orbitals) on nw energies of (0, 4]."""
    return helix_problem(ni, nt, U, order=order, n_omega=nw, wlo=0.0, whi=4.0, E_CC=2.0, t_hop=-0.4, lamb=0.05,
                         gamma=1.0, mu=2.0, tolf=1e-10, nhist=8, voltages=[V], **kw)


def helix(ni, nt, nw, U, V, **kw):
    """The benchmark helices' physics (E_CC = 0, t = -1, SOC 0.01, Gamma 0.1) on nw
    energies of (-12, 12]."""
    return helix_problem(ni, nt, U, order=2, n_omega=nw, wlo=-12.0, whi=12.0, E_CC=0.0, t_hop=-1.0, lamb=0.01,
                         gamma=0.1, mu=0.0, tolf=1e-8, nhist=16, voltages=[V], **kw)


def alone(prob, V, fresh=True, before=None):
    """qs_solve_scf on a handle of its own (after solving `before` on it, if given)."""
    e = gpu_engine(prob)
    if before is not None:
        e.solve_scf(before, True)
    r = e.solve_scf(V, fresh)
    out = (r, e.current(), e.g_less_obs())
    e.close()
    return out


def same(a, b):
    ra, Ja, Ga = a
    rb, Jb, Gb = b
    assert (ra.converged, ra.n_iter, ra.runaway, ra.unphysical, ra.i_worst) == \
           (rb.converged, rb.n_iter, rb.runaway, rb.unphysical, rb.i_worst)
    assert ra.fmax == rb.fmax
    assert Ja == Jb
    assert np.array_equal(Ga, Gb)


def drive(group, engines, joins=None, max_steps=400):
    """Step the group until it is empty.  joins: {step: [(index, V, fresh)]} adds
    members before that step.  Returns {index: (result, J, G)} and the Dyson
    launch counts of every step."""
    joins = joins or {}
    byh = {id(e): k for k, e in enumerate(engines)}
    out, launches = {}, []
    for step in range(max_steps):
        for k, V, fresh in joins.get(step, []):
            group.add(engines[k], V, fresh)
        if len(group) == 0 and not any(s > step for s in joins):
            break
        for eng, r, st in group.step():
            assert st == 0
            out[byh[id(eng)]] = (r, eng.current(), eng.g_less_obs())
        launches.append(group.info()[:2])
    return out, launches


@pytest.mark.parametrize("variant,sizes", [("2", [(4, 2, 256), (3, 2, 384), (4, 2, 512)]),
                                           ("4", [(8, 5, 2048), (6, 6, 2048)])])
def test_members_of_one_tile_class_share_a_launch_and_keep_their_bits(monkeypatch, variant, sizes):
    """16 and 12 spin-orbitals (8 x 8 tile, one-kernel shape), 80 and 72 (16 x 16
    tile, split shape), uneven grids: one Dyson launch per step for all members."""
    monkeypatch.setenv("QS_DYSON_VARIANT", variant)
    pkg = __import__("qs_testlib").product()
    if variant == "2":
        Vs = [0.15 * (k + 1) for k in range(len(sizes))]
        probs = [chain(ni, nt, nw, 0.6 + 0.1 * k, V) for k, ((ni, nt, nw), V) in enumerate(zip(sizes, Vs))]
    else:
        # fresh starts at V = -1 converge on this grid (at V near 0 these helices need finer grids)
        Vs = [-1.0] * len(sizes)
        probs = [helix(ni, nt, nw, 2.0 - k, V) for k, ((ni, nt, nw), V) in enumerate(zip(sizes, Vs))]
    ref = [alone(p, V) for p, V in zip(probs, Vs)]
    engines = [gpu_engine(p) for p in probs]
    g = pkg.ScfGroup(0)
    out, launches = drive(g, engines, joins={0: [(k, V, True) for k, V in enumerate(Vs)]})
    g.close()
    assert launches[0] == (1, len(sizes)), launches[0]
    for k in range(len(sizes)):
        assert ref[k][0].converged
        same(out[k], ref[k])
    for e in engines:
        e.close()


def test_solves_join_and_leave_between_steps(monkeypatch):
    """Continuous batching: members of three kinds (16-orbital register tile,
    80-orbital register tile, and a 48-orbital solve on the 512-thread shape, which
    has no grouped form and sweeps on its own inside the group step) join at
    different steps, one continues from its handle's last solution and one stops at
    maxscf; every result is its solve alone."""
    pkg = __import__("qs_testlib").product()
    pA = chain(4, 2, 256, 0.96, 0.24)
    pB = helix(8, 5, 2048, 2.0, -1.0)
    pC = chain(6, 4, 256, 0.35, -0.74)
    pD = chain(4, 2, 384, 0.9, 0.4, maxscf=5)
    monkeypatch.setenv("QS_DYSON_VARIANT", "1")
    refC = alone(pC, -0.74)
    eC = gpu_engine(pC)
    monkeypatch.delenv("QS_DYSON_VARIANT")
    assert eC.dyson_choice()[0] == 1
    ref = {0: alone(pA, 0.24), 1: alone(pB, -1.0), 2: refC, 3: alone(pD, 0.4),
           4: alone(pA, 0.3, fresh=False, before=0.24)}
    engines = [gpu_engine(pA), gpu_engine(pB), eC, gpu_engine(pD)]
    eA2 = gpu_engine(pA)
    eA2.solve_scf(0.24, True)
    engines.append(eA2)
    g = pkg.ScfGroup(0)
    out, launches = drive(g, engines, joins={0: [(0, 0.24, True)], 2: [(1, -1.0, True), (3, 0.4, True)],
                                             5: [(2, -0.74, True), (4, 0.3, False)]})
    g.close()
    for k in range(5):
        print("member %d alone: converged %s in %d steps" % (k, ref[k][0].converged, ref[k][0].n_iter))
    assert max(n for n, _ in launches) >= 3
    assert not ref[3][0].converged and ref[3][0].n_iter == 5
    for k in (0, 1, 2, 4):
        assert ref[k][0].n_iter > 5 and not ref[k][0].runaway
    for k in range(5):
        same(out[k], ref[k])
    for e in engines:
        e.close()


def test_group_member_matches_the_oracle():
    pkg = __import__("qs_testlib").product()
    p = chain(4, 2, 512, 0.8, 0.25)
    q = chain(4, 2, 256, 0.5, -0.2)
    e, f = gpu_engine(p), gpu_engine(q)
    g = pkg.ScfGroup(0)
    out, _ = drive(g, [e, f], joins={0: [(0, 0.25, True), (1, -0.2, True)]})
    g.close()
    r, J, _ = out[0]
    o = Oracle(p)
    ro = o.solve(0.25, True)
    Jo = np.array(o.current())
    print("group member: %d steps, J = %s; oracle %d steps, J = %s" % (r.n_iter, J, ro["n_iter"], Jo))
    assert r.converged and ro["converged"] and r.n_iter == ro["n_iter"]
    assert np.all(np.abs(np.array(J) - Jo) <= 1e-10 * np.abs(Jo))
    e.close()
    f.close()


def test_group_rejects_what_it_cannot_step():
    pkg = __import__("qs_testlib").product()
    g = pkg.ScfGroup(0)
    e1 = gpu_engine(chain(4, 2, 256, 0.5, 0.0, order=1))
    with pytest.raises(pkg.QsError) as ei:
        g.add(e1, 0.0)
    assert ei.value.code == QS_ERR_UNSUPPORTED
    e2 = gpu_engine(chain(4, 2, 256, 0.5, 0.0))
    g.add(e2, 0.1)
    with pytest.raises(pkg.QsError) as ei:
        g.add(e2, 0.1)
    assert ei.value.code == QS_ERR_STATE
    # destroying a member takes it out of the group
    e2.close()
    assert len(g) == 0
    assert g.step() == []
    g.close()
    e1.close()


def test_two_groups_stepped_half_a_step_apart():
    """Two groups whose steps are begun while the other's runs (step_begin /
    step_end): every member still equals its solve alone; joining during a step is
    refused."""
    pkg = __import__("qs_testlib").product()
    specs = [(4, 2, 256, 0.96, 0.24), (4, 2, 512, 0.72, -0.94), (6, 4, 256, 0.35, -0.74), (4, 2, 384, 0.6, 0.15)]
    probs = [chain(ni, nt, nw, U, V) for ni, nt, nw, U, V in specs]
    ref = [alone(p, s[4]) for p, s in zip(probs, specs)]
    engines = [gpu_engine(p) for p in probs]
    gs = [pkg.ScfGroup(0), pkg.ScfGroup(0)]
    for k, (e, s) in enumerate(zip(engines, specs)):
        gs[k % 2].add(e, s[4], True)
    out = {}
    for g in gs:
        g.step_begin()
    extra = gpu_engine(probs[0])
    with pytest.raises(pkg.QsError) as ei:
        gs[0].add(extra, 0.1)
    assert ei.value.code == QS_ERR_STATE
    live = [True, True]
    while any(live):
        for j, g in enumerate(gs):
            if not live[j]:
                continue
            for eng, r, st in g.step_end():
                assert st == 0
                out[engines.index(eng)] = (r, eng.current(), eng.g_less_obs())
            if len(g):
                g.step_begin()
            else:
                live[j] = False
    for g in gs:
        g.close()
    for k in range(len(specs)):
        same(out[k], ref[k])
    for e in engines + [extra]:
        e.close()
