"""The serving loop's scheduling on the CPU: stand-in engines and groups replace
the HIP library (ScfEngine, ScfGroup, footprint_bytes, device_memory), so these
tests check only what the Server decides: every request answered once, the HBM
budget and the concurrency limits never exceeded, groupable requests stepped in
the group with fuse=True and on threads otherwise, handles reused from the pool,
and a failing request answered with its error while the others go on."""
import importlib
import threading
import time

import numpy as np
import pytest

serving = importlib.import_module("quichotte-and-sancho_amd.serving")
engine_mod = importlib.import_module("quichotte-and-sancho_amd.engine")


class Tracker:
    def __init__(self):
        self.lock = threading.Lock()
        self.live_bytes = 0
        self.max_live_bytes = 0
        self.solo = 0
        self.max_solo = 0
        self.members = 0
        self.max_members = 0
        self.created = 0


class FakeResult:
    def __init__(self, n_iter, converged=True):
        self.n_iter = n_iter
        self.converged = converged


def make_fakes(tr, fail_norb=None, group_fail_step=None):
    class FakeEngine:
        def __init__(self, prob):
            if fail_norb is not None and prob.H.shape[0] == fail_norb:
                raise engine_mod.QsError(5, "stand-in failure")
            self.prob = prob
            self.fp = fake_footprint(prob)
            with tr.lock:
                tr.created += 1
                tr.live_bytes += self.fp
                tr.max_live_bytes = max(tr.max_live_bytes, tr.live_bytes)
            self.closed = False
            self.V = None

        def reset(self, prob):
            if prob.H.shape != self.prob.H.shape:
                raise engine_mod.QsError(6, "other sizes")
            self.prob = prob

        def step_info(self):
            return True, 0

        def dyson_choice(self):
            return 2, None

        def solve_scf(self, V, fresh):
            with tr.lock:
                tr.solo += 1
                tr.max_solo = max(tr.max_solo, tr.solo)
            time.sleep(0.002 * self.prob.H.shape[0] / 8)
            with tr.lock:
                tr.solo -= 1
            self.V = V
            return FakeResult(10 + self.prob.H.shape[0])

        def current(self):
            return self.V, -self.V

        def close(self):
            assert not self.closed
            self.closed = True
            with tr.lock:
                tr.live_bytes -= self.fp

    class FakeGroup:
        steps = 0

        def __init__(self, device=0):
            self.m = []

        def add(self, eng, V, fresh=True):
            if eng.prob.H.shape[0] > 82:
                raise engine_mod.QsError(5, "too large for group steps")
            eng.V = V
            self.m.append([eng, 3 + eng.prob.H.shape[0] // 8])
            with tr.lock:
                tr.members += 1
                tr.max_members = max(tr.max_members, tr.members)

        def __len__(self):
            return len(self.m)

        def step(self):
            time.sleep(0.001)
            FakeGroup.steps += 1
            if group_fail_step is not None and FakeGroup.steps == group_fail_step:
                with tr.lock:
                    tr.members -= len(self.m)
                self.m = []
                raise engine_mod.QsError(1, "stand-in group failure")
            done = []
            for item in list(self.m):
                item[1] -= 1
                if item[1] == 0:
                    self.m.remove(item)
                    with tr.lock:
                        tr.members -= 1
                    done.append((item[0], FakeResult(10 + item[0].prob.H.shape[0]), 0))
            return done

        def close(self):
            pass

    return FakeEngine, FakeGroup


def fake_footprint(prob):
    return int(prob.H.shape[0]) * 1000


class P:
    """the ScfProblem fields the Server reads"""

    def __init__(self, norb, nhist=16):
        self.H = np.zeros((norb, norb))
        self.U_site = np.zeros(norb)
        self.U_site[0::2] = 1.0
        self.n_omega = 256
        self.order = 2
        self.nhist = nhist
        self.single_history = True
        self.atom_norb = None
        self.device = 0


@pytest.fixture
def fakes(monkeypatch):
    def install(fail_norb=None, budget_bytes=10 ** 9, group_fail_step=None):
        tr = Tracker()
        FE, FG = make_fakes(tr, fail_norb, group_fail_step)
        monkeypatch.setattr(serving, "ScfEngine", FE)
        monkeypatch.setattr(serving, "ScfGroup", FG)
        monkeypatch.setattr(serving, "footprint_bytes", fake_footprint)
        monkeypatch.setattr(serving, "device_memory", lambda device=0: (budget_bytes, budget_bytes))
        return tr
    return install


def stream(sizes, seed=0):
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    for rid, n in enumerate(sizes):
        t += rng.exponential(0.001)
        reqs.append(serving.Request(rid, P(n), float(rng.uniform(-1, 1)), arrival=t))
    return reqs


SIZES = [16, 48, 80, 160, 240, 16, 48, 16, 80, 240, 16, 16, 48, 160, 80, 16]


@pytest.mark.parametrize("fuse", [False, True])
def test_every_request_answered_once_within_the_limits(fakes, fuse):
    tr = fakes()
    reqs = stream(SIZES)
    srv = serving.Server(max_running=3, fuse=fuse, max_grouped=5)
    out = srv.run(reqs)
    srv.shutdown()
    assert [r.rid for r in out] == list(range(len(reqs)))
    for r, q in zip(out, reqs):
        assert r.error == "" and r.converged and r.n_iter == 10 + q.problem.H.shape[0]
        assert (r.J_up, r.J_down) == (q.V, -q.V)
        assert r.grouped == (fuse and q.problem.H.shape[0] <= 82)
    assert tr.max_solo <= 3 and tr.max_members <= 5
    assert tr.live_bytes == 0                      # every handle closed after shutdown
    if fuse:
        assert tr.max_members >= 2                 # the group did hold several requests at once


def test_budget_is_never_exceeded(fakes):
    budget = 300_000
    tr = fakes(budget_bytes=budget)
    reqs = stream(SIZES, seed=1)
    srv = serving.Server(max_running=4, fuse=True, max_grouped=8, hbm_fraction=1.0)
    out = srv.run(reqs)
    srv.shutdown()
    assert len(out) == len(reqs) and all(r.error == "" for r in out)
    assert tr.max_live_bytes <= budget


def test_pool_reuses_handles_of_equal_sizes(fakes):
    tr = fakes()
    reqs = stream([16] * 10, seed=2)
    srv = serving.Server(max_running=1, fuse=False)
    out = srv.run(reqs)
    srv.shutdown()
    assert all(r.error == "" for r in out)
    assert tr.created == 1 and srv.reused == 9
    assert sum(r.reused for r in out) == 9


def test_a_failing_request_does_not_stop_the_others(fakes):
    fakes(fail_norb=48)
    reqs = stream(SIZES, seed=3)
    srv = serving.Server(max_running=2, fuse=True, max_grouped=4)
    out = srv.run(reqs)
    srv.shutdown()
    for r, q in zip(out, reqs):
        if q.problem.H.shape[0] == 48:
            assert "stand-in failure" in r.error and not r.converged
        else:
            assert r.error == "" and r.converged


def test_groupable_follows_the_library_limits():
    assert serving.groupable(P(80)) and serving.groupable(P(16))
    assert not serving.groupable(P(84)) and not serving.groupable(P(160))
    assert not serving.groupable(P(16, nhist=32))
    q = P(16)
    q.order = 1
    assert not serving.groupable(q)


def test_a_failing_group_step_answers_its_members_and_serving_goes_on(fakes):
    tr = fakes(group_fail_step=3)
    reqs = stream(SIZES, seed=4)
    srv = serving.Server(max_running=2, fuse=True, max_grouped=4)
    out = srv.run(reqs)
    srv.shutdown()
    assert [r.rid for r in out] == list(range(len(reqs)))
    failed = [r for r in out if r.error]
    assert failed and all("stand-in group failure" in r.error for r in failed)
    assert all(r.converged for r in out if not r.error)
    assert any(r.grouped and not r.error for r in out)     # later requests went through a new group
    assert tr.live_bytes == 0
