"""Continuous serving of independent bias-point SCF requests on one GPU.

Each request is one `solve_scf` (a system, a voltage, a fresh start) plus its
current; requests of different orbital counts and grid lengths arrive over
time.  A request is admitted as soon as its device footprint
(`qs_footprint_bytes`) fits in the HBM budget next to the requests already
running.  Its G, Sigma and Broyden state then stay resident in its own handle
until it finishes.  Every running request has its own HIP stream, so the
kernels of different requests overlap on the device.

With fuse=True the requests that group steps can take (order 2 with
device-controlled steps: FP32 history of at most 16 columns and at most 82 active
spin-orbitals) are stepped together in one ScfGroup by a thread of their own: each
step runs one shared Dyson launch for every member of a register-tile class, and
requests join and leave the group between steps (continuous batching).  The other
requests run on their own threads as before.

The reference has no serving layer: it solves one voltage after another in one
process (src/quichotte.f90:125-143).  This loop only schedules; every number
comes from the same C ABI path as a single solve (a group member's steps are
those of its solve alone, bit for bit), so each result is the same alone and in
the mix.
"""
import threading
import time
from dataclasses import dataclass

import numpy as np

from .engine import QsError, ScfEngine, ScfGroup, ScfProblem, device_memory, footprint_bytes

QS_ERR_NOMEM = 4


@dataclass
class Request:
    rid: int
    problem: ScfProblem
    V: float
    arrival: float = 0.0          # seconds after the start of the run


@dataclass
class Response:
    rid: int
    J_up: float
    J_down: float
    converged: bool
    n_iter: int
    arrival: float
    start: float
    done: float
    footprint: int
    error: str = ""
    device_steps: bool = False    # the solve's scalar logic ran on the device (step modes 1/2)
    dyson_shape: int = -1         # the handle's Dyson launch shape
    reused: bool = False          # the request ran on a pooled handle
    grouped: bool = False         # the request was stepped as a member of the fused group

    @property
    def latency(self):
        return self.done - self.arrival


def groupable(p: ScfProblem):
    """Whether group steps can take the request: order 2, FP32 history of at most 16
    columns, at most 82 active spin-orbitals (the sites with U; the library checks
    the rest and the request falls back to a thread of its own if it refuses)."""
    U = np.asarray(p.U_site)
    n_active = 2 * int(np.count_nonzero(U[0::2]))
    return int(p.order) == 2 and bool(p.single_history) and 0 < int(p.nhist) <= 16 and 0 < n_active <= 82


def pool_key(p: ScfProblem):
    """Requests with equal keys can share a handle (qs_reset_problem checks the rest)."""
    an = None if p.atom_norb is None else tuple(int(a) for a in p.atom_norb)
    return (p.H.shape[0], int(p.n_omega), int(p.order), int(p.nhist), bool(p.single_history), an, int(p.device))


class Server:
    """Admission by HBM footprint, up to `max_running` requests on the GPU at once.

    Finished requests leave their handle in a pool keyed by the request's sizes;
    a later request of the same sizes takes it (qs_reset_problem) instead of
    allocating and tuning a new one.  Idle handles count against the budget and
    are released, oldest first, when a request of another size needs the memory."""

    def __init__(self, max_running=4, hbm_fraction=0.85, device=0, pool=True, on_done=None, fuse=False,
                 max_grouped=16):
        free, total = device_memory(device)
        self.budget = int(free * hbm_fraction)
        self.max_running = max_running
        self.device = device
        self.fuse = fuse           # groupable requests go to the fused group
        self.max_grouped = max_grouped
        self.grouped = 0           # admitted groupable requests (creating, queued or in the group)
        self.gq = []               # handed to the group thread: (request, engine, footprint, start, reused)
        self.on_done = on_done     # called with each Response as it finishes (from a worker thread)
        self.mem_waiting = 0       # requests waiting for device memory to create their handle
        self.cv = threading.Condition()
        self.used = 0              # footprints of the running requests
        self.idle = []             # pooled handles: (key, engine, footprint), oldest first
        self.idle_bytes = 0
        self.use_pool = pool
        self.running = 0
        self.pending = []          # arrived, not admitted (arrival order)
        self.closed = False
        self.reused = 0

    def _admit_one(self, evict):
        """Under the lock: the first pending request that fits, with a pooled
        handle if one matches; idle handles to release go to `evict`.  With fuse,
        groupable requests count against max_grouped, the others against
        max_running."""
        for i, (req, fp) in enumerate(self.pending):
            g = self.fuse and groupable(req.problem)
            if (self.grouped >= self.max_grouped) if g else (self.running >= self.max_running):
                continue
            key = pool_key(req.problem)
            for j, (k, eng, efp) in enumerate(self.idle):
                if k == key:
                    self.idle.pop(j)
                    self.idle_bytes -= efp
                    self.pending.pop(i)
                    self.used += efp
                    self._take_slot(g)
                    return req, efp, eng
            if self.used + fp > self.budget:
                continue
            while self.idle and self.used + self.idle_bytes + fp > self.budget:
                k, eng, efp = self.idle.pop(0)
                self.idle_bytes -= efp
                evict.append(eng)
            if self.used + self.idle_bytes + fp <= self.budget:
                self.pending.pop(i)
                self.used += fp
                self._take_slot(g)
                return req, fp, None
        return None

    def _take_slot(self, grouped):
        if grouped:
            self.grouped += 1
        else:
            self.running += 1

    def _release(self, req, eng, fp, resp, keep, grouped, out):
        """A finished request: its budget back, its handle to the pool or closed."""
        if eng is not None and not keep:
            eng.close()
        with self.cv:
            self.used -= fp
            if keep:
                self.idle.append((pool_key(req.problem), eng, fp))
                self.idle_bytes += fp
            if grouped:
                self.grouped -= 1
            else:
                self.running -= 1
            out.append(resp)
            self.cv.notify_all()
        if self.on_done is not None:
            self.on_done(resp)

    def _create(self, problem):
        """A handle for the problem.  If the device cannot hold it although the budget
        admitted it (the runtime's own allocations, fragmentation), the idle handles
        are released, then the request waits for a running one to finish, and tries
        again; it fails only when every other running request is waiting as well."""
        while True:
            try:
                return ScfEngine(problem)
            except QsError as e:
                if e.code != QS_ERR_NOMEM:
                    raise
            with self.cv:
                idle, self.idle, self.idle_bytes = self.idle, [], 0
            for _, eng, _ in idle:
                eng.close()
            if idle:
                continue
            with self.cv:
                # requests that are solving (not this one, not others waiting here)
                if self.running + self.grouped - 1 - self.mem_waiting <= 0:
                    raise QsError(QS_ERR_NOMEM, "qs_create failed: the device cannot hold this request")
                target = self.running + self.grouped
                self.mem_waiting += 1
                try:
                    while (self.running + self.grouped >= target and
                           self.running + self.grouped - self.mem_waiting > 0):
                        self.cv.wait(0.05)
                finally:
                    self.mem_waiting -= 1

    def _worker(self, t0, out):
        while True:
            evict = []
            with self.cv:
                item = self._admit_one(evict)
                while item is None:
                    if self.closed and not self.pending:
                        return
                    self.cv.wait(0.05)
                    item = self._admit_one(evict)
            for e in evict:
                e.close()
            req, fp, eng = item
            start = time.perf_counter() - t0
            keep = False
            reused = False
            grouped = self.fuse and groupable(req.problem)
            try:
                if eng is not None:
                    try:
                        eng.reset(req.problem)
                        self.reused += 1
                        reused = True
                    except Exception:   # another active structure: a handle of its own
                        eng.close()
                        eng = None
                if eng is None:
                    eng = self._create(req.problem)
                if grouped and eng.step_info()[0]:
                    with self.cv:
                        self.gq.append((req, eng, fp, start, reused))
                        self.cv.notify_all()
                    continue
                r = eng.solve_scf(req.V, True)
                ju, jd = eng.current()
                resp = Response(req.rid, ju, jd, r.converged, r.n_iter, req.arrival, start,
                                time.perf_counter() - t0, fp, "", bool(eng.step_info()[0]), eng.dyson_choice()[0],
                                reused)
                keep = self.use_pool
            except Exception as e:  # the request fails, the server goes on
                resp = Response(req.rid, float("nan"), float("nan"), False, 0, req.arrival, start,
                                time.perf_counter() - t0, fp, str(e))
            # its device memory is free (closed) before the budget says so
            self._release(req, eng, fp, resp, keep, grouped, out)

    def _group_loop(self, t0, out):
        """The fused group: members join between steps, every step runs one shared
        Dyson launch per tile class, finished members leave and are answered."""
        g = ScfGroup(self.device)
        members = {}
        try:
            while True:
                with self.cv:
                    while not self.gq and len(g) == 0:
                        if self.closed and not self.pending and self.grouped == 0:
                            return
                        self.cv.wait(0.05)
                    joins, self.gq = self.gq, []
                for req, eng, fp, start, reused in joins:
                    try:
                        g.add(eng, req.V, True)
                        members[eng] = (req, fp, start, reused)
                    except QsError:
                        # not a group member after all: the solve runs here, alone
                        self._solo(req, eng, fp, start, reused, t0, out)
                if len(g) == 0:
                    continue
                try:
                    finished = g.step()
                except QsError as e:
                    # the group step itself failed: every member is answered with the
                    # error and closed, and a new group takes the next requests
                    for eng, (req, fp, start, reused) in list(members.items()):
                        resp = Response(req.rid, float("nan"), float("nan"), False, 0, req.arrival, start,
                                        time.perf_counter() - t0, fp, "group step: %s" % e)
                        self._release(req, eng, fp, resp, False, True, out)
                    members.clear()
                    g.close()
                    g = ScfGroup(self.device)
                    continue
                for eng, r, st in finished:
                    req, fp, start, reused = members.pop(eng)
                    if st == 0:
                        ju, jd = eng.current()
                        resp = Response(req.rid, ju, jd, r.converged, r.n_iter, req.arrival, start,
                                        time.perf_counter() - t0, fp, "", True, eng.dyson_choice()[0], reused, True)
                        self._release(req, eng, fp, resp, self.use_pool, True, out)
                    else:
                        resp = Response(req.rid, float("nan"), float("nan"), False, 0, req.arrival, start,
                                        time.perf_counter() - t0, fp, "group step: status %d" % st)
                        self._release(req, eng, fp, resp, False, True, out)
        finally:
            g.close()

    def _solo(self, req, eng, fp, start, reused, t0, out):
        try:
            r = eng.solve_scf(req.V, True)
            ju, jd = eng.current()
            resp = Response(req.rid, ju, jd, r.converged, r.n_iter, req.arrival, start, time.perf_counter() - t0, fp,
                            "", bool(eng.step_info()[0]), eng.dyson_choice()[0], reused)
            keep = self.use_pool
        except Exception as e:
            resp = Response(req.rid, float("nan"), float("nan"), False, 0, req.arrival, start,
                            time.perf_counter() - t0, fp, str(e))
            keep = False
        self._release(req, eng, fp, resp, keep, True, out)

    def shutdown(self):
        """Release the pooled handles."""
        with self.cv:
            idle, self.idle, self.idle_bytes = self.idle, [], 0
        for _, eng, _ in idle:
            eng.close()

    def run(self, requests):
        """Release each request at its arrival time; return the responses in rid order."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            fp = footprint_bytes(r.problem)
            if fp > self.budget:
                raise ValueError("request %d needs %.1f GB, more than the budget %.1f GB" %
                                 (r.rid, fp / 1e9, self.budget / 1e9))
        out = []
        t0 = time.perf_counter()
        workers = [threading.Thread(target=self._worker, args=(t0, out), daemon=True)
                   for _ in range(self.max_running)]
        if self.fuse:
            workers.append(threading.Thread(target=self._group_loop, args=(t0, out), daemon=True))
        for w in workers:
            w.start()
        for r in reqs:
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            with self.cv:
                self.pending.append((r, footprint_bytes(r.problem)))
                self.cv.notify_all()
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for w in workers:
            w.join()
        return sorted(out, key=lambda x: x.rid)
