"""Continuous serving of independent bias-point SCF requests on one GPU.

Each request is one `solve_scf` (a system, a voltage, a fresh start) plus its
current; requests of different orbital counts and grid lengths arrive over
time.  A request is admitted as soon as its device footprint
(`qs_footprint_bytes`) fits in the HBM budget next to the requests already
running.  Its G, Sigma and Broyden state then stay resident in its own handle
until it finishes.  Every running request has its own HIP stream, so the
kernels of different requests overlap on the device.

The reference has no serving layer: it solves one voltage after another in one
process (src/quichotte.f90:125-143).  This loop only schedules; every number
comes from the same C ABI path as a single solve, so each result is the same
alone and in the mix.
"""
import threading
import time
from dataclasses import dataclass

from .engine import ScfEngine, ScfProblem, device_memory, footprint_bytes


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

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    """Admission by HBM footprint, up to `max_running` requests on the GPU at once."""

    def __init__(self, max_running=4, hbm_fraction=0.85, device=0):
        free, total = device_memory(device)
        self.budget = int(free * hbm_fraction)
        self.max_running = max_running
        self.cv = threading.Condition()
        self.used = 0
        self.running = 0
        self.pending = []          # arrived, not admitted (arrival order)
        self.closed = False

    def _admit_one(self):
        """Under the lock: the first pending request that fits, or None."""
        if self.running >= self.max_running:
            return None
        for i, (req, fp) in enumerate(self.pending):
            if self.used + fp <= self.budget:
                self.pending.pop(i)
                self.used += fp
                self.running += 1
                return req, fp
        return None

    def _worker(self, t0, out):
        while True:
            with self.cv:
                item = self._admit_one()
                while item is None:
                    if self.closed and not self.pending:
                        return
                    self.cv.wait(0.05)
                    item = self._admit_one()
            req, fp = item
            start = time.perf_counter() - t0
            eng = None
            try:
                eng = ScfEngine(req.problem)
                r = eng.solve_scf(req.V, True)
                ju, jd = eng.current()
                resp = Response(req.rid, ju, jd, r.converged, r.n_iter, req.arrival, start,
                                time.perf_counter() - t0, fp)
            except Exception as e:  # the request fails, the server goes on
                resp = Response(req.rid, float("nan"), float("nan"), False, 0, req.arrival, start,
                                time.perf_counter() - t0, fp, str(e))
            finally:
                if eng is not None:
                    eng.close()   # its device memory is free before the budget says so
            with self.cv:
                self.used -= fp
                self.running -= 1
                out.append(resp)
                self.cv.notify_all()

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
