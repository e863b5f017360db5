#!/usr/bin/env python3
"""Benchmark of the order-2 SCF path on one MI355X: converged bias points/s.

One step = one bias point: solve_scf to max|F| < tolf (Broyden, FP32 history
of 16), then the final sweep and the current (src/scf.f90:78-151,
src/observables.f90:22-39).  Steps follow the voltage sweep with continuation,
as the reference's voltage loop (src/quichotte.f90:125-143).

Workloads (synthetic Hamiltonians from the model; no input files):
  helix80 (default): helix N_ions = 8, N_turns = 5 -> 80 spin-orbitals,
      hel_radius 1.4, hel_length 3.4, hand +1, E_CC 0, t_hop -1, SOC 0.01,
      U = 2 on every site, T = 300 K, mu = 0, Gamma = 0.1 on the first and last
      site, 2^16 energies on (-12, 12], V = -1 ... 1 in steps of 0.1, tolf 1e-8.
  helix16: the same model with N_ions = 4, N_turns = 2 (16 spin-orbitals), U = 1,
      2^14 energies on (-10, 10], V = 0 ... 1 in steps of 0.1.

Printed: one JSON line with the metric, per-step latency percentiles, the
roofline of the dominant kernel (algorithmic count per launch / average launch
time from HIP events on the library's stream) and the CPU baseline (the FP64
oracle timed on a bounded sample of the same workload on the host cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

KB = 8.6173303e-5
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def helix_H(N_ions, N_turns, R=1.4, L=3.4, hand=1.0, E=0.0, t=-1.0, lamb=0.01):
    """Helix model Hamiltonian (src/hamiltonian.f90:83-163)."""
    nat = N_ions * N_turns
    n = 2 * nat
    H = np.zeros((n, n), complex)

    def pos(i):
        a = hand * (i - 1) * 2 * np.pi / N_ions
        return np.array([R * np.cos(a), R * np.sin(a), L * (i - 1) / (N_turns * (N_ions - 1))])

    for i in range(1, nat + 1):
        ii = 2 * i
        H[ii - 2, ii - 2] = E
        H[ii - 1, ii - 1] = E
        if i <= nat - 1:
            H[ii - 2, ii] = t
            H[ii - 1, ii + 1] = t
        if i <= nat - 2:
            r, r1, r2 = pos(i), pos(i + 1), pos(i + 2)
            d1 = (r - r1) / np.linalg.norm(r - r1)
            d2 = (r - r2) / np.linalg.norm(r - r2)
            v = np.cross(d1, d2)
            jj = 2 * (i + 2)
            H[ii - 2, jj - 2] = 1j * lamb * v[2]
            H[ii - 2, jj - 1] = 1j * lamb * (v[0] - 1j * v[1])
            H[ii - 1, jj - 2] = 1j * lamb * (v[0] + 1j * v[1])
            H[ii - 1, jj - 1] = -1j * lamb * v[2]
    iu = np.triu_indices(n, 1)
    H[(iu[1], iu[0])] = np.conj(H[iu])
    return H


def workload(name):
    if name == "helix80":
        w = dict(N_ions=8, N_turns=5, U=2.0, log2nw=16, wmax=12.0, volts=[round(-1.0 + 0.1 * k, 10) for k in range(21)])
    elif name == "helix16":
        w = dict(N_ions=4, N_turns=2, U=1.0, log2nw=14, wmax=10.0, volts=[round(0.1 * k, 10) for k in range(11)])
    elif name == "helix240":
        # the 240-orbital size of the dftb_hs configuration, at 2^12 energies so that
        # x, F, Fold, xold and the nhist 16 history fit one GPU
        w = dict(N_ions=12, N_turns=10, U=1.5, log2nw=12, wmax=12.0, volts=[round(-1.0 + 0.1 * k, 10) for k in range(21)])
    else:
        raise SystemExit("unknown workload " + name)
    H = helix_H(w["N_ions"], w["N_turns"])
    n = H.shape[0]
    GL = np.zeros((n, n), complex)
    GR = np.zeros((n, n), complex)
    for i in (0, 1):
        GL[i, i] = 0.1
        GR[n - 2 + i, n - 2 + i] = 0.1
    U = np.zeros(n)
    U[0::2] = w["U"]
    nw = 1 << w["log2nw"]
    delta = 2 * w["wmax"] / nw
    prob = dict(H=H, gammaL=GL, gammaR=GR, U_site=U, order=2, n_omega=nw, w_init=-w["wmax"], delta=delta,
                beta=1.0 / (KB * 300.0), mu=0.0, alpha_mix=0.5, nhist=16, maxscf=300, tolf=1e-8,
                single_history=True)
    desc = ("helix %dx%d (%d spin-orbitals), U=%g on every site, order 2, 2^%d energies on (-%g,%g], "
            "V %g..%g step 0.1, tolf 1e-8, Broyden nhist 16 FP32 history"
            % (w["N_ions"], w["N_turns"], n, w["U"], w["log2nw"], w["wmax"], w["wmax"], w["volts"][0], w["volts"][-1]))
    return prob, w["volts"], desc


def sweep_points(nv, steps):
    """Indices of the voltages the timed steps solve: the voltage loop with
    continuation (src/quichotte.f90:125-143) from a fresh start at its first
    point; past the last point it starts over (fresh) at the first."""
    return [k % nv for k in range(steps)]


def stage_counts(eng, nh):
    """Algorithmic work per launch of each stage (flops or HBM bytes)."""
    nw, norb, nA, nv, Nf = eng.prob.n_omega, eng.norb, eng.nA, eng.n_vec, eng.Nf
    nAs = nA // 2
    ncol = float((12 * (nAs * (nAs + 1) // 2) + 7) // 8 * 8)   # site-pair layout columns (12 per pair, padded to 8)
    return {
        "dyson_keldysh": ("mfma", nw * (8.0 * norb ** 3 + 16.0 * nA ** 3), "flop"),
        # each pass reads its input once and writes its output once
        "sigma2_fft_f1": ("hbm", 16.0 * ncol * (nw + Nf), "byte"),
        "sigma2_fft_mid": ("hbm", 32.0 * ncol * Nf, "byte"),
        "sigma2_fft_i2": ("hbm", 16.0 * ncol * (Nf + nw), "byte"),
        "broyden_history": ("hbm", nv * (32.0 + 4.0 * (nh - 1) + 8.0), "byte"),
        "broyden_update": ("hbm", nv * (16.0 + 4.0 * nh + 8.0), "byte"),
        "permute_in": ("hbm", nv * 8.0 + 16.0 * ncol * nw, "byte"),
        "equal_time_sigma1": ("hbm", 8.0 * 4 * (nA // 2) * nw, "byte"),
    }


def measured_traffic(workload_name, stage):
    """HBM bytes per launch of a stage's kernel from the committed counter summary
    (profiles/<workload>_hbm_traffic.md: rocprofv3 FETCH_SIZE and WRITE_SIZE
    passes scaled by a plain copy kernel of known size), or None."""
    path = os.path.join(REPO, "profiles", "%s_hbm_traffic.md" % workload_name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == stage:
            try:
                return float(cells[5]) * 1e9
            except ValueError:
                return None
    return None


def host_cores():
    """Physical cores this process may use: the distinct (package, core) pairs of
    the CPUs in its affinity mask, capped by a cgroup CPU quota and by
    OMP_NUM_THREADS when either is set (the host's share of a shared node)."""
    cpus = sorted(os.sched_getaffinity(0))
    phys, cur = set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in list(f) + [""]:
                if not line.strip():
                    if cur.get("processor") in cpus:
                        phys.add((cur.get("physical id", 0), cur.get("core id", cur["processor"])))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("processor", "physical id", "core id"):
                    cur[k] = int(v)
    except (OSError, ValueError):
        pass
    n = len(phys) or len(cpus)
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = min(n, max(1, int(q) // int(per)))
    except (OSError, ValueError):
        pass
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, n)


def cpu_baseline_child(args):
    """Runs in its own process (OpenMP placement set before libgomp loads):
    the FP64 oracle on the workload's host cores.  Two timings:
      * the workload itself, extrapolated from bounded samples of one SCF
        iteration (a full bias point of 80 spin-orbitals x 2^16 energies takes
        the oracle tens of minutes);
      * complete bias points of the 16-spin-orbital helix at 2^14 energies
        (V = 0 fresh, then V = 0.1 with continuation), timed end to end."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import ctypes

    from qs_testlib import Oracle, _OrcProblem, helix_problem, oracle_lib, ptr

    threads = args.cpu_threads
    prob, volts, _ = workload(args.workload)
    L = oracle_lib()
    L.orc_bench_dyson.restype = ctypes.c_double
    L.orc_bench_dyson.argtypes = [ctypes.c_void_p]
    L.orc_bench_sigma2.restype = ctypes.c_double
    L.orc_bench_sigma2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_bench_broyden.restype = ctypes.c_double
    L.orc_bench_broyden.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    n = prob["H"].shape[0]
    nw = prob["n_omega"]
    nA, nAs = n, n // 2
    n_vec = 3 * nA * nA * nw
    n_iter_avg = args.cpu_iters
    # Dyson/Keldysh: a sample of energies
    e_sample = max(threads * 8, 64)
    H = np.asfortranarray(prob["H"])
    GL = np.asfortranarray(prob["gammaL"])
    GR = np.asfortranarray(prob["gammaR"])
    U = np.ascontiguousarray(prob["U_site"])
    p = _OrcProblem()
    p.norb, p.n_omega, p.w_init, p.delta = n, e_sample, prob["w_init"], prob["delta"]
    p.H, p.gammaL, p.gammaR, p.U_site = ptr(H), ptr(GL), ptr(GR), ptr(U)
    p.order, p.beta, p.mu = 2, prob["beta"], prob["mu"]
    p.n_atoms, p.atom_norb = 0, None
    p.alpha_mix, p.nhist, p.maxscf, p.tolf, p.tolf_order1 = 0.5, 16, 1, 1e-8, 1e-8
    p.single_history, p.nthreads = 1, threads
    t_dys = L.orc_bench_dyson(ctypes.byref(p)) / e_sample
    # Sigma2: zero-padded forward transforms of 2 nA^2 columns and nAs^2 site pairs
    cols, pairs = max(threads * 4, 32), max(threads, 16)
    t_s2 = L.orc_bench_sigma2(nw, cols, pairs, threads)
    t_cols = L.orc_bench_sigma2(nw, cols, 0, threads)
    t_col = t_cols / cols
    t_pair = max(t_s2 - t_cols, 0.0) / pairs
    # Broyden + unpack on a sample of the vector
    nb = 20_000_000
    t_br = L.orc_bench_broyden(nb, 16, threads) / nb
    t_iter = t_dys * nw + t_col * 2 * nA * nA + t_pair * nAs * nAs + t_br * n_vec
    t_point = t_iter * n_iter_avg + t_dys * nw   # + the final sweep
    out = {"value": 1.0 / t_point, "unit": "bias points/s", "cores": threads, "kind": "port",
           "extrapolated": True,
           "sample": ("oracle (C, OpenMP, %d threads, OMP_PROC_BIND=close, OMP_PLACES=cores) timed on %d energies "
                      "of the Dyson/Keldysh sweep, %d zero-padded length-%d FFT columns and %d site pairs of "
                      "Sigma2, and one Broyden step on %d elements; EXTRAPOLATED to one SCF iteration (%.1f s) x "
                      "%.1f iterations per bias point (the GPU run's mean) + the final sweep"
                      % (threads, e_sample, cols, 2 * nw, pairs, nb, t_iter, n_iter_avg)),
           "s_per_iteration": t_iter}
    # complete bias points of the 16-spin-orbital helix (not extrapolated)
    p16, v16, _ = workload("helix16")
    pr = helix_problem(4, 2, 1.0, order=2, n_omega=p16["n_omega"], wlo=p16["w_init"],
                       whi=p16["w_init"] + p16["n_omega"] * p16["delta"], E_CC=0.0, t_hop=-1.0, lamb=0.01,
                       gamma=0.1, mu=0.0, tolf=1e-8, nhist=16, voltages=v16[:2])
    o = Oracle(pr, nthreads=threads)
    t0 = time.perf_counter()
    steps, fresh = [], True
    for V in v16[:2]:
        r = o.solve(V, fresh)
        fresh = not r["converged"]
        steps.append(r["n_iter"])
    tc = time.perf_counter() - t0
    out["helix16_full_solve"] = {"value": 2.0 / tc, "unit": "bias points/s", "cores": threads,
                                 "sample": "two complete bias points (V = 0 fresh, V = 0.1 continued; %d and %d SCF "
                                           "steps) of the 16-spin-orbital helix at 2^14 energies, timed end to end"
                                           % tuple(steps)}
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(args, n_iter_avg):
    """The oracle in a child process, so that OpenMP placement is set before
    its runtime starts; returns the cpu_baseline object (or an error note)."""
    import subprocess
    threads = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", "--workload", args.workload,
           "--cpu-threads", str(threads), "--cpu-iters", "%.3f" % n_iter_avg]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return {"value": None, "error": (r.stderr or r.stdout)[-400:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


# largest grid per size for the serving stream: one request's footprint (x, F, Fold,
# xold, x - xold and 16 FP32 history columns per packed element) stays near 1/4 of HBM
SERVE_SIZES = {16: ((4, 2), 16), 48: ((8, 3), 15), 80: ((8, 5), 14), 160: ((8, 10), 12), 240: ((12, 10), 11)}


def serve_bench(args):
    """Poisson arrivals of single-bias solves (the serving stream): helix systems of
    16/48/80/160/240 spin-orbitals drawn uniformly, grids of 2^13..2^16 energies on
    (-12, 12] drawn uniformly and capped per size (SERVE_SIZES), U ~ U[0.5, 3] eV,
    V ~ U[-2, 2] V, fresh starts, tolf 1e-8, nhist 16."""
    import importlib
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    serving = importlib.import_module("quichotte-and-sancho_amd.serving")
    rates = [float(r) for r in args.serve.split(",")]
    sizes = [int(x) for x in args.serve_sizes.split(",")]
    out = {"metric": "serving: converged bias points/s and latency", "unit": "bias points/s", "n_gpus": 1,
           "dtype": "f64", "data": "synthetic", "higher_is_better": True,
           "config": {"workload": ("helix systems of %s spin-orbitals, grids 2^13..2^16 on (-12,12] capped per size "
                                   "at %s, U~U[0.5,3], V~U[-2,2], fresh starts, tolf 1e-8, nhist 16 FP32 history, "
                                   "Poisson arrivals" % (sizes, {n: "2^%d" % SERVE_SIZES[n][1] for n in sizes})),
                      "requests_per_rate": args.requests, "max_running": args.serve_streams,
                      "fused_group": bool(args.serve_fuse)},
           "rates": {}}
    for rate in rates:
        rng = np.random.default_rng(int(rate))
        t, reqs = 0.0, []
        for rid in range(args.requests):
            t += rng.exponential(1.0 / rate)
            norb = int(rng.choice(sizes))
            lg = min(int(rng.choice([13, 14, 15, 16])), SERVE_SIZES[norb][1])
            U = float(rng.uniform(0.5, 3.0))
            V = float(rng.uniform(-2.0, 2.0))
            (ni, nt), _ = SERVE_SIZES[norb]
            H = helix_H(ni, nt)
            n = H.shape[0]
            GL = np.zeros((n, n), complex)
            GR = np.zeros((n, n), complex)
            for i in (0, 1):
                GL[i, i] = 0.1
                GR[n - 2 + i, n - 2 + i] = 0.1
            Us = np.zeros(n)
            Us[0::2] = U
            nw = 1 << lg
            prob = pkg.ScfProblem(H=H, gammaL=GL, gammaR=GR, U_site=Us, order=2, n_omega=nw, w_init=-12.0,
                                  delta=24.0 / nw, beta=1.0 / (KB * 300.0), mu=0.0, nhist=16, tolf=1e-8)
            reqs.append(serving.Request(rid, prob, V, arrival=t))
        def progress(r, tag):
            print("bench: serve rate %g/s %s: request %d done (%s, %d steps, %.2f s after arrival; device steps %s, "
                  "Dyson shape %d)" % (rate, tag, r.rid, "converged" if r.converged else ("error" if r.error else
                                       "not converged"), r.n_iter, r.latency, r.device_steps, r.dyson_shape),
                  file=sys.stderr, flush=True)

        # with --serve-fuse the requests of at most 82 active spin-orbitals are stepped in one group
        # (shared Dyson launches, joining and leaving between steps); the others run on their own threads
        srv = serving.Server(max_running=args.serve_streams, fuse=args.serve_fuse,
                             on_done=lambda r: progress(r, "mix"))
        t0 = time.perf_counter()
        res = srv.run(reqs)
        wall = time.perf_counter() - t0
        srv.shutdown()
        lat = np.array([r.latency for r in res])
        ok = [r for r in res if r.converged and not r.error]
        errs = [r.error for r in res if r.error]
        by_size = {}
        for r, q in zip(sorted(res, key=lambda r: r.rid), reqs):
            n = q.problem.H.shape[0]
            by_size.setdefault(n, []).append(r)
        per_size = {str(n): {"requests": len(v), "converged": sum(r.converged for r in v),
                             "median_steps": float(np.median([r.n_iter for r in v])),
                             "median_service_ms": 1e3 * float(np.median([r.done - r.start for r in v])),
                             "median_latency_ms": 1e3 * float(np.median([r.latency for r in v]))}
                    for n, v in sorted(by_size.items())}
        out["rates"][str(rate)] = {"per_size": per_size, "bias_points_per_s": len(ok) / wall, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                                   "p99_ms": 1e3 * float(np.percentile(lat, 99)), "converged": len(ok),
                                   "requests": len(res), "wall_s": wall, "handles_reused": srv.reused,
                                   "grouped_requests": sum(1 for r in res if r.grouped), "errors": errs[:3]}
        print("bench: serve rate %g/s: %d/%d converged, %.2f points/s, p50 %.0f ms, p99 %.0f ms, %d handles reused"
              % (rate, len(ok), len(res), len(ok) / wall, 1e3 * np.percentile(lat, 50), 1e3 * np.percentile(lat, 99),
                 srv.reused), file=sys.stderr, flush=True)
        if args.serve_plain:
            # the plain FP64 path on the same stream: one request on the GPU at a time
            srv1 = serving.Server(max_running=1, on_done=lambda r: progress(r, "plain"))
            t0 = time.perf_counter()
            res1 = srv1.run(reqs)
            wall1 = time.perf_counter() - t0
            srv1.shutdown()
            lat1 = np.array([r.latency for r in res1])
            ok1 = [r for r in res1 if r.converged and not r.error]
            same = all(a.J_up == b.J_up and a.J_down == b.J_down and a.n_iter == b.n_iter
                       for a, b in zip(sorted(res, key=lambda r: r.rid), sorted(res1, key=lambda r: r.rid)))
            d = out["rates"][str(rate)]
            d["plain_one_at_a_time"] = {"bias_points_per_s": len(ok1) / wall1, "wall_s": wall1, "converged": len(ok1),
                                        "p50_ms": 1e3 * float(np.percentile(lat1, 50)),
                                        "p99_ms": 1e3 * float(np.percentile(lat1, 99))}
            d["speedup_vs_plain"] = (len(ok) / wall) / (len(ok1) / wall1) if ok1 else None
            d["results_equal_to_plain"] = same
            print("bench: serve rate %g/s, plain (one at a time): %.2f points/s, p50 %.0f ms, p99 %.0f ms; same results %s"
                  % (rate, len(ok1) / wall1, 1e3 * np.percentile(lat1, 50), 1e3 * np.percentile(lat1, 99), same),
                  file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0


# batched bias points of one system: (helix ions, turns, U, energy window, V window, default log2 grid, points)
BATCH_WORKLOADS = {"helix16": ((4, 2), 1.0, 10.0, (0.0, 1.0), 14, 64),
                   "helix80": ((8, 5), 2.0, 12.0, (-1.0, 1.0), 12, 32)}


def batch_bench(args):
    """Many independent bias points of one system (seeded V, fresh starts, tolf 1e-8,
    nhist 16), solved three ways on the same points: one at a time on one handle (the
    plain path), by B host threads with a handle and stream each, and by a group of B
    handles stepped together (qs_group_*: one Dyson launch per step for all members,
    a finished point's handle takes the next point at once)."""
    import importlib
    import threading
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    (ni, nt), U, wmax, (v0, v1), lg_def, np_def = BATCH_WORKLOADS[args.batch]
    lg = args.batch_log2nw or lg_def
    npts = args.batch_points or np_def
    B = args.batch_size
    H = helix_H(ni, nt)
    n = H.shape[0]
    GL = np.zeros((n, n), complex)
    GR = np.zeros((n, n), complex)
    for i in (0, 1):
        GL[i, i] = 0.1
        GR[n - 2 + i, n - 2 + i] = 0.1
    Us = np.zeros(n)
    Us[0::2] = U
    nw = 1 << lg
    prob = pkg.ScfProblem(H=H, gammaL=GL, gammaR=GR, U_site=Us, order=2, n_omega=nw, w_init=-wmax,
                          delta=2 * wmax / nw, beta=1.0 / (KB * 300.0), mu=0.0, nhist=16, tolf=1e-8)
    Vs = [float(v) for v in np.random.default_rng(3).uniform(v0, v1, npts)]
    desc = ("%d bias points of the helix %dx%d (%d spin-orbitals), U=%g, order 2, 2^%d energies on (-%g,%g], "
            "V ~ U[%g,%g] (seed 3), fresh starts, tolf 1e-8, Broyden nhist 16 FP32 history"
            % (npts, ni, nt, n, U, lg, wmax, wmax, v0, v1))
    print("bench: batch: %s; %d in flight" % (desc, B), file=sys.stderr, flush=True)
    engines = [pkg.ScfEngine(prob) for _ in range(B)]
    print("bench: batch: %d handles, %.1f GB each, Dyson launch shape %s" %
          (B, engines[0].device_bytes() / 1e9, engines[0].dyson_shape_name()), file=sys.stderr, flush=True)
    # warm-up: code loading and graph-free first launches on every handle
    for e in engines:
        e.solve_scf(Vs[0], True)

    def plain():
        e = engines[0]
        res = []
        t0 = time.perf_counter()
        for V in Vs:
            r = e.solve_scf(V, True)
            res.append((r, e.current(), time.perf_counter() - t0))
        return time.perf_counter() - t0, res

    def threads():
        res = [None] * npts
        nxt = [0]
        lock = threading.Lock()
        t0 = time.perf_counter()

        def worker(e):
            while True:
                with lock:
                    k = nxt[0]
                    nxt[0] += 1
                if k >= npts:
                    return
                r = e.solve_scf(Vs[k], True)
                res[k] = (r, e.current(), time.perf_counter() - t0)

        ts = [threading.Thread(target=worker, args=(e,)) for e in engines]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return time.perf_counter() - t0, res

    def group():
        g = pkg.ScfGroup(0)
        res = [None] * npts
        point = {}
        k = 0
        t0 = time.perf_counter()
        for e in engines:
            if k < npts:
                g.add(e, Vs[k], True)
                point[id(e)] = k
                k += 1
        steps = 0
        while len(g):
            for e, r, st in g.step():
                if st != 0:
                    raise RuntimeError("group member failed: %d" % st)
                res[point[id(e)]] = (r, e.current(), time.perf_counter() - t0)
                if k < npts:
                    g.add(e, Vs[k], True)
                    point[id(e)] = k
                    k += 1
            steps += 1
        total = time.perf_counter() - t0
        launches, fused, dyson_ms, gsteps = g.info()
        g.close()
        return total, res, {"group_steps": steps, "dyson_ms": dyson_ms}

    def two_groups():
        # two groups of B/2, each step begun while the other group's step runs
        gs = [pkg.ScfGroup(0), pkg.ScfGroup(0)]
        res = [None] * npts
        point = {}
        k = 0
        t0 = time.perf_counter()
        for j, e in enumerate(engines):
            if k < npts:
                gs[j % 2].add(e, Vs[k], True)
                point[id(e)] = (k, j % 2)
                k += 1
        for g in gs:
            g.step_begin()
        live = [True, True]
        while any(live):
            for j, g in enumerate(gs):
                if not live[j]:
                    continue
                for e, r, st in g.step_end():
                    if st != 0:
                        raise RuntimeError("group member failed: %d" % st)
                    res[point[id(e)][0]] = (r, e.current(), time.perf_counter() - t0)
                    if k < npts:
                        g.add(e, Vs[k], True)
                        point[id(e)] = (k, j)
                        k += 1
                if len(g):
                    g.step_begin()
                else:
                    live[j] = False
        total = time.perf_counter() - t0
        for g in gs:
            g.close()
        return total, res

    out = {}
    for name, fn in (("plain", plain), ("threads", threads), ("group", group), ("two_groups", two_groups)):
        r = fn()
        total, res = r[0], r[1]
        ok = [x for x in res if x[0].converged]
        lat = np.array([x[2] for x in res])
        out[name] = {"value": len(ok) / total, "wall_s": total, "converged": len(ok), "points": npts,
                     "scf_steps_per_point": float(np.mean([x[0].n_iter for x in res])),
                     "completion_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                       "p99": 1e3 * float(np.percentile(lat, 99))},
                     "J": [list(x[1]) for x in res]}
        if len(r) > 2:
            out[name].update(r[2])
        print("bench: batch %s: %d/%d converged in %.2f s, %.2f points/s, %.1f steps per point" %
              (name, len(ok), npts, total, len(ok) / total, out[name]["scf_steps_per_point"]),
              file=sys.stderr, flush=True)
    Jp = np.array(out["plain"]["J"])
    same = {nm: bool(np.array_equal(np.array(out[nm]["J"]), Jp)) for nm in ("threads", "group", "two_groups")}
    for nm in out:
        del out[nm]["J"]
    line = {"metric": "batched bias points: converged bias points/s", "value": out["group"]["value"],
            "unit": "bias points/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": desc, "in_flight": B,
                       "form": "group of %d handles stepped together, one Dyson launch per step" % B},
            "plain": out["plain"], "threads": out["threads"], "group": out["group"], "two_groups": out["two_groups"],
            "speedup_vs_plain": out["group"]["value"] / out["plain"]["value"],
            "threads_speedup_vs_plain": out["threads"]["value"] / out["plain"]["value"],
            "currents_equal_to_plain": same}
    for e in engines:
        e.close()
    print(json.dumps(line), flush=True)
    return 0


def chain_gf(N, dw, up, delta, E=2.0, t=-0.4, gamma=1.0, mu=2.0, T=300.0):
    """Non-interacting G^r, G^< of SANCHO's chain (E_CC on every site, t_hop
    between neighbours, GammaL on site 1, GammaR on site N) on its grid
    w_init = Ev(1) - dw, n_omega = int((Ev(N) + up - w_init) / delta)
    (SANCHO/src/sancho.f90:84-91), V = 0."""
    H = np.diag(np.full(N, E)) + np.diag(np.full(N - 1, t), 1) + np.diag(np.full(N - 1, t), -1)
    ev = np.linalg.eigvalsh(H)
    w_init = ev[0] - dw
    nw = int((ev[-1] + up - w_init) / delta)
    GamL = np.zeros((N, N))
    GamR = np.zeros((N, N))
    GamL[0, 0] = gamma
    GamR[-1, -1] = gamma
    beta = 1.0 / (KB * T)
    GR = np.empty((N, N, nw), complex, order="F")
    GL = np.empty((N, N, nw), complex, order="F")
    for k in range(nw):
        w = w_init + (k + 1) * delta
        r = np.linalg.inv(w * np.eye(N) - H + 0.5j * (GamL + GamR))
        f = 1.0 / (np.exp(min((w - mu) * beta, 700.0)) + 1.0)
        GR[:, :, k] = r
        GL[:, :, k] = r @ (1j * f * (GamL + GamR)) @ r.conj().T
    return GR, GL, delta, nw


SANCHO_WORKLOADS = {
    # name: (N_sites, dw, up, order): the reference's 8-site order-3 run and 20-site order-2 example
    "chain_N8_order3": (8, 3.0, 3.0, 3),
    "chain_N20_order2": (20, 2.0, 2.0, 2),
}


def sancho_bench(args):
    """SANCHO's self-energies (complete_input_gf + build_self_energies,
    SANCHO/src/self_energies.f90:38-207) on device arrays: builds per second.
    Algorithmic flops per build, na sites, nw energies (complex multiply-add = 8):
    chi and phi 48 na^2 nw^2 each, Sigma2 32 na^2 nw^2, Sigma3 64 na^2 nw^2,
    the internal-site sums 56 na^3 (2nw - 1) each (order 3)."""
    import importlib
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    N, dw, up, order = SANCHO_WORKLOADS[args.sancho]
    GR, GL, delta, nw = chain_gf(N, dw, up, 0.01)
    U = np.full(N, 0.8)
    flops = 48 * N * N * nw * nw + 32 * N * N * nw * nw
    if order == 3:
        flops += 48 * N * N * nw * nw + 64 * N * N * nw * nw + 2 * 56 * N ** 3 * (2 * nw - 1)
    s = pkg.SanchoSelfEnergies(N, nw, order)
    nbytes = 16 * N * N * nw
    gr, gl = pkg.DeviceBuffer(nbytes), pkg.DeviceBuffer(nbytes)
    gr.upload(GR.ravel(order="F"))
    gl.upload(GL.ravel(order="F"))
    outs = [pkg.DeviceBuffer(nbytes) for _ in range(s.n_outputs)]
    # inputs resident in HBM; each build is synchronous (ends with a stream synchronize)
    for _ in range(args.warmup):
        s.build_device(gr, gl, outs, U, delta)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.build_device(gr, gl, outs, U, delta)
    dt = (time.perf_counter() - t0) / args.steps
    # the two summation stages timed with HIP events on the library's stream, in a second loop
    s.profile(True)
    for _ in range(args.steps):
        s.build_device(gr, gl, outs, U, delta)
    prof = s.profile_read()
    s.profile(False)
    stage_flops = {"bubbles": 48 * N * N * nw * nw * (2 if order == 3 else 1),
                   "self_energies": (32 + (64 if order == 3 else 0)) * N * N * nw * nw}
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items()}
    dom = max(stage_ms, key=stage_ms.get)
    ach = stage_flops[dom] / (stage_ms[dom] * 1e-3) / 1e12
    line = {"metric": "SANCHO self-energy builds/s", "value": 1.0 / dt, "unit": "builds/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "SANCHO %s: %d sites, %d energies, order %d, non-interacting chain G" %
                                   (args.sancho, N, nw, order)},
            "roofline": {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / FP64_PEAK_TFLOPS, "traffic": None, "kernel": dom,
                         "avg_launch_ms": stage_ms[dom],
                         "note": "FP64 (vector FMA; the peak is the same for MFMA) over the stage's HIP-event time"},
            "stages_ms": stage_ms, "whole_call_tflops": flops / dt / 1e12}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from qs_testlib import sancho_oracle
        t0 = time.perf_counter()
        sancho_oracle(GR, GL, U, delta, order)
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": 1.0 / tc, "unit": "builds/s", "cores": args.cpu_threads, "kind": "port",
                                "sample": "one full build of the same workload (oracle/sancho_oracle.c, OpenMP)"}
    print(json.dumps(line), flush=True)
    return 0


def copy_calibration():
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch scale
    the same counters read on the kernels of the path to HBM bytes."""
    import torch
    n = 1 << 28                      # 2^28 doubles = 2 GiB each way
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}), flush=True)
    return 0


def make_parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="helix80", choices=["helix80", "helix16", "helix240"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: the host's physical cores)")
    ap.add_argument("--cpu-iters", type=float, default=0.0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: time the mixed-precision form after the plain FP64 path on the same points")
    ap.add_argument("--mixed-switch", type=float, default=1e-5,
                    help="max|F| below which the mixed form leaves FP32 sweeps")
    ap.add_argument("--serve", default="", help="comma-separated Poisson arrival rates (requests/s)")
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--serve-sizes", default="16,48,80,160,240", help="spin-orbital counts of the serving stream")
    ap.add_argument("--serve-fuse", action="store_true",
                    help="--serve: step the requests of at most 82 active spin-orbitals in one group (fused Dyson launches)")
    ap.add_argument("--serve-plain", action="store_true",
                    help="also serve each rate's stream one request at a time (the plain path) for comparison")
    ap.add_argument("--maxscf", type=int, default=0,
                    help="cap on SCF steps per bias point (counter-collection passes; 0 = the workload's 300)")
    ap.add_argument("--batch", default="", choices=[""] + list(BATCH_WORKLOADS),
                    help="time independent bias points of this system: one at a time, threads, and a group")
    ap.add_argument("--batch-points", type=int, default=0, help="bias points of --batch (default per system)")
    ap.add_argument("--batch-log2nw", type=int, default=0, help="log2 of the --batch grid (default per system)")
    ap.add_argument("--batch-size", type=int, default=16, help="solves in flight for --batch")
    ap.add_argument("--sancho", default="", choices=[""] + list(SANCHO_WORKLOADS),
                    help="time SANCHO's order-2/3 self-energy build on this workload instead")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain device copy of 2 GiB five times (scale for HBM byte counters) and exit")
    return ap


def main():
    args = make_parser().parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (one bias point is one "
              "GPU's work; several GPUs would each take their own bias points or energy ranges)" % args.gpus,
              file=sys.stderr)
        return 2
    if args.cpu_baseline_child:
        return cpu_baseline_child(args)
    if args.cpu_threads <= 0:
        args.cpu_threads = host_cores()
    if args.serve:
        return serve_bench(args)
    if args.sancho:
        return sancho_bench(args)
    if args.batch:
        return batch_bench(args)
    if args.calibrate_copy:
        return copy_calibration()
    import importlib
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    prob, volts, desc = workload(args.workload)
    if args.maxscf > 0:
        prob["maxscf"] = args.maxscf
    nv = len(volts)
    print("bench: %s" % desc, file=sys.stderr, flush=True)
    t0 = time.time()
    eng = pkg.ScfEngine(pkg.ScfProblem(**prob))
    eng.set_step_mode(2)   # one handle in this process: steps replayed from captured graphs
    dv, dms = eng.dyson_choice()
    print("bench: device memory %.1f GB, setup %.1f s, Dyson launch shape %d (tuned ms %s)" %
          (eng.device_bytes() / 1e9, time.time() - t0, dv, dms), file=sys.stderr, flush=True)
    # warmup: fresh solves of sweep points; they leave no state the timed sweep uses
    for k in range(args.warmup):
        t1 = time.time()
        r = eng.solve_scf(volts[k % nv], True)
        print("bench: warmup (fresh) V=%+.2f converged=%s steps=%d %.2f s" % (volts[k % nv], r.converged, r.n_iter,
                                                                              time.time() - t1),
              file=sys.stderr, flush=True)
    def timed_sweep(tag):
        eng.profile(True)
        lat, iters, conv, Js, n32 = [], [], [], [], []
        graph_steps.clear()
        fresh = True
        tstart = time.perf_counter()
        for k, i in enumerate(sweep_points(nv, args.steps)):
            if i == 0:
                fresh = True
            t1 = time.perf_counter()
            r = eng.solve_scf(volts[i], fresh)
            J = eng.current()
            lat.append(time.perf_counter() - t1)
            fresh = not r.converged
            iters.append(r.n_iter)
            conv.append(r.converged)
            Js.append(J)
            n32.append(eng.precision_steps()[0])
            graph_steps.append(eng.step_info()[1])
            print("bench: %s V=%+.2f %s converged=%s steps=%d (FP32 %d) J=(%.9e, %.9e) %.3f s" %
                  (tag, volts[i], "fresh" if (i == 0 or (k and not conv[-2])) else "cont.", r.converged, r.n_iter,
                   n32[-1], J[0], J[1], lat[-1]), file=sys.stderr, flush=True)
        total = time.perf_counter() - tstart
        prof = eng.profile_read()
        eng.profile(False)
        return total, lat, iters, conv, np.array(Js), n32, prof

    graph_steps = []
    plain = None
    if args.precision == "mixed":
        plain = timed_sweep("plain")
        eng.set_precision(1, args.mixed_switch)
    total, lat, iters, conv, Js, n32, prof = timed_sweep(args.precision)
    counts = stage_counts(eng, prob["nhist"])
    counts["dyson_keldysh_fp32"] = counts["dyson_keldysh"]
    dom = max((s for s in prof if s in counts and prof[s][1] > 0), key=lambda s: prof[s][0])
    bound, work, kind = counts[dom]
    ms, nl = prof[dom]
    avg_s = ms / 1e3 / nl
    if kind == "flop":
        ach = work / avg_s / 1e12
        # "mfma" only where the kernel issues matrix instructions; the register-resident
        # Dyson shapes do their FP64 arithmetic as vector FMAs (same 78.6 TFLOP/s peak)
        fp32 = dom == "dyson_keldysh_fp32"
        peak = FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS
        roof = {"bound": "mfma" if (fp32 or eng.dyson_uses_mfma()) else "fp64", "achieved": ach, "peak": peak,
                "unit": "TFLOP/s", "frac": ach / peak, "traffic": None}
    else:
        ach = work / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None}
    roof["traffic"] = measured_traffic(args.workload, dom)
    roof["kernel"] = dom
    roof["avg_launch_ms"] = avg_s * 1e3
    stages = {s: {"ms": round(v[0], 3), "launches": v[1]} for s, v in prof.items() if v[1] > 0}
    line = {
        "metric": "converged bias points/s",
        "value": args.steps / total,
        "unit": "bias points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "form": "plain FP64 path, one bias point at a time",
                   "sweep": ("%d timed bias points walking the %d-point voltage sweep with continuation from a fresh "
                             "start at V=%g%s; %d warmup points solved fresh beforehand"
                             % (args.steps, nv, volts[0], ", starting over fresh after the last point"
                                if args.steps > nv else "", args.warmup)),
                   "n_spin_orbitals": eng.norb, "n_omega": eng.prob.n_omega, "fft_length": eng.Nf,
                   "dyson_launch_shape": eng.dyson_shape_name()},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "scf_steps_per_point": float(np.mean(iters)),
        "graph_replayed_steps_per_point": float(np.mean(graph_steps)) if graph_steps else 0.0,
        "all_converged": bool(all(conv)),
        "stages": stages,
        "roofline": roof,
    }
    if plain is not None:
        p_total, p_lat, p_iters, p_conv, p_J = plain[:5]
        rel = np.abs(Js - p_J) / np.maximum(np.abs(p_J), 1e-300)
        line["dtype"] = "c64 MFMA Dyson/Keldysh + f64"
        line["config"]["form"] = ("mixed precision: complex-FP32 MFMA Dyson/Keldysh while max|F| >= %g, FP64 steps "
                                  "to tolf/10, FP64 final sweep; the plain FP64 path timed on the same points first"
                                  % args.mixed_switch)
        line["plain_fp64"] = {"value": args.steps / p_total, "ms_per_step": 1e3 * p_total / args.steps,
                              "latency_ms": {"p50": 1e3 * float(np.percentile(p_lat, 50)),
                                             "p99": 1e3 * float(np.percentile(p_lat, 99))},
                              "scf_steps_per_point": float(np.mean(p_iters)), "all_converged": bool(all(p_conv))}
        line["speedup_vs_plain"] = p_total / total
        line["fp32_steps_per_point"] = float(np.mean(n32))
        line["rel_dI_vs_plain"] = {"max": float(rel[np.abs(p_J) > 1e-9].max()) if np.any(np.abs(p_J) > 1e-9) else 0.0,
                                   "per_point": [[float(volts[i]), float(rel[k, 0]), float(rel[k, 1])]
                                                 for k, i in enumerate(sweep_points(nv, args.steps))]}
    if not args.no_cpu_baseline:
        print("bench: cpu baseline (oracle, %d threads)" % args.cpu_threads, file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args, float(np.mean(iters)))
    eng.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
