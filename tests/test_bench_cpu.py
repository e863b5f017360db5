"""bench.py's host-side logic (no GPU): the driver's command line, the timed
sweep over the workload's voltages, and the host core count of the CPU
baseline."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402
import numpy as np  # noqa: E402


def test_driver_command_line_parses():
    a = bench.make_parser().parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    assert (a.gpus, a.steps, a.warmup, a.workload) == (1, 20, 5, "helix80")


def test_timed_steps_walk_the_sweep_and_wrap():
    prob, volts, desc = bench.workload("helix80")
    assert len(volts) == 21 and volts[0] == -1.0 and volts[-1] == 1.0
    assert prob["n_omega"] == 2 ** 16 and prob["H"].shape == (80, 80)
    # 20 timed points fit in one sweep; warmups never consume sweep points
    assert bench.sweep_points(len(volts), 20) == list(range(20))
    # more steps than points: the sweep starts over at its first point
    pts = bench.sweep_points(len(volts), 25)
    assert pts[:21] == list(range(21)) and pts[21:] == [0, 1, 2, 3]


def test_helix16_workload_is_the_small_benchmark_system():
    prob, volts, _ = bench.workload("helix16")
    assert prob["H"].shape == (16, 16) and prob["n_omega"] == 2 ** 14
    assert volts == [round(0.1 * k, 10) for k in range(11)]
    assert prob["w_init"] == -10.0 and prob["delta"] == 20.0 / 2 ** 14


def test_host_cores_is_positive_and_bounded():
    n = bench.host_cores()
    assert 1 <= n <= len(os.sched_getaffinity(0))


def test_more_than_one_gpu_is_refused(capsys, monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py", "--gpus", "2"])
    assert bench.main() == 2
    assert "one GPU" in capsys.readouterr().err


def test_batch_mode_arguments_and_workloads():
    a = bench.make_parser().parse_args(["--batch", "helix80", "--batch-size", "12", "--batch-log2nw", "12"])
    assert (a.batch, a.batch_size, a.batch_log2nw, a.batch_points) == ("helix80", 12, 12, 0)
    for name, ((ni, nt), U, wmax, (v0, v1), lg, npts) in bench.BATCH_WORKLOADS.items():
        H = bench.helix_H(ni, nt)
        assert H.shape == (2 * ni * nt, 2 * ni * nt) and np.allclose(H, H.conj().T)
        assert v0 < v1 and 10 <= lg <= 16 and npts > 0 and U > 0 and wmax > 0
    # the batch systems are the benchmark helices: 16 and 80 spin-orbitals
    assert bench.helix_H(*bench.BATCH_WORKLOADS["helix16"][0]).shape[0] == 16
    assert bench.helix_H(*bench.BATCH_WORKLOADS["helix80"][0]).shape[0] == 80
