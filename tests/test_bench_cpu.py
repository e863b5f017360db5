"""bench.py's host-side logic (no GPU): the driver's command line, the timed
sweep over the workload's voltages, and the host core count of the CPU
baseline."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


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
