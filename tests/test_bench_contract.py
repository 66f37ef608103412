"""bench.py's contract on the CPU: it measures one GPU and refuses more; the roofline record's
algorithmic bytes (864 B per TVD-RK3 cell-update, 496 for the predictor-corrector, 144 per cell
for each CFL reduction in the timed steps); the resident-tape budget of a device that runs an
unroll alone."""
import os
import subprocess
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _run(args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True,
                          text=True, timeout=120, env=dict(os.environ, **(env or {})))


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2"])
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_a_multi_rank_launch_is_refused():
    r = _run([], env={"WORLD_SIZE": "4"})
    assert r.returncode == 2 and "one GPU" in r.stderr


def _jobs(time_scheme=0, viscous=False, cfl_count=10):
    return [types.SimpleNamespace(cfl_count=cfl_count, viscous=viscous, time_scheme=time_scheme)]


@pytest.mark.parametrize("time_scheme,per_step", [(0, 864), (1, 496)])
def test_roofline_record_counts_the_algorithmic_bytes(time_scheme, per_step):
    cells, steps, warmup = 4194304, 20, 5
    # Timed steps 6..25 hold the CFL steps 10, 20: two reductions.
    kernel_ms = 33.0
    launches = steps * (2 if time_scheme else 3)
    rec = bench.roofline_record(_jobs(time_scheme), {}, steps, warmup, "none", kernel_ms, launches, cells)
    alg = cells * (per_step * steps + 144 * 2)
    assert rec["alg_bytes_per_launch"] == pytest.approx(alg / launches, rel=1e-15)
    assert rec["achieved"] == pytest.approx(alg / 33.0e-3 / 1e9, rel=1e-15)
    assert rec["frac"] == pytest.approx(alg / 33.0e-3 / 8.0e12, rel=1e-15)
    assert rec["peak"] == 8000.0 and rec["unit"] == "GB/s" and rec["bound"] == "hbm"
    assert rec["avg_launch_us"] == pytest.approx(33.0e3 / launches)


def test_roofline_record_counts_reductions_in_the_timed_steps_only():
    # Warmup 5, 4 timed steps (6..9): no CFL step among them.
    rec = bench.roofline_record(_jobs(), {}, 4, 5, "none", 1.0, 12, 1000)
    assert rec["alg_bytes_per_launch"] == pytest.approx(1000 * 864 * 4 / 12)
    # Warmup 9, 1 timed step (10): one.
    rec = bench.roofline_record(_jobs(), {}, 1, 9, "none", 1.0, 3, 1000)
    assert rec["alg_bytes_per_launch"] == pytest.approx(1000 * (864 + 144) / 3)


def test_dedicated_tape_budget(monkeypatch):
    torch = pytest.importorskip("torch")
    sim = types.SimpleNamespace(active_cells=1000000)
    seg = 8 * 6 * 1000000 * (3 * 50 + 1)  # one 50-step TVD-RK3 segment: 151 stage states
    free = {"bytes": 0}
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda *a: (free["bytes"], 2 * free["bytes"]))
    # Every segment fits beside the 8 GB margin: all of them, no recompute tape.
    free["bytes"] = int(10 * seg + 8e9)
    assert bench.dedicated_tape_budget(sim, _jobs(), 500, 50) == 10 * seg
    # They do not: as many as fit beside one recompute tape.
    free["bytes"] = int(4.5 * seg + 8e9)
    assert bench.dedicated_tape_budget(sim, _jobs(), 500, 50) == 3 * seg
    free["bytes"] = int(0.5 * seg)
    assert bench.dedicated_tape_budget(sim, _jobs(), 500, 50) == 0
    # Predictor-corrector: 2C + 1 stage states per segment.
    free["bytes"] = int(100 * seg)
    assert bench.dedicated_tape_budget(sim, _jobs(time_scheme=1), 300, 50) == 6 * 8 * 6 * 1000000 * 101
