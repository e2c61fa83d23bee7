"""GPU: every bench.py mode runs to its one JSON line with the contract's keys.

Each mode runs once as a child process on small inputs: the plain renderer on a
512 x 512 scene, a 48-scene batch and the differentiable mode with a few steps.  The
numbers are not checked here, only that each mode completes and prints the fields its
consumers read.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMMON = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
          "scaling", "vs_baseline", "dtype", "data", "config"}


def run_bench(*args):
    print("bench", args, flush=True)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    line = p.stdout.strip().splitlines()[-1]
    return json.loads(line)


def test_plain_mode():
    d = run_bench("--steps", "2", "--warmup", "1", "--size", "512", "--paths", "400", "--no-cpu-baseline")
    assert COMMON <= set(d)
    assert d["n_gpus"] == 1 and d["value"] > 0 and d["unit"] == "Mfragments/s"
    assert {"bound", "achieved", "peak", "unit", "frac", "traffic"} <= set(d["roofline"])
    assert d["count_pixel"] > 0 and d["frame_event_ms_median"] > 0


def test_batch_mode():
    d = run_bench("--mode", "batch", "--batch-scenes", "48", "--steps", "2", "--warmup", "1")
    assert COMMON <= set(d)
    assert d["spot_check_mismatches"] == 0 and d["value"] > 0


def test_diff_mode():
    d = run_bench("--mode", "diff", "--steps", "2", "--warmup", "1")
    assert COMMON <= set(d)
    assert d["forward_ms"] > 0 and d["backward_ms"] > 0


def test_animation_mode():
    d = run_bench("--mode", "animation", "--steps", "3", "--warmup", "1", "--no-cpu-baseline")
    assert COMMON <= set(d)
    assert d["unit"] == "frames/s" and d["value"] > 0 and d["generate_us"] > 0
    assert {"bound", "achieved", "peak", "unit", "frac", "traffic"} <= set(d["roofline"])
    assert d["counts"]["path"] > 0 and d["counts"]["curve"] == 6 * d["counts"]["path"]
