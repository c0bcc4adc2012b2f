"""bench.py's host logic, no GPU: the committed counter summaries that the
roofline's `traffic` comes from cover every tile shape the collide-stream
autotuner can pick, the --gpus guard, and the timed region's placement
between two FP32 re-centrings."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import ekpnp  # noqa: E402


@pytest.mark.parametrize("config,nodes", [("c4", 8192 * 2046), ("c1", 4096 * 1022)])
def test_traffic_summaries_cover_every_tile_shape(config, nodes):
    with open(os.path.join(ROOT, "profiles", f"k_lbm_{config}_fp32_traffic.json")) as f:
        d = json.load(f)
    assert d["bytes_per_launch_by_tile_rows"]["0"] == d["bytes_per_launch"]
    for wrows in (0, 2, 4):
        t = bench.measured_traffic(config, "fp32", wrows)
        # at least the 232 algorithmic bytes per interior node, at most a few
        # per cent above them
        assert 232 * nodes <= t < 1.1 * 232 * nodes, (wrows, t)
    assert bench.measured_traffic("c2", "fp32", 0) is None  # no summary for the sweep


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


@pytest.mark.parametrize("warmup,steps", [(20, 200), (10, 40), (5, 7), (0, 1), (2, 4)])
def test_timed_region_lies_between_two_recentrings(warmup, steps):
    # bench.py's timed steps start at timed_start() and must not contain a
    # re-centring step of FP32 storage; the warm-up and the graph-building
    # calls before them must fit in front
    start = bench.timed_start(warmup, steps, "fp32")
    pre = 1 + (steps & 1)
    assert start - pre * steps - warmup >= 0
    rc = ekpnp.rebase_steps(start + steps + 1)
    # none inside the timed steps, nor inside the untimed calls that capture
    # their graphs just before them
    assert not any(start - pre * steps < r <= start + steps for r in rc), (start, rc)
