"""bench.py host logic (no GPU): the one-GPU contract, the algorithmic bytes
behind roofline.achieved, and the committed traffic summaries behind
roofline.traffic."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from diffuse_vs_gfm_mms_amd import workloads as W  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


@pytest.mark.parametrize("precision, grid, expected", [
    # FP64: 6 fields x 8 B, read and written: 96 B per cell
    ("fp64", (2048, 1024), 96 * 2048 * 1024),
    # FP32-mixed: 2 FP64 partial masses + 4 FP32 fields, read and written: 64 B per cell
    ("fp32_mixed", (4096, 2048), 64 * 4096 * 2048),
])
def test_sweep_bytes(precision, grid, expected):
    case = W.helium_shock_bubble_2d(16, 8)
    case.N = grid  # only the cell count and the field counts enter
    assert bench.sweep_bytes(case, precision) == expected


def test_3d_sweep_bytes():
    case = W.bubble_collapse_3d(8, build_state=False)
    case.N = (512, 512, 512)
    # FP32-mixed 3D: 2 x (2 x 8 B + 5 x 4 B) per cell
    assert bench.sweep_bytes(case, "fp32_mixed") == 2 * (2 * 8 + 5 * 4) * 512 ** 3


@pytest.mark.parametrize("workload, algorithmic", [("helium_fp64", 201326592), ("helium_fp32", 536870912)])
def test_traffic_summaries(workload, algorithmic):
    """The committed counter summaries give each sweep axis's HBM bytes per
    launch, at least the algorithmic bytes, and bench.py reads the axis it
    is asked for."""
    with open(os.path.join(ROOT, "profiles", f"traffic_{workload}.json")) as f:
        t = json.load(f)
    assert t["algorithmic_bytes_per_sweep_launch"] == algorithmic
    for axis in "xy":
        traffic = bench.traffic_for(workload, axis)
        assert isinstance(traffic, int)
        assert traffic == t[f"{axis}_sweep_bytes"]
        assert traffic >= algorithmic
    assert bench.traffic_for(workload, "z") is None


def test_dominant_sweep_roofline():
    """The line's roofline is the slowest sweep launch's, with every axis beside it."""
    case = W.helium_shock_bubble_2d(64, 32)
    r = bench.sweep_roofline(case, "fp64", [0.5, 0.01, 0.0], [0.4, 0.6, 0.0])
    b = bench.sweep_bytes(case, "fp64")
    assert r["kernel"].startswith("y sweep")
    assert abs(r["achieved"] - b / 0.6e-3 / 1e9) < 1e-9
    assert r["frac"] == r["per_axis"]["y"]["frac"] < r["per_axis"]["x"]["frac"]


def test_batch_tuning_cache(tmp_path):
    """The autotuner cache is keyed by the instance set and keeps other keys."""
    from diffuse_vs_gfm_mms_amd import dimhip
    a = W.batch_instances(count=3)
    b = W.batch_instances(count=4)
    ka, kb = dimhip.batch_tuning_key(a), dimhip.batch_tuning_key(b)
    assert ka == dimhip.batch_tuning_key(W.batch_instances(count=3)) and ka != kb
    path = str(tmp_path / "tune.json")
    assert dimhip.load_tuned(path, ka) is None
    entry = {"best": {"x": 1, "y": 2, "schedule": 0, "blocks_per_cu": 4, "ms": 0.5}}
    dimhip.store_tuned(path, ka, entry)
    dimhip.store_tuned(path, kb, dict(entry, best=dict(entry["best"], x=0)))
    assert dimhip.load_tuned(path, ka) == entry
    assert dimhip.load_tuned(path, kb)["best"]["x"] == 0
