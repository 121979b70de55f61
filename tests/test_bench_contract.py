"""bench.py's contract pieces that need no GPU: the single-GPU refusal, the
algorithmic byte model and the committed traffic record it reports."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single" in r.stderr


def test_algorithmic_bytes_per_site():
    # B = (9c + 5r) + 4c (N/2+1)/N per site: 108.0625 at N = 512 in FP32
    assert bench.algorithmic_bytes_per_site(512) == pytest.approx(108.0625)
    assert bench.algorithmic_bytes_per_site(512, c=16, r=8) == pytest.approx(216.125)


def test_committed_traffic_record_covers_every_pass():
    path = os.path.join(ROOT, "profiles", "kdk512_traffic.json")
    with open(path) as f:
        rec = json.load(f)
    passes = rec["passes"]
    for name in ("c2r_kick_drift2", "drift1_tiles", "drift0", "density_r2c", "vt_fwd1", "vt_poisson0", "vt_inv1"):
        assert name in passes, name
        assert passes[name]["fetch_bytes"] > 0 and passes[name]["write_bytes"] > 0
    assert bench.measured_traffic("c2r_kick_drift2", path) == pytest.approx(
        passes["c2r_kick_drift2"]["fetch_bytes"] + passes["c2r_kick_drift2"]["write_bytes"])
    assert bench.measured_traffic("no_such_pass", path) is None


def test_cpu_leg_threads_follow_the_share():
    info = bench.host_cpu()
    assert 1 <= info["threads"] <= info["cpu_share"]
    if info["physical_cores"]:
        assert info["threads"] <= info["physical_cores"]


def test_spectrum_errors_covers_the_bins_inside_the_cut():
    import numpy as np
    Po = np.array([1.0, 0.5, 0.25, 0.125, 1e-30])
    P = Po * np.array([1.0, 1.0 + 1e-6, 1.0 - 2e-6, 1.0, 5.0])
    # N = 16, two shells per bin: bins 0, 1 (shells up to 3) lie inside |n| <= 4
    assert bench.spectrum_errors(P, Po, 16, 2) == pytest.approx(1e-6)
    assert bench.spectrum_errors(P, Po, 32, 2) == pytest.approx(2e-6)
