"""Host-side logic of bench.py (no GPU): the multi-GPU refusal and the
traffic summary parser."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_traffic_summary_parser(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("# rocprofv3 --pmc FETCH_SIZE -- python3 bench.py\n"
                 "calibration FETCH_SIZE 2.0\n"
                 "calibration WRITE_SIZE 1.0\n"
                 "kernel dispatches FETCH_SIZE_KiB WRITE_SIZE_KiB\n"
                 "k_lists_t 8 100.0 50.0\n")
    assert bench.traffic_from_summary(str(f), "lists") == pytest.approx(1024 * (200.0 + 50.0))
    assert bench.traffic_from_summary(str(f), "force") is None
    assert bench.traffic_from_summary(str(tmp_path / "missing.txt"), "lists") is None
