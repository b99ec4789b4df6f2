"""bench.py's contract pieces that need no GPU: the refusal of more than one
GPU, the line surviving an error in one of its further measurements, the
HBM traffic figure with its source, and the CPU thread count."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_an_error_in_a_further_measurement_is_recorded():
    line = {"value": 1.0}

    def fails():
        raise RuntimeError("device lost")

    bench.optional_leg(line, "batch", fails)
    bench.optional_leg(line, "serving", lambda: {"p99_ms": 2.0})
    assert line["batch"] == {"error": "RuntimeError: device lost"}
    assert line["serving"] == {"p99_ms": 2.0} and line["value"] == 1.0


def test_traffic_comes_with_its_source():
    traffic, source = bench.traffic_bytes()
    assert traffic is not None and traffic > 0
    assert "k_fine_large2048_pmc.json" in source and "not in this run" in source


def test_cpu_threads_within_the_host_share():
    assert 1 <= bench.cpu_threads() <= 16
    assert bench.cpu_model()
