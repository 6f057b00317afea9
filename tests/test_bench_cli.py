"""bench.py on the CPU: the one-GPU rule for every workload, the committed traffic
summary it prices the roofline with, and the CPU model string of its baseline."""
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("workload", ["terrain1m", "dynamic4m", "serve64"])
def test_more_than_one_gpu_exits_with_a_message(workload):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--workload", workload],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""   # no JSON line


def test_traffic_summary_is_parsed():
    t = bench.profiled_traffic()
    assert isinstance(t, int) and t > 0


def test_cpu_model_is_named():
    assert isinstance(bench.cpu_model(), str) and bench.cpu_model()
