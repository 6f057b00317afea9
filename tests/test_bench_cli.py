"""bench.py on the CPU: the one-GPU rule for every workload, the committed traffic
summary it prices the roofline with, and its float64 texel directions (the FP64
reference of the convolution study) against the oracle's binary32 ones."""
import os
import subprocess
import sys

import numpy as np
import pytest

import bench
from oracle import oracle as O

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


def test_texel_directions_float64_match_oracle():
    np.testing.assert_allclose(bench.texel_dirs64(), O.texel_dirs().astype(np.float64), atol=1e-7)
