"""Shared test setup: marker registration, oracle/library builds, helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libane_ddgi.so)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    yield


def rel_err(a, b, floor=1e-3):
    """max |a - b| / max(|b|, floor) (floor keeps near-zero values from dominating)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)
