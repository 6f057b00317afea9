"""The oracle against the committed small-room fixtures (regression pins)."""
import json
import os

import numpy as np

from tests.golden.make_golden import run_room8, sha  # noqa: F401

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_oracle_matches_room8_fixture():
    meta = json.load(open(os.path.join(HERE, "room8.json")))
    steps, atlas = run_room8()
    for got, want in zip(steps, meta["updates"]):
        assert got["hits_sha256"] == want["hits_sha256"]
        assert got["shadow_sha256"] == want["shadow_sha256"]
        assert got["counters"] == want["counters"]
        assert np.isclose(got["radiance_sum"], want["radiance_sum"], rtol=1e-6)
    ref = np.load(os.path.join(HERE, "room8_atlas.npz"))["atlas"]
    assert np.allclose(atlas[::3, ::3], ref, rtol=1e-6, atol=1e-7)


def test_oracle_matches_room8_lights_fixture():
    from tests.golden.make_golden import run_room8_lights
    want = json.load(open(os.path.join(HERE, "room8_lights.json")))
    got = run_room8_lights()
    for k in ("dist_sha256", "moments_sha256", "counters"):
        assert got[k] == want[k], k
    assert np.isclose(got["mean_sum"], want["mean_sum"], rtol=1e-9)
