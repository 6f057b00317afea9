"""GPU parity of the point-light shadow moments (LightShadowRayTracingPass +
LightShadowSumPass: light_shadow.rgen/.rchit/.rmiss, light_shadow_sum.comp)
against the oracle: ray distances and traversal counters bit-exact; the moments
bit-exact too, since every texel sums its rays in the shader's order with the
same binary32 operations and weights."""
import numpy as np
import pytest

from anerend_amd import native, scenegen
from oracle import oracle as O
from tests.conftest import bits

pytestmark = pytest.mark.gpu


def _lights(rng, n, lo, hi, rmin, rmax):
    pos = rng.uniform(lo, hi, (n, 3))
    rad = rng.uniform(rmin, rmax, n)
    col = rng.uniform(0.2, 1.0, (n, 3))
    L = np.concatenate([pos, rad[:, None], col, np.ones((n, 1))], 1).astype(np.float32)
    return L


def _check(ds, osc, L, tag):
    print(f"light shadows {tag}: {len(L)} lights", flush=True)
    dg, mg, cg = ds.light_shadows(L, count=True, moments_fill=-7.0)
    do, mo, co = osc.light_shadows(L, moments_fill=-7.0)
    assert np.array_equal(bits(dg), bits(do)), f"distances differ ({tag})"
    diff = np.abs(mg.astype(np.float64) - mo) / np.maximum(np.abs(mo), 1e-6)
    print(f"  max rel moment difference {diff.max():.3e}", flush=True)
    assert np.array_equal(bits(mg), bits(mo)), f"moments differ ({tag})"
    assert np.array_equal(cg, co), f"counters differ ({tag})"
    # nothing is written outside the tiles
    n = len(L)
    for l in range(n):
        assert np.all(mg[:, 10 * l, :] == -7.0) and np.all(mg[:, 10 * l + 9, :] == -7.0)
    return dg, mg


def test_room_lights_match_oracle():
    scene, _, _ = scenegen.workload("room8")
    ds, osc = native.DeviceScene(scene), O.OracleScene(scene)
    rng = np.random.default_rng(11)
    L = _lights(rng, 16, [-4.8, -0.4, -4.8], [4.8, 3.4, 4.8], 1.0, 12.0)
    L[3, 7] = 0.0                   # disabled
    L[7, 7] = 0.2                   # disabled (color.w < 0.5)
    L[7, :3] = np.nan               # a disabled light's contents are not read
    L[9, 3] = 0.05                  # range below t_min: every ray misses, dist = range
    L[12, :3] = (5.0, 1.0, 0.0)     # on the room's +x wall
    L[14, 7] = np.nan               # the shaders test color.w < 0.5: NaN is enabled
    dg, mg = _check(ds, osc, L, "room")
    assert np.all(dg[:, 32 * 3:32 * 4] == 0.0) and np.all(dg[:, 32 * 9:32 * 10] == np.float32(0.05))
    assert np.all(dg[:, 32 * 14:32 * 15] > 0.0) and np.all(mg[:, 141:149, 0] > 0.0)
    # the same scene again with fewer lights (workspace reuse)
    _check(ds, osc, L[:5], "room, 5 lights")


def test_max_lights_on_terrain():
    scene = scenegen.terrain_scene(9, cells=48, num_instances=200, spread=18.0)
    ds, osc = native.DeviceScene(scene), O.OracleScene(scene)
    rng = np.random.default_rng(12)
    L = _lights(rng, 1024, [-20.0, -2.0, -20.0], [20.0, 10.0, 20.0], 4.0, 30.0)
    L[rng.uniform(size=1024) < 0.1, 7] = 0.0
    _check(ds, osc, L, "terrain, 1024 lights")


def test_invalid_arguments():
    scene, _, _ = scenegen.workload("room8")
    ds = native.DeviceScene(scene)
    rng = np.random.default_rng(13)
    L = _lights(rng, 2, [-1, 0, -1], [1, 2, 1], 2.0, 5.0)
    bad = L.copy()
    bad[1, 3] = np.inf               # enabled light with a non-finite range
    with pytest.raises(native.AneError):
        ds.light_shadows(bad)
    with pytest.raises(native.AneError):
        ds.light_shadows(np.zeros((1025, 8), np.float32))


def test_room8_lights_match_golden_fixture():
    import json
    import os
    from tests.golden.make_golden import room8_lights, sha
    want = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "room8_lights.json")))
    scene, _, _ = scenegen.workload("room8")
    dist, mom, ctr = native.DeviceScene(scene).light_shadows(room8_lights(), count=True)
    assert sha(dist) == want["dist_sha256"]
    assert sha(mom) == want["moments_sha256"]
    assert [int(c) for c in ctr] == want["counters"]
