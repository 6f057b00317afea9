"""ane_scene_trace_rays_noalloc (the allocation-free variant of ane_scene_trace_rays)
through the C-ABI: its hits (id, t, u, v) and work counters equal the allocating
entry point's bit for bit, on two calls in a row with the second one smaller,
for closest and any hit, on scenes whose hierarchy is deep enough that the
traversal stack spills into the scratch array (depth 23 and 28 > 16 LDS entries)."""
import ctypes

import numpy as np
import pytest

from anerend_amd import native, scenegen
from anerend_amd.types import HIT_DTYPE, ptr

pytestmark = pytest.mark.gpu


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    org = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[::97] = 0.0                     # zero directions: no ray (a miss, no traversal)
    return org, d


def _noalloc(ds, org, dirs, any_hit, counters=True, tmin=0.1, tmax=50.0):
    n = len(org)
    bufs = [native.DeviceBuffer(org.nbytes), native.DeviceBuffer(dirs.nbytes),
            native.DeviceBuffer(n * HIT_DTYPE.itemsize)]
    try:
        bufs[0].upload(org)
        bufs[1].upload(dirs)
        ctr = np.zeros(2, np.uint64)
        native.check(native.load().ane_scene_trace_rays_noalloc(
            ds.handle, bufs[0].ptr, bufs[1].ptr, n, ctypes.c_float(tmin), ctypes.c_float(tmax), int(any_hit),
            bufs[2].ptr, ptr(ctr) if counters else None, None))
        hits = bufs[2].download(np.zeros(n, HIT_DTYPE))
    finally:
        for b in bufs:
            b.close()
    return hits, ctr


@pytest.mark.parametrize("cells,instances,depth", [(48, 200, 23), (128, 2000, 28)])
def test_noalloc_trace_equals_trace_rays(cells, instances, depth):
    scene = scenegen.terrain_scene(2, cells=cells, num_instances=instances)
    ds = native.DeviceScene(scene)
    assert ds.info()["bvh_depth"] == depth
    half = cells / 2
    for call, (n, any_hit) in enumerate(((150_000, False), (20_000, False), (9_000, True))):
        print(f"[noalloc] cells {cells} call {call}: {n} rays any_hit={any_hit}", flush=True)
        org, dirs = _rays(n, 10 * cells + call, (-half, 0.0, -half), (half, 14.0, half))
        want, wctr = ds.trace_rays(org, dirs, any_hit=any_hit, count=True)
        got, gctr = _noalloc(ds, org, dirs, any_hit)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), call
        assert np.array_equal(gctr, wctr), call
        assert 0.1 < (got["tri"] >= 0).mean() < 0.95
    # without counters (no counting kernels), the same hits
    got, _ = _noalloc(ds, org, dirs, any_hit, counters=False)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_noalloc_argument_checks():
    scene = scenegen.terrain_scene(2, cells=16, num_instances=10)
    ds = native.DeviceScene(scene)
    L = native.load()
    assert L.ane_scene_trace_rays_noalloc(ds.handle, None, None, 10, ctypes.c_float(0.1), ctypes.c_float(50.0),
                                          0, None, None, None) == 1
    buf = native.DeviceBuffer(64)
    try:
        assert L.ane_scene_trace_rays_noalloc(ds.handle, buf.ptr, buf.ptr, 1, ctypes.c_float(1.0),
                                              ctypes.c_float(0.5), 0, buf.ptr, None, None) == 1
        assert L.ane_scene_trace_rays_noalloc(ds.handle, buf.ptr, buf.ptr, 1 << 31, ctypes.c_float(0.1),
                                              ctypes.c_float(50.0), 0, buf.ptr, None, None) == 1
        assert L.ane_scene_trace_rays_noalloc(ds.handle, buf.ptr, buf.ptr, 0, ctypes.c_float(0.1),
                                              ctypes.c_float(50.0), 0, buf.ptr, None, None) == 0
    finally:
        buf.close()
