"""Request loop (include/ane_serve.h) on the GPU: every session's atlas after a
mixed, overlapping burst of requests equals, bit for bit, the atlas of the same
session updated alone through the plain per-scene path; admission by HBM
budget; argument errors reported without touching the device."""
import numpy as np
import pytest

from anerend_amd import native, scenegen, serving
from anerend_amd.types import PRECISION_BF16, Config
from tests.conftest import bits

pytestmark = pytest.mark.gpu


def _sessions():
    """Four small scenes of uneven sizes with uneven grids and ray counts."""
    specs = [
        (scenegen.room_scene(1), Config.make((8, 8, 8), step=(1, 1, 1), rays=256, hysteresis=0.97),
         (0.5, 1.5, 0.5)),
        (scenegen.room_scene(3, num_boxes=20), Config.make((4, 4, 4), step=(1, 1, 1), rays=128),
         (0.2, 1.0, -0.3)),
        (scenegen.terrain_scene(7, cells=64, num_instances=400),
         Config.make((16, 8, 16), step=(1, 2, 1), rays=128, precision=PRECISION_BF16), (0.5, 10.0, 0.5)),
        (scenegen.terrain_scene(8, cells=40, num_instances=100),
         Config.make((8, 4, 8), step=(1, 2, 1), rays=512, precision=PRECISION_BF16), (1.5, 9.0, -2.5)),
    ]
    return specs


@pytest.mark.parametrize("graphs,nodes,fused", [(True, "full", False), (False, "full", False),
                                               (True, "quantized", False), (False, "full", True),
                                               (False, "quantized", True)])
def test_mixed_equals_alone(graphs, nodes, fused):
    """With captured graphs every session's graph is replayed (updated in place)
    once per request, with the eager per-scene updates of the other sessions and
    host synchronizations in between; each replay feeds the next update through
    the atlas, so the final atlas checks them all.  The quantized case runs every
    scene on 32-byte nodes.  Fused: the requests launched together share one
    shading and one convolution launch per precision (FP32, BF16 here)."""
    specs = _sessions()
    scenes = [native.DeviceScene(s) for s, _, _ in specs]
    if nodes == "quantized":
        for ds in scenes:
            ds.set_node_format(native.DeviceScene.NODES_QUANTIZED)
    srv = serving.Server(hbm_budget_bytes=64 << 30, max_inflight=4, num_streams=3, capture_graphs=graphs,
                         fuse_stages=fused)
    ids = [srv.add_session(ds, cfg) for ds, (_, cfg, _) in zip(scenes, specs)]
    # interleaved requests: 3 full updates per session, submitted back to back
    frames = {}
    tickets = []
    for it in range(3):
        for k, sid in enumerate(ids):
            cfg, cam = specs[k][1], specs[k][2]
            fr = scenegen.make_frame(11 + k, camera=cam, config=cfg, iteration=it)
            frames.setdefault(k, []).append(fr)
            tickets.append(srv.submit(sid, fr))
    print(f"serve (graphs {graphs}, {nodes} nodes, fused {fused}): {len(tickets)} requests submitted",
          flush=True)
    res = [srv.poll(t, timeout_ms=-1) for t in tickets]
    assert all(r["status"] == 0 for r in res)
    st = srv.stats()
    print(f"serve stats {st}", flush=True)
    assert st["completed"] == len(tickets) and st["failed"] == 0
    mixed = [srv.atlas(sid) for sid in ids]
    srv.close()
    for k, (ds, (_, cfg, _)) in enumerate(zip(scenes, specs)):
        grid = native.ProbeGrid(cfg)
        for fr in frames[k]:
            grid.update(ds, fr)
        alone = grid.atlas
        assert np.array_equal(bits(alone), bits(mixed[k])), f"session {k} differs alone vs mixed"
        grid.close()


def test_admission_by_budget_and_errors():
    scene, cfg, frame = scenegen.workload("room8")
    ds = native.DeviceScene(scene)
    scene_bytes = ds.info()["device_bytes"]
    probe = native.ProbeGrid(cfg)
    grid_bytes = probe.stats()["device_bytes"]
    probe.close()
    # room for the scene and exactly one grid
    srv = serving.Server(hbm_budget_bytes=scene_bytes + grid_bytes + grid_bytes // 2,
                         max_inflight=2, num_streams=2)
    a = srv.add_session(ds, cfg)
    with pytest.raises(native.AneError) as e:
        srv.add_session(ds, cfg)
    assert e.value.status == 3  # ANE_ERR_ADMISSION_REFUSED
    assert srv.stats()["sessions"] == 1
    bad = frame.copy()
    bad.layer_begin, bad.layer_count = 6, 4
    with pytest.raises(native.AneError):
        srv.submit(a, bad)
    with pytest.raises(native.AneError):
        srv.submit(a + 100, frame)
    t = srv.submit(a, frame)
    r = srv.poll(t, timeout_ms=-1)
    assert r["status"] == 0 and r["rays"] == 512 * 256
    srv.remove_session(a)
    assert srv.stats()["used_bytes"] == 0
    b = srv.add_session(ds, cfg)  # admitted again once the memory is free
    assert srv.poll(srv.submit(b, frame), timeout_ms=-1)["status"] == 0
    srv.close()


@pytest.mark.parametrize("fused", [True, False])
def test_serving_c3_scenes_mixed_equals_alone(fused):
    """The serving workload itself: the 64 seeded scenes (10k-2M triangles, grids
    4^3 .. 32x16x32, 128-512 rays, BF16 convolution) resident under a 200 GB
    budget, 192 requests drawn uniformly over the sessions and submitted in
    real time as a Poisson process at 200/s.  Each session's atlas equals, bit for
    bit, the same frames applied alone through the plain per-scene path, and the
    bytes charged at admission equal the bytes the sessions hold afterwards."""
    import time
    specs = scenegen.serving_specs()
    dss = []
    for k, sp in enumerate(specs):
        dss.append(native.DeviceScene(scenegen.serving_scene(4, k, sp)))
    print(f"[c3] {len(dss)} scenes, {sum(d.info()['num_triangles'] for d in dss)} triangles", flush=True)
    cfgs = [scenegen.serving_config(sp, PRECISION_BF16) for sp in specs]
    srv = serving.Server(hbm_budget_bytes=200 << 30, max_inflight=8, num_streams=4, capture_graphs=not fused,
                         fuse_stages=fused)
    sids = [srv.add_session(d, c) for d, c in zip(dss, cfgs)]
    admitted = srv.stats()
    assert admitted["used_bytes"] == admitted["held_bytes"]
    rng = np.random.default_rng(44)
    R = 192
    arrive = np.cumsum(rng.exponential(1.0 / 200.0, R))
    which = rng.integers(0, len(specs), R)
    frames = {}
    tickets = []
    t0 = time.perf_counter()
    for i in range(R):
        k = int(which[i])
        fr = scenegen.serving_frame(4, k, len(frames.get(k, [])), cfgs[k])
        frames.setdefault(k, []).append(fr)
        while time.perf_counter() - t0 < arrive[i]:
            time.sleep(0.0005)
        tickets.append(srv.submit(sids[k], fr))
    res = [srv.poll(t, timeout_ms=-1) for t in tickets]
    lat = np.array([r["latency_ms"] for r in res])
    st = srv.stats()
    print(f"[c3] fused {fused}: {R} requests p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms, "
          f"peak in flight {st['peak_inflight']}", flush=True)
    assert all(r["status"] == 0 for r in res) and st["completed"] == R + 0 and st["failed"] == 0
    assert st["used_bytes"] == st["held_bytes"] == admitted["used_bytes"]
    mixed = {k: srv.atlas(sids[k]) for k in frames}
    srv.close()
    for k in sorted(frames):
        grid = native.ProbeGrid(cfgs[k])
        for fr in frames[k]:
            grid.update(dss[k], fr)
        assert np.array_equal(bits(grid.atlas), bits(mixed[k])), f"session {k} differs alone vs mixed"
        grid.close()
    assert len(frames) > 40
