"""GPU checks of the drop-in boundary itself: a scene built from the reference's
own GPUMeshInfo array (GpuBuffers.h:38-44, _blasRef set) through
ane_scene_create_mesh_info is the scene ane_scene_create builds from the
library's mesh records: same world triangles, same LBVH, same probe update."""
import numpy as np
import pytest

from anerend_amd import native, scenegen
from tests.conftest import bits
from tests.test_scenegen_abi import gpu_mesh_info

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["room", "terrain"])
def test_scene_from_gpu_mesh_info_equals_scene_from_mesh_records(name):
    scene = (scenegen.room_scene(1) if name == "room"
             else scenegen.terrain_scene(2, cells=48, num_instances=200))
    print(f"[boundary] {name}: {scene.num_tris} triangles", flush=True)
    a = native.DeviceScene(scene)
    b = native.DeviceScene(scene, mesh_info=gpu_mesh_info(scene))
    assert np.array_equal(bits(a.world_tris()), bits(b.world_tris()))
    ba, bb = a.bvh(), b.bvh()
    for k in ba:
        assert np.array_equal(ba[k].view(np.uint8), bb[k].view(np.uint8)), k
    cfg = scenegen.small_room_config()
    fr = scenegen.make_frame(1, config=cfg)
    ga, gb = native.ProbeGrid(cfg), native.ProbeGrid(cfg)
    for it in range(2):
        fr = scenegen.make_frame(1, config=cfg, iteration=it)
        ga.update(a, fr)
        gb.update(b, fr)
    ha, sa = ga.debug_rays(fr)
    hb, sb = gb.debug_rays(fr)
    assert np.array_equal(ha.view(np.uint8), hb.view(np.uint8)) and np.array_equal(sa, sb)
    assert np.array_equal(bits(ga.atlas), bits(gb.atlas))


def test_traversal_fault_is_reported_and_cleared():
    """A probe ray that ends without a result raises the context's traversal fault.
    The condition is forced from the host with a step limit of 3 node visits (the
    kernel's own bounded loop ends every such ray as a miss): stats_get then fails
    with the stats filled in, ane_ddgi_check fails once and clears the flag, and
    without the limit the same update is fault-free and equals the oracle's hits."""
    from oracle import oracle as O
    scene, cfg, frame = scenegen.workload("room8")
    ds = native.DeviceScene(scene)
    grid = native.ProbeGrid(cfg)
    grid.set_traversal_limit(3)
    print("[boundary] update with a 3-visit step limit", flush=True)
    grid.update(ds, frame)
    with pytest.raises(native.AneError) as e:
        grid.stats()
    assert e.value.status == 4 and e.value.stats["traversal_fault"] == 1
    hits, _ = grid.debug_rays(frame)
    assert (hits["tri"] < 0).mean() > 0.9          # cut rays read as misses
    with pytest.raises(native.AneError):
        grid.check()
    grid.check()                                   # cleared by the failed check
    assert grid.stats()["traversal_fault"] == 0
    grid.set_traversal_limit(0)
    grid.atlas = O.new_atlas(cfg)
    grid.update(ds, frame)
    grid.check()
    hits, _ = grid.debug_rays(frame)
    ho, _, _, _ = O.OracleScene(scene).update(cfg, frame, O.new_atlas(cfg))
    assert np.array_equal(hits["tri"], ho["tri"])
    # the scene-level passes' flags: a normal shadow mask leaves none
    depth = np.full((8, 8), 0.99, np.float32)
    ds.shadow_mask(depth, np.eye(4, dtype=np.float32).ravel(), np.float32([0.3, -1, 0.2]))
    ds.check()
