"""GPU parity of the quantized 32-byte traversal nodes (ANE_NODES_QUANTIZED)
against the oracle traversing the same decoded boxes (OracleScene.set_node32):
hit ids, t/u/v, shadow decisions, light-shadow distances and moments, and the
traversal counters bit-exact; after refits and rebuilds too."""
import numpy as np
import pytest

from anerend_amd import native, scenegen
from anerend_amd.types import PRECISION_FP32, Config
from oracle import oracle as O
from tests.conftest import bits, rel_err
from tests.test_gpu_edge_cases import _soup_scene, _tris

pytestmark = pytest.mark.gpu

Q = native.DeviceScene.NODES_QUANTIZED
RTOL = 1e-4


def _pair(scene):
    ds, osc = native.DeviceScene(scene), O.OracleScene(scene)
    ds.set_node_format(Q)
    osc.set_node32(True)
    assert ds.info()["node_format"] == Q
    return ds, osc


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    return o, (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def _check_rays(ds, osc, o, d, tag):
    for any_hit in (False, True):
        hg, cg = ds.trace_rays(o, d, any_hit=any_hit, count=True)
        ho, co = osc.trace(o, d, any_hit=any_hit)
        assert np.array_equal(hg["tri"], ho["tri"]), tag
        for f in ("t", "u", "v"):
            assert np.array_equal(bits(hg[f]), bits(ho[f])), (tag, f)
        assert np.array_equal(cg, co), (tag, cg, co)


@pytest.mark.parametrize("family", ["room", "terrain", "degenerate", "two", "three"])
def test_caller_rays(family):
    if family == "room":
        scene = scenegen.room_scene(2)
    elif family == "terrain":
        scene = scenegen.terrain_scene(9, cells=48, num_instances=300, spread=18.0)
    else:
        scene = _soup_scene(_tris(family))
    print(f"quantized nodes, caller rays: {family}", flush=True)
    ds, osc = _pair(scene)
    o, d = _rays(30000, 5, [-6, -1, -6], [6, 6, 6])
    _check_rays(ds, osc, o, d, family)
    # back to the 64-byte nodes: the exact boxes' counters again
    ds.set_node_format(native.DeviceScene.NODES_FULL)
    osc.set_node32(False)
    assert ds.info()["node_format"] == native.DeviceScene.NODES_FULL
    _check_rays(ds, osc, o, d, family + " full")


def test_probe_updates_with_counters():
    scene, cfg, _ = scenegen.workload("room8")
    ds, osc = _pair(scene)
    grid = native.ProbeGrid(cfg)
    grid.set_counting(True)
    atlas = O.new_atlas(cfg)
    for it in range(3):
        fr = scenegen.make_frame(1, camera=(0.5, 1.5, 0.5), config=cfg, iteration=it)
        print(f"quantized nodes, room update {it}", flush=True)
        ho, sho, rado, ctro = osc.update(cfg, fr, atlas)
        grid.update(ds, fr)
        hg, shg = grid.debug_rays(fr)
        assert np.array_equal(hg["tri"], ho["tri"])
        for f in ("t", "u", "v"):
            assert np.array_equal(bits(hg[f]), bits(ho[f])), f
        assert np.array_equal(shg, sho)
        st = grid.stats()
        assert [st["primary_nodes"], st["primary_tris"], st["shadow_nodes"], st["shadow_tris"]] == \
            [int(x) for x in ctro]
        assert rel_err(grid.atlas, atlas) <= RTOL


def test_dynamic_refits_and_rebuild():
    scene, vel = scenegen.dynamic_scene(seed=5, cells=48, num_instances=300, spread=20.0)
    cfg = Config.make((16, 4, 16), step=(1.0, 2.0, 1.0), rays=128, precision=PRECISION_FP32)
    ds, osc = _pair(scene)
    grid = native.ProbeGrid(cfg)
    grid.set_counting(True)
    atlas = O.new_atlas(cfg)
    for f in range(5):
        inst = scenegen.dynamic_instances(scene, vel, 3 * f)
        print(f"quantized nodes, dynamic frame {f}", flush=True)
        if f == 3:
            ds.rebuild(inst)
            osc = O.OracleScene(scenegen.Scene(scene.vertices, scene.indices, scene.meshes, scene.materials, inst))
            osc.set_node32(True)
        else:
            ds.refit(inst)
            osc.refit(inst)
        fr = scenegen.make_frame(5, camera=(-3.5 + f, 6.0, 0.5), config=cfg, iteration=f)
        ho, sho, _, ctro = osc.update(cfg, fr, atlas)
        grid.update(ds, fr)
        hg, shg = grid.debug_rays(fr)
        assert np.array_equal(hg["tri"], ho["tri"]), f"hit ids differ at frame {f}"
        assert np.array_equal(shg, sho)
        st = grid.stats()
        assert [st["primary_nodes"], st["primary_tris"], st["shadow_nodes"], st["shadow_tris"]] == \
            [int(x) for x in ctro]


def test_shadow_mask_and_light_shadows():
    from tests.test_oracle_shading import _view_proj
    scene, _, _ = scenegen.workload("room8")
    ds, osc = _pair(scene)
    _, ivp = _view_proj((0.3, 1.8, 4.2), (0.0, 0.5, 0.0))
    ivp16 = ivp.T.ravel().astype(np.float32)
    depth = np.random.default_rng(2).uniform(0.95, 0.9995, (30, 40)).astype(np.float32)
    light = np.array([0.3, -0.85, 0.4], np.float32)
    vg, cg = ds.shadow_mask(depth, ivp16, light, count=True)
    vo, co = osc.shadow_mask(depth, ivp16, light)
    assert np.array_equal(vg, vo) and np.array_equal(cg, co)
    rng = np.random.default_rng(6)
    L = np.zeros((24, 8), np.float32)
    L[:, :3] = rng.uniform([-4.5, 0.0, -4.5], [4.5, 3.0, 4.5], (24, 3))
    L[:, 3] = rng.uniform(2.0, 12.0, 24)
    L[:, 4:8] = 1.0
    dg, mg, cg = ds.light_shadows(L, count=True)
    do, mo, co = osc.light_shadows(L)
    assert np.array_equal(bits(dg), bits(do)) and np.array_equal(bits(mg), bits(mo))
    assert np.array_equal(cg, co)


def test_default_layout_by_size():
    scene, _, _ = scenegen.workload("room8")
    ds = native.DeviceScene(scene)
    assert ds.info()["node_format"] == native.DeviceScene.NODES_FULL   # far below the threshold
    ds.set_node_format(native.DeviceScene.NODES_AUTO)
    assert ds.info()["node_format"] == native.DeviceScene.NODES_FULL
    with pytest.raises(native.AneError):
        ds.set_node_format(7)
