"""GPU parity of the baked-tile branch (pbr_light.glsl:251-275, bindless.glsl:291-308,
deferred_tiled.comp:264-298) against the oracle.

The room workload with 4 m tiles (the room spans tiles -2..1 on x and z around
camera tile (0, 0)): six of them carry a baked RGBA16F atlas, the others and
everything beyond the radius-2 window keep the live placement.
* probe updates: probe hits in a baked tile place their probes on the tile's grid
  and sample the live atlas (rchit:99-106); hits, shadow flags bit-exact,
  radiance and atlas within 1e-4 over 2 updates;
* deferred lookup: texels in baked tiles sample their tile's atlas (also without
  ANE_FLAG_DDGI and beyond half the grid); within 1e-4;
* mirror reflections: the reflection hits' indirect term with the tile placement;
* ane_ddgi_tiles_set argument checks and clearing.
"""
import numpy as np
import pytest

from anerend_amd import native, scenegen
from anerend_amd.types import FLAG_DDGI, FLAG_DDGI_MULTI, GBUFFER_DTYPE
from oracle import oracle as O
from tests.conftest import bits, rel_err

pytestmark = pytest.mark.gpu

RTOL = 1e-4
TS = 4
XZ = [(0, 0), (-1, 0), (-2, -1), (1, 1), (0, -2), (-1, 1), (3, 0)]   # (3, 0) lies outside the window


def _baked(cfg, seed=17):
    rng = np.random.default_rng(seed)
    a = np.empty((len(XZ),) + tuple(cfg.atlas_shape), np.float32)
    a[..., :3] = rng.uniform(0.02, 2.0, a[..., :3].shape)
    a[..., 3] = 1.0
    return a.astype(np.float16)


@pytest.fixture
def room():
    scene, cfg, frame = scenegen.workload("room8")
    baked = _baked(cfg)
    O.tiles_set(cfg, XZ, baked, TS)
    yield scene, cfg, frame, baked
    O.tiles_set(cfg, [], np.zeros((0,) + tuple(cfg.atlas_shape), np.float32), TS)


def test_probe_updates_with_baked_tiles(room):
    scene, cfg, _, baked = room
    ds, osc = native.DeviceScene(scene), O.OracleScene(scene)
    grid = native.ProbeGrid(cfg)
    grid.set_tiles(XZ, baked, TS)
    rng = np.random.default_rng(3)
    atlas = O.new_atlas(cfg)
    atlas[..., :3] = rng.uniform(0.05, 1.0, atlas[..., :3].shape)
    grid.atlas = atlas
    for it in range(2):
        fr = scenegen.make_frame(1, camera=(0.5, 1.5, 0.5), config=cfg, iteration=it,
                                 flags=FLAG_DDGI | FLAG_DDGI_MULTI)
        print(f"[tiles] update {it}", flush=True)
        ho, sho, rado, _ = osc.update(cfg, fr, atlas)
        grid.update(ds, fr)
        hg, shg = grid.debug_rays(fr)
        assert np.array_equal(hg["tri"], ho["tri"]) and np.array_equal(shg, sho)
        assert rel_err(grid.radiance(fr), rado) <= RTOL
        assert rel_err(grid.atlas, atlas) <= RTOL
    # the tiles changed the result: the same update without them differs
    plain = native.ProbeGrid(cfg)
    plain.atlas = grid.atlas
    fr = scenegen.make_frame(1, camera=(0.5, 1.5, 0.5), config=cfg, iteration=5)
    grid.update(ds, fr)
    plain.update(ds, fr)
    assert rel_err(plain.radiance(fr), grid.radiance(fr)) > 1e-2


def test_gi_lookup_with_baked_tiles(room):
    _, cfg, frame, baked = room
    grid = native.ProbeGrid(cfg)
    grid.set_tiles(XZ, baked, TS)
    rng = np.random.default_rng(21)
    atlas = O.new_atlas(cfg)
    atlas[..., :3] = rng.uniform(0, 2.0, atlas[..., :3].shape)
    grid.atlas = atlas
    n = 40000
    g = np.zeros(n, GBUFFER_DTYPE)
    cam = np.array(frame.camera_pos[:3], np.float32)
    g["pos"] = cam + rng.uniform(-14, 14, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    g["normal"] = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    g["albedo"] = rng.uniform(0, 1, (n, 3))
    g["metallic"] = rng.choice([0.0, 1.0], n)
    g["roughness"] = rng.uniform(0.05, 1, n)
    g["shadow"] = rng.uniform(0, 1, n)
    for flags in (FLAG_DDGI, 0):
        fr = frame.copy()
        fr.flags = flags
        got = grid.gi_lookup(fr, g)
        want = O.gi_lookup(cfg, fr, atlas, g.view(np.float32).reshape(n, 12))
        assert rel_err(got, want) <= RTOL, flags
    tile = np.floor(g["pos"][:, [0, 2]] / TS).astype(int)
    inwin = np.array([tuple(t) in set(XZ[:6]) for t in tile])
    assert 0.05 < inwin.mean() < 0.8          # both branches are exercised


def test_mirror_reflections_with_baked_tiles(room):
    from tests.test_oracle_shading import _mirror_case
    scene, cfg, fr, atlas, depth, normal4, ivp16 = _mirror_case()
    baked = _baked(cfg, seed=5)
    O.tiles_set(cfg, XZ, baked, TS)
    grid = native.ProbeGrid(cfg)
    grid.set_tiles(XZ, baked, TS)
    grid.atlas = atlas
    ds = native.DeviceScene(scene)
    osc = O.OracleScene(scene)
    want, _ = osc.mirror_reflections(cfg, fr, atlas, depth, normal4, ivp16)
    got = grid.mirror_reflections(ds, fr, depth, normal4, ivp16)
    assert rel_err(got, want) <= RTOL


def test_tiles_set_validation():
    _, cfg, _ = scenegen.workload("room8")
    grid = native.ProbeGrid(cfg)
    baked = np.zeros((2,) + tuple(cfg.atlas_shape), np.float16)
    with pytest.raises(native.AneError):
        grid.set_tiles([(1, 1), (1, 1)], baked, TS)      # duplicate tile
    with pytest.raises(native.AneError):
        grid.set_tiles([(1, 1), (0, 1)], baked, 0)       # tile size
    before = grid.stats()["device_bytes"]
    grid.set_tiles([(1, 1), (0, 1)], baked, TS)
    assert grid.stats()["device_bytes"] == before + baked.nbytes
    grid.set_tiles([], [], TS)
    assert grid.stats()["device_bytes"] == before
