"""Frame-graph / pass state rules of the host mirror (CPU, oracle-backed double)."""
import numpy as np
import pytest

from anerend_amd import passes, scenegen
from anerend_amd.types import Config
from tests.fakes import OracleGrid


def make_rc(ny=4, **kw):
    scene = scenegen.room_scene(1, num_boxes=8)
    cfg = Config.make((8, ny, 8), step=(1, 1, 1), rays=64, hysteresis=0.97)
    grid = OracleGrid(cfg, scene)
    rc = passes.RenderContext(grid=grid, scene=None, light_dir=(-0.3, -0.8, 0.2), **kw)
    return rc, grid


def test_submission_order_and_layer_round_robin():
    rc, grid = make_rc(ny=4)
    fgb, _ = passes.ddgi_frame_graph(rc, seed=5)
    assert fgb.order == ["IrradianceProbeTrans", "IrradianceProbeRT", "IrradianceProbeConvolve"]
    layers = []
    for _ in range(10):
        fgb.execute_graph(rc)
        layers.append(rc.blackboard_value_int("ProbeLayer"))
    assert layers == [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]
    kinds = [c[0] for c in grid.calls]
    assert kinds == ["translate", "trace", "convolve"] * 10
    traced = [c[1:] for c in grid.calls if c[0] == "trace"]
    assert traced[:4] == [(0, 1), (1, 1), (2, 1), (3, 1)]


def test_rotations_are_seeded_and_vary():
    p1, p2 = passes.IrradianceProbeRayTracingPass(3), passes.IrradianceProbeRayTracingPass(3)
    r1 = [p1.random_rotation() for _ in range(4)]
    r2 = [p2.random_rotation() for _ in range(4)]
    assert all(np.array_equal(a, b) for a, b in zip(r1, r2))
    assert not np.array_equal(r1[0], r1[1])
    for r in r1:
        m = np.asarray(r, np.float64).reshape(4, 4)[:3, :3]
        assert np.allclose(m @ m.T, np.eye(3), atol=1e-6)


def test_option_gates():
    rc, grid = make_rc()
    fgb, _ = passes.ddgi_frame_graph(rc)
    rc.options.ddgi_enabled = False
    fgb.execute_graph(rc)
    assert grid.calls == []
    rc.options.ddgi_enabled = True
    rc.options.raytracing_enabled = False     # translation only checks ddgiEnabled
    fgb.execute_graph(rc)
    assert [c[0] for c in grid.calls] == ["translate"]
    rc.options.raytracing_enabled = True
    rc.baking = True                           # no translation while baking
    grid.calls.clear()
    fgb.execute_graph(rc)
    assert [c[0] for c in grid.calls] == ["trace", "convolve"]


def test_translation_follows_camera_cells():
    rc, grid = make_rc()
    fgb, _ = passes.ddgi_frame_graph(rc)
    moves = []
    for x in [0.2, 0.7, 1.1, 1.9, 2.0, -0.5, -1.2]:
        rc.camera.position = (x, 1.0, 0.3)
        fgb.execute_graph(rc)
        moves.append([c[1] for c in grid.calls if c[0] == "translate"][-1])
    # cells: 0, 0, 1, 1, 2, 0 (trunc toward zero), -1
    assert moves == [False, False, True, False, True, True, True]
    assert tuple(grid.last) == (-1, 0)


def test_full_grid_mode_and_registration_errors():
    rc, grid = make_rc(ny=3)
    fgb, _ = passes.ddgi_frame_graph(rc, full_grid=True)
    fgb.execute_graph(rc)
    assert ("trace", 0, 3) in grid.calls
    with pytest.raises(KeyError):
        fgb.register_render_pass_exe("NoSuchPass", lambda rc: None)
    with pytest.raises(ValueError):
        fgb.register_render_pass("IrradianceProbeRT")
