"""C++ frame-graph host (anerend_amd/host): the three passes registered on a
FrameGraphBuilder and run frame by frame over the C-ABI.

CPU: the scene-file writer round trip and the driver's argument and input errors
(no device call is reached).  GPU: a moving-camera run of the room workload
(layer round robin, translation on cell crossings) replayed frame by frame on
the oracle from the frames the driver logged; atlas within 1e-4 relative."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from anerend_amd import build as hip_build
from anerend_amd import scenegen, types
from tests.conftest import rel_err

BIN = hip_build.HOST_BIN


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(BIN):
        hip_build.build_host(quiet=True)
    return BIN


def test_scene_file_layout(tmp_path):
    scene, _, _ = scenegen.workload("room8")
    p = tmp_path / "room.bin"
    scenegen.save_scene(scene, p)
    raw = p.read_bytes()
    assert raw[:8] == b"ANESCN01"
    nv, ni = np.frombuffer(raw, "<i8", 2, 8)
    nm, nmat, ninst, zero = np.frombuffer(raw, "<i4", 4, 24)
    assert (nv, ni, nm, nmat, ninst, zero) == (len(scene.vertices), len(scene.indices),
                                              len(scene.meshes), len(scene.materials),
                                              len(scene.instances), 0)
    off = 40
    v = np.frombuffer(raw, np.float32, nv * types.VERTEX_WORDS, off)
    assert np.array_equal(v.reshape(scene.vertices.shape), scene.vertices)
    off += v.nbytes
    idx = np.frombuffer(raw, np.uint32, ni, off)
    assert np.array_equal(idx, scene.indices)
    off += idx.nbytes
    off += scene.meshes.nbytes + scene.materials.nbytes
    inst = np.frombuffer(raw, types.INSTANCE_DTYPE, ninst, off)
    assert inst.tobytes() == np.ascontiguousarray(scene.instances).tobytes()
    assert off + inst.nbytes == len(raw)


def test_driver_rejects_bad_input(driver, tmp_path):
    r = subprocess.run([driver], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([driver, str(tmp_path / "missing.bin"), str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTASCENE" + bytes(64))
    r = subprocess.run([driver, str(bad), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1 and "not a scene file" in r.stderr
    r = subprocess.run([driver, str(bad), str(tmp_path), "--precision", "fp16"],
                       capture_output=True, text=True)
    assert r.returncode == 2


@pytest.mark.gpu
def test_driver_frames_match_oracle(driver, tmp_path):
    from tests.fakes import OracleGrid

    scene, cfg, _ = scenegen.workload("room8")
    sp = tmp_path / "room.bin"
    scenegen.save_scene(scene, sp)
    frames = 6
    cmd = [driver, str(sp), str(tmp_path), "--probes", "8", "8", "8", "--step", "1", "1", "1",
           "--rays", "256", "--hysteresis", "0.97", "--frames", str(frames), "--seed", "5",
           "--camera", "0.5", "1.5", "0.5", "--velocity", "0.6", "0", "0.35"]
    print("host driver:", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    logged = np.fromfile(tmp_path / "frames.bin", np.uint8)
    fsz = ctypes.sizeof(types.Frame)
    assert len(logged) == frames * fsz
    og = OracleGrid(cfg, scene)
    moves = 0
    for i in range(frames):
        f = types.Frame.from_buffer_copy(logged[i * fsz:(i + 1) * fsz].tobytes())
        assert (f.layer_begin, f.layer_count) == (i % cfg.probes_y, 1)
        moves += bool(og.translate(list(f.camera_pos)[:3]))
        og.trace(None, f)
        og.convolve(f)
    assert moves >= 2  # the camera path crosses cells in x and z
    atlas = np.fromfile(tmp_path / "atlas.f32", np.float32).reshape(cfg.atlas_shape)
    assert rel_err(atlas, og.atlas) <= 1e-4
