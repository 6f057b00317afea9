"""Host-side checks with no GPU: the PCG32 generator against its published
known answers, scene determinism and sizes, the C-ABI library (exports, record
layouts, argument validation) and host-side parity of the ray-direction table."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from anerend_amd import native, scenegen
from anerend_amd.types import (GPU_MESH_INFO_DTYPE, INSTANCE_DTYPE, MATERIAL_DTYPE, MESH_DTYPE, Config,
                               Frame, PRECISION_BF16, ptr)
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ane_ddgi.h")
HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                 if f.endswith(".h"))


def test_pcg32_known_answers():
    # pcg32-demo (O'Neill, pcg-c-basic): pcg32_srandom_r(&rng, 42u, 54u)
    g = scenegen.PCG32(42, [54])
    got = [int(g.next_u32()[0]) for _ in range(6)]
    assert got == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_scenes_are_deterministic_and_sized():
    a, b = scenegen.room_scene(1), scenegen.room_scene(1)
    for f in ("vertices", "indices", "meshes", "materials", "instances"):
        assert getattr(a, f).tobytes() == getattr(b, f).tobytes()
    assert a.num_tris == 12 + 64 * 12
    assert scenegen.room_scene(2).instances.tobytes() != a.instances.tobytes()
    t = scenegen.terrain_scene(2, cells=32, num_instances=100)
    assert t.num_tris == 2 * 32 * 32 + 50 * 12 + 50 * 80
    mats = a.materials
    assert np.all((mats["base_col_fac"][:, :3] >= 0.1) & (mats["base_col_fac"][:, :3] <= 0.9))
    assert np.all(np.isin(mats["met_rough"][:, 0], [0.0, 1.0]))


def test_library_exports_every_declared_symbol():
    L = native.load()
    text = "".join(open(h).read() for h in HEADERS)
    declared = sorted(set(re.findall(r"^\w[\w\s\*]*?\b(ane_\w+)\(", text, re.M)))
    assert len(declared) >= 33
    assert "ane_serve_submit" in declared
    missing = [d for d in declared if not hasattr(L, d)]
    assert not missing, missing
    assert set(native.EXPORTED) <= set(declared)
    assert L.ane_abi_version() == 1


def test_record_layouts_match_header(tmp_path):
    """Compile a probe against include/ane_ddgi.h with gcc and compare sizes/offsets."""
    src = tmp_path / "probe.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "ane_ddgi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(ane_ddgi_config), sizeof(ane_ddgi_frame),
         sizeof(ane_mesh), sizeof(ane_material), sizeof(ane_instance), sizeof(ane_hit));
  printf("%zu %zu %zu %zu\n", offsetof(ane_ddgi_frame, rotation), offsetof(ane_ddgi_frame, bake_tile),
         offsetof(ane_instance, custom_index_and_mask), offsetof(ane_ddgi_config, precision));
  printf("%zu %zu %zu\n", sizeof(ane_gpu_mesh_info), offsetof(ane_gpu_mesh_info, vertex_offset),
         offsetof(ane_gpu_mesh_info, blas_ref));
  return 0;
}''')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes[:6] == [ctypes.sizeof(Config), ctypes.sizeof(Frame), MESH_DTYPE.itemsize,
                         MATERIAL_DTYPE.itemsize, INSTANCE_DTYPE.itemsize, 16]
    assert sizes[6:10] == [Frame.rotation.offset, Frame.bake_tile.offset,
                           INSTANCE_DTYPE.fields["custom_index_and_mask"][1], Config.precision.offset]
    # GPUMeshInfo (GpuBuffers.h:38-44): two vec4, two uint32, uint64 _blasRef = 48 B
    assert sizes[10:] == [48, 32, 40] and GPU_MESH_INFO_DTYPE.itemsize == 48
    assert GPU_MESH_INFO_DTYPE.fields["blas_ref"][1] == 40


def test_serve_layouts_and_validation(tmp_path):
    """Serving records match include/ane_serve.h; invalid configs are refused
    before any device work."""
    from anerend_amd import serving
    src = tmp_path / "probe_serve.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "ane_serve.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(ane_serve_config), sizeof(ane_serve_result),
         sizeof(ane_serve_stats), offsetof(ane_serve_result, latency_ms), offsetof(ane_serve_stats, rays_done));
  return 0;
}''')
    exe = tmp_path / "probe_serve"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert out == [ctypes.sizeof(serving.ServeConfig), ctypes.sizeof(serving.ServeResult),
                   ctypes.sizeof(serving.ServeStats), serving.ServeResult.latency_ms.offset,
                   serving.ServeStats.rays_done.offset]
    L = native.load()
    h = ctypes.c_void_p()
    for cfg in (serving.ServeConfig(0, 4, 2), serving.ServeConfig(1 << 30, 0, 2),
                serving.ServeConfig(1 << 30, 4, 0), serving.ServeConfig(1 << 30, 4, 17),
                serving.ServeConfig(1 << 30, 4, 2, 2, 0), serving.ServeConfig(1 << 30, 4, 2, 1, 7)):
        assert L.ane_serve_open(ctypes.byref(cfg), ctypes.byref(h)) == 1
        assert not h.value


@pytest.mark.parametrize("n,angle,axis", [(256, 0.0, 0), (512, 123.4, 1), (100, 359.0, 2)])
def test_host_ray_directions_bit_exact_with_oracle(n, angle, axis):
    rot = scenegen.rotation_matrix(angle, axis)
    lib = native.ray_directions(n, rot)
    ora = O.rotate_dirs(rot, O.fibonacci_dirs(n))
    assert np.array_equal(lib.view(np.uint32), ora.view(np.uint32))


def _create(cfg):
    h = ctypes.c_void_p()
    return native.load().ane_ddgi_create(ctypes.byref(cfg), ctypes.byref(h))


def test_config_validation_without_device():
    L = native.load()
    bad = [Config.make((0, 8, 8)), Config.make((8, 8, 8), rays=0), Config.make((8, 8, 8), rays=5000),
           Config.make((8, 8, 8), step=(0, 1, 1)), Config.make((8, 8, 8), hysteresis=1.5),
           Config.make((8, 8, 8), t_min=1.0, t_max=0.5)]
    for c in bad:
        assert _create(c) == 1, L.ane_last_error()
    assert _create(Config.make((8, 8, 8), rays=100, precision=PRECISION_BF16)) == 6
    assert _create(Config.make((8, 8, 8), quirks=1)) == 6
    assert L.ane_ddgi_create(ctypes.byref(Config.make((8, 8, 8))), None) == 1
    assert b"out is NULL" in L.ane_last_error()


def test_scene_validation_without_device():
    L = native.load()
    s = scenegen.room_scene(1, num_boxes=2)

    def create(scene):
        h = ctypes.c_void_p()
        arrs = [np.ascontiguousarray(a) for a in
                (scene.vertices, scene.indices, scene.meshes, scene.materials, scene.instances)]
        v, i, m, mt, inst = arrs
        return L.ane_scene_create(ptr(v), len(v), ptr(i), len(i), ptr(m), len(m), ptr(mt), len(mt),
                                  ptr(inst), len(inst), None, ctypes.byref(h))

    bad_idx = scenegen.Scene(s.vertices, s.indices.copy(), s.meshes, s.materials, s.instances)
    bad_idx.indices[5] = 10_000
    assert create(bad_idx) == 1 and b"out of the vertex array" in L.ane_last_error()
    tex = scenegen.Scene(s.vertices, s.indices, s.meshes, s.materials.copy(), s.instances)
    tex.materials["bindless_indices"][0, 1] = 3
    assert create(tex) == 6
    bad_mesh = scenegen.Scene(s.vertices, s.indices, s.meshes, s.materials, s.instances.copy())
    bad_mesh.instances["custom_index_and_mask"][1] |= (9 << 10)
    assert create(bad_mesh) == 1 and b"references mesh" in L.ane_last_error()
    bad_x = scenegen.Scene(s.vertices, s.indices, s.meshes, s.materials, s.instances.copy())
    bad_x.instances["transform"][0, 0] = np.nan
    assert create(bad_x) == 1


def gpu_mesh_info(scene, seed=7):
    """The scene's meshes as the reference fills GPUMeshInfo (VulkanRenderer.cpp:3486-3493):
    bounds, offsets and a BLAS device address (random, non-zero in both 32-bit halves),
    plus the index counts the reference keeps on the host (InternalMesh::_numIndices)."""
    gm = np.zeros(len(scene.meshes), GPU_MESH_INFO_DTYPE)
    for f in ("min_pos", "max_pos", "vertex_offset", "index_offset"):
        gm[f] = scene.meshes[f]
    rng = np.random.default_rng(seed)
    gm["blas_ref"] = rng.integers(1 << 32, 1 << 63, len(gm), dtype=np.uint64) | np.uint64(0x5)
    return gm, scene.meshes["index_count"].astype(np.uint32).copy()


def test_gpu_mesh_info_validation_without_device():
    """ane_scene_create_mesh_info takes GPUMeshInfo as it is; a GPUMeshInfo array
    passed to ane_scene_create is refused (its _blasRef lands in index_count and
    the reserved word, which must be 0), and every check runs before device work."""
    L = native.load()
    s = scenegen.room_scene(1, num_boxes=2)
    gm, cnt = gpu_mesh_info(s)
    v, i, mt, inst = (np.ascontiguousarray(a) for a in (s.vertices, s.indices, s.materials, s.instances))
    h = ctypes.c_void_p()

    def create_info(gm_, cnt_):
        return L.ane_scene_create_mesh_info(ptr(v), len(v), ptr(i), len(i), ptr(gm_), ptr(cnt_), len(gm_),
                                            ptr(mt), len(mt), ptr(inst), len(inst), None, ctypes.byref(h))

    as_mesh = np.ascontiguousarray(gm).view(MESH_DTYPE)   # the cast the old binding made
    assert as_mesh["reserved"].min() > 0
    rc = L.ane_scene_create(ptr(v), len(v), ptr(i), len(i), ptr(as_mesh), len(as_mesh), ptr(mt), len(mt),
                            ptr(inst), len(inst), None, ctypes.byref(h))
    assert rc == 1 and not h.value
    assert L.ane_scene_create_mesh_info(ptr(v), len(v), ptr(i), len(i), ptr(gm), None, len(gm),
                                        ptr(mt), len(mt), ptr(inst), len(inst), None, ctypes.byref(h)) == 1
    bad = cnt.copy()
    bad[0] += 1
    assert create_info(gm, bad) == 1 and b"not a multiple of 3" in L.ane_last_error()
    bad = cnt.copy()
    bad[-1] += 3 * 10_000
    assert create_info(gm, bad) == 1 and b"exceed the index array" in L.ane_last_error()
    assert not h.value


def test_status_strings():
    L = native.load()
    assert L.ane_status_string(0) == b"ok"
    assert L.ane_status_string(3) == b"admission refused"
    err = native.AneError(1, "x")
    assert "invalid argument" in str(err)
