"""Record layouts shared by the C-ABI (include/ane_ddgi.h) and the Python side.

Each numpy dtype / ctypes structure has the byte layout of the reference's
GPU-side struct it stands for:

* ``VERTEX_WORDS`` packed vertex, 6 x vec4 = 96 B        (bindless.glsl:60-68, Vertex.h:7-19)
* ``MESH_DTYPE``  this library's mesh record: GPUMeshInfo's fields with the index
                  count where GPUMeshInfo keeps blasRef (GpuBuffers.h:38-44)
* ``GPU_MESH_INFO_DTYPE`` GPUMeshInfo itself, 48 B, blasRef ignored
* ``MATERIAL_DTYPE`` GPUMaterialInfo, 64 B                 (GpuBuffers.h:29-36)
* ``INSTANCE_DTYPE`` VkAccelerationStructureInstanceKHR, 64 B (tlas_update.comp:10-21)
* ``Config``  probe grid constants                        (probe_helpers.glsl:1-7, probe_conv.comp:85)
* ``Frame``   the scene UBO subset + push constants        (scene_ubo.glsl:14-37, rgen:26-29)
"""
import ctypes

import numpy as np

VERTEX_WORDS = 24

MESH_DTYPE = np.dtype([
    ("min_pos", "<f4", (4,)),
    ("max_pos", "<f4", (4,)),
    ("vertex_offset", "<u4"),
    ("index_offset", "<u4"),
    ("index_count", "<u4"),
    ("reserved", "<u4"),
])
assert MESH_DTYPE.itemsize == 48

GPU_MESH_INFO_DTYPE = np.dtype([   # GPUMeshInfo verbatim (GpuBuffers.h:38-44), ane_gpu_mesh_info
    ("min_pos", "<f4", (4,)),
    ("max_pos", "<f4", (4,)),
    ("vertex_offset", "<u4"),
    ("index_offset", "<u4"),
    ("blas_ref", "<u8"),
])
assert GPU_MESH_INFO_DTYPE.itemsize == 48

MATERIAL_DTYPE = np.dtype([
    ("base_col_fac", "<f4", (4,)),
    ("emissive", "<f4", (4,)),
    ("bindless_indices", "<i4", (4,)),
    ("met_rough", "<f4", (4,)),
])
assert MATERIAL_DTYPE.itemsize == 64

INSTANCE_DTYPE = np.dtype([
    ("transform", "<f4", (12,)),
    ("custom_index_and_mask", "<u4"),
    ("sbt_offset_and_flags", "<u4"),
    ("reserved", "<u8"),
])
assert INSTANCE_DTYPE.itemsize == 64

GBUFFER_DTYPE = np.dtype([  # ane_gbuffer_texel, 48 B (deferred_tiled.comp:264-298 inputs)
    ("pos", "<f4", 3), ("shadow", "<f4"), ("normal", "<f4", 3), ("metallic", "<f4"),
    ("albedo", "<f4", 3), ("roughness", "<f4"),
])
assert GBUFFER_DTYPE.itemsize == 48

HIT_DTYPE = np.dtype([("tri", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])

# scene_ubo.glsl:1-12
FLAG_DDGI = 1 << 4
FLAG_DDGI_MULTI = 1 << 5
FLAG_BAKE = 1 << 11

PRECISION_FP32 = 0
PRECISION_BF16 = 1
PRECISION_FP8 = 2
PRECISION_FP16 = 3


class Config(ctypes.Structure):
    """ane_ddgi_config."""
    _fields_ = [
        ("probes_x", ctypes.c_int32),
        ("probes_y", ctypes.c_int32),
        ("probes_z", ctypes.c_int32),
        ("probe_step", ctypes.c_float * 3),
        ("rays_per_probe", ctypes.c_int32),
        ("hysteresis", ctypes.c_float),
        ("t_min", ctypes.c_float),
        ("t_max", ctypes.c_float),
        ("precision", ctypes.c_int32),
        ("quirks", ctypes.c_uint32),
    ]

    @classmethod
    def make(cls, grid, step=(1.0, 2.0, 1.0), rays=256, hysteresis=0.96,
             t_min=0.1, t_max=50.0, precision=PRECISION_FP32, quirks=0):
        c = cls()
        c.probes_x, c.probes_y, c.probes_z = (int(g) for g in grid)
        for i in range(3):
            c.probe_step[i] = float(step[i])
        c.rays_per_probe = int(rays)
        c.hysteresis = float(hysteresis)
        c.t_min = float(t_min)
        c.t_max = float(t_max)
        c.precision = int(precision)
        c.quirks = int(quirks)
        return c

    @property
    def num_probes(self):
        return self.probes_x * self.probes_y * self.probes_z

    @property
    def atlas_shape(self):
        """(height, width, 4) of the FP32 RGBA atlas (IrradianceProbeConvolvePass.cpp:89-90)."""
        return (10 * self.probes_z * self.probes_y, 10 * self.probes_x, 4)


class Frame(ctypes.Structure):
    """ane_ddgi_frame."""
    _fields_ = [
        ("camera_pos", ctypes.c_float * 4),
        ("light_dir", ctypes.c_float * 4),
        ("sky_color", ctypes.c_float * 4),
        ("sun_intensity", ctypes.c_float),
        ("sky_intensity", ctypes.c_float),
        ("flags", ctypes.c_uint32),
        ("layer_begin", ctypes.c_int32),
        ("layer_count", ctypes.c_int32),
        ("rotation", ctypes.c_float * 16),
        ("bake_tile", ctypes.c_int32 * 4),
    ]

    def copy(self):
        f = Frame()
        ctypes.memmove(ctypes.byref(f), ctypes.byref(self), ctypes.sizeof(Frame))
        return f


assert ctypes.sizeof(Config) == 48
assert ctypes.sizeof(Frame) == 4 * 4 * 3 + 4 * 5 + 64 + 16


def ptr(a):
    """ctypes void* of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)
