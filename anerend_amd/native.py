"""ctypes binding of libane_ddgi.so (the C-ABI declared in include/ane_ddgi.h).

The library is built in-tree by ``python -m anerend_amd.build``.  There is no
fallback: if the shared library is missing every call raises.
"""
import ctypes
import os

import numpy as np

from .types import HIT_DTYPE, Config, Frame, ptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ANE_DDGI_LIB") or os.path.join(HERE, "libane_ddgi.so")

ANE_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "out of device memory", 3: "admission refused",
          4: "device error", 5: "not ready", 6: "unsupported", 7: "cancelled"}


class AneError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class SceneInfo(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_int64), ("num_instances", ctypes.c_int32),
                ("bvh_depth", ctypes.c_int32), ("device_bytes", ctypes.c_int64),
                ("build_ms", ctypes.c_float), ("node_format", ctypes.c_int32)]


STAGES = ("setup", "trace", "sungrid", "shade", "convolve", "translate", "shadow")


class Stats(ctypes.Structure):
    _fields_ = [("last_ms", ctypes.c_float * 7), ("mean_ms", ctypes.c_float * 7),
                ("calls", ctypes.c_uint64 * 7),
                ("primary_nodes", ctypes.c_uint64), ("primary_tris", ctypes.c_uint64),
                ("shadow_nodes", ctypes.c_uint64), ("shadow_tris", ctypes.c_uint64),
                ("rays", ctypes.c_uint64), ("device_bytes", ctypes.c_int64),
                ("sun_grid_state", ctypes.c_int32), ("sun_grid_candidates", ctypes.c_float),
                ("traversal_fault", ctypes.c_uint32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_[3:]}
        d["last_ms"] = dict(zip(STAGES, self.last_ms))
        d["mean_ms"] = dict(zip(STAGES, self.mean_ms))
        d["calls"] = dict(zip(STAGES, self.calls))
        return d


# (name, restype, argtypes); every status-returning entry point is int
_vp, _i32, _i64, _f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
_SIGS = [
    ("ane_abi_version", ctypes.c_int, []),
    ("ane_last_error", ctypes.c_char_p, []),
    ("ane_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("ane_device", ctypes.c_int, [_vp]),
    ("ane_stream_synchronize", ctypes.c_int, [_vp]),
    ("ane_scene_create", ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _vp, _i32, _vp, _i32, _vp, _vp]),
    ("ane_scene_create_mesh_info", ctypes.c_int,
     [_vp, _i64, _vp, _i64, _vp, _vp, _i32, _vp, _i32, _vp, _i32, _vp, _vp]),
    ("ane_scene_refit", ctypes.c_int, [_vp, _vp, _i32, _vp]),
    ("ane_scene_rebuild", ctypes.c_int, [_vp, _vp, _i32, _vp]),
    ("ane_scene_info_get", ctypes.c_int, [_vp, _vp]),
    ("ane_scene_check", ctypes.c_int, [_vp, _vp]),
    ("ane_scene_destroy", ctypes.c_int, [_vp]),
    ("ane_ddgi_create", ctypes.c_int, [_vp, _vp]),
    ("ane_ddgi_destroy", ctypes.c_int, [_vp]),
    ("ane_ddgi_translate", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("ane_ddgi_trace", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("ane_ddgi_convolve", ctypes.c_int, [_vp, _vp, _vp]),
    ("ane_ddgi_update", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("ane_ddgi_atlas_get", ctypes.c_int, [_vp, _vp, _vp]),
    ("ane_ddgi_atlas_set", ctypes.c_int, [_vp, _vp, _vp]),
    ("ane_ddgi_atlas_device", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("ane_ddgi_translation_state", ctypes.c_int, [_vp, _vp, _i32]),
    ("ane_ddgi_atlas_export_rgba16f", ctypes.c_int, [_vp, _vp, _vp]),
    ("ane_ddgi_gi_lookup", ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
    ("ane_ddgi_mirror_reflections", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp]),
    ("ane_ddgi_atlas_import_rgba16f", ctypes.c_int, [_vp, _vp, _vp]),
    ("ane_ddgi_radiance_set", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("ane_ddgi_radiance_get", ctypes.c_int, [_vp, _vp, _vp]),
    ("ane_ddgi_debug_rays", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("ane_ddgi_set_counting", ctypes.c_int, [_vp, _i32]),
    ("ane_ddgi_stats_get", ctypes.c_int, [_vp, _vp]),
    ("ane_ddgi_stats_reset", ctypes.c_int, [_vp]),
    ("ane_ddgi_check", ctypes.c_int, [_vp, _vp]),
    ("ane_ddgi_tiles_set", ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("ane_ddgi_conv_error_fp64", ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, _vp, _vp]),
    ("ane_ddgi_set_traversal_limit", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("ane_scene_debug_bvh", ctypes.c_int, [_vp] * 7),
    ("ane_scene_debug_world_tris", ctypes.c_int, [_vp, _vp]),
    ("ane_scene_shadow_mask", ctypes.c_int, [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    ("ane_scene_light_shadows", ctypes.c_int, [_vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    ("ane_scene_set_node_format", ctypes.c_int, [_vp, _i32, _vp]),
    ("ane_scene_trace_rays", ctypes.c_int, [_vp, _vp, _vp, _i64, _f32, _f32, _i32, _vp, _vp, _vp]),
    ("ane_ddgi_ray_directions", ctypes.c_int, [_i32, _vp, _vp]),
    ("ane_scene_trace_rays_noalloc", ctypes.c_int, [_vp, _vp, _vp, _i64, _f32, _f32, _i32, _vp, _vp, _vp]),
]

EXPORTED = [name for name, _, _ in _SIGS]

_lib = None


def load():
    """Load libane_ddgi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m anerend_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != ANE_OK:
        raise AneError(status, load().ane_last_error().decode())
    return status


class DeviceBuffer:
    """Plain device allocation for entry points that take device pointers."""

    _hip = None

    def __init__(self, nbytes):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = ctypes.CDLL("libamdhip64.so")
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        if DeviceBuffer._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 1))) != 0:
            raise AneError(2, f"hipMalloc of {self.nbytes} bytes failed")

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if DeviceBuffer._hip.hipMemcpy(self.ptr, ptr(a), ctypes.c_size_t(a.nbytes), 1) != 0:
            raise AneError(4, "hipMemcpy to device failed")

    def download(self, a):
        assert a.flags["C_CONTIGUOUS"] and a.nbytes <= self.nbytes
        if DeviceBuffer._hip.hipMemcpy(ptr(a), self.ptr, ctypes.c_size_t(a.nbytes), 2) != 0:
            raise AneError(4, "hipMemcpy to host failed")
        return a

    def close(self):
        if self.ptr.value:
            DeviceBuffer._hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    __del__ = close


def ray_directions(n, rotation16):
    """Host-computed rotated Fibonacci directions (no device work)."""
    rot = np.ascontiguousarray(rotation16, np.float32)
    out = np.zeros((n, 3), np.float32)
    check(load().ane_ddgi_ray_directions(n, ptr(rot), ptr(out)))
    return out


class DeviceScene:
    """Geometry + LBVH resident in HBM (replaces the BLAS/TLAS builds)."""

    def __init__(self, scene, stream=None, mesh_info=None):
        """``mesh_info`` = (GPUMeshInfo array of GPU_MESH_INFO_DTYPE, uint32 index counts)
        builds through ane_scene_create_mesh_info instead of scene.meshes."""
        L = load()
        self._h = ctypes.c_void_p()
        arrs = [np.ascontiguousarray(a) for a in
                (scene.vertices, scene.indices, scene.meshes, scene.materials, scene.instances)]
        v, i, m, mt, inst = arrs
        if mesh_info is None:
            check(L.ane_scene_create(ptr(v), len(v), ptr(i), len(i), ptr(m), len(m), ptr(mt), len(mt),
                                     ptr(inst), len(inst), stream, ctypes.byref(self._h)))
        else:
            gm = np.ascontiguousarray(mesh_info[0])
            cnt = np.ascontiguousarray(mesh_info[1], np.uint32)
            assert len(cnt) == len(gm)
            check(L.ane_scene_create_mesh_info(ptr(v), len(v), ptr(i), len(i), ptr(gm), ptr(cnt), len(gm),
                                               ptr(mt), len(mt), ptr(inst), len(inst), stream,
                                               ctypes.byref(self._h)))
        self.instances = inst.copy()

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().ane_scene_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    def info(self):
        s = SceneInfo()
        check(load().ane_scene_info_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def check(self, stream=None):
        """Raises AneError if the last shadow-mask or light-shadow pass had a ray end
        without a result (clears the flags)."""
        check(load().ane_scene_check(self._h, stream))

    NODES_AUTO, NODES_FULL, NODES_QUANTIZED = 0, 1, 2

    def set_node_format(self, fmt, stream=None):
        """Traversal node layout: NODES_AUTO, NODES_FULL or NODES_QUANTIZED."""
        check(load().ane_scene_set_node_format(self._h, int(fmt), stream))

    def refit(self, instances, stream=None):
        inst = np.ascontiguousarray(instances)
        check(load().ane_scene_refit(self._h, ptr(inst), len(inst), stream))

    def rebuild(self, instances, stream=None):
        inst = np.ascontiguousarray(instances)
        check(load().ane_scene_rebuild(self._h, ptr(inst), len(inst), stream))

    def bvh(self):
        n = self.info()["num_triangles"]
        m = max(n - 1, 0)
        out = dict(keys=np.zeros(n, np.uint64), ids=np.zeros(n, np.uint32),
                   left=np.zeros(m, np.int32), right=np.zeros(m, np.int32),
                   parent=np.zeros(m, np.int32), boxes=np.zeros((m, 6), np.float32))
        check(load().ane_scene_debug_bvh(self._h, ptr(out["keys"]), ptr(out["ids"]), ptr(out["left"]),
                                         ptr(out["right"]), ptr(out["parent"]), ptr(out["boxes"])))
        return out

    def world_tris(self):
        n = self.info()["num_triangles"]
        out = np.zeros((n, 9), np.float32)
        check(load().ane_scene_debug_world_tris(self._h, ptr(out)))
        return out

    def shadow_mask(self, depth, inv_view_proj16, light_dir3, count=False):
        """RT hard shadows of a host depth buffer [h][w] -> visibility [h][w] (and counters)."""
        d = np.ascontiguousarray(depth, np.float32)
        h, w = d.shape
        ivp = np.ascontiguousarray(inv_view_proj16, np.float32)
        light = np.ascontiguousarray(light_dir3, np.float32)
        src, dst = DeviceBuffer(d.nbytes), DeviceBuffer(d.nbytes)
        ctr = np.zeros(2, np.uint64)
        try:
            src.upload(d)
            check(load().ane_scene_shadow_mask(self._h, src.ptr, w, h, ptr(ivp), ptr(light), dst.ptr,
                                               ptr(ctr) if count else None, None))
            synchronize()
            vis = dst.download(np.zeros((h, w), np.float32))
        finally:
            src.close()
            dst.close()
        return (vis, ctr) if count else vis

    def light_shadows(self, lights, count=False, moments_fill=0.0):
        """Point-light shadow moments of ``lights`` [n][8] = {x, y, z, range, r, g, b,
        enabled} -> (distance image [32][32 n], moments image [8][10 n][2]) (and
        counters).  Border texels of the moments image keep ``moments_fill``."""
        L = np.ascontiguousarray(lights, np.float32).reshape(-1, 8)
        n = len(L)
        dist = np.zeros((32, 32 * n), np.float32)
        mom = np.full((8, 10 * n, 2), moments_fill, np.float32)
        ctr = np.zeros(2, np.uint64)
        dd, dm = DeviceBuffer(max(dist.nbytes, 4)), DeviceBuffer(max(mom.nbytes, 4))
        try:
            dm.upload(mom)
            check(load().ane_scene_light_shadows(self._h, ptr(L), n, dd.ptr, dm.ptr,
                                                 ptr(ctr) if count else None, None))
            synchronize()
            dist = dd.download(dist)
            mom = dm.download(mom)
        finally:
            dd.close()
            dm.close()
        return (dist, mom, ctr) if count else (dist, mom)

    def trace_rays(self, orig, dirs, tmin=0.1, tmax=50.0, any_hit=False, count=False, stream=None):
        orig = np.ascontiguousarray(orig, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        hits = np.zeros(len(orig), HIT_DTYPE)
        ctr = np.zeros(2, np.uint64) if count else None
        check(load().ane_scene_trace_rays(self._h, ptr(orig), ptr(dirs), len(orig), tmin, tmax,
                                          int(any_hit), ptr(hits), ptr(ctr), stream))
        return (hits, ctr) if count else hits


class ProbeGrid:
    """One DDGI probe grid (atlas, radiance, per-pass state) on the device."""

    def __init__(self, config: Config):
        self.config = config
        self._h = ctypes.c_void_p()
        check(load().ane_ddgi_create(ctypes.byref(config), ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().ane_ddgi_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    # -- passes --
    def translate(self, camera_pos, stream=None):
        cam = np.ascontiguousarray(camera_pos[:3], np.float32)
        shifted = ctypes.c_int32(0)
        check(load().ane_ddgi_translate(self._h, ptr(cam), ctypes.byref(shifted), stream))
        return bool(shifted.value)

    def trace(self, scene: DeviceScene, frame: Frame, stream=None):
        check(load().ane_ddgi_trace(self._h, scene.handle, ctypes.byref(frame), stream))

    def convolve(self, frame: Frame, stream=None):
        check(load().ane_ddgi_convolve(self._h, ctypes.byref(frame), stream))

    def update(self, scene: DeviceScene, frame: Frame, stream=None):
        check(load().ane_ddgi_update(self._h, scene.handle, ctypes.byref(frame), stream))

    # -- state --
    @property
    def atlas(self):
        a = np.zeros(self.config.atlas_shape, np.float32)
        check(load().ane_ddgi_atlas_get(self._h, ptr(a), None))
        return a

    @atlas.setter
    def atlas(self, value):
        a = np.ascontiguousarray(value, np.float32)
        assert a.shape == self.config.atlas_shape
        check(load().ane_ddgi_atlas_set(self._h, ptr(a), None))

    def atlas_device(self):
        p = ctypes.c_void_p()
        w = ctypes.c_int64()
        h = ctypes.c_int64()
        check(load().ane_ddgi_atlas_device(self._h, ctypes.byref(p), ctypes.byref(w), ctypes.byref(h)))
        return p.value, w.value, h.value

    def atlas_rgba16f(self):
        """Atlas as the reference's baked-tile texture: [H][W][4] IEEE binary16."""
        a = np.zeros(self.config.atlas_shape, np.float16)
        check(load().ane_ddgi_atlas_export_rgba16f(self._h, ptr(a), None))
        return a

    def set_atlas_rgba16f(self, value):
        a = np.ascontiguousarray(value, np.float16)
        assert a.shape == tuple(self.config.atlas_shape)
        check(load().ane_ddgi_atlas_import_rgba16f(self._h, ptr(a), None))

    def gi_lookup(self, frame, gbuf):
        """Deferred GI lookup of G-buffer texels (GBUFFER_DTYPE or n x 12 float32) -> n x 4."""
        g = np.ascontiguousarray(gbuf).view(np.float32).reshape(-1, 12)
        n = len(g)
        src, dst = DeviceBuffer(g.nbytes), DeviceBuffer(16 * n)
        try:
            src.upload(g)
            check(load().ane_ddgi_gi_lookup(self._h, ctypes.byref(frame), src.ptr, n, dst.ptr, None))
            synchronize()
            return dst.download(np.zeros((n, 4), np.float32))
        finally:
            src.close()
            dst.close()

    def mirror_reflections(self, scene, frame, depth, normal4, inv_view_proj16, count=False):
        """Mirror reflections of a host depth [h][w] + normal [h][w][4] buffer ->
        [h][w][4] (irradiance, hit distance) (and counters[4])."""
        d = np.ascontiguousarray(depth, np.float32)
        h, w = d.shape
        nrm = np.ascontiguousarray(normal4, np.float32).reshape(h, w, 4)
        ivp = np.ascontiguousarray(inv_view_proj16, np.float32)
        bufs = [DeviceBuffer(d.nbytes), DeviceBuffer(nrm.nbytes), DeviceBuffer(nrm.nbytes)]
        ctr = np.zeros(4, np.uint64)
        try:
            bufs[0].upload(d)
            bufs[1].upload(nrm)
            check(load().ane_ddgi_mirror_reflections(self._h, scene.handle, ctypes.byref(frame), bufs[0].ptr,
                                                     bufs[1].ptr, w, h, ptr(ivp), bufs[2].ptr,
                                                     ptr(ctr) if count else None, None))
            synchronize()
            out = bufs[2].download(np.zeros((h, w, 4), np.float32))
        finally:
            for b in bufs:
                b.close()
        return (out, ctr) if count else out

    def translation_state(self, cell=None):
        c = np.zeros(2, np.int32) if cell is None else np.ascontiguousarray(cell, np.int32)
        check(load().ane_ddgi_translation_state(self._h, ptr(c), int(cell is not None)))
        return c

    def set_radiance(self, frame, radiance):
        r = np.ascontiguousarray(radiance, np.float32)
        check(load().ane_ddgi_radiance_set(self._h, ctypes.byref(frame), ptr(r), None))

    def radiance(self, frame):
        c = self.config
        n = c.probes_x * c.probes_z * frame.layer_count
        r = np.zeros((n, 3, c.rays_per_probe), np.float32)
        check(load().ane_ddgi_radiance_get(self._h, ptr(r), None))
        return r

    def debug_rays(self, frame):
        c = self.config
        n = c.probes_x * c.probes_z * frame.layer_count * c.rays_per_probe
        hits = np.zeros(n, HIT_DTYPE)
        sh = np.zeros(n, np.uint8)
        check(load().ane_ddgi_debug_rays(self._h, ptr(hits), ptr(sh), None))
        return hits, sh

    def set_counting(self, mode=1):
        """0 off; 1 (or True) every ray traverses; 2 shadow rays take the sun grid when current."""
        check(load().ane_ddgi_set_counting(self._h, int(mode)))

    def stats(self):
        """Stats dict; raises AneError (its ``stats`` holds the dict) while a traversal
        fault is pending."""
        s = Stats()
        rc = load().ane_ddgi_stats_get(self._h, ctypes.byref(s))
        if rc != ANE_OK:
            err = AneError(rc, load().ane_last_error().decode())
            err.stats = s.as_dict() if rc == 4 else None
            raise err
        return s.as_dict()

    def reset_stats(self):
        check(load().ane_ddgi_stats_reset(self._h))

    def check(self, stream=None):
        """Raises AneError if a ray since the last check ended without a result (clears it)."""
        check(load().ane_ddgi_check(self._h, stream))

    def set_tiles(self, tile_xz, atlases_rgba16f, tile_size=32):
        """Baked tile atlases: tile_xz [n][2], atlases [n][H][W][4] float16 ([] clears)."""
        xz = np.ascontiguousarray(tile_xz, np.int32).reshape(-1, 2)
        a = np.ascontiguousarray(atlases_rgba16f, np.float16)
        if len(xz):
            assert a.shape == (len(xz),) + tuple(self.config.atlas_shape)
        check(load().ane_ddgi_tiles_set(self._h, ptr(xz), ptr(a) if len(xz) else None, len(xz),
                                        int(tile_size), None))

    def conv_error_fp64(self, ref, rel_floor=1e-3):
        """(max, rms) relative error of this grid's last convolution against float64 over
        the FP32 radiance ``ref`` holds, reduced on the GPU (see ane_ddgi_conv_error_fp64)."""
        mx, rms = ctypes.c_double(), ctypes.c_double()
        check(load().ane_ddgi_conv_error_fp64(self._h, ref.handle, float(rel_floor), ctypes.byref(mx),
                                              ctypes.byref(rms), None))
        return mx.value, rms.value

    def set_traversal_limit(self, max_node_visits):
        check(load().ane_ddgi_set_traversal_limit(self._h, int(max_node_visits)))


def synchronize(stream=None):
    check(load().ane_stream_synchronize(stream))
