"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.  See ddgi_oracle.h for what each function
restates from the reference and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

from anerend_amd.types import HIT_DTYPE, Config, Frame, ptr

HERE = os.path.dirname(os.path.abspath(__file__))
# DDGI_ORACLE_LIB names another build of ddgi_oracle.c to load instead (the
# planted-mistake copies of tests/test_oracle_mutants.py); unset, the in-tree one.
LIB_PATH = os.environ.get("DDGI_ORACLE_LIB") or os.path.join(HERE, "libddgi_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        L.ora_scene_create.restype = vp
        L.ora_scene_create.argtypes = [vp, i64, vp, i64, vp, i32, vp, i32, vp, i32]
        L.ora_scene_destroy.argtypes = [vp]
        L.ora_scene_refit.argtypes = [vp, vp, i32]
        L.ora_scene_num_tris.restype = i64
        L.ora_scene_num_tris.argtypes = [vp]
        L.ora_scene_max_depth.restype = i32
        L.ora_scene_max_depth.argtypes = [vp]
        L.ora_scene_world_tris.argtypes = [vp, vp]
        L.ora_scene_bvh.argtypes = [vp] * 7
        L.ora_gi_lookup.argtypes = [vp, vp, vp, vp, i64, vp]
        L.ora_shadow_mask.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp]
        L.ora_light_shadows.argtypes = [vp, vp, i32, vp, vp, vp]
        L.ora_scene_set_node32.argtypes = [vp, i32]
        L.ora_scene_qboxes.argtypes = [vp, vp]
        L.ora_light_shadow_weights.argtypes = [vp]
        L.ora_mirror_reflections.argtypes = [vp, vp, vp, vp, vp, vp, i32, i32, vp, vp, vp]
        L.ora_trace.argtypes = [vp, vp, vp, i64, f32, f32, i32, vp, vp]
        L.ora_trace_bruteforce.argtypes = [vp, vp, vp, i64, f32, f32, i32, vp]
        L.ora_fibonacci_dirs.argtypes = [i32, vp]
        L.ora_rotate_dirs.argtypes = [i32, vp, vp, vp]
        L.ora_probe_origin.argtypes = [vp, vp, i32, i32, i32, vp]
        L.ora_oct_encode.argtypes = [vp, vp]
        L.ora_oct_decode.argtypes = [vp, vp]
        L.ora_texel_dirs.argtypes = [vp]
        L.ora_conv_weights.argtypes = [i32, vp, vp, vp]
        L.ora_sky_radiance.argtypes = [vp, vp]
        L.ora_shade_hit.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ora_sample_probe.argtypes = [vp, vp, vp, vp, vp, vp]
        L.ora_convolve_blend.argtypes = [vp, vp, vp, vp, vp]
        L.ora_translate.restype = ctypes.c_int
        L.ora_translate.argtypes = [vp, vp, vp, vp]
        L.ora_update.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.ora_atlas_clear.argtypes = [vp, vp]
        L.ora_f16_round.argtypes = [vp, vp, i64]
        L.ora_tiles_set.argtypes = [vp, vp, vp, i32, i32]
        L.ora_last_update_stage_ms.argtypes = [vp]
        L.ora_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ref(s):
    return ctypes.byref(s)


class OracleScene:
    def __init__(self, scene):
        self.scene = scene
        L = lib()
        self._keep = [np.ascontiguousarray(a) for a in
                      (scene.vertices, scene.indices, scene.meshes, scene.materials, scene.instances)]
        v, i, m, mt, inst = self._keep
        self.h = L.ora_scene_create(ptr(v), len(v), ptr(i), len(i), ptr(m), len(m), ptr(mt), len(mt),
                                    ptr(inst), len(inst))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_scene_destroy(self.h)
            self.h = None

    @property
    def num_tris(self):
        return lib().ora_scene_num_tris(self.h)

    @property
    def max_depth(self):
        return lib().ora_scene_max_depth(self.h)

    def refit(self, instances):
        instances = np.ascontiguousarray(instances)
        rc = lib().ora_scene_refit(self.h, ptr(instances), len(instances))
        if rc != 0:
            raise ValueError("refit rejected")

    def world_tris(self):
        out = np.zeros((self.num_tris, 9), np.float32)
        lib().ora_scene_world_tris(self.h, ptr(out))
        return out

    def bvh(self):
        n = self.num_tris
        keys = np.zeros(n, np.uint64)
        ids = np.zeros(n, np.uint32)
        m = max(n - 1, 0)
        left = np.zeros(m, np.int32)
        right = np.zeros(m, np.int32)
        parent = np.zeros(m, np.int32)
        boxes = np.zeros((m, 6), np.float32)
        lib().ora_scene_bvh(self.h, ptr(keys), ptr(ids), ptr(left), ptr(right), ptr(parent), ptr(boxes))
        return dict(keys=keys, ids=ids, left=left, right=right, parent=parent, boxes=boxes)

    def shadow_mask(self, depth, inv_view_proj16, light_dir3):
        """RT hard shadows of a depth buffer [h][w] -> (visibility [h][w], counters[2])."""
        d = np.ascontiguousarray(depth, np.float32)
        h, w = d.shape
        ivp = np.ascontiguousarray(inv_view_proj16, np.float32)
        light = np.ascontiguousarray(light_dir3, np.float32)
        out = np.zeros((h, w), np.float32)
        ctr = np.zeros(2, np.uint64)
        lib().ora_shadow_mask(self.h, ptr(d), w, h, ptr(ivp), ptr(light), ptr(out), ptr(ctr))
        return out, ctr

    def set_node32(self, on=True):
        """Traverse the decoded boxes of the library's quantized 32-byte nodes."""
        lib().ora_scene_set_node32(self.h, int(bool(on)))

    def qboxes(self):
        """Decoded quantized child boxes [ntri-1][2][2][3] (child, lo/hi, xyz)."""
        out = np.zeros((max(self.num_tris - 1, 0), 12), np.float32)
        lib().ora_scene_qboxes(self.h, ptr(out))
        return out.reshape(-1, 2, 2, 3)

    def light_shadows(self, lights, moments_fill=0.0):
        """Point-light shadow moments (light_shadow.rgen, light_shadow_sum.comp) of
        ``lights`` [n][8] = {x, y, z, range, r, g, b, enabled} -> (distance image
        [32][32 n], moments image [8][10 n][2], counters[2]).  Border texels of the
        moments image keep ``moments_fill``."""
        L = np.ascontiguousarray(lights, np.float32).reshape(-1, 8)
        n = len(L)
        dist = np.zeros((32, 32 * n), np.float32)
        mom = np.full((8, 10 * n, 2), moments_fill, np.float32)
        ctr = np.zeros(2, np.uint64)
        lib().ora_light_shadows(self.h, ptr(L), n, ptr(dist), ptr(mom), ptr(ctr))
        return dist, mom, ctr

    def mirror_reflections(self, cfg, frame, atlas, depth, normal4, inv_view_proj16):
        """Mirror reflections of a depth + normal buffer -> ([h][w][4] (rgb, dist), counters[4])."""
        d = np.ascontiguousarray(depth, np.float32)
        h, w = d.shape
        nrm = np.ascontiguousarray(normal4, np.float32).reshape(h, w, 4)
        ivp = np.ascontiguousarray(inv_view_proj16, np.float32)
        out = np.zeros((h, w, 4), np.float32)
        ctr = np.zeros(4, np.uint64)
        lib().ora_mirror_reflections(self.h, _ref(cfg), _ref(frame), ptr(atlas), ptr(d), ptr(nrm), w, h,
                                     ptr(ivp), ptr(out), ptr(ctr))
        return out, ctr

    def trace(self, orig, dirs, tmin=0.1, tmax=50.0, any_hit=False, brute=False):
        orig = np.ascontiguousarray(orig, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(orig)
        out = np.zeros(n, HIT_DTYPE)
        if brute:
            lib().ora_trace_bruteforce(self.h, ptr(orig), ptr(dirs), n, tmin, tmax, int(any_hit), ptr(out))
            return out
        ctr = np.zeros(2, np.uint64)
        lib().ora_trace(self.h, ptr(orig), ptr(dirs), n, tmin, tmax, int(any_hit), ptr(out), ptr(ctr))
        return out, ctr

    def update(self, cfg, frame, atlas, want_rays=True):
        """One probe update in place on ``atlas``; returns hits, shadow flags, radiance, counters."""
        assert atlas.dtype == np.float32 and atlas.flags["C_CONTIGUOUS"]
        np_ = cfg.probes_x * cfg.probes_z * frame.layer_count
        n = cfg.rays_per_probe
        hits = np.zeros(np_ * n, HIT_DTYPE) if want_rays else None
        sh = np.zeros(np_ * n, np.uint8) if want_rays else None
        rad = np.zeros((np_, 3, n), np.float32) if want_rays else None
        ctr = np.zeros(4, np.uint64)
        lib().ora_update(self.h, _ref(cfg), _ref(frame), ptr(atlas), ptr(hits), ptr(sh), ptr(rad), ptr(ctr))
        return hits, sh, rad, ctr

    def shade_hit(self, cfg, frame, atlas, origin, hit):
        out = np.zeros(3, np.float32)
        shadowed = ctypes.c_int32(0)
        h = np.zeros(1, HIT_DTYPE)
        h[0] = hit
        o = np.ascontiguousarray(origin, np.float32)
        lib().ora_shade_hit(self.h, _ref(cfg), _ref(frame), ptr(atlas), ptr(o), ptr(h), ptr(out),
                            ctypes.byref(shadowed), None)
        return out, shadowed.value


def light_shadow_weights():
    """The sum's [64][1024] weights (light_shadow_sum.comp:56-61)."""
    out = np.zeros((64, 1024), np.float32)
    lib().ora_light_shadow_weights(ptr(out))
    return out


def fibonacci_dirs(n):
    out = np.zeros((n, 3), np.float32)
    lib().ora_fibonacci_dirs(n, ptr(out))
    return out


def rotate_dirs(rot16, dirs):
    rot = np.ascontiguousarray(rot16, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    out = np.zeros_like(dirs)
    lib().ora_rotate_dirs(len(dirs), ptr(rot), ptr(dirs), ptr(out))
    return out


def probe_origin(cfg, frame, px, py, pz):
    out = np.zeros(3, np.float32)
    lib().ora_probe_origin(_ref(cfg), _ref(frame), px, py, pz, ptr(out))
    return out


def oct_encode(v):
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(2, np.float32)
    lib().ora_oct_encode(ptr(v), ptr(out))
    return out


def oct_decode(o):
    o = np.ascontiguousarray(o, np.float32)
    out = np.zeros(3, np.float32)
    lib().ora_oct_decode(ptr(o), ptr(out))
    return out


def texel_dirs():
    out = np.zeros((64, 3), np.float32)
    lib().ora_texel_dirs(ptr(out))
    return out


def conv_weights(dirs):
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = len(dirs)
    w = np.zeros((64, n), np.float32)
    s = np.zeros(64, np.float32)
    lib().ora_conv_weights(n, ptr(dirs), ptr(w), ptr(s))
    return w, s


def sky_radiance(frame):
    out = np.zeros(3, np.float32)
    lib().ora_sky_radiance(_ref(frame), ptr(out))
    return out


def sample_probe(cfg, frame, atlas, pos, normal):
    out = np.zeros(3, np.float32)
    p = np.ascontiguousarray(pos, np.float32)
    n = np.ascontiguousarray(normal, np.float32)
    lib().ora_sample_probe(_ref(cfg), _ref(frame), ptr(atlas), ptr(p), ptr(n), ptr(out))
    return out


def gi_lookup(cfg, frame, atlas, gbuf):
    """Deferred GI lookup of G-buffer texels (n x 12 float32) -> n x 4."""
    g = np.ascontiguousarray(gbuf, np.float32).reshape(-1, 12)
    out = np.zeros((len(g), 4), np.float32)
    lib().ora_gi_lookup(_ref(cfg), _ref(frame), ptr(atlas), ptr(g), len(g), ptr(out))
    return out


def convolve_blend(cfg, frame, dirs, radiance, atlas):
    dirs = np.ascontiguousarray(dirs, np.float32)
    radiance = np.ascontiguousarray(radiance, np.float32)
    lib().ora_convolve_blend(_ref(cfg), _ref(frame), ptr(dirs), ptr(radiance), ptr(atlas))


def translate(cfg, camera_pos, last_idx, atlas):
    cam = np.ascontiguousarray(camera_pos, np.float32)
    last = np.ascontiguousarray(last_idx, np.int32)
    moved = lib().ora_translate(_ref(cfg), ptr(cam), ptr(last), ptr(atlas))
    return bool(moved), last


QUIRK_FP16 = 1 << 0   # ORA_QUIRK_FP16: emulate the reference's RGBA16F images


def f16_round(x):
    """binary32 -> binary16 (nearest even) -> binary32, the oracle's own conversion."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().ora_f16_round(ptr(x), ptr(out), x.size)
    return out


def tiles_set(cfg, tile_xz, atlases, tile_size=32):
    """Baked tile atlases: tile_xz [n][2], atlases [n][H][W][4] (RGBA16F or float32;
    binary16 input is widened exactly).  An empty list clears the set."""
    xz = np.ascontiguousarray(tile_xz, np.int32).reshape(-1, 2)
    a = np.ascontiguousarray(np.asarray(atlases).astype(np.float32)).reshape(len(xz), *cfg.atlas_shape)
    lib().ora_tiles_set(_ref(cfg), ptr(xz), ptr(a), len(xz), int(tile_size))


def new_atlas(cfg):
    a = np.zeros(cfg.atlas_shape, np.float32)
    lib().ora_atlas_clear(_ref(cfg), ptr(a))
    return a


def last_update_stage_ms():
    """Wall-clock ms of the last update: (trace + shadow + shade, convolve + blend)."""
    out = np.zeros(2, np.float64)
    lib().ora_last_update_stage_ms(ptr(out))
    return float(out[0]), float(out[1])


def num_threads():
    return lib().ora_num_threads()
