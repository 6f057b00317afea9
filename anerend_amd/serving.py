"""ctypes binding of the request loop in include/ane_serve.h (many resident
scenes, continuous admission, overlapping updates on several HIP streams)."""
import ctypes

import numpy as np

from . import native
from .types import Config, Frame

ANE_ERR_NOT_READY = 5


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64),
                ("max_inflight", ctypes.c_int32),
                ("num_streams", ctypes.c_int32),
                ("capture_graphs", ctypes.c_int32),
                ("fuse_stages", ctypes.c_int32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32),
                ("session", ctypes.c_int32),
                ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double),
                ("rays", ctypes.c_int64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("sessions", ctypes.c_int32), ("queued", ctypes.c_int32),
                ("inflight", ctypes.c_int32), ("peak_inflight", ctypes.c_int32),
                ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("rays_done", ctypes.c_int64),
                ("used_bytes", ctypes.c_int64), ("budget_bytes", ctypes.c_int64),
                ("held_bytes", ctypes.c_int64)]


assert ctypes.sizeof(ServeConfig) == 24
assert ctypes.sizeof(ServeResult) == 32


class Server:
    """Owns the request loop; sessions reference DeviceScenes the caller keeps alive."""

    def __init__(self, hbm_budget_bytes, max_inflight=8, num_streams=4, capture_graphs=True, fuse_stages=False):
        """fuse_stages: the requests launched together share one shading launch and one
        convolution launch (capture_graphs then does not apply)."""
        L = native.load()
        c = ServeConfig(int(hbm_budget_bytes), int(max_inflight), int(num_streams), int(bool(capture_graphs)),
                        int(bool(fuse_stages)))
        self._h = ctypes.c_void_p()
        native.check(L.ane_serve_open(ctypes.byref(c), ctypes.byref(self._h)))
        self.configs = {}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            native.load().ane_serve_close(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    def add_session(self, scene: "native.DeviceScene", config: Config):
        sid = ctypes.c_int32(-1)
        native.check(native.load().ane_serve_add_session(self._h, scene.handle, ctypes.byref(config),
                                                         ctypes.byref(sid)))
        self.configs[sid.value] = config
        return sid.value

    def remove_session(self, session):
        native.check(native.load().ane_serve_remove_session(self._h, ctypes.c_int32(session)))
        self.configs.pop(session, None)

    def submit(self, session, frame: Frame):
        t = ctypes.c_uint64()
        native.check(native.load().ane_serve_submit(self._h, ctypes.c_int32(session),
                                                    ctypes.byref(frame), ctypes.byref(t)))
        return t.value

    def poll(self, ticket, timeout_ms=-1):
        """Result dict, or None when not finished within ``timeout_ms``."""
        r = ServeResult()
        st = native.load().ane_serve_poll(self._h, ctypes.c_uint64(ticket), ctypes.c_int32(timeout_ms),
                                          ctypes.byref(r))
        if st == ANE_ERR_NOT_READY:
            return None
        native.check(st)
        return {k: getattr(r, k) for k, _ in r._fields_}

    def atlas(self, session):
        a = np.zeros(self.configs[session].atlas_shape, np.float32)
        native.check(native.load().ane_serve_session_atlas(self._h, ctypes.c_int32(session),
                                                           native.ptr(a)))
        return a

    def stats(self):
        s = ServeStats()
        native.check(native.load().ane_serve_stats_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}
