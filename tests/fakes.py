"""Test double of native.ProbeGrid backed by the CPU oracle (tests only)."""
import numpy as np

from oracle import oracle as O


class OracleGrid:
    def __init__(self, config, scene):
        self.config = config
        self.osc = O.OracleScene(scene)
        self.atlas = O.new_atlas(config)
        self.last = np.zeros(2, np.int32)
        self._pending = None
        self.calls = []

    def translate(self, cam):
        moved, self.last = O.translate(self.config, np.asarray(cam, np.float32)[:3], self.last, self.atlas)
        self.calls.append(("translate", bool(moved)))
        return moved

    def trace(self, scene, frame):
        self._pending = frame.copy()
        self.calls.append(("trace", frame.layer_begin, frame.layer_count))

    def convolve(self, frame):
        assert self._pending is not None and self._pending.layer_begin == frame.layer_begin
        self.osc.update(self.config, self._pending, self.atlas, want_rays=False)
        self.calls.append(("convolve", frame.layer_begin, frame.layer_count))
