"""Host-side mirror of AneRend's probe-update passes over the C-ABI.

The reference registers three passes into its frame graph and runs them in
submission order every frame (VulkanRenderer.cpp:4471-4475):

    IrradianceProbeTrans     IrradianceProbeTranslationPass.cpp:16-187
    IrradianceProbeRT        IrradianceProbeRayTracingPass.cpp:28-166
    IrradianceProbeConvolve  IrradianceProbeConvolvePass.cpp:56-144

This module keeps that plugin interface -- ``RenderPass.register_to_graph``,
``FrameGraphBuilder.register_render_pass_exe`` / ``execute_graph``, a
``RenderContext`` with the queries the passes make -- and the passes' state
rules: the Y layer round robin (``_currentLayer``), the per-dispatch random
rotation about a random axis, the ``ProbeLayer`` blackboard value handed from
RT to Convolve, and the camera-cell translation state (``_lastProbeIdx``,
kept by the device context).  Each exe calls the C-ABI where the reference
recorded vkCmdCopyImage / vkCmdTraceRaysKHR / vkCmdDispatch.

Differences from the reference, all explicit inputs here: the rotation is
drawn from a seeded PCG32 stream instead of ``std::random_device`` (Q8), and
``full_grid=True`` updates every layer per frame instead of one.
"""
from dataclasses import dataclass, field

import numpy as np

from .scenegen import PCG32, rotation_matrix
from .types import FLAG_BAKE, FLAG_DDGI, FLAG_DDGI_MULTI, Frame


@dataclass
class RenderOptions:
    """RenderOptions.h:7-28 (the fields the probe passes read)."""
    raytracing_enabled: bool = True
    ddgi_enabled: bool = True
    multi_bounce_ddgi_enabled: bool = True
    sun_intensity: float = 5.0
    sky_intensity: float = 1.0
    sky_color: tuple = (0.2, 0.5, 0.7)


@dataclass
class Camera:
    position: tuple = (0.0, 0.0, 0.0)

    def get_position(self):
        return np.asarray(self.position, np.float32)


@dataclass
class RenderContext:
    """The RenderContext queries of the probe passes (RenderContext.h:85-188)
    plus the scene-UBO fields they read (GpuBuffers.h:125-148)."""
    grid: object                      # native.ProbeGrid (or a test double with the same methods)
    scene: object                     # native.DeviceScene
    options: RenderOptions = field(default_factory=RenderOptions)
    camera: Camera = field(default_factory=Camera)
    light_dir: tuple = (0.0, -1.0, 0.0)
    baking: bool = False
    bake_tile: tuple = (0, 0, 32)
    blackboard: dict = field(default_factory=dict)

    def get_num_irradiance_probes_xz(self):
        return self.grid.config.probes_x

    def get_num_irradiance_probes_y(self):
        return self.grid.config.probes_y

    def get_render_options(self):
        return self.options

    def get_camera(self):
        return self.camera

    def is_baking(self):
        return self.baking

    def set_blackboard_value_int(self, key, value):
        self.blackboard[key] = int(value)

    def blackboard_value_int(self, key):
        return self.blackboard[key]

    def scene_frame(self, layer_begin, layer_count, rotation16):
        """Scene UBO subset + push constants (VulkanRenderer.cpp:1201-1243)."""
        f = Frame()
        f.camera_pos[:] = (*self.camera.position, 1.0)
        f.light_dir[:] = (*self.light_dir, 0.0)
        f.sky_color[:] = (*self.options.sky_color, 1.0)
        f.sun_intensity = self.options.sun_intensity
        f.sky_intensity = self.options.sky_intensity
        flags = 0
        if self.options.ddgi_enabled:
            flags |= FLAG_DDGI
        if self.options.multi_bounce_ddgi_enabled:
            flags |= FLAG_DDGI_MULTI
        if self.baking:
            flags |= FLAG_BAKE
        f.flags = flags
        f.layer_begin = layer_begin
        f.layer_count = layer_count
        f.rotation[:] = [float(x) for x in rotation16]
        f.bake_tile[:] = (*self.bake_tile, 0)
        return f


class RenderPass:
    """passes/RenderPass.h:14-35."""

    def init(self, rc):
        return True

    def register_to_graph(self, fgb, rc):
        raise NotImplementedError

    def cleanup(self, rc):
        pass


class FrameGraphBuilder:
    """The subset of FrameGraphBuilder (FrameGraphBuilder.h:150-185) the probe
    passes use: named exe callbacks run in registration (submission) order."""

    def __init__(self):
        self._order = []
        self._exes = {}

    def register_render_pass(self, name, group="IrradianceProbe"):
        if name in self._exes:
            raise ValueError(f"render pass {name} registered twice")
        self._order.append(name)
        self._exes[name] = None

    def register_render_pass_exe(self, name, fn):
        if name not in self._exes:
            raise KeyError(f"no render pass named {name}")   # FrameGraphBuilder.cpp:184-187
        self._exes[name] = fn

    def execute_graph(self, rc):
        for name in self._order:
            fn = self._exes[name]
            if fn is not None:
                fn(rc)

    @property
    def order(self):
        return list(self._order)


class IrradianceProbeTranslationPass(RenderPass):
    def register_to_graph(self, fgb, rc):
        fgb.register_render_pass("IrradianceProbeTrans")

        def exe(rc):
            if not rc.get_render_options().ddgi_enabled:
                return
            if rc.is_baking():
                return
            # camera cell delta and the shifted copies: IrradianceProbeTranslationPass.cpp:60-186
            rc.grid.translate(rc.get_camera().get_position())
        fgb.register_render_pass_exe("IrradianceProbeTrans", exe)


class IrradianceProbeRayTracingPass(RenderPass):
    def __init__(self, seed=0, full_grid=False):
        self._current_layer = 0
        self._rng = PCG32(seed, [31337])
        self.full_grid = full_grid

    def random_rotation(self):
        """angle U[0, 360) degrees about x, y or z (IrradianceProbeRayTracingPass.cpp:114-133)."""
        angle = float(self._rng.uniform(0.0, 360.0)[0])
        axis = int(self._rng.next_u32()[0] % 3)
        return rotation_matrix(angle, axis)

    def register_to_graph(self, fgb, rc):
        fgb.register_render_pass("IrradianceProbeRT")
        ny = rc.get_num_irradiance_probes_y()
        self._current_layer = 0

        def exe(rc):
            opts = rc.get_render_options()
            if not opts.raytracing_enabled or not opts.ddgi_enabled:
                return
            layer = self._current_layer
            begin, count = (0, ny) if self.full_grid else (layer, 1)
            frame = rc.scene_frame(begin, count, self.random_rotation())
            rc.grid.trace(rc.scene, frame)
            rc.set_blackboard_value_int("ProbeLayer", layer)
            rc.blackboard["ProbeFrame"] = frame
            self._current_layer = 0 if layer + 1 >= ny else layer + 1
        fgb.register_render_pass_exe("IrradianceProbeRT", exe)


class IrradianceProbeConvolvePass(RenderPass):
    def register_to_graph(self, fgb, rc):
        fgb.register_render_pass("IrradianceProbeConvolve")

        def exe(rc):
            opts = rc.get_render_options()
            if not opts.raytracing_enabled or not opts.ddgi_enabled:
                return
            rc.blackboard_value_int("ProbeLayer")      # IrradianceProbeConvolvePass.cpp:125
            rc.grid.convolve(rc.blackboard["ProbeFrame"])
        fgb.register_render_pass_exe("IrradianceProbeConvolve", exe)


def ddgi_frame_graph(rc, seed=0, full_grid=False):
    """Register the three probe passes in the reference's submission order."""
    fgb = FrameGraphBuilder()
    passes = [IrradianceProbeTranslationPass(), IrradianceProbeRayTracingPass(seed, full_grid),
              IrradianceProbeConvolvePass()]
    for p in passes:
        p.init(rc)
        p.register_to_graph(fgb, rc)
    return fgb, passes
