#!/usr/bin/env python3
"""DDGI probe-update benchmark on one MI355X.

Workload (default): 32 x 16 x 32 = 16,384 probes x 512 rays per probe over a
seeded ~1M-triangle scene (512 x 512 value-noise terrain + 10,000 rotated and
scaled box / icosphere instances), LBVH built on the GPU.  One step = one full
probe update: ray directions + weights, primary traversal, shadow traversal,
shading (direct + indirect from the previous atlas + emissive), convolution and
hysteresis blend of all 16,384 probes.  Inputs are resident in HBM before the
timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp32|bf16|fp8]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
NODE_BYTES, TRI_BYTES, HIT_BYTES = 64, 48, 16


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "fp8"],
                   help="convolution precision of the headline update (fp32: the plain path)")
    p.add_argument("--workload", default="terrain1m",
                   choices=["terrain1m", "room8", "dynamic4m", "serve64"])
    p.add_argument("--requests", type=int, default=2000, help="serve64: open-loop requests")
    p.add_argument("--burst", type=int, default=256, help="serve64: requests of the saturation burst")
    p.add_argument("--inflight", type=int, default=8, help="serve64: requests on the device at once")
    p.add_argument("--no-graphs", action="store_true", help="serve64: direct launches instead of captured graphs")
    p.add_argument("--streams", type=int, default=4, help="serve64: HIP streams of the server")
    p.add_argument("--no-fuse", action="store_true",
                   help="serve64: launch every request on its own instead of fusing the per-probe stages")
    p.add_argument("--frames", type=int, default=300, help="frames of the dynamic4m workload")
    p.add_argument("--rebuild-every", type=int, default=4,
                   help="dynamic4m: full LBVH rebuild every K frames, refit in between (0 = refit only)")
    p.add_argument("--node-format", default="auto", choices=["auto", "full", "quantized"],
                   help="traversal node layout (auto: quantized from 2M triangles up)")
    p.add_argument("--cpu-runs", type=int, default=5,
                   help="CPU oracle baseline: timed whole updates (after one warm-up); the median is reported")
    p.add_argument("--legs", default="dynamic,serve",
                   help="extra legs of the default line: dynamic (compact 4M-triangle run), serve "
                        "(64-scene request loop); empty for none")
    p.add_argument("--leg-frames", type=int, default=60, help="frames of the compact dynamic leg")
    p.add_argument("--leg-requests", type=int, default=600, help="open-loop requests of the serving leg")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-conv-study", action="store_true",
                   help="skip the FP32/BF16/FP8 convolution timing and FP64 error study")
    return p.parse_args()


def dynamic_cpu_baseline(scene, vel, cfg, frames=1):
    """The oracle on a bounded sample of the dynamic workload: ``frames`` whole frames
    (instance refit + full FP32 probe update of all 131,072 probes), timed on this
    host's cores.  The per-frame figure extrapolates to the 300-frame run."""
    from anerend_amd import scenegen
    from oracle import oracle as O
    t0 = time.time()
    osc = O.OracleScene(scene)
    build_s = time.time() - t0
    atlas = O.new_atlas(cfg)
    t0 = time.time()
    for f in range(frames):
        osc.refit(scenegen.dynamic_instances(scene, vel, f))
        osc.update(cfg, scenegen.dynamic_frame(5, f, cfg), atlas, want_rays=False)
    dt = time.time() - t0
    return {"value": round(dt * 1000.0 / frames, 1), "unit": "ms/frame", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"{frames} frame(s) of the same workload (refit + full probe update, "
                      f"{frames * cfg.probes_x * cfg.probes_y * cfg.probes_z * cfg.rays_per_probe} rays) in "
                      f"{dt:.2f} s, extrapolated per frame; CPU LBVH build {build_s:.2f} s not counted"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, cfg, frame_at, runs):
    """The oracle (C, OpenMP over rays and probes) on the same workload: one warm-up
    update, then ``runs`` whole probe updates, each timed per stage (trace + shadow
    + shade, convolve + blend); the medians are reported."""
    from oracle import oracle as O
    t0 = time.time()
    osc = O.OracleScene(scene)
    build_s = time.time() - t0
    atlas = O.new_atlas(cfg)
    per = cfg.probes_x * cfg.probes_y * cfg.probes_z * cfg.rays_per_probe
    osc.update(cfg, frame_at(0), atlas, want_rays=False)
    wall, stages = [], []
    for i in range(runs):
        t0 = time.time()
        osc.update(cfg, frame_at(1 + i), atlas, want_rays=False)
        wall.append(time.time() - t0)
        stages.append(O.last_update_stage_ms())
    med = float(np.median(wall))
    st = np.median(np.array(stages), axis=0)
    return {"value": per / med, "unit": "probe rays/s", "cores": O.num_threads(), "kind": "port",
            "cpu_model": cpu_model(), "ms_per_update": round(med * 1000.0, 1),
            "stages_ms": {"trace_shadow_shade": round(float(st[0]), 1), "convolve_blend": round(float(st[1]), 1)},
            "sample": f"median of {runs} full probe updates of the same workload ({per} rays each: trace, "
                      f"shadow, shade, convolve, blend) after one warm-up; CPU LBVH build {build_s:.2f} s "
                      f"(single thread) not counted"}


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "trace_terrain1m_traffic.md")


def profiled_traffic():
    """HBM bytes per launch of the trace kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the 1 GiB copy calibration), or
    None when there is no summary."""
    import re
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", open(TRAFFIC_SUMMARY).read(), re.M)
    return int(m.group(1)) if m else None


def copy_calibration(nbytes):
    """One device-to-device copy of ``nbytes`` through the HIP runtime (a copy
    kernel of known traffic for scaling the profiler's byte counters)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) != 0:
        return
    try:
        if hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) != 0:
            return
        try:
            hip.hipMemset(a, 0, ctypes.c_size_t(nbytes))
            hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
            hip.hipDeviceSynchronize()
        finally:
            hip.hipFree(b)
    finally:
        hip.hipFree(a)


def conv_study(cfg0, frame, reps=20):
    """Convolution alone at the production shape (16,384 probes x 512 rays x RGB) on
    seeded radiance, log-uniform on [1e-3, 50] (seed 3): FP32 / BF16 / FP16 / FP8
    (MFMA) kernel times and their error against a float64 GEMM of the same
    weights, evaluated and reduced in float64 on the GPU.  Hysteresis 0, so the
    atlas interior is the convolution itself."""
    from anerend_amd import native
    from anerend_amd.types import Config, PRECISION_BF16, PRECISION_FP8, PRECISION_FP16, PRECISION_FP32
    P = cfg0.probes_x * cfg0.probes_y * cfg0.probes_z
    N = cfg0.rays_per_probe
    rng = np.random.default_rng(3)
    rad = np.exp(rng.uniform(np.log(1e-3), np.log(50.0), (P, 3, N))).astype(np.float32)
    fr = frame.copy()
    fr.layer_begin, fr.layer_count = 0, cfg0.probes_y
    nx, ny, nz = cfg0.probes_x, cfg0.probes_y, cfg0.probes_z
    # the FP32 radiance the errors are measured against (float64 on the GPU)
    ref = native.ProbeGrid(Config.make((nx, ny, nz), step=tuple(cfg0.probe_step), rays=N, hysteresis=0.0))
    ref.set_radiance(fr, rad)
    flops = 2.0 * 64 * N * 3 * P
    peaks = {"fp32": 157.3e12, "bf16": 2.5e15, "fp16": 2.5e15, "fp8": 5.0e15}
    rad_bytes = {"fp32": 4, "bf16": 2, "fp16": 2, "fp8": 1}
    out = {}
    for name, prec in (("fp32", PRECISION_FP32), ("bf16", PRECISION_BF16), ("fp16", PRECISION_FP16),
                       ("fp8", PRECISION_FP8)):
        cfg = Config.make((nx, ny, nz), step=tuple(cfg0.probe_step), rays=N, hysteresis=0.0,
                          precision=prec)
        grid = native.ProbeGrid(cfg)
        grid.set_radiance(fr, rad)
        grid.convolve(fr)
        mx, rms = grid.conv_error_fp64(ref, rel_floor=1e-3)  # radiance >= 1e-3: the floor is inert
        native.synchronize()
        grid.reset_stats()
        for _ in range(reps):
            grid.convolve(fr)
        native.synchronize()
        ms = grid.stats()["mean_ms"]["convolve"]
        hbm = P * 3 * N * rad_bytes[name] + 2 * P * 64 * 16
        out[name] = {"ms": round(ms, 5), "max_rel_err_vs_fp64": float(f"{mx:.3e}"),
                     "rms_rel_err_vs_fp64": float(f"{rms:.3e}"),
                     "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                     "compute_frac": round(flops / (ms * 1e-3) / peaks[name], 4),
                     "hbm_gbs": round(hbm / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(hbm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        grid.close()
        log(f"[bench] conv {name}: {json.dumps(out[name])}")
    out["bf16_speedup_vs_fp32"] = round(out["fp32"]["ms"] / out["bf16"]["ms"], 3)
    out["fp16_speedup_vs_fp32"] = round(out["fp32"]["ms"] / out["fp16"]["ms"], 3)
    out["fp8_speedup_vs_fp32"] = round(out["fp32"]["ms"] / out["fp8"]["ms"], 3)
    out["shape"] = f"{P} probes x {N} rays x RGB, GEMM 64 x {N} x {3 * P}"
    return out


def gi_lookup_study(grid, frame, width=1920, height=1080, reps=20):
    """Deferred GI lookup of a 1920x1080 G-buffer (seeded texels within 24 m of the
    camera, so that both the probe and the far-field branch run) on the current
    atlas; kernel time from HIP events over ``reps`` launches."""
    import ctypes
    from anerend_amd import native
    from anerend_amd.types import GBUFFER_DTYPE
    n = width * height
    rng = np.random.default_rng(6)
    g = np.zeros(n, GBUFFER_DTYPE)
    cam = np.array(frame.camera_pos[:3], np.float32)
    g["pos"] = cam + rng.uniform(-24, 24, (n, 3)).astype(np.float32) * np.float32([1, 0.3, 1])
    nrm = rng.normal(size=(n, 3))
    g["normal"] = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    g["albedo"] = rng.uniform(0, 1, (n, 3))
    g["roughness"] = rng.uniform(0.05, 1, n)
    g["shadow"] = rng.uniform(0, 1, n)
    src, dst = native.DeviceBuffer(g.nbytes), native.DeviceBuffer(16 * n)
    hip = native.DeviceBuffer._hip
    L = native.load()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    try:
        src.upload(g)
        native.check(L.ane_ddgi_gi_lookup(grid.handle, ctypes.byref(frame), src.ptr, n, dst.ptr, None))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            native.check(L.ane_ddgi_gi_lookup(grid.handle, ctypes.byref(frame), src.ptr, n, dst.ptr, None))
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        t = ms.value / reps
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        src.close()
        dst.close()
    algo = n * (48 + 16)
    return {"texels": n, "ms": round(t, 5), "texels_per_s": round(n / (t * 1e-3), 1),
            "gbs": round(algo / (t * 1e-3) / 1e9, 1), "hbm_frac": round(algo / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_texel": "48 B G-buffer in + 16 B out (atlas taps served from cache)"}


def view_proj(eye, target, fovy_deg=60.0, aspect=16 / 9, near=0.1, far=500.0):
    """Column-major Vulkan-style (depth 0..1) view-projection and its inverse, float64."""
    eye, target = np.asarray(eye, float), np.asarray(target, float)
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, [0.0, 1.0, 0.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    V = np.eye(4)
    V[0, :3], V[1, :3], V[2, :3] = s, u, -f
    V[:3, 3] = -V[:3, :3] @ eye
    t = 1.0 / np.tan(np.radians(fovy_deg) / 2)
    P = np.zeros((4, 4))
    P[0, 0], P[1, 1] = t / aspect, t
    P[2, 2], P[2, 3], P[3, 2] = far / (near - far), -far * near / (far - near), -1.0
    VP = P @ V
    return VP, np.linalg.inv(VP)


def camera_gbuffer(ds, width, height):
    """Depth buffer and normals of a width x height view of the scene from
    (0.5, 20, 0.5) toward (60, 0, 60), from camera rays cast through the library.
    Normals are the depth buffer's world-position gradients, facing the camera."""
    eye = np.array([0.5, 20.0, 0.5])
    VP, IVP = view_proj(eye, (60.0, 0.0, 60.0))
    y, x = np.mgrid[0:height, 0:width]
    uv = np.stack([(2 * x + 1) / (2 * width), (2 * y + 1) / (2 * height)], -1).reshape(-1, 2)
    clip = np.concatenate([uv * 2 - 1, np.ones((len(uv), 1)), np.ones((len(uv), 1))], 1)
    far = clip @ IVP.T
    far = far[:, :3] / far[:, 3:]
    d = far - eye
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    org = np.tile(eye.astype(np.float32), (len(d), 1))
    hits = ds.trace_rays(org, d, tmin=0.0, tmax=490.0)
    hit = hits["tri"] >= 0
    pos = eye + d.astype(float) * hits["t"][:, None].astype(float)
    ndc = np.concatenate([pos, np.ones((len(pos), 1))], 1) @ VP.T
    with np.errstate(divide="ignore", invalid="ignore"):   # misses: w is 0 or inf, masked below
        depth = np.where(hit, ndc[:, 2] / ndc[:, 3], 1.0).astype(np.float32).reshape(height, width)
    p = pos.reshape(height, width, 3)
    gx = np.gradient(p, axis=1)
    gy = np.gradient(p, axis=0)
    nrm = np.cross(gy, gx)
    nrm *= np.where(np.sum(nrm * (eye - p), -1) < 0, -1.0, 1.0)[..., None]
    ln = np.linalg.norm(nrm, axis=-1, keepdims=True)
    nrm = np.where(ln > 1e-12, nrm / np.maximum(ln, 1e-12), np.array([0.0, 1.0, 0.0]))
    normal4 = np.zeros((height, width, 4), np.float32)
    normal4[..., :3] = nrm
    return depth, normal4, IVP.T.ravel().astype(np.float32)


def shadow_mask_study(ds, frame, width=1920, height=1080, reps=10):
    """RT hard shadows (shadow.rgen) of a 1920x1080 view of the scene: the depth
    buffer comes from camera rays cast through the library, then the shadow mask
    is timed with HIP events over ``reps`` launches.  The work counters come from
    a traversal call; the timed calls go through the scene's sun grid."""
    import ctypes
    from anerend_amd import native
    depth, _, ivp16 = camera_gbuffer(ds, width, height)
    light = np.array(frame.light_dir[:3], np.float32)
    _, ctr = ds.shadow_mask(depth, ivp16, light, count=True)
    n = width * height
    src, dst = native.DeviceBuffer(4 * n), native.DeviceBuffer(4 * n)
    hip = native.DeviceBuffer._hip
    L = native.load()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    try:
        src.upload(depth)
        lp = light.ctypes.data_as(ctypes.c_void_p)
        ip = ivp16.ctypes.data_as(ctypes.c_void_p)
        # two untimed calls: the second with this scene and light builds its sun grid
        for _ in range(2):
            native.check(L.ane_scene_shadow_mask(ds.handle, src.ptr, width, height, ip, lp, dst.ptr, None, None))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            native.check(L.ane_scene_shadow_mask(ds.handle, src.ptr, width, height, ip, lp, dst.ptr, None, None))
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        t = ms.value / reps
        vis = dst.download(np.zeros((height, width), np.float32))
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        src.close()
        dst.close()
    rays = int((depth != 1.0).sum())
    return {"pixels": n, "shadow_rays": rays, "ms": round(t, 5), "rays_per_s": round(rays / (t * 1e-3), 1),
            "shadowed_fraction": round(float((vis == 0).mean()), 4), "path": "sun grid",
            "traversal_node_visits": int(ctr[0]), "traversal_triangle_tests": int(ctr[1])}


def mirror_study(grid, ds, frame, width=1920, height=1080, reps=10):
    """Mirror reflections (mirror_reflections.rgen/.rchit) of a 1920x1080 view of
    the scene with the probe atlas as indirect light, timed with HIP events over
    ``reps`` calls (reflection rays, their shadow rays through the context's sun
    grid, the hit shading).  The counters come from a traversal call."""
    import ctypes
    from anerend_amd import native
    depth, normal4, ivp16 = camera_gbuffer(ds, width, height)
    out, ctr = grid.mirror_reflections(ds, frame, depth, normal4, ivp16, count=True)
    n = width * height
    bufs = [native.DeviceBuffer(4 * n), native.DeviceBuffer(16 * n), native.DeviceBuffer(16 * n)]
    hip = native.DeviceBuffer._hip
    L = native.load()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    try:
        bufs[0].upload(depth)
        bufs[1].upload(normal4)
        ip = ivp16.ctypes.data_as(ctypes.c_void_p)
        args = (grid.handle, ds.handle, ctypes.byref(frame), bufs[0].ptr, bufs[1].ptr, width, height, ip,
                bufs[2].ptr, None, None)
        for _ in range(2):  # untimed; the context's sun grid is current after these
            native.check(L.ane_ddgi_mirror_reflections(*args))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            native.check(L.ane_ddgi_mirror_reflections(*args))
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        t = ms.value / reps
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        for b in bufs:
            b.close()
    rays = int((depth != 1.0).sum())
    hit = (depth != 1.0) & (out[..., 3] < 1e6)
    return {"pixels": n, "reflection_rays": rays, "reflection_hits": int(hit.sum()), "ms": round(t, 5),
            "rays_per_s": round(rays / (t * 1e-3), 1), "node_visits": int(ctr[0]), "triangle_tests": int(ctr[1]),
            "shadow_path": "sun grid", "shadow_traversal_node_visits": int(ctr[2]),
            "shadow_traversal_triangle_tests": int(ctr[3])}


def light_shadow_study(ds, num_lights=1024, reps=10, seed=21):
    """Point-light shadow moments (light_shadow.rgen + light_shadow_sum.comp) of
    ``num_lights`` seeded lights (the reference's full light budget) over the
    scene: positions within +-200 m in x, z and 1-12 m above the terrain's base,
    ranges U[5, 30] m.  Timed with HIP events over ``reps`` calls (upload of the
    light records, 1024 rays per light, distances, 8x8 moments per light)."""
    import ctypes
    from anerend_amd import native
    rng = np.random.default_rng(seed)
    n = num_lights
    L = np.zeros((n, 8), np.float32)
    L[:, 0] = rng.uniform(-200.0, 200.0, n)
    L[:, 1] = rng.uniform(1.0, 12.0, n)
    L[:, 2] = rng.uniform(-200.0, 200.0, n)
    L[:, 3] = rng.uniform(5.0, 30.0, n)
    L[:, 4:7] = 1.0
    L[:, 7] = 1.0
    dist, _, ctr = ds.light_shadows(L, count=True)
    dd, dm = native.DeviceBuffer(4 * 32 * 32 * n), native.DeviceBuffer(4 * 8 * 10 * n * 2)
    hip = native.DeviceBuffer._hip
    lib = native.load()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    try:
        lp = L.ctypes.data_as(ctypes.c_void_p)
        native.check(lib.ane_scene_light_shadows(ds.handle, lp, n, dd.ptr, dm.ptr, None, None))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            native.check(lib.ane_scene_light_shadows(ds.handle, lp, n, dd.ptr, dm.ptr, None, None))
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        t = ms.value / reps
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        dd.close()
        dm.close()
    rays = n * 1024
    return {"lights": n, "rays": rays, "ms": round(t, 5), "rays_per_s": round(rays / (t * 1e-3), 1),
            "hit_fraction": round(float((dist < L[:, 3].repeat(32)[None, :]).mean()), 4),
            "node_visits": int(ctr[0]), "triangle_tests": int(ctr[1])}


def fp64_conv_error(grid_conv, grid_ref, scene_dev, fr):
    """Error of one convolution of ``grid_conv`` (its precision) against float64
    over the FP32 radiance of the same frame traced by ``grid_ref`` from the same
    atlas, evaluated and reduced in float64 on the GPU (ane_ddgi_conv_error_fp64)."""
    grid_ref.atlas = grid_conv.atlas
    grid_ref.trace(scene_dev, fr)
    grid_conv.update(scene_dev, fr)
    mx, rms = grid_conv.conv_error_fp64(grid_ref, rel_floor=1e-3)
    return {"max_rel": float(f"{mx:.3e}"), "rms_rel": float(f"{rms:.3e}"), "floor": 1e-3,
            "reduction": "float64 on the GPU"}


def run_dynamic(frames, warmup, rebuild_every, node_format, with_cpu):
    """Dynamic large scene: 64 x 32 x 64 probes x 256 rays over ~4M triangles whose
    20,000 instances move every frame (refit, a rebuild every ``rebuild_every``),
    a camera crossing an integer cell every 3 frames (translation), one full probe
    update per frame with FP8 convolution.  Returns its figures as a dict."""
    from anerend_amd import native, scenegen
    from anerend_amd.types import PRECISION_FP32
    cfg = scenegen.dynamic_config()
    log("[bench] building the dynamic scene")
    scene, vel = scenegen.dynamic_scene()
    insts = [scenegen.dynamic_instances(scene, vel, f) for f in range(frames + 1)]
    frs = [scenegen.dynamic_frame(5, f, cfg) for f in range(frames + 1)]
    ds = native.DeviceScene(scene)
    ds.set_node_format({"auto": 0, "full": 1, "quantized": 2}[node_format])
    first_build_ms = ds.info()["build_ms"]
    # warm: a rebuild and a refit of frame 0's instances (the same hierarchy and boxes)
    ds.rebuild(insts[0])
    info = ds.info()
    ds.refit(insts[0])
    warm_refit_ms = ds.info()["build_ms"]
    log(f"[bench] {scene.num_tris} triangles, LBVH build {info['build_ms']:.3f} ms (first {first_build_ms:.2f}), "
        f"refit {warm_refit_ms:.3f} ms, depth {info['bvh_depth']}")
    grid = native.ProbeGrid(cfg)
    for f in range(warmup):  # warm-up frames outside the timed run, then restart the path
        grid.update(ds, frs[f])
    native.synchronize()
    grid.translation_state((0, 0))
    grid.reset_stats()
    moves = 0
    log(f"[bench] timing {frames} frames")
    t0 = time.perf_counter()
    K = rebuild_every
    for f in range(frames):
        if K > 0 and f % K == 0:
            ds.rebuild(insts[f])
        else:
            ds.refit(insts[f])
        moves += bool(grid.translate(frs[f].camera_pos[:3]))
        grid.update(ds, frs[f])
    native.synchronize()
    t1 = time.perf_counter()
    refit_ms = ds.info()["build_ms"]
    st = grid.stats()
    ms = (t1 - t0) * 1000.0 / frames
    copy_calibration(1 << 30)   # scale for the profiler's byte counters, outside the timed run
    log(f"[bench] {ms:.3f} ms/frame, {moves} translations, stages "
        f"{json.dumps({k: round(v, 4) for k, v in st['mean_ms'].items()})}")
    # FP8 convolution error against float64 on the frame after the run (GPU reduction)
    ds.refit(insts[frames])
    ref_grid = native.ProbeGrid(scenegen.dynamic_config(PRECISION_FP32))
    err = fp64_conv_error(grid, ref_grid, ds, frs[frames])
    ref_grid.close()
    log(f"[bench] FP8 error vs FP64: {err}")
    cpu = None
    if with_cpu:
        log("[bench] CPU oracle on one frame of the same workload")
        cpu = dynamic_cpu_baseline(scene, vel, scenegen.dynamic_config(PRECISION_FP32))
        log(f"[bench] cpu baseline {json.dumps(cpu)}")
    P = cfg.probes_x * cfg.probes_y * cfg.probes_z
    out = {"ms_per_frame": round(ms, 4), "frames": frames, "warmup": warmup,
           "probe_rays_per_s": round(P * cfg.rays_per_probe / (ms / 1000.0), 1),
           "probes": P, "rays_per_probe": cfg.rays_per_probe, "triangles": scene.num_tris,
           "conv_precision": "fp8",
           "bvh_nodes": "quantized 32 B" if ds.info()["node_format"] == 2 else "full 64 B",
           "translations": moves,
           "bvh_update": (f"rebuild every {K} frames, refit in between" if K > 1 else
                          "rebuild every frame" if K == 1 else "refit every frame"),
           "stages_ms": {k: round(v, 4) for k, v in st["mean_ms"].items()},
           "last_bvh_update_ms": round(refit_ms, 4),
           "lbvh_build_ms": round(info["build_ms"], 3),
           "lbvh_refit_ms": round(warm_refit_ms, 3),
           "fp8_error_vs_fp64": err,
           "cpu_baseline": cpu}
    grid.close()
    ds.close()
    return out


def bench_dynamic(args):
    """--workload dynamic4m: the dynamic scene over ``--frames`` frames, one JSON line
    (ms per frame)."""
    d = run_dynamic(args.frames, args.warmup, args.rebuild_every, args.node_format,
                    not args.no_cpu_baseline)
    out = {
        "metric": "ms per frame (dynamic 4M-triangle scene: refit + translation + full probe update)",
        "value": d["ms_per_frame"], "unit": "ms/frame", "n_gpus": 1, "steps": d["frames"],
        "warmup": d["warmup"], "ms_per_step": d["ms_per_frame"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+fp8",
        "data": "synthetic (seeded scene generator, seed 5)",
        "config": {"workload": "dynamic4m: 64x32x64 probes x 256 rays, 600^2 terrain + 20k moving "
                               "box/icosphere instances, camera +1/3 m per frame",
                   "rebuild_every": args.rebuild_every, "probes": d["probes"],
                   "rays_per_probe": d["rays_per_probe"], "triangles": d["triangles"],
                   "conv_precision": "fp8", "bvh_nodes": d["bvh_nodes"],
                   "probe_rays_per_s": d["probe_rays_per_s"]},
        "translations": d["translations"], "bvh_update": d["bvh_update"], "stages_ms": d["stages_ms"],
        "last_bvh_update_ms": d["last_bvh_update_ms"], "fp8_error_vs_fp64": d["fp8_error_vs_fp64"],
        "cpu_baseline": d["cpu_baseline"],
    }
    print(json.dumps(out), flush=True)


def run_serve(requests, burst, inflight, streams, graphs, fuse):
    """Serving: 64 resident scenes of uneven sizes (10k-2M triangles, grids 4^3 to
    32x16x32, 128-512 rays), BF16 convolution, requests arriving as a Poisson
    process at 200/s (open loop), admitted continuously under a 200 GB budget.
    Returns p50/p99 latency and aggregate rays/s, then the throughput of a burst
    of requests through the server against the plain per-scene FP32 path."""
    from anerend_amd import native, scenegen, serving
    from anerend_amd.types import PRECISION_BF16, PRECISION_FP32
    specs = scenegen.serving_specs()
    log(f"[bench] building {len(specs)} scenes")
    dss, tris = [], 0
    for k, sp in enumerate(specs):
        sc = scenegen.serving_scene(4, k, sp)
        tris += sc.num_tris
        dss.append(native.DeviceScene(sc))
    cfgs = [scenegen.serving_config(sp, PRECISION_BF16) for sp in specs]
    budget = 200 << 30
    srv = serving.Server(hbm_budget_bytes=budget, max_inflight=inflight, num_streams=streams,
                         capture_graphs=graphs and not fuse, fuse_stages=fuse)
    sids = [srv.add_session(d, c) for d, c in zip(dss, cfgs)]
    admitted = srv.stats()
    log(f"[bench] {tris} triangles resident, {admitted['used_bytes'] / 2**30:.2f} GiB admitted")
    R, rate = requests, 200.0
    rng = np.random.default_rng(4)
    arrive = np.cumsum(rng.exponential(1.0 / rate, R))
    which = rng.integers(0, len(specs), R)
    frames = [scenegen.serving_frame(4, int(k), i, cfgs[k]) for i, k in enumerate(which)]
    rays = np.array([np.prod(specs[k][3]) * specs[k][4] for k in which], np.int64)
    # warm-up: one update per session, outside the measurement
    for t in [srv.submit(sid, scenegen.serving_frame(4, k, 10 ** 6, cfgs[k]))
              for k, sid in enumerate(sids)]:
        srv.poll(t, -1)
    log(f"[bench] open loop: {R} requests at {rate:.0f}/s")
    tickets = []
    t0 = time.perf_counter()
    for i in range(R):
        target = t0 + arrive[i]
        while True:
            dt = target - time.perf_counter()
            if dt <= 0:
                break
            time.sleep(dt if dt > 0.002 else 0.0)
        tickets.append(srv.submit(sids[which[i]], frames[i]))
    res = [srv.poll(t, -1) for t in tickets]
    t1 = time.perf_counter()
    assert all(r["status"] == 0 for r in res)
    lat = np.array([r["latency_ms"] for r in res])
    que = np.array([r["queue_ms"] for r in res])
    st = srv.stats()
    open_loop = {"requests": R, "rate_per_s": rate, "p50_ms": round(float(np.percentile(lat, 50)), 4),
                 "p99_ms": round(float(np.percentile(lat, 99)), 4),
                 "mean_ms": round(float(lat.mean()), 4), "max_ms": round(float(lat.max()), 4),
                 "queue_p99_ms": round(float(np.percentile(que, 99)), 4),
                 "rays_per_s": round(float(rays.sum()) / (t1 - t0), 1),
                 "peak_inflight": st["peak_inflight"], "wall_s": round(t1 - t0, 3)}
    log(f"[bench] open loop {json.dumps(open_loop)}")
    # saturation: a burst of B requests at once through the server (BF16) ...
    B = min(burst, R)
    native.synchronize()
    t0 = time.perf_counter()
    tk = [srv.submit(sids[which[i]], frames[i]) for i in range(B)]
    for t in tk:
        srv.poll(t, -1)
    t1 = time.perf_counter()
    serve_rays_s = float(rays[:B].sum()) / (t1 - t0)
    held = srv.stats()
    srv.close()
    # ... and the same requests through the plain per-scene FP32 path, one stream
    plain = {}
    for k in sorted(set(int(x) for x in which[:B])):
        plain[k] = native.ProbeGrid(scenegen.serving_config(specs[k], PRECISION_FP32))
        plain[k].update(dss[k], scenegen.serving_frame(4, k, 10 ** 6, cfgs[k]))
    native.synchronize()
    t0 = time.perf_counter()
    for i in range(B):
        k = int(which[i])
        plain[k].update(dss[k], frames[i])
    native.synchronize()
    t1 = time.perf_counter()
    plain_rays_s = float(rays[:B].sum()) / (t1 - t0)
    for g in plain.values():
        g.close()
    for d in dss:
        d.close()
    burst_d = {"requests": B, "serve_bf16_rays_per_s": round(serve_rays_s, 1),
               "plain_fp32_rays_per_s": round(plain_rays_s, 1),
               "speedup": round(serve_rays_s / plain_rays_s, 3)}
    log(f"[bench] burst {json.dumps(burst_d)}")
    return {"open_loop": open_loop, "burst": burst_d, "triangles": tris, "scenes": len(specs),
            "hbm_budget_bytes": budget, "admitted_bytes": admitted["used_bytes"],
            "held_bytes_after": held["held_bytes"], "max_inflight": inflight, "streams": streams,
            "capture_graphs": graphs and not fuse, "fused_stages": fuse}


def bench_serve(args):
    """--workload serve64: the request loop, one JSON line (p99 latency)."""
    d = run_serve(args.requests, args.burst, args.inflight, args.streams, not args.no_graphs, not args.no_fuse)
    ol = d["open_loop"]
    out = {
        "metric": "serving p99 latency (ms), Poisson 200 req/s over 64 resident scenes",
        "value": ol["p99_ms"], "unit": "ms", "n_gpus": 1, "steps": ol["requests"], "warmup": d["scenes"],
        "ms_per_step": ol["mean_ms"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic (seeded scene generator, seed 4)",
        "config": {"workload": "serve64: 64 scenes, 10k-2M triangles log-uniform, grids "
                               "{4^3, 8^3, 16x8x16, 32x16x32} x rays {128, 256, 512}, BF16 conv",
                   "triangles": d["triangles"], "hbm_budget_bytes": d["hbm_budget_bytes"],
                   "max_inflight": d["max_inflight"], "streams": d["streams"],
                   "capture_graphs": d["capture_graphs"], "fused_stages": d["fused_stages"]},
        "open_loop": ol,
        "burst": d["burst"],
        "admitted_bytes": d["admitted_bytes"], "held_bytes_after": d["held_bytes_after"],
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.workload == "serve64":
        if args.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            sys.exit(2)
        bench_serve(args)
        return
    if args.workload == "dynamic4m":
        if args.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            sys.exit(2)
        bench_dynamic(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); the multi-GPU "
              "sharding of the probe grid is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    from anerend_amd import native, scenegen
    from anerend_amd.types import PRECISION_BF16, PRECISION_FP8, PRECISION_FP16, PRECISION_FP32, Config
    PRECS = {"fp32": PRECISION_FP32, "bf16": PRECISION_BF16, "fp16": PRECISION_FP16, "fp8": PRECISION_FP8}
    prec = PRECS[args.precision]

    log(f"[bench] building workload {args.workload}")
    scene, cfg, frame = scenegen.workload(args.workload)
    cfg.precision = prec
    log(f"[bench] scene: {scene.num_tris} triangles, {len(scene.instances)} instances")
    ds = native.DeviceScene(scene)
    ds.set_node_format({"auto": 0, "full": 1, "quantized": 2}[args.node_format])
    info = ds.info()
    first_build_ms = info["build_ms"]  # includes the first launch of every build kernel in the process
    # warm figures: a rebuild and a refit from the same instances (the same hierarchy and boxes)
    ds.rebuild(scene.instances)
    info = ds.info()
    ds.refit(scene.instances)
    refit_ms = ds.info()["build_ms"]
    log(f"[bench] LBVH on GPU: build {info['build_ms']:.3f} ms (first {first_build_ms:.2f}), "
        f"refit {refit_ms:.3f} ms, depth {info['bvh_depth']}")
    grid = native.ProbeGrid(cfg)
    P = cfg.probes_x * cfg.probes_y * cfg.probes_z
    N = cfg.rays_per_probe

    def frame_at(i):
        return scenegen.make_frame(2, camera=tuple(frame.camera_pos[:3]),
                                   config=cfg, iteration=i)

    # work counts (counting kernels, outside the timed region).  Mode 1: every ray
    # traverses the hierarchy (the counts the oracle reproduces).  Mode 2: the
    # update's own shadow path; its second update with this scene and sun builds
    # the sun grid, after which shadow rays count cell lookups and candidate tests.
    grid.set_counting(1)
    grid.update(ds, frame_at(0))
    work = grid.stats()
    grid.set_counting(2)
    grid.update(ds, frame_at(0))
    grid.update(ds, frame_at(0))
    gwork = grid.stats()
    grid.set_counting(0)
    sungrid_ms = gwork["mean_ms"]["sungrid"]   # the one build of this scene and sun
    log(f"[bench] work: primary nodes {work['primary_nodes']} tris {work['primary_tris']}, "
        f"shadow nodes {work['shadow_nodes']} tris {work['shadow_tris']}")
    log(f"[bench] sun grid: shadow-ray cell lookups {gwork['shadow_nodes']}, candidate tests "
        f"{gwork['shadow_tris']}, build {sungrid_ms:.3f} ms")

    def time_updates(g, frames, warmup):
        """Warm-up updates, then the timed ones; wall ms per update and the stage means."""
        for fr in frames[:warmup]:
            g.update(ds, fr)
        native.synchronize()
        g.reset_stats()
        t0 = time.perf_counter()
        for fr in frames[warmup:]:
            g.update(ds, fr)
        native.synchronize()
        t1 = time.perf_counter()
        return (t1 - t0) * 1000.0 / (len(frames) - warmup), g.stats()

    log(f"[bench] timing {args.steps} updates ({args.precision} convolution)")
    ms, st = time_updates(grid, [frame_at(1 + i) for i in range(args.warmup)] +
                          [frame_at(100 + i) for i in range(args.steps)], args.warmup)
    value = P * N / (ms / 1000.0)
    mean = st["mean_ms"]
    log(f"[bench] {ms:.3f} ms/update, stages {json.dumps({k: round(v, 4) for k, v in mean.items()})}")

    # the other convolution precisions in the same run, each with its error against
    # float64 on this scene's traced radiance (reduced on the GPU)
    def grid_of(name, h=cfg.hysteresis):
        return native.ProbeGrid(Config.make((cfg.probes_x, cfg.probes_y, cfg.probes_z), step=tuple(cfg.probe_step),
                                            rays=N, hysteresis=h, precision=PRECS[name]))
    ref32 = grid_of("fp32")
    variants = {}
    for name in ("fp32", "bf16", "fp16"):
        g = grid if name == args.precision else grid_of(name)
        if name != args.precision:
            vms, vst = time_updates(g, [frame_at(1 + i) for i in range(args.warmup)] +
                                    [frame_at(200 + i) for i in range(args.steps)], args.warmup)
            variants[name] = {"ms_per_update": round(vms, 4), "probe_rays_per_s": round(P * N / (vms / 1000.0), 1),
                              "stages_ms": {k: round(v, 4) for k, v in vst["mean_ms"].items()
                                            if k not in ("translate", "sungrid")}}
        else:
            variants[name] = {"ms_per_update": round(ms, 4), "probe_rays_per_s": round(value, 1), "headline": True}
        if name != "fp32":
            variants[name]["conv_error_vs_fp64_on_traced_radiance"] = fp64_conv_error(g, ref32, ds, frame_at(999))
        log(f"[bench] {name} update {json.dumps(variants[name])}")
        if g is not grid:
            g.close()
    ref32.close()
    # a sun that moves every update over the static scene: each update builds the grid of
    # its sun (the scene's grid was valid and there are many rays per triangle)
    def moving_sun_frame(i):
        return scenegen.make_frame(500 + i, camera=tuple(frame.camera_pos[:3]), config=cfg, iteration=i)
    ms_ms, st_ms = time_updates(grid, [moving_sun_frame(i) for i in range(args.warmup + args.steps)],
                                args.warmup)
    moving = {"ms_per_update": round(ms_ms, 4), "stages_ms": {k: round(v, 4) for k, v in st_ms["mean_ms"].items()
                                                             if k != "translate"},
              "sun_grid_builds": int(st_ms["calls"]["sungrid"])}
    log(f"[bench] moving sun {json.dumps(moving)}")

    # calibration kernels outside the timed region, for the profiler's FETCH_SIZE /
    # WRITE_SIZE counters: one camera-cell move (the translation remap reads and
    # writes the atlas once) and a 1 GiB device-to-device copy (the runtime's copy
    # kernel reads and writes exactly 1 GiB)
    atlas_bytes = cfg.atlas_shape[0] * cfg.atlas_shape[1] * 16
    cam = np.array(frame.camera_pos[:3], np.float32)
    grid.translate(cam + np.array([1.0, 0.0, 0.0], np.float32))
    native.synchronize()
    copy_calibration(1 << 30)

    # roofline of the dominant kernel (algorithmic bytes per launch / mean launch time)
    rays = P * N
    grid_shadows = gwork.get("sun_grid_state") == 1
    if grid_shadows:
        # trace: the primary rays (node and triangle records), a 16 B hit record out
        # per ray.  shadow: per ray the hit record in and the 1 B flag out; per hit
        # (one cell lookup each) the 64 B triangle record and 48 B instance transform
        # of the hit position and 8 B of cell bounds; per candidate test a 4 B list
        # entry and the 48 B triangle
        algo = {
            "trace": work["primary_nodes"] * NODE_BYTES + work["primary_tris"] * TRI_BYTES + rays * HIT_BYTES,
            "shadow": rays * (HIT_BYTES + 1) + gwork["shadow_nodes"] * (64 + 48 + 8)
                      + gwork["shadow_tris"] * (4 + TRI_BYTES),
        }
    else:
        # the shadow rays traverse in the primary rays' lanes: one launch
        algo = {
            "trace": (work["primary_nodes"] + gwork["shadow_nodes"]) * NODE_BYTES
                     + (work["primary_tris"] + gwork["shadow_tris"]) * TRI_BYTES + rays * (HIT_BYTES + 1),
        }
    dom = max(("trace", "shade", "convolve"), key=lambda k: mean[k])
    roof = None
    if dom in algo and mean[dom] > 0:
        ach = algo[dom] / (mean[dom] / 1000.0) / 1e9
        traffic = profiled_traffic() if dom == "trace" else None
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/trace_terrain1m_traffic.md" if traffic else None,
                "algorithmic_bytes": int(algo[dom]), "mean_ms": round(mean[dom], 4)}
        if ach > HBM_PEAK_GBS:
            roof["note"] = ("algorithmic bytes above the HBM peak: the node and triangle reads are served "
                            "by L2 and the Infinity Cache (traffic: the HBM bytes per launch); the kernel "
                            "is bound by VALU issue")
        if "shadow" in algo and mean.get("shadow", 0.0) > 0.0:
            sh = algo["shadow"] / (mean["shadow"] / 1000.0) / 1e9
            roof["shadow_kernel"] = {"achieved": round(sh, 1), "frac": round(sh / HBM_PEAK_GBS, 4),
                                     "algorithmic_bytes": int(algo["shadow"]), "mean_ms": round(mean["shadow"], 4)}
        elif "shadow" in algo:
            roof["shadow_rays"] = ("sun-grid shadow rays inside the shading launch (shade stage); "
                                   f"{int(algo['shadow'])} algorithmic bytes per update")

    log("[bench] deferred GI lookup, 1920x1080")
    gi = gi_lookup_study(grid, frame_at(0))
    log(f"[bench] gi lookup {json.dumps(gi)}")

    log("[bench] RT hard shadows, 1920x1080")
    shadows = shadow_mask_study(ds, frame_at(0))
    log(f"[bench] shadow mask {json.dumps(shadows)}")

    log("[bench] mirror reflections, 1920x1080")
    mirror = mirror_study(grid, ds, frame_at(0))
    log(f"[bench] mirror reflections {json.dumps(mirror)}")

    log("[bench] point-light shadow moments, 1024 lights")
    lights = light_shadow_study(ds)
    log(f"[bench] light shadows {json.dumps(lights)}")

    conv = None
    if not args.no_conv_study:
        log("[bench] convolution study (FP32 / BF16 / FP8 vs FP64)")
        conv = conv_study(cfg, frame_at(0))

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU oracle baseline: median of {args.cpu_runs} updates")
        cpu = cpu_baseline(scene, cfg, frame_at, args.cpu_runs)
        log(f"[bench] CPU {cpu['value']:.4g} rays/s on {cpu['cores']} threads ({cpu['cpu_model']})")

    legs = [x for x in args.legs.split(",") if x]
    node_format = ds.info()["node_format"]
    grid.close()
    ds.close()
    dyn = serve = None
    if "dynamic" in legs:
        log(f"[bench] dynamic 4M-triangle leg, {args.leg_frames} frames")
        dyn = run_dynamic(args.leg_frames, 3, args.rebuild_every, args.node_format, False)
    if "serve" in legs:
        log(f"[bench] serving leg, {args.leg_requests} requests")
        serve = run_serve(args.leg_requests, args.burst, args.inflight, args.streams, not args.no_graphs,
                          not args.no_fuse)

    out = {
        "metric": "probe rays/s (16k probes x 512 rays DDGI update)",
        "value": round(value, 1),
        "unit": "probe rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "f32+bf16", "fp16": "f32+fp16", "fp8": "f32+fp8"}[args.precision],
        "data": "synthetic (seeded scene generator, seed 2)",
        "config": {"workload": "terrain1m: 32x16x32 probes x 512 rays, 512^2 value-noise terrain "
                               "+ 10k box/icosphere instances",
                   "probes": P, "rays_per_probe": N, "triangles": scene.num_tris,
                   "bvh_nodes": "quantized 32 B" if node_format == 2 else "full 64 B",
                   "conv_precision": args.precision, "bvh_depth": info["bvh_depth"],
                   "lbvh_build_ms": round(info["build_ms"], 3),
                   "lbvh_first_build_ms": round(first_build_ms, 3),
                   "lbvh_refit_ms": round(refit_ms, 3),
                   "sun_grid_build_ms": round(sungrid_ms, 3),
                   "atlas_bytes": atlas_bytes,
                   "precision_note": "trace, shading and hysteresis blend in FP32; the convolution "
                                     f"GEMM in {args.precision} (FP32 accumulate)"},
        "stages_ms": {k: round(v, 4) for k, v in mean.items() if k not in ("translate", "sungrid")},
        "updates_by_conv_precision": variants,
        "moving_sun_update": moving,
        "work": {"primary_node_visits": work["primary_nodes"], "primary_triangle_tests": work["primary_tris"],
                 "shadow_rays_traversing": {"node_visits": work["shadow_nodes"],
                                            "triangle_tests": work["shadow_tris"]},
                 "shadow_rays_sun_grid": {"cell_lookups": gwork["shadow_nodes"],
                                          "candidate_tests": gwork["shadow_tris"],
                                          "state": st["sun_grid_state"],
                                          "estimated_candidates": round(st["sun_grid_candidates"], 2)}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "conv_study": conv,
        "gi_lookup": gi,
        "shadow_mask": shadows,
        "mirror_reflections": mirror,
        "light_shadows": lights,
        "dynamic4m": dyn,
        "serve64": serve,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
