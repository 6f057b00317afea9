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
    p.add_argument("--precision", default="bf16", choices=["fp32", "bf16", "fp8"])
    p.add_argument("--workload", default="terrain1m")
    p.add_argument("--cpu-layers", type=int, default=2,
                   help="probe layers of the same workload traced by the CPU oracle baseline")
    p.add_argument("--no-cpu-baseline", action="store_true")
    return p.parse_args()


def cpu_baseline(scene, cfg, frame, layers):
    """The oracle (C, OpenMP) on a bounded sample of the same workload."""
    from oracle import oracle as O
    t0 = time.time()
    osc = O.OracleScene(scene)
    build_s = time.time() - t0
    fr = frame.copy()
    fr.layer_begin, fr.layer_count = 0, layers
    atlas = O.new_atlas(cfg)
    t0 = time.time()
    osc.update(cfg, fr, atlas, want_rays=False)
    dt = time.time() - t0
    rays = layers * cfg.probes_x * cfg.probes_z * cfg.rays_per_probe
    return {"value": rays / dt, "unit": "probe rays/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{layers} of {cfg.probes_y} probe layers ({rays} rays: trace, shadow, shade, "
                      f"convolve, blend) of the same workload in {dt:.2f} s; CPU LBVH build "
                      f"{build_s:.2f} s not counted"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); the multi-GPU "
              "sharding of the probe grid is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    from anerend_amd import native, scenegen
    from anerend_amd.types import PRECISION_BF16, PRECISION_FP8, PRECISION_FP32
    prec = {"fp32": PRECISION_FP32, "bf16": PRECISION_BF16, "fp8": PRECISION_FP8}[args.precision]

    log(f"[bench] building workload {args.workload}")
    scene, cfg, frame = scenegen.workload(args.workload)
    cfg.precision = prec
    log(f"[bench] scene: {scene.num_tris} triangles, {len(scene.instances)} instances")
    ds = native.DeviceScene(scene)
    info = ds.info()
    log(f"[bench] LBVH on GPU: {info['build_ms']:.2f} ms, depth {info['bvh_depth']}")
    grid = native.ProbeGrid(cfg)
    P = cfg.probes_x * cfg.probes_y * cfg.probes_z
    N = cfg.rays_per_probe

    def frame_at(i):
        return scenegen.make_frame(2, camera=tuple(frame.camera_pos[:3]),
                                   config=cfg, iteration=i)

    # traversal work counts (separate, counting kernels; outside the timed region)
    grid.set_counting(True)
    grid.update(ds, frame_at(0))
    work = grid.stats()
    grid.set_counting(False)
    log(f"[bench] work: primary nodes {work['primary_nodes']} tris {work['primary_tris']}, "
        f"shadow nodes {work['shadow_nodes']} tris {work['shadow_tris']}")

    for i in range(args.warmup):
        grid.update(ds, frame_at(1 + i))
    native.synchronize()
    grid.reset_stats()
    log(f"[bench] timing {args.steps} updates")
    t0 = time.perf_counter()
    for i in range(args.steps):
        grid.update(ds, frame_at(100 + i))
    native.synchronize()
    t1 = time.perf_counter()
    st = grid.stats()
    ms = (t1 - t0) * 1000.0 / args.steps
    value = P * N / (ms / 1000.0)
    mean = st["mean_ms"]
    log(f"[bench] {ms:.3f} ms/update, stages {json.dumps({k: round(v, 4) for k, v in mean.items()})}")

    # calibration kernel outside the timed region: one camera-cell move makes the
    # translation remap read and write the whole atlas once (2 x atlas bytes),
    # the known-size copy that scales the profiler's FETCH_SIZE / WRITE_SIZE
    atlas_bytes = cfg.atlas_shape[0] * cfg.atlas_shape[1] * 16
    cam = np.array(frame.camera_pos[:3], np.float32)
    grid.translate(cam + np.array([1.0, 0.0, 0.0], np.float32))
    native.synchronize()

    # roofline of the dominant kernel (algorithmic bytes per launch / mean launch time)
    rays = P * N
    algo = {
        "trace": work["primary_nodes"] * NODE_BYTES + work["primary_tris"] * TRI_BYTES + rays * HIT_BYTES,
        "shadow": work["shadow_nodes"] * NODE_BYTES + work["shadow_tris"] * TRI_BYTES + rays * (HIT_BYTES + 1),
    }
    dom = max(("trace", "shadow", "shade", "convolve"), key=lambda k: mean[k])
    roof = None
    if dom in algo and mean[dom] > 0:
        ach = algo[dom] / (mean[dom] / 1000.0) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "algorithmic_bytes": int(algo[dom]), "mean_ms": round(mean[dom], 4)}

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU oracle baseline on {args.cpu_layers} layers")
        cpu = cpu_baseline(scene, cfg, frame_at(0), args.cpu_layers)
        log(f"[bench] CPU {cpu['value']:.4g} rays/s on {cpu['cores']} threads")

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
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed 2)",
        "config": {"workload": "terrain1m: 32x16x32 probes x 512 rays, 512^2 value-noise terrain "
                               "+ 10k box/icosphere instances",
                   "probes": P, "rays_per_probe": N, "triangles": scene.num_tris,
                   "conv_precision": args.precision, "bvh_depth": info["bvh_depth"],
                   "lbvh_build_ms": round(info["build_ms"], 3),
                   "atlas_bytes": atlas_bytes},
        "stages_ms": {k: round(v, 4) for k, v in mean.items() if k != "translate"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
