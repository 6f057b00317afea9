"""Writes tests/golden/room8.json + room8_atlas.npz + room8_lights.json from the CPU oracle.

Small-room workload: 8^3 probes x 256 rays, 780 triangles, 4 updates with
seeded rotations.  Recorded per update: SHA-256 of the hit records and of the
shadow decisions (bit-exact pins), traversal counters, radiance/atlas sums;
the final atlas is stored subsampled (every 3rd texel row and column).  Point-light
shadow moments of 8 seeded lights in the same scene (one disabled, one with a NaN
enable flag, which the shaders treat as enabled): SHA-256 of the distance and
moments images, traversal counters, moment sums.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from anerend_amd import scenegen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_room8():
    scene, cfg, _ = scenegen.workload("room8")
    osc = O.OracleScene(scene)
    atlas = O.new_atlas(cfg)
    steps = []
    for it in range(4):
        fr = scenegen.make_frame(1, camera=(0.5, 1.5, 0.5), config=cfg, iteration=it)
        hits, sh, rad, ctr = osc.update(cfg, fr, atlas)
        steps.append({"hits_sha256": sha(hits), "shadow_sha256": sha(sh),
                      "counters": [int(c) for c in ctr],
                      "misses": int((hits["tri"] < 0).sum()),
                      "radiance_sum": float(rad.astype(np.float64).sum()),
                      "atlas_sum": float(atlas.astype(np.float64).sum())})
    return steps, atlas


def room8_lights():
    """The 8 lights {x, y, z, range, r, g, b, enabled} of the light fixture."""
    g = scenegen.PCG32(21, np.arange(8) + 100)
    L = np.zeros((8, 8), np.float32)
    L[:, 0] = g.uniform(-4.5, 4.5)
    L[:, 1] = g.uniform(-0.3, 3.3)
    L[:, 2] = g.uniform(-4.5, 4.5)
    L[:, 3] = g.uniform(1.5, 14.0)
    L[:, 4:7] = 1.0
    L[:, 7] = 1.0
    L[2, 7] = 0.0
    L[5, 7] = np.nan
    return L


def run_room8_lights():
    scene, _, _ = scenegen.workload("room8")
    dist, mom, ctr = O.OracleScene(scene).light_shadows(room8_lights())
    return {"dist_sha256": sha(dist), "moments_sha256": sha(mom), "counters": [int(c) for c in ctr],
            "mean_sum": float(mom[..., 0].astype(np.float64).sum()),
            "mean_square_sum": float(mom[..., 1].astype(np.float64).sum())}


def main():
    lights = {"workload": "room8 scene, 8 PCG32 lights (stream 21), light 2 disabled, light 5 with a "
                          "NaN enable flag; 1024 rays per light; moments image borders 0",
              **run_room8_lights()}
    with open(os.path.join(HERE, "room8_lights.json"), "w") as f:
        json.dump(lights, f, indent=1)
    steps, atlas = run_room8()
    meta = {"workload": "room8: seed 1, 8x8x8 probes at 1 m, 256 rays, hysteresis 0.97, "
                        "camera (0.5, 1.5, 0.5), rotations from iterations 0..3",
            "triangles": 780, "updates": steps, "atlas_subsample": "rows[::3], cols[::3]"}
    with open(os.path.join(HERE, "room8.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "room8_atlas.npz"), atlas=atlas[::3, ::3])
    print("wrote room8.json, room8_atlas.npz, room8_lights.json")


if __name__ == "__main__":
    main()
