"""Writes tests/golden/room8.json + room8_atlas.npz from the CPU oracle.

Small-room workload: 8^3 probes x 256 rays, 780 triangles, 4 updates with
seeded rotations.  Recorded per update: SHA-256 of the hit records and of the
shadow decisions (bit-exact pins), traversal counters, radiance/atlas sums;
the final atlas is stored subsampled (every 3rd texel row and column).

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


def main():
    steps, atlas = run_room8()
    meta = {"workload": "room8: seed 1, 8x8x8 probes at 1 m, 256 rays, hysteresis 0.97, "
                        "camera (0.5, 1.5, 0.5), rotations from iterations 0..3",
            "triangles": 780, "updates": steps, "atlas_subsample": "rows[::3], cols[::3]"}
    with open(os.path.join(HERE, "room8.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "room8_atlas.npz"), atlas=atlas[::3, ::3])
    print("wrote room8.json, room8_atlas.npz")


if __name__ == "__main__":
    main()
