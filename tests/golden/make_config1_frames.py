"""Regression fixture of the small plain-DDGI workload: the room + 64 boxes (780
triangles), 8x4x8 probes x 128 rays, hysteresis 0.97, probe init then 10 frames of the FP64 oracle
from zeroed atlases, every probe updated every frame.

Writes config1_frames.json beside this script: per frame the SHA-256 of the hit records (count x n
x 6 uint32: t bits, instance, primitive, back face, u bits, v bits) and of the probe states, and
for frames 0 and 9 the atlas sums and 64 fixed irradiance and 64 fixed depth values.  These pin the
oracle (and, through tests/test_gpu_golden.py, the GPU path) against the repository's own earlier
results: a regression pin, not a reference pin (the reference holds no golden vectors for this
path).

Run: python tests/golden/make_config1_frames.py"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import py as orc  # noqa: E402
from vulkanexp_amd import scenes  # noqa: E402

FRAMES, INIT_SEED, FRAME_SEED, SAMPLES = 10, 99, 100, 64
OUT = os.path.join(HERE, "config1_frames.json")


def workload():
    scene, grid = scenes.config_small(seed=1)
    return scene, grid, scenes.orientations(INIT_SEED, 1)[0], scenes.orientations(FRAME_SEED, FRAMES)


def sample_positions(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.choice(int(np.prod(shape)), SAMPLES, replace=False)


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_oracle():
    """The oracle's frames: a list of (hit records, states, irradiance, depth) per frame."""
    scene, grid, r_init, frames = workload()
    og, osc = orc.Grid(grid), orc.OracleScene(scene)
    states, _ = osc.init_probes(og, r_init)
    irr, dep = og.zero_atlases()
    out = []
    for R in frames:
        res = osc.update(og, scenes.sun_light(), scenes.SKY_RADIANCE.astype(np.float64), R, np.arange(og.count),
                         states, irr, dep, want_hits=True)
        out.append((res["hits"].astype(np.uint32), states.copy(), irr.copy(), dep.copy()))
    return out


def main():
    frames = run_oracle()
    irr_pos = sample_positions(frames[0][2].shape, 1)
    dep_pos = sample_positions(frames[0][3].shape, 2)
    doc = {"workload": "config_small(seed=1): 780 tris, 8x4x8 probes x 128 rays, hysteresis 0.97; "
                       f"init orientations({INIT_SEED}, 1), frames orientations({FRAME_SEED}, {FRAMES})",
           "irradiance_positions": irr_pos.tolist(), "depth_positions": dep_pos.tolist(), "frames": []}
    for f, (hits, states, irr, dep) in enumerate(frames):
        rec = {"hits_sha256": digest(hits), "states_sha256": digest(states)}
        if f in (0, FRAMES - 1):
            rec.update({"irradiance_sum": float(irr.sum()), "depth_sum": float(dep.sum()),
                        "irradiance_samples": irr.reshape(-1)[irr_pos].tolist(),
                        "depth_samples": dep.reshape(-1)[dep_pos].tolist()})
        doc["frames"].append(rec)
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
