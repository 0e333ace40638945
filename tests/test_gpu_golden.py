"""The GPU path against the committed 10-frame history of the small plain-DDGI workload
(tests/golden/config1_frames.json, the FP64 oracle's frames from probe init and zeroed atlases),
and the save/resume invariant of its state.

  * every frame: the update's hit records (slot -> instance / primitive) hash to the stored
    digest, bit for bit;
  * frames 0 and 9: the stored irradiance / depth values within 1e-4 relative (max(|v|, 1e-2)
    denominators), after 10 frames of FP32 on the GPU against FP64 without teacher forcing;
  * N frames, save, M frames == N frames, save, load, M frames (atlases, states, hysteresis)."""
import hashlib
import json

import numpy as np
import pytest

from tests.golden.make_config1_frames import OUT, workload
from vulkanexp_amd import IrradianceProbes, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu


def _probes(scene, grid):
    p = IrradianceProbes(grid, max_rays_per_probe=int(grid["rays_per_probe"]))
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    return p


def test_config1_frames_match_the_fixture():
    want = json.load(open(OUT))
    scene, grid, r_init, frames = workload()
    ip, dp = np.array(want["irradiance_positions"]), np.array(want["depth_positions"])
    p = _probes(scene, grid)
    try:
        p.init_probes(r_init)
        ids = p.slot_ids()
        for f, (R, w) in enumerate(zip(frames, want["frames"])):
            p.update(R, _lib.UPDATE_ALL_PROBES)
            p.synchronize()
            hits = np.ascontiguousarray(p.last_update_hit_records(ids))
            assert hashlib.sha256(hits.tobytes()).hexdigest() == w["hits_sha256"], f"frame {f}: hit records differ"
            if "irradiance_samples" in w:
                gi = p.irradiance()[..., :3].astype(np.float64).reshape(-1)[ip]
                gd = p.depth().astype(np.float64).reshape(-1)[dp]
                ei = np.abs(gi - w["irradiance_samples"]) / np.maximum(np.abs(w["irradiance_samples"]), 1e-2)
                ed = np.abs(gd - w["depth_samples"]) / np.maximum(np.abs(w["depth_samples"]), 1e-2)
                print(f"frame {f}: irradiance max {ei.max():.2e}, depth max {ed.max():.2e}")
                assert ei.max() <= 1e-4 and ed.max() <= 1e-4
    finally:
        p.destroy()


def test_save_resume_equals_uninterrupted_frames():
    scene, grid, r_init, frames = workload()
    p = _probes(scene, grid)
    try:
        p.init_probes(r_init)
        flags = _lib.UPDATE_ALL_PROBES | _lib.UPDATE_HYSTERESIS_RAMP
        for R in frames[:4]:
            p.update(R, flags)
        state = p.save_state()
        for R in frames[4:]:
            p.update(R, flags)
        p.synchronize()
        first = (p.irradiance().copy(), p.depth().copy(), p.states().copy())
        p.load_state(state)
        for R in frames[4:]:
            p.update(R, flags)
        p.synchronize()
        again = (p.irradiance(), p.depth(), p.states())
        for a, b in zip(first, again):
            assert np.array_equal(a, b)
    finally:
        p.destroy()
