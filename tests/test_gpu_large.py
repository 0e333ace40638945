"""GPU parity at the large workload: 262,144 triangles (a floor grid + 255 clustered 1024-triangle
instances), 32 x 16 x 32 probes x 256 rays = 4,194,304 probe rays per update, 8x8 / 16x16 FP32
atlases.  This is the workload the update's throughput is quoted on, and the one place where the
4-wide tree is deep enough (13 levels) for the traversal stack to spill past its 16-entry LDS ring,
so every decision here is compared with the oracle rather than inferred from the small scenes.

  * every probe ray of two frames through ddgi_trace_rays (traceProbes.rgen:29-49 rays, mask 0x3):
    full hit records (t, u, v bits, instance, primitive, facing) bit-exact against the oracle's
    FP32 intersection replica (src/Renderer.cpp:532-551 instance semantics);
  * the update's own hit records (slot -> instance / primitive through ddgi_get_slot_ids) bit-exact
    against the same oracle hits;
  * teacher-forced frames (the oracle starts from the GPU's atlases and states): ray radiance,
    irradiance and depth texels against the FP64 oracle at the small-workload bars;
  * 30 frames with the reference schedule's inputs (probe init, hysteresis ramp) and property
    checks: states in {0, 1..8} with the disabled set unchanged, finite non-negative atlases, depth
    means within the cell diagonal, and a late teacher-forced frame within the bars.
"""
import os

import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu

# relative-error bars of the teacher-forced comparison (FP32 GPU vs FP64 oracle), denominators
# max(|oracle|, 1e-2): at least RAY_FRACTION of the rays within RAY_BAR and every ray within
# RAY_MAX (of ~10^6 rays a handful sit where FP32 and FP64 take different sides of a sampleProbes
# threshold: measured 10 rays of 1.2 M, at most 4.6e-4), every irradiance / depth texel within
# TEXEL_BAR (measured <= 1e-6).
RAY_BAR, RAY_FRACTION, RAY_MAX = 1e-4, 0.9999, 1e-3
TEXEL_BAR = 1e-5


@pytest.fixture(scope="module")
def large():
    orc.set_threads(min(16, os.cpu_count() or 1))
    scene, grid = scenes.config_large(seed=2)
    osc = orc.OracleScene(scene)
    probes = IrradianceProbes(grid, max_rays_per_probe=256)
    probes.set_scene(scene)
    probes.set_light(scenes.sun_light())
    probes.set_sky(scenes.SKY_RADIANCE)
    info = probes.bvh_info()
    print(f"large: {osc.tri_count} triangles, BVH {info['nodes']} nodes", flush=True)
    yield scene, grid, osc, probes
    probes.destroy()


def _gpu_trace(probes, rays, mask):
    import torch
    out = []
    for lo in range(0, len(rays), 1 << 21):
        t = torch.from_numpy(np.ascontiguousarray(rays[lo:lo + (1 << 21)], np.float32)).cuda()
        h = probes.trace_rays(t, mask=mask).cpu().numpy().view(np.uint32)
        out.append(np.stack([h[:, 0], h[:, 3], h[:, 4], h[:, 5], h[:, 1], h[:, 2]], 1))
    return np.concatenate(out)


def _assert_same_hits(g, o, what):
    bad = np.any(g != o, axis=-1)
    assert not bad.any(), (f"{what}: {bad.sum()} of {bad.size} records differ; first gpu {g[bad][0]} "
                           f"oracle {o[bad][0]}")


def test_large_probe_rays_bit_exact(large):
    """ddgi_trace_rays over all 4,194,304 probe rays of two frames."""
    scene, grid, osc, probes = large
    og = orc.Grid(grid)
    for f, R in enumerate(scenes.orientations(31, 2)):
        rays = osc.probe_rays(og, R, np.arange(og.count))
        o, _ = osc.trace(rays, mask=0x3)
        g = _gpu_trace(probes, rays, 0x3)
        hit = o[:, 1] != 0xFFFFFFFF
        print(f"frame {f}: {len(rays)} rays, {hit.mean():.3f} hit, {np.mean(o[hit, 3] == 1):.3f} of hits back faces",
              flush=True)
        _assert_same_hits(g, o, f"frame {f}")


def _update_and_compare(probes, osc, og, grid, R, flags=_lib.UPDATE_ALL_PROBES):
    """One update teacher-forced from the GPU state: returns the per-frame error statistics."""
    irr = probes.irradiance()[..., :3].astype(np.float64)
    dep = probes.depth().astype(np.float64)
    states = probes.states()
    probes.update(R, flags)
    probes.synchronize()
    gray, gidx = probes.last_update_rays()
    ghit = probes.last_update_hit_records()
    res = osc.update(og, scenes.sun_light(), scenes.SKY_RADIANCE.astype(np.float64), R, gidx, states, irr, dep,
                     want_hits=True)
    _assert_same_hits(ghit, res["hits"], "update hit records")
    ray_err = np.abs(gray - res["rays"]) / np.maximum(np.abs(res["rays"]), 1e-2)
    gi, gd = probes.irradiance()[..., :3], probes.depth()
    ie = np.abs(gi - irr) / np.maximum(np.abs(irr), 1e-2)
    de = np.abs(gd - dep) / np.maximum(np.abs(dep), 1e-2)
    st = probes.states()
    return {"ray_ok": float(np.mean(ray_err.max(-1) <= RAY_BAR)), "ray_max": float(ray_err.max()),
            "irr_max": float(ie.max()), "dep_max": float(de.max()), "states_equal": float(np.mean(st == states)),
            "count": len(gidx)}


def test_large_update_teacher_forced(large):
    """Three warm-up updates (so the probes sample non-zero atlases), then two updates compared with
    the oracle: the update's hit records bit-exact, radiance and texels within the bars."""
    scene, grid, osc, probes = large
    og = orc.Grid(grid)
    probes.init_probes(scenes.orientations(40, 1)[0])
    frames = scenes.orientations(41, 5)
    for R in frames[:3]:
        probes.update(R, _lib.UPDATE_ALL_PROBES)
    for f, R in enumerate(frames[3:]):
        s = _update_and_compare(probes, osc, og, grid, R)
        print(f"frame {f}: {s}", flush=True)
        assert s["count"] == og.count
        assert s["ray_ok"] >= RAY_FRACTION and s["ray_max"] <= RAY_MAX
        assert s["irr_max"] <= TEXEL_BAR and s["dep_max"] <= TEXEL_BAR
        assert s["states_equal"] >= 0.999


def test_large_thirty_frames_properties(large):
    """30 updates of the reference's per-frame inputs (init, then the reference schedule with the
    hysteresis ramp), then one teacher-forced frame: states stay in {0, 1..8}, atlases stay finite
    and non-negative, and the late frame still meets the oracle bars."""
    scene, grid, osc, probes = large
    og = orc.Grid(grid)
    probes.init_probes(scenes.orientations(50, 1)[0])
    init_states = probes.states()
    assert set(np.unique(init_states)) <= {0, 1, 8}
    for R in scenes.orientations(51, 30):
        probes.update(R, _lib.UPDATE_HYSTERESIS_RAMP)
    probes.synchronize()
    st = probes.states()
    assert st.max() <= 8
    assert np.array_equal(st == 0, init_states == 0), "the state machine never enters or leaves state 0"
    irr, dep = probes.irradiance(), probes.depth()
    assert np.isfinite(irr).all() and np.isfinite(dep).all()
    assert irr[..., :3].min() >= 0.0 and dep[..., 0].min() >= 0.0
    cell = (grid["extent_max"] - grid["extent_min"]) / (np.array(grid["resolution"]) - 1)
    assert dep[..., 0].max() <= np.linalg.norm(cell) * (1 + 1e-5), "depth means are clamped to the cell diagonal"
    # the ramp left some hysteresis in the uniforms; the oracle's grid carries 0.97: upload it
    probes.update_uniforms()
    s = _update_and_compare(probes, osc, og, grid, scenes.orientations(52, 1)[0], flags=0)
    print(f"frame 31: {s}", flush=True)
    assert 0 < s["count"] < og.count
    assert s["ray_ok"] >= RAY_FRACTION and s["ray_max"] <= RAY_MAX
    assert s["irr_max"] <= TEXEL_BAR and s["dep_max"] <= TEXEL_BAR


# sampleProbes is discontinuous where its weight sum crosses the normalisation threshold and where
# a probe's depth variance is ~0 at the Chebyshev branch: a ray whose shading point lies within
# FP32 rounding of such a point may take the other branch than FP64 and change by up to ~1 %.
# Over 8.4 M rays a handful do (measured: 11 rays, at most 7.1e-3), so the 512-ray comparison bounds
# their share (1 in 10^5) and their size (2e-2) separately.
MAX_RAY_FRACTION, MAX_RAY_JUMP = 0.99999, 2e-2


@pytest.mark.parametrize("mode,flag,texel_bar", [("fp32", 0, TEXEL_BAR), ("bf16", _lib.UPDATE_BLEND_BF16, 1e-4)])
def test_large_maximum_ray_count(large, mode, flag, texel_bar):
    """The largest ray count a probe takes (512) on the large workload: 8,388,608 probe rays per
    update, twice the shade list, sort keys and blend K steps of 256.  One warm-up update, then one
    teacher-forced update compared with the oracle (hit records bit-exact, rays and texels within
    the bars; the BF16 blend at its own 1e-4)."""
    scene, grid, osc, _ = large
    g = np.array(grid)
    g["rays_per_probe"] = 512
    og = orc.Grid(g)
    probes = IrradianceProbes(g, max_rays_per_probe=512)
    try:
        probes.set_scene(scene)
        probes.set_light(scenes.sun_light())
        probes.set_sky(scenes.SKY_RADIANCE)
        frames = scenes.orientations(71, 2)
        probes.update(frames[0], _lib.UPDATE_ALL_PROBES | flag)
        s = _update_and_compare(probes, osc, og, g, frames[1], _lib.UPDATE_ALL_PROBES | flag)
        print(f"512 rays, {mode}: {s}", flush=True)
        assert s["count"] == og.count
        assert s["irr_max"] <= texel_bar and s["dep_max"] <= texel_bar
        assert s["ray_ok"] >= MAX_RAY_FRACTION and s["ray_max"] <= MAX_RAY_JUMP
    finally:
        probes.destroy()
