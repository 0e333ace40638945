"""Differentiable probes at the full G-buffer size of the workload: 1920x1080 camera rays over the
room + boxes scene (open top), 8x4x8 probes x 128 rays, 256 albedo slots + sun, against the FP64
oracle (the same checks as the 192x108 test, at the size the forward/backward times are quoted
on).

  * the G-buffer's ray queries: every hit record of the 2,073,600 camera rays bit-exact against
    the oracle's FP32 intersection replica; the G-buffer pass (back faces culled) against the
    oracle's G-buffer;
  * forward: loss within 1e-4 relative of FP64; shaded texels within 1e-4 (relative to
    max(|FP64|, 1e-3 of the largest)) but for at most 1 in 100,000, each of which must be a point
    where the oracle's FP32 sampling departs from FP64 as well;
  * backward: every gradient component within 2e-3 of the FP64 adjoint and of FP64 central
    differences (eps 1e-4, 32 parameters: the sun's three channels and 29 random albedo
    components of materials the G-buffer shows), relative to the largest gradient component."""
import os

import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu
SKY = scenes.SKY_RADIANCE.astype(np.float64)
W, H = 1920, 1080


@pytest.fixture(scope="module")
def full():
    import torch
    from vulkanexp_amd.autograd import DiffContext
    orc.set_threads(min(16, os.cpu_count() or 1))
    scene, grid = scenes.config_small(seed=1, open_top=True)
    light = scenes.sun_light()
    p = IrradianceProbes(grid, max_rays_per_probe=128)
    p.set_scene(scene)
    p.set_light(light)
    p.set_sky(scenes.SKY_RADIANCE)
    Rs = scenes.orientations(41, 4)
    for R in Rs[:3]:
        p.update(R, _lib.UPDATE_ALL_PROBES)
    p.synchronize()
    og = orc.Grid(grid)
    osc = orc.OracleScene(scene)
    irr_prev = p.irradiance()[..., :3].astype(np.float64)
    dep = p.depth().astype(np.float64)
    states = p.states()
    res = osc.update(og, light, SKY, Rs[3], np.arange(og.count), states, irr_prev.copy(), dep, want_aux=True)
    rays = gbuffer.camera_rays(W, H, eye=(2.0, 7.0, 2.0), look_at=(12.0, 1.0, 12.0))
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).cuda().to(dt)
    gpu_hits = gbuffer.hits_from_device(p.trace_rays(t(rays)).cpu().numpy())
    hits, _ = osc.trace(rays)
    gb = gbuffer.build(scene, rays, hits)
    prob = orc.DiffProblem(og, states, irr_prev, dep, orc.frame_dirs(Rs[3], 128), np.arange(og.count),
                           res["aux"].reshape(-1), gb["position"], gb["normal"], gb["to_camera"], gb["material"], 256)
    albedo, albedo_t, sun, sun_t = gbuffer.perturbed_params(scene.materials["base_color"], light["color"][:3], seed=5)
    target = prob.run(albedo_t, sun_t, np.zeros((prob.N, 3)))["shade"].astype(np.float32)
    dctx = DiffContext(p, Rs[3], t(gb["position"]), t(gb["normal"]), t(gb["to_camera"]),
                       t(gb["material"].astype(np.int32), torch.int32), 256)
    yield {"p": p, "prob": prob, "dctx": dctx, "albedo": albedo, "sun": sun, "target": target, "t": t,
           "gpu_hits": gpu_hits, "hits": hits, "scene": scene, "texels": len(gb["pixel"])}
    p.destroy()


def test_full_gbuffer_ray_queries_bit_exact(full):
    got, want = full["gpu_hits"], full["hits"]
    assert got.shape == (W * H, want.shape[1])
    assert np.array_equal(got, want), f"{np.sum(np.any(got != want, 1))} of {W * H} hit records differ"


def test_full_gbuffer_pass_matches_oracle(full):
    """ddgi_gbuffer over the same camera (back faces culled) against the oracle's G-buffer: hit
    pattern and material ids exact, geometry within FP32 rounding of FP64."""
    s = full
    rays = gbuffer.camera_rays(W, H, eye=(2.0, 7.0, 2.0), look_at=(12.0, 1.0, 12.0))
    gb = s["p"].gbuffer(s["t"](rays)).cpu().numpy()
    hits, _ = orc.OracleScene(s["scene"]).trace(rays, cull_back=True)
    ref = orc.gbuffer(s["scene"], rays, hits)
    mat = gb[:, 15].view(np.uint32).astype(np.int64)
    mat[mat == 0xFFFFFFFF] = -1
    assert np.array_equal(mat, ref[:, 15].astype(np.int64))
    hit = ref[:, 15] >= 0
    assert hit.mean() > 0.9
    scale = 1.0 + np.abs(ref[hit, 0:3]).max()
    assert np.abs(gb[hit, 0:3] - ref[hit, 0:3]).max() <= 1e-5 * scale
    assert np.allclose(gb[hit, 3], ref[hit, 3], rtol=1e-5, atol=1e-5)
    assert np.abs(gb[hit, 4:7] - ref[hit, 4:7]).max() <= 1e-5


def test_full_forward_matches_fp64(full):
    s = full
    t = s["t"]
    assert s["dctx"].N == s["texels"] > 0.9 * W * H  # the front-facing hits of 2,073,600 pixels
    loss, shade = s["dctx"].forward(t(s["albedo"]), t(s["sun"]), t(s["target"]))
    ref = s["prob"].run(s["albedo"], s["sun"], s["target"].astype(np.float64))
    lg = float(loss.cpu()[0])
    err = np.abs(shade.cpu().numpy() - ref["shade"]) / np.maximum(np.abs(ref["shade"]), 1e-3 * np.abs(ref["shade"]).max())
    e = err.max(1)
    bad = np.nonzero(e > 1e-4)[0]
    print(f"{W}x{H}: loss gpu {lg:.6e} fp64 {ref['loss']:.6e}; shade max rel err {e.max():.2e}; "
          f"{len(bad)} of {len(e)} texels above 1e-4")
    assert abs(lg - ref["loss"]) <= 1e-4 * abs(ref["loss"])
    # Of ~2 M texels a few may sit on a discontinuity of sampleProbes, where FP32 and FP64 land on
    # different sides (measured: one texel, 1.0e-2; the FP64 oracle's own result there jumps by
    # 1 % when the point moves 3e-7 m along its normal).  Each one must be such a point: the
    # oracle's own FP32 sampling of the same FP64 atlases moves there by at least half the GPU's
    # deviation.
    assert len(bad) <= 1e-5 * len(e)
    if len(bad):
        pr = s["prob"]
        args = (pr.grid, pr.states, ref["irradiance"], pr.dep, pr.pos[bad], pr.nrm[bad], pr.tocam[bad])
        s64, s32 = orc.sample(*args, fp32=False), orc.sample(*args, fp32=True)
        d32 = (np.abs(s32 - s64) / np.maximum(np.abs(s64), 1e-30)).max(1)
        print(f"  texels {bad.tolist()}: GPU {e[bad].tolist()}, oracle FP32 sampling {d32.tolist()}")
        assert np.all(e[bad] <= 2.0 * d32 + 1e-4)


def test_full_gradients_match_fp64(full):
    from vulkanexp_amd.autograd import shaded_gbuffer_loss
    s = full
    t = s["t"]
    albedo = t(s["albedo"]).requires_grad_(True)
    sun = t(s["sun"]).requires_grad_(True)
    loss, _ = shaded_gbuffer_loss(s["dctx"], albedo, sun, t(s["target"]))
    loss.backward()
    ga, gs = albedo.grad.cpu().numpy().astype(np.float64), sun.grad.cpu().numpy().astype(np.float64)
    ref = s["prob"].run(s["albedo"], s["sun"], s["target"].astype(np.float64), grad=True)
    scale = max(np.abs(ref["g_albedo"]).max(), np.abs(ref["g_sun"]).max())
    ea = np.abs(ga - ref["g_albedo"]).max() / scale
    es = np.abs(gs - ref["g_sun"]).max() / scale
    print(f"{W}x{H}: gradient vs FP64 adjoint: albedo {ea:.2e}, sun {es:.2e} (scale {scale:.3e})")
    assert ea <= 2e-3 and es <= 2e-3
    rng = np.random.Generator(np.random.PCG64(13))
    used = np.unique(s["prob"].gmat)
    picks = [("sun", c) for c in range(3)] + [("albedo", (int(m), int(c))) for m, c in
                                              zip(rng.choice(used, 29), rng.integers(0, 3, 29))]
    worst = 0.0
    for kind, k in picks:
        a, sn = s["albedo"].copy(), s["sun"].copy()
        arr = sn if kind == "sun" else a
        arr[k] += 1e-4
        lp = s["prob"].run(a, sn, s["target"].astype(np.float64))["loss"]
        arr[k] -= 2e-4
        lm = s["prob"].run(a, sn, s["target"].astype(np.float64))["loss"]
        fd = (lp - lm) / 2e-4
        g = gs[k] if kind == "sun" else ga[k]
        worst = max(worst, abs(g - fd) / scale)
    print(f"{W}x{H}: gradient vs FP64 finite differences: worst {worst:.2e} over {len(picks)} parameters")
    assert worst <= 2e-3
