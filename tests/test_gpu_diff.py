"""Differentiable probes on the GPU against the FP64 oracle: room + boxes scene (open top, so the
sun reaches the interior), 8x4x8 probes x 128 rays, a 192x108 G-buffer, 256 albedo slots + sun.

Bars: loss and shaded texels within 1e-4 relative of FP64; every gradient component within 2e-3
of the FP64 adjoint and of FP64 central differences (eps 1e-4, 32 parameters), relative to the
largest gradient component; gradients bit-identical across repeated backward calls."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu
SKY = scenes.SKY_RADIANCE.astype(np.float64)


@pytest.fixture(scope="module")
def setup():
    import torch
    from vulkanexp_amd.autograd import DiffContext
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
    # oracle starts from the GPU's state before the prepared update
    og = orc.Grid(grid)
    osc = orc.OracleScene(scene)
    irr_prev = p.irradiance()[..., :3].astype(np.float64)
    dep = p.depth().astype(np.float64)
    states = p.states()
    irr = irr_prev.copy()
    res = osc.update(og, light, SKY, Rs[3], np.arange(og.count), states, irr, dep, want_aux=True)
    rays = gbuffer.camera_rays(192, 108, eye=(2.0, 7.0, 2.0), look_at=(12.0, 1.0, 12.0))
    hits, _ = osc.trace(rays)
    gb = gbuffer.build(scene, rays, hits)
    dirs = orc.frame_dirs(Rs[3], 128)
    prob = orc.DiffProblem(og, states, irr_prev, dep, dirs, np.arange(og.count), res["aux"].reshape(-1),
                           gb["position"], gb["normal"], gb["to_camera"], gb["material"], 256)
    albedo, albedo_t, sun, sun_t = gbuffer.perturbed_params(scene.materials["base_color"], light["color"][:3], seed=5)
    target = prob.run(albedo_t, sun_t, np.zeros((prob.N, 3)))["shade"].astype(np.float32)
    dev = "cuda"
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    dctx = DiffContext(p, Rs[3], t(gb["position"]), t(gb["normal"]), t(gb["to_camera"]),
                       t(gb["material"].astype(np.int32), torch.int32), 256)
    yield {"p": p, "prob": prob, "dctx": dctx, "albedo": albedo, "sun": sun, "target": target, "t": t}
    p.destroy()


def test_forward_matches_fp64(setup):
    s = setup
    t = s["t"]
    loss, shade = s["dctx"].forward(t(s["albedo"]), t(s["sun"]), t(s["target"]))
    ref = s["prob"].run(s["albedo"], s["sun"], s["target"].astype(np.float64))
    lg = float(loss.cpu()[0])
    err = np.abs(shade.cpu().numpy() - ref["shade"]) / np.maximum(np.abs(ref["shade"]), 1e-3 * np.abs(ref["shade"]).max())
    print(f"loss gpu {lg:.6e} fp64 {ref['loss']:.6e}; shade max rel err {err.max():.2e}")
    assert abs(lg - ref["loss"]) <= 1e-4 * abs(ref["loss"])
    assert err.max() <= 1e-4


def test_gradients_match_fp64_adjoint_and_finite_differences(setup):
    import torch
    from vulkanexp_amd.autograd import shaded_gbuffer_loss
    s = setup
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
    print(f"gradient vs FP64 adjoint: albedo {ea:.2e}, sun {es:.2e} (scale {scale:.3e})")
    assert ea <= 2e-3 and es <= 2e-3
    # FP64 central differences on 32 random parameters
    rng = np.random.Generator(np.random.PCG64(11))
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
    print(f"gradient vs FP64 finite differences: worst {worst:.2e} over {len(picks)} parameters")
    assert worst <= 2e-3


def test_backward_is_deterministic(setup):
    import torch
    s = setup
    t = s["t"]
    a, sn, tg = t(s["albedo"]), t(s["sun"]), t(s["target"])
    s["dctx"].forward(a, sn, tg)
    one = torch.ones(1, device="cuda")
    g1 = s["dctx"].backward(a, sn, tg, one)
    g2 = s["dctx"].backward(a, sn, tg, one)
    assert torch.equal(g1[0], g2[0]) and torch.equal(g1[1], g2[1])


def test_optimisation_recovers_the_target(setup):
    import torch
    from vulkanexp_amd.autograd import shaded_gbuffer_loss
    s = setup
    t = s["t"]
    albedo = t(s["albedo"]).requires_grad_(True)
    sun = t(s["sun"]).requires_grad_(True)
    target = t(s["target"])
    opt = torch.optim.Adam([albedo, sun], lr=0.02)
    first = None
    for it in range(40):
        opt.zero_grad()
        loss, _ = shaded_gbuffer_loss(s["dctx"], albedo, sun, target)
        loss.backward()
        opt.step()
        first = float(loss) if first is None else first
    assert float(loss) < 0.2 * first
