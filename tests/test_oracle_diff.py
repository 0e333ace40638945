"""Pins the oracle's differentiable update (orc_diff): its forward reproduces the oracle's own
probe update, and its hand-written FP64 adjoint matches central finite differences."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import gbuffer, scenes

SKY = np.array([0.5, 0.5, 0.5])


@pytest.fixture(scope="module")
def problem():
    scene, grid = scenes.config_small(seed=1, open_top=True)
    osc = orc.OracleScene(scene)
    g = orc.Grid(grid)
    light = scenes.sun_light()
    irr, dep = g.zero_atlases()
    states = np.ones(g.count, np.uint32)
    Rs = scenes.orientations(41, 4)
    for R in Rs[:3]:
        osc.update(g, light, SKY, R, np.arange(g.count), states, irr, dep)
    irr_prev = irr.copy()
    res = osc.update(g, light, SKY, Rs[3], np.arange(g.count), states, irr, dep, want_aux=True)
    rays = gbuffer.camera_rays(64, 36, eye=(2.0, 7.0, 2.0), look_at=(12.0, 1.0, 12.0))
    hits, _ = osc.trace(rays)
    gb = gbuffer.build(scene, rays, hits)
    dirs = orc.frame_dirs(Rs[3], int(grid["rays_per_probe"]))
    prob = orc.DiffProblem(g, states, irr_prev, dep, dirs, np.arange(g.count), res["aux"].reshape(-1),
                           gb["position"], gb["normal"], gb["to_camera"], gb["material"], 256)
    albedo, albedo_t, sun, sun_t = gbuffer.perturbed_params(scene.materials["base_color"], light["color"][:3], seed=5)
    target = prob.run(albedo_t, sun_t, np.zeros((prob.N, 3)))["shade"]
    return prob, irr, albedo, sun, target, scene


def test_forward_reproduces_the_probe_update(problem):
    prob, irr_after, albedo, sun, target, scene = problem
    r = prob.run(albedo, sun, target)
    assert np.allclose(r["irradiance"], irr_after, rtol=1e-12, atol=1e-15)
    assert r["loss"] > 0
    assert prob.N > 1500


def test_adjoint_matches_finite_differences(problem):
    prob, _, albedo, sun, target, scene = problem
    r = prob.run(albedo, sun, target, grad=True)
    ga, gs = r["g_albedo"], r["g_sun"]
    used = np.unique(np.concatenate([prob.gmat, prob.aux["material"][prob.aux["material"] != 0xFFFFFFFF]]))
    assert np.all(ga[np.setdiff1d(np.arange(256), used)] == 0)        # unused material slots
    rng = np.random.Generator(np.random.PCG64(7))
    eps = 1e-4
    checks = [("sun", c) for c in range(3)] + [("albedo", (int(m), int(c))) for m, c in
                                               zip(rng.choice(used, 29), rng.integers(0, 3, 29))]
    worst = 0.0
    for kind, k in checks:
        a, s = albedo.copy(), sun.copy()
        if kind == "sun":
            s[k] += eps
            lp = prob.run(a, s, target)["loss"]
            s[k] -= 2 * eps
            lm = prob.run(a, s, target)["loss"]
            an = gs[k]
        else:
            a[k] += eps
            lp = prob.run(a, s, target)["loss"]
            a[k] -= 2 * eps
            lm = prob.run(a, s, target)["loss"]
            an = ga[k]
        fd = (lp - lm) / (2 * eps)
        err = abs(fd - an) / max(abs(fd), 1e-3 * np.abs(ga).max())
        worst = max(worst, err)
        assert err < 1e-5, (kind, k, fd, an)
    print(f"FD vs FP64 adjoint: worst relative error {worst:.2e} over {len(checks)} parameters")
