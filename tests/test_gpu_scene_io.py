"""A scene loaded from the reference's own glTF fixture (tests/golden/gltf/sphere.gltf) on the GPU:
ray queries bit-exact against the oracle and against the closed form of a tessellated unit sphere,
and a probe update around it under the atmospheric sky against the FP64 oracle (the sphere seen
through the reference's loader, material defaults and sky model)."""
import os

import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu

SPHERE = os.path.join(os.path.dirname(__file__), "golden", "gltf", "sphere.gltf")


@pytest.fixture(scope="module")
def loaded():
    scene = scenes.load(SPHERE)
    grid = scenes.grid_params((-3, -3, -3), (3, 3, 3), (4, 4, 4), 64)
    p = IrradianceProbes(grid, max_rays_per_probe=64)
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    yield {"p": p, "scene": scene, "grid": grid, "osc": orc.OracleScene(scene)}
    p.destroy()


def test_rays_at_the_loaded_sphere(loaded):
    import torch
    rng = np.random.Generator(np.random.PCG64(9))
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((5000, 8), np.float32)
    rays[:, :3] = 5.0 * d
    rays[:, 4:7] = -d
    rays[:, 7] = 100.0
    got = gbuffer.hits_from_device(loaded["p"].trace_rays(torch.from_numpy(rays).cuda()).cpu().numpy())
    want, _ = loaded["osc"].trace(rays)
    assert np.array_equal(got, want)
    t = got[:, 0].view(np.float32)
    assert np.all(got[:, 1] == 0)                    # every ray hits the one instance
    assert np.all((t >= 4.0 - 1e-5) & (t <= 4.03))   # vertices on the unit sphere, faces just inside
    assert np.all(got[:, 3] == 0)                    # front faces: the normals point outward


def test_probe_update_around_the_loaded_sphere(loaded):
    p, grid, osc = loaded["p"], loaded["grid"], loaded["osc"]
    og = orc.Grid(grid)
    p.set_sky_model(_lib.SKY_ATMOSPHERE)
    try:
        p.init_probes(scenes.orientations(4, 1)[0])
        R = scenes.orientations(23, 1)[0]
        irr = p.irradiance()[..., :3].astype(np.float64)
        dep = p.depth().astype(np.float64)
        states = p.states()
        p.update(R, _lib.UPDATE_ALL_PROBES)
        p.synchronize()
        gray, gidx = p.last_update_rays()
        res = osc.update(og, scenes.sun_light(), scenes.SKY_RADIANCE.astype(np.float64), R, gidx, states, irr, dep,
                         sky_model=1)
        want = res["rays"]
        assert np.array_equal(gray[..., 3] == -1, want[..., 3] == -1)
        assert 0.5 < np.mean(want[..., 3] == -1) < 1.0
        err = np.abs(gray - want) / np.maximum(np.abs(want), 1e-2)
        assert np.mean(err.max(-1) <= 1e-4) >= 0.995
    finally:
        p.set_sky_model(_lib.SKY_CONSTANT)
