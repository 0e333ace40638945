"""Ray counts that are not multiples of the kernels' tiles, against the oracle.

The traversal runs one thread per probe ray in 256-thread workgroups, and the blends pad the rays
of a probe to K steps of 32 (16-bit blends) or chunks of 64 (FP32 blend): a count of 1, 33, 255
or 511 leaves partial workgroups, partial K steps and padding that must contribute nothing.

  * the whole update (trace + shade + blend + border), FP32: hit records bit-exact, every ray
    within 1e-4 and every texel within 1e-5 of the FP64 oracle, as in the small-workload parity;
  * the 16-bit blends of one ray buffer at 33 and 100 rays per probe: within their bars of the
    FP64 blend (BF16 1e-4, FP16 1e-3), hysteresis 0."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, scenes
from vulkanexp_amd import _lib

from test_gpu_parity import _check, _frame_compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 33, 255, 511])
def test_update_ragged_ray_counts(n):
    scene, _ = scenes.config_small(seed=2, open_top=True)
    lo, hi = scenes.inset_extent(scene)
    grid = scenes.grid_params(lo, hi, (6, 3, 5), n)
    _check(_frame_compare(scene, grid, orc.OracleScene(scene), 2, seed=40 + n), f"{n} rays")


def _rel(g, o):
    floor = 1e-3 * max(np.abs(o).max(), 1e-12)
    return np.abs(g - o) / np.maximum(np.abs(o), floor)


@pytest.mark.parametrize("n", [33, 100])
@pytest.mark.parametrize("mode,flag,bar", [("bf16", _lib.UPDATE_BLEND_BF16, 1e-4), ("fp16", _lib.UPDATE_BLEND_FP16, 1e-3)])
def test_16bit_blends_ragged_ray_counts(n, mode, flag, bar):
    import torch
    scene, grid = scenes.config_small(seed=1, rays=n, open_top=True)
    grid = np.array(grid)
    grid["hysteresis"] = 0.0
    p = IrradianceProbes(grid, max_rays_per_probe=n)
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    for R in scenes.orientations(50 + n, 2):
        p.update(R, _lib.UPDATE_ALL_PROBES)
    R = scenes.orientations(60 + n, 1)[0]
    p.update(R, _lib.UPDATE_ALL_PROBES)
    p.synchronize()
    rays, idx = p.last_update_rays()
    irr = p.irradiance()[..., :3].astype(np.float64)
    dep = p.depth().astype(np.float64)
    states = p.states()
    p.blend_rays(torch.from_numpy(np.ascontiguousarray(rays)).cuda(),
                 torch.from_numpy(idx.view(np.int32).copy()).cuda(), R, flag)
    p.synchronize()
    gi, gd = p.irradiance()[..., :3], p.depth()
    orc.blend(orc.Grid(grid), orc.frame_dirs(R, n), rays.astype(np.float64), idx, states, irr, dep)
    ei, ed = _rel(gi, irr), _rel(gd, dep)
    print(f"{mode} {n} rays: irradiance max {ei.max():.2e}, depth max {ed.max():.2e}")
    p.destroy()
    assert ei.max() <= bar and ed.max() <= bar
