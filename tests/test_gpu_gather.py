"""Full-frame consumer on the GPU against the oracle: back-face-culled ray queries (bit-exact hit
records), the G-buffer of a seeded camera (hit pattern and material ids exact, geometry within FP32
rounding of the FP64 oracle) and the final gather over that G-buffer (within 1e-4 of FP64, with
the hard-shadow visibility decisions identical).  Room + boxes scene with an open top (so the sun
reaches the floor), 8x4x8 probes x 128 rays after three updates, 192x108 pixels."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu

W, H = 192, 108
EYE = (2.0, 7.0, 2.0)


@pytest.fixture(scope="module")
def setup():
    import torch
    scene, grid = scenes.config_small(seed=1, open_top=True)
    light = scenes.sun_light()
    p = IrradianceProbes(grid, max_rays_per_probe=128)
    p.set_scene(scene)
    p.set_light(light)
    p.set_sky(scenes.SKY_RADIANCE)
    p.init_probes(scenes.orientations(3, 1)[0])
    for R in scenes.orientations(41, 3):
        p.update(R, _lib.UPDATE_ALL_PROBES)
    p.synchronize()
    rays = gbuffer.camera_rays(W, H, eye=EYE, look_at=(12.0, 1.0, 12.0))
    osc = orc.OracleScene(scene)
    yield {"p": p, "scene": scene, "grid": grid, "light": light, "rays": rays, "osc": osc,
           "t": lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()}
    p.destroy()


def _box_interior_rays(scene, osc, k=4000, seed=7):
    """Rays starting at the centres of the boxes (whose faces point outward): the first surface they
    meet is a back face."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tris = osc.world_triangles()                      # instance-major
    counts = scene.instances["triangle_count"].astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)])
    centres = np.array([tris[starts[i]:starts[i + 1]].reshape(-1, 3).mean(0) for i in range(1, len(counts))])
    o = centres[rng.integers(0, len(centres), k)] + rng.normal(0, 0.02, (k, 3))
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((k, 8), np.float32)
    r[:, :3], r[:, 4:7], r[:, 7] = o, d, 100.0
    return r


def test_culled_ray_queries_bit_exact(setup):
    s = setup
    rays = np.concatenate([s["rays"], _box_interior_rays(s["scene"], s["osc"])])
    for cull in (False, True):
        got = gbuffer.hits_from_device(s["p"].trace_rays(s["t"](rays), cull_back=cull).cpu().numpy())
        want, _ = s["osc"].trace(rays, cull_back=cull)
        assert np.array_equal(got, want), f"cull={cull}: {np.sum(np.any(got != want, 1))} records differ"
    # the interior rays see back faces without culling and pass through them with it
    inner = slice(len(s["rays"]), None)
    plain, _ = s["osc"].trace(rays[inner])
    assert np.mean(plain[:, 3] == 1) > 0.9


def test_gbuffer_matches_oracle(setup):
    s = setup
    gb = s["p"].gbuffer(s["t"](s["rays"])).cpu().numpy()
    hits, _ = s["osc"].trace(s["rays"], cull_back=True)
    ref = orc.gbuffer(s["scene"], s["rays"], hits)
    mat = gb[:, 15].view(np.uint32).astype(np.int64)
    mat[mat == 0xFFFFFFFF] = -1
    assert np.array_equal(mat, ref[:, 15].astype(np.int64))
    hit = ref[:, 15] >= 0
    assert hit.mean() > 0.9
    assert np.all(gb[~hit, 3] == 0)
    scale = 1.0 + np.abs(ref[hit, 0:3]).max()
    assert np.abs(gb[hit, 0:3] - ref[hit, 0:3]).max() <= 1e-5 * scale
    assert np.allclose(gb[hit, 3], ref[hit, 3], rtol=1e-5, atol=1e-5)
    assert np.abs(gb[hit, 4:7] - ref[hit, 4:7]).max() <= 1e-5
    assert np.allclose(gb[hit, 8:11], ref[hit, 8:11], rtol=1e-6, atol=1e-7)
    for c in (7, 11, 12, 13, 14):
        assert np.array_equal(gb[hit, c], ref[hit, c].astype(np.float32))


@pytest.mark.parametrize("inputs", ["shadow_ray", "given_direct_and_reflection"])
def test_gather_matches_oracle(setup, inputs):
    s = setup
    p = s["p"]
    gbt = p.gbuffer(s["t"](s["rays"]))
    gb = gbt.cpu().numpy()
    direct = reflection = None
    if inputs != "shadow_ray":
        rng = np.random.Generator(np.random.PCG64(12))
        direct = rng.uniform(0, 1, len(gb)).astype(np.float32)
        reflection = rng.uniform(0, 2, (len(gb), 3)).astype(np.float32)
    out = p.gather(gbt, EYE, direct=None if direct is None else s["t"](direct),
                   reflection=None if reflection is None else s["t"](reflection)).cpu().numpy()
    og = orc.Grid(s["grid"])
    irr = p.irradiance()[..., :3].astype(np.float64)
    dep = p.depth().astype(np.float64)
    ref = orc.gather(og, p.states(), irr, dep, s["osc"], s["light"], scenes.SKY_RADIANCE, gb, EYE,
                     direct=direct, reflection=reflection)
    err = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    print(f"gather ({inputs}): max rel err {err.max():.2e}, mean {err.mean():.2e}")
    assert err.max() <= 1e-4
    if inputs == "shadow_ray":
        # sunlit and shadowed texels both occur, so the visibility decisions are exercised
        lit = orc.gather(og, p.states(), irr, dep, s["osc"], s["light"], scenes.SKY_RADIANCE, gb, EYE,
                         direct=np.zeros(len(gb), np.float32))
        sunlit = np.abs(ref - lit).max(1) > 1e-3
        assert 0.05 < sunlit.mean() < 0.95
