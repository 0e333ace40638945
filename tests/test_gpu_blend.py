"""GPU blend modes against the FP64 oracle blend (src/shaders/probesUpdate.glsl) on the same ray
buffer: the room + boxes scene and 8x4x8 grid with 512 rays per probe, in two forms:
  * "cos50": the mixed-precision workload as specified: the closed 780-triangle room and a depth
    lobe of cos^50 (depthSharpness 50, probesUpdate.glsl:78-81), where most depth weights of a
    texel fall below FP16's normal range (cos < 0.82);
  * "open12": the open-top room at the reference's default depthSharpness 12.

Per-texel relative error |gpu - fp64| / max(|fp64|, 1e-3 * max|fp64|), stated per output:
  FP32 MFMA (f32 operands) irradiance and depth <= 2e-5
  BF16 MFMA, hi/lo split irradiance and depth <= 1e-4
  FP16 MFMA              irradiance <= 1e-3, depth <= 1e-3
measured on the new-frame estimate (hysteresis 0) and after hysteresis 0.97."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu

BARS = {"fp32": (2e-5, 2e-5), "bf16": (1e-4, 1e-4), "fp16": (1e-3, 1e-3)}
FLAGS = {"fp32": 0, "bf16": _lib.UPDATE_BLEND_BF16, "fp16": _lib.UPDATE_BLEND_FP16}


def _rel(g, o):
    floor = 1e-3 * max(np.abs(o).max(), 1e-12)
    return np.abs(g - o) / np.maximum(np.abs(o), floor)


WORKLOADS = {"cos50": (False, 50.0), "open12": (True, 12.0)}


@pytest.fixture(scope="module", params=sorted(WORKLOADS))
def ray_buffer(request):
    open_top, sharpness = WORKLOADS[request.param]
    scene, grid = scenes.config_small(seed=1, rays=512, open_top=open_top)
    grid = np.array(grid)
    grid["depth_sharpness"] = sharpness
    p = IrradianceProbes(grid, max_rays_per_probe=512)
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    for R in scenes.orientations(31, 3):
        p.update(R, _lib.UPDATE_ALL_PROBES)
    R = scenes.orientations(32, 1)[0]
    state = p.save_state()
    p.update(R, _lib.UPDATE_ALL_PROBES)
    p.synchronize()
    rays, idx = p.last_update_rays()
    p.load_state(state)
    yield p, grid, rays, idx, R, state
    p.destroy()


@pytest.mark.parametrize("mode", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("hysteresis", [0.0, 0.97])
def test_blend_modes_against_fp64(ray_buffer, mode, hysteresis):
    import torch
    p, grid, rays, idx, R, state = ray_buffer
    g = np.array(grid)
    g["hysteresis"] = hysteresis
    p.load_state(state)        # restores atlases, states and the saved hysteresis ...
    p.GridParameters = g
    p.update_uniforms()        # ... which this then overrides
    og = orc.Grid(g)
    irr = p.irradiance()[..., :3].astype(np.float64)
    dep = p.depth().astype(np.float64)
    states = p.states()
    t_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    t_idx = torch.from_numpy(idx.view(np.int32).copy()).cuda()
    p.blend_rays(t_rays, t_idx, R, FLAGS[mode])
    p.synchronize()
    gi, gd = p.irradiance()[..., :3], p.depth()
    dirs = orc.frame_dirs(R, 512)
    orc.blend(og, dirs, rays.astype(np.float64), idx, states, irr, dep)
    ei, ed = _rel(gi, irr), _rel(gd, dep)
    print(f"{mode} sharpness={float(g['depth_sharpness']):g} h={hysteresis}: irradiance max {ei.max():.2e} mean {ei.mean():.2e}; depth max {ed.max():.2e} mean {ed.mean():.2e}")
    bi, bd = BARS[mode]
    assert ei.max() <= bi and ed.max() <= bd
    assert np.array_equal(p.states(), states)


def test_bf16_and_fp32_select_same_states(ray_buffer):
    import torch
    p, grid, rays, idx, R, state = ray_buffer
    t_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    t_idx = torch.from_numpy(idx.view(np.int32).copy()).cuda()
    out = []
    for mode in ("fp32", "bf16"):
        p.load_state(state)
        p.blend_rays(t_rays, t_idx, R, FLAGS[mode])
        p.synchronize()
        out.append(p.states())
    assert np.mean(out[0] == out[1]) >= 0.99


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
def test_full_probe_list_skips_the_copy_forward(ray_buffer, mode):
    """ddgi_blend_rays with UPDATE_ALL_PROBES and every probe listed skips copying the exposed
    atlas forward: atlases and states identical to the call without the flag.  With a partial list
    the flag changes nothing (the copy still runs: unlisted probes keep their tiles)."""
    import torch
    p, grid, rays, idx, R, state = ray_buffer
    half = len(idx) // 2
    for count in (len(idx), half):
        t_rays = torch.from_numpy(np.ascontiguousarray(rays[:count])).cuda()
        t_idx = torch.from_numpy(idx[:count].view(np.int32).copy()).cuda()
        out = []
        for extra in (0, _lib.UPDATE_ALL_PROBES):
            p.load_state(state)
            p.blend_rays(t_rays, t_idx, R, FLAGS[mode] | extra)
            p.synchronize()
            out.append((p.irradiance(_lib.LAYOUT_TILED), p.depth(_lib.LAYOUT_TILED), p.states()))
        for a, b in zip(*out):
            assert np.array_equal(a, b), f"{mode}, {count} probes listed"


@pytest.mark.parametrize("rays", [33, 64, 256, 512])
def test_fp32_workgroup_shapes_give_the_same_sums(rays):
    """The FP32 blend runs in 64-probe workgroups (k_blend_wide) once they fill the device and in
    16-probe blocks (k_blend_fp32) below that; both issue the same MFMAs per output element in the
    same order, so their atlases and states are bit-identical.  A grid of 64 probes per compute unit
    is blended whole (64-probe workgroups) and as its first half (16-probe blocks), the probes
    listed in a shuffled order; the first half's tiles and states must agree exactly (ray counts
    with a partial K step, 2, 8 and 16 K steps)."""
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    scene, _ = scenes.config_small(seed=3, rays=rays, open_top=True)
    lo, hi = scenes.inset_extent(scene)
    grid = np.array(scenes.grid_params(lo, hi, (32, 2 * cus // 32, 32), rays))
    p = IrradianceProbes(grid, max_rays_per_probe=rays)
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    P = p.get_probe_count()
    assert P >= 64 * cus
    for R in scenes.orientations(70 + rays, 2):
        p.update(R, _lib.UPDATE_ALL_PROBES)
    R = scenes.orientations(80 + rays, 1)[0]
    state = p.save_state()
    p.update(R, _lib.UPDATE_ALL_PROBES)
    p.synchronize()
    ray_buf, idx = p.last_update_rays()
    perm = np.random.Generator(np.random.PCG64(rays)).permutation(P)  # slots in any order
    ray_buf, idx = ray_buf[perm], idx[perm]
    out = []
    for count, flags in ((P // 2, 0), (P, _lib.UPDATE_ALL_PROBES)):
        p.load_state(state)
        p.blend_rays(torch.from_numpy(np.ascontiguousarray(ray_buf[:count])).cuda(),
                     torch.from_numpy(idx[:count].view(np.int32).copy()).cuda(), R, flags)
        p.synchronize()
        sel = idx[:P // 2].astype(np.int64)
        out.append((p.irradiance(_lib.LAYOUT_TILED)[sel], p.depth(_lib.LAYOUT_TILED)[sel], p.states()[sel]))
    p.destroy()
    for a, b, what in zip(*out, ("irradiance", "depth", "states")):
        assert np.array_equal(a, b), f"{rays} rays: {what} differs"
