"""GPU parity: the HIP kernels (through the C-ABI) against the CPU oracle on the same seeded inputs.

Bars
  * ray/triangle decisions: bit-exact (t, barycentrics as FP32 bits; instance, primitive, facing),
    for ddgi_trace_rays and for the update's own hit records (BVH slot -> instance / primitive)
  * probe states from init: identical; after an update: >= 99 % identical (a state moves when the
    probe's max change crosses 0.02 / state or 0.04 / state, which FP32 and FP64 may see apart)
  * ray radiance/depth (FP32 GPU vs FP64 oracle): per-ray relative error <= 1e-4 for every ray
  * irradiance and depth texels: relative error <= 1e-5 (measured: <= 1.4e-6)
  (relative error = |gpu - oracle| / max(|oracle|, 1e-2))
"""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _gpu_trace(probes, rays, mask, any_hit=False):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    h = probes.trace_rays(t, mask=mask, any_hit=any_hit).cpu().numpy().view(np.uint32)
    # (t, u, v, instance, primitive, flags) -> oracle record order (t, inst, prim, flags, u, v)
    return np.stack([h[:, 0], h[:, 3], h[:, 4], h[:, 5], h[:, 1], h[:, 2]], 1)


def _make(scene, grid, max_rays=None):
    p = IrradianceProbes(grid, max_rays_per_probe=max_rays or int(grid["rays_per_probe"]))
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    return p


@pytest.fixture(scope="module", params=["closed", "open_top"])
def small(request):
    scene, grid = scenes.config_small(seed=1, open_top=request.param == "open_top")
    return scene, grid, orc.OracleScene(scene)


def _compare_hits(g, o, what):
    miss_g, miss_o = g[:, 1] == 0xFFFFFFFF, o[:, 1] == 0xFFFFFFFF
    assert np.array_equal(miss_g, miss_o), f"{what}: hit/miss differs on {np.sum(miss_g != miss_o)} rays"
    bad = np.any(g != o, axis=1)
    assert not bad.any(), f"{what}: {bad.sum()} of {len(g)} records differ, first {np.argmax(bad)}: gpu {g[bad][0]} oracle {o[bad][0]}"


def test_probe_rays_hit_ids_bit_exact(small):
    scene, grid, osc = small
    og = orc.Grid(grid)
    probes = _make(scene, grid)
    for frame, R in enumerate(scenes.orientations(11, 3)):
        rays = osc.probe_rays(og, R, np.arange(og.count))
        o, _ = osc.trace(rays, mask=0x3)
        ob, _ = osc.trace(rays, mask=0x3, brute=True)
        assert np.array_equal(o, ob), "oracle BVH and brute force disagree"
        g = _gpu_trace(probes, rays, 0x3)
        _compare_hits(g, o, f"frame {frame}")
    probes.destroy()


def test_shadow_rays_any_hit(small):
    scene, grid, osc = small
    rng = np.random.Generator(np.random.PCG64(5))
    n = 20000
    origin = rng.uniform([0.2, 0.2, 0.2], [19.8, 9.8, 19.8], (n, 3)).astype(np.float32)
    d = -np.array([0.3, -1.0, 0.2])
    d /= np.linalg.norm(d)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = origin
    rays[:, 3] = 0.1
    rays[:, 4:7] = d
    rays[:, 7] = 10000.0
    probes = _make(scene, grid)
    o, _ = osc.trace(rays, mask=0xFF, any_hit=True)
    g = _gpu_trace(probes, rays, 0xFF, any_hit=True)
    assert np.array_equal(o[:, 1] != 0xFFFFFFFF, g[:, 1] != 0xFFFFFFFF)
    # closest-hit variant of the same rays is bit-exact too
    o, _ = osc.trace(rays, mask=0xFF)
    g = _gpu_trace(probes, rays, 0xFF)
    _compare_hits(g, o, "shadow closest")
    probes.destroy()


def _pad(v, down):
    """The build's box padding of a coordinate: the float just below v - (|v| 1e-6 + 1e-6) (or just
    above v + ...), computed in double."""
    v = np.asarray(v, np.float64)
    d = v - (np.abs(v) * 1e-6 + 1e-6) if down else v + (np.abs(v) * 1e-6 + 1e-6)
    f = d.astype(np.float32)
    step = np.float32(-np.inf) if down else np.float32(np.inf)
    off = (f.astype(np.float64) > d) if down else (f.astype(np.float64) < d)
    return np.where(off, np.nextafter(f, step), f).astype(np.float32)


def test_axis_aligned_rays_bit_exact(small):
    """Rays along the coordinate axes, their other direction components +0 or -0 (1/d infinite on
    two axes), a third of them from origins that lie on the planes of triangles (the room and the
    boxes are axis-aligned) or on the padded planes of BVH boxes, where a slab distance is 0 x inf:
    the traversal must keep every box such a ray meets, so its closest hits equal the brute-force
    scan's bit for bit and its any-hit decisions agree."""
    scene, grid, osc = small
    rng = np.random.Generator(np.random.PCG64(17))
    n = 6000
    origin = rng.uniform([0.2, 0.2, 0.2], [19.8, 9.8, 19.8], (n, 3)).astype(np.float32)
    coords = [np.unique(scene.vertices["pos"][:, k]) for k in range(3)]
    for i in range(0, n, 3):
        for k in rng.choice(3, size=int(rng.integers(1, 3)), replace=False):
            v = coords[k][rng.integers(len(coords[k]))]
            origin[i, k] = [v, _pad(v, True), _pad(v, False)][int(rng.integers(3))]
    axis = rng.integers(0, 3, n)
    sign = np.where(rng.integers(0, 2, n) == 0, 1.0, -1.0).astype(np.float32)
    d = np.where(rng.integers(0, 2, (n, 3)) == 0, np.float32(0.0), np.float32(-0.0)).astype(np.float32)
    d[np.arange(n), axis] = sign
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = origin
    rays[:, 3] = 0.01
    rays[:, 4:7] = d
    rays[:, 7] = 1000.0
    assert np.signbit(d).sum() > n // 2 and (d == 0).sum() == 2 * n  # -0 and +0 components both present
    probes = _make(scene, grid)
    want, _ = osc.trace(rays, mask=0xFF, brute=True)
    _compare_hits(_gpu_trace(probes, rays, 0xFF), want, "axis-aligned closest")
    want_any, _ = osc.trace(rays, mask=0xFF, any_hit=True, brute=True)
    got_any = _gpu_trace(probes, rays, 0xFF, any_hit=True)
    assert np.array_equal(want_any[:, 1] != 0xFFFFFFFF, got_any[:, 1] != 0xFFFFFFFF)
    probes.destroy()


def test_edge_fallback_and_ties_bit_exact():
    """The two decisions of the intersection that the random scenes never reach: a ray whose FP32
    edge function is exactly 0 but whose exact one is negative (the FP64 recomputation makes it a
    miss), and two coincident triangles met at the same t (the lower triangle id wins)."""
    from test_oracle_pins import edge_and_tie_scene
    scene, rays = edge_and_tie_scene()
    p = _make(scene, scenes.grid_params((-1, -1, -1), (11, 1, 5), (2, 2, 2), 64))
    got = _gpu_trace(p, rays, 0xFF)
    p.destroy()
    want, _ = orc.OracleScene(scene).trace(rays)
    _compare_hits(got, want, "edge / tie rays")
    assert got[0, 1] == 0xFFFFFFFF and (got[1, 1], got[1, 2]) == (0, 0) and (got[2, 1], got[2, 2]) == (1, 0)


@pytest.mark.parametrize("height,front_up", [(2.0 - 0.5 * np.sqrt(3.0), False), (-0.4, True)])
def test_init_probes_threshold_scenes(height, front_up):
    """Probe init on the two scenes that sit between its thresholds (a back face seen by 3/8 of
    the rays; geometry within reach but beyond one cell): states and back-face counts identical."""
    from test_oracle_probes import _floor_quad_scene
    scene = _floor_quad_scene(height, 8.0, front_up)
    grid = scenes.grid_params((0, 0, 0), (6, 6, 6), (4, 4, 4), 512)
    R = scenes.orientations(1, 1)[0]
    p = _make(scene, grid)
    p.init_probes(R)
    got = p.states()
    p.destroy()
    want, _ = orc.OracleScene(scene).init_probes(orc.Grid(grid), R)
    assert np.array_equal(got, want), f"{np.sum(got != want)} probe states differ"
    assert got[1 + 4 + 16] == (1 if not front_up else 8)


def test_init_probes_states_identical(small):
    scene, grid, osc = small
    og = orc.Grid(grid)
    probes = _make(scene, grid)
    for R in scenes.orientations(3, 2):
        probes.init_probes(R)
        got = probes.states()
        want, _ = osc.init_probes(og, R)
        assert np.array_equal(got, want), f"{np.sum(got != want)} probe states differ"
    assert set(np.unique(got)) <= {0, 1, 8}
    probes.destroy()


RAY_BAR = 1e-4
TEXEL_BAR = 1e-5


def _frame_compare(scene, grid, osc, frames, seed, flags=_lib.UPDATE_ALL_PROBES, probes_per_update=0, states=None):
    """Updates teacher-forced from the GPU state (the oracle starts every frame from the GPU's
    atlases and states).  Per frame: the update's hit records must equal the oracle's; returns the
    fraction of rays within RAY_BAR, the max texel errors over the WHOLE atlases (probes that were
    not updated must keep their tiles) and the fraction of equal states."""
    og = orc.Grid(grid)
    probes = IrradianceProbes(grid, max_rays_per_probe=int(grid["rays_per_probe"]),
                              probes_per_update=probes_per_update)
    probes.set_scene(scene)
    probes.set_light(scenes.sun_light())
    probes.set_sky(scenes.SKY_RADIANCE)
    if states is not None:
        probes.set_states(states)
    slot_ids = probes.slot_ids()
    light = scenes.sun_light()
    stats = []
    for f, R in enumerate(scenes.orientations(seed, frames)):
        irr = probes.irradiance()[..., :3].astype(np.float64)
        dep = probes.depth().astype(np.float64)
        states = probes.states()
        probes.update(R, flags)
        probes.synchronize()
        gray, gidx = probes.last_update_rays()
        ghit = probes.last_update_hit_records(slot_ids)
        res = osc.update(og, light, scenes.SKY_RADIANCE.astype(np.float64), R, gidx, states, irr, dep, want_hits=True)
        bad = np.any(ghit != res["hits"], axis=-1)
        assert not bad.any(), f"frame {f}: {bad.sum()} hit records differ, first gpu {ghit[bad][0]} oracle {res['hits'][bad][0]}"
        ray_err = np.abs(gray - res["rays"]) / np.maximum(np.abs(res["rays"]), 1e-2)
        gi, gd = probes.irradiance(), probes.depth()
        ie = np.abs(gi[..., :3] - irr) / np.maximum(np.abs(irr), 1e-2)
        de = np.abs(gd - dep) / np.maximum(np.abs(dep), 1e-2)
        stats.append((np.mean(ray_err.max(-1) <= RAY_BAR), ie.max(), np.quantile(ie, 0.999), de.max(),
                      np.mean(probes.states() == states), len(gidx)))
    probes.destroy()
    return stats


def _check(stats, what):
    for f, (ok_frac, ie_max, ie_999, de_max, st_eq, count) in enumerate(stats):
        print(f"{what} frame {f}: {count} probes, rays within {RAY_BAR:g} {ok_frac:.5f}, irr max {ie_max:.2e} "
              f"p99.9 {ie_999:.2e}, depth max {de_max:.2e}, states eq {st_eq:.4f}", flush=True)
        assert ok_frac == 1.0 and ie_max <= TEXEL_BAR and de_max <= TEXEL_BAR and st_eq >= 0.99


def test_update_long_grid_parity():
    """A grid of 70 x 2 x 3 probes: the shade list's half-cell keys (140 half cells along x) are
    shifted right to stay below 64 per axis, a path no cubic test grid reaches.  The ordering
    changes no output, so the bars are the update parity's."""
    scene, _ = scenes.config_small(seed=3, open_top=True)
    lo, hi = scenes.inset_extent(scene)
    grid = scenes.grid_params(lo, hi, (70, 2, 3), 64)
    _check(_frame_compare(scene, grid, orc.OracleScene(scene), 2, seed=5), "70x2x3")


def test_update_with_no_front_face_hits():
    """Every probe ray misses (the scene's one quad lies 500 m above the grid, beyond the rays'
    tmax of the grid diagonal): the shade list is empty, so the key sort and the shading run on a
    list of length 0, and every ray is sky."""
    sb = scenes.SceneBuilder()
    m = sb.add_material((0.5, 0.5, 0.5))
    pos = np.array([[0, 500, 0], [1, 500, 0], [1, 500, 1], [0, 500, 1]], np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    sb.add_instance(sb.add_mesh(pos, nrm, np.array([[0, 2, 1], [0, 3, 2]]), m))
    scene = sb.build((-10, -10, -10), (10, 10, 10))
    grid = scenes.grid_params((-8, -8, -8), (8, 8, 8), (4, 2, 4), 64)
    _check(_frame_compare(scene, grid, orc.OracleScene(scene), 2, seed=9), "all-miss")


def test_update_frames_small_parity(small):
    scene, grid, osc = small
    _check(_frame_compare(scene, grid, osc, 4, seed=21), "all probes")


def test_ray_order_follows_each_frame(small):
    """Each update's thread -> ray order (`ray_order` of the device view) lists its rays by the
    world elevation (y) of their directions, ties by ray index, and changes with the rotation; the
    entries past the ray count stay the identity."""
    import torch
    from vulkanexp_amd.probes import _device_copy
    scene, grid, _ = small
    p = _make(scene, grid)
    n = int(grid["rays_per_probe"])
    orders = []
    for R in scenes.orientations(23, 2):
        p.update(R, _lib.UPDATE_ALL_PROBES)
        p.synchronize()
        v = p.device_view()
        dirs = _device_copy(v.ray_directions, n * 4, torch.float32).numpy().reshape(n, 4)
        m = int(v.max_rays_per_probe)
        order = _device_copy(v.ray_order, m, torch.int16).numpy().view(np.uint16).astype(np.int64)
        assert np.array_equal(order[:n], np.argsort(dirs[:, 1], kind="stable")), "not sorted by world y"
        assert np.array_equal(order[n:], np.arange(n, m))
        orders.append(order[:n])
    p.destroy()
    assert not np.array_equal(orders[0], orders[1])


def test_update_reference_schedule_parity(small):
    """The reference's per-frame path: update() with the probes selectProbesToUpdate picks
    (src/IrradianceProbes.cpp:396-424) from random states, without and with a ProbesPerUpdate cap
    below P.  Rays, hit records, the whole atlases (tiles of probes not picked are carried over)
    and the states are compared with the oracle."""
    scene, grid, osc = small
    og = orc.Grid(grid)
    rng = np.random.Generator(np.random.PCG64(17))
    states = rng.choice(np.array([0, 1, 2, 3, 5, 8], np.uint32), og.count)
    stats = _frame_compare(scene, grid, osc, 4, seed=23, flags=0, states=states)
    assert all(0 < s[-1] < og.count for s in stats)
    _check(stats, "reference schedule")
    stats = _frame_compare(scene, grid, osc, 4, seed=24, flags=0, probes_per_update=37, states=states)
    assert all(s[-1] == 37 for s in stats), [s[-1] for s in stats]
    _check(stats, "ProbesPerUpdate 37")


def test_reference_schedule_selection_matches_oracle(small):
    scene, grid, osc = small
    og = orc.Grid(grid)
    probes = _make(scene, grid)
    rng = np.random.Generator(np.random.PCG64(9))
    states = rng.choice(np.array([0, 1, 2, 3, 5, 8], np.uint32), og.count)
    probes.set_states(states)
    last, loop = 0, 0
    for R in scenes.orientations(4, 5):
        want, last, loop = orc.select(probes.states(), 0, last, loop)
        probes.update(R, 0)
        probes.synchronize()
        _, gidx = probes.last_update_rays()
        assert np.array_equal(np.sort(gidx), np.sort(want)) and np.array_equal(gidx, want)
    probes.destroy()


def test_sample_probes_parity(small):
    scene, grid, osc = small
    torch = _torch()
    og = orc.Grid(grid)
    probes = _make(scene, grid)
    for R in scenes.orientations(5, 3):
        probes.update(R, _lib.UPDATE_ALL_PROBES)
    probes.synchronize()
    rng = np.random.Generator(np.random.PCG64(13))
    n = 50000
    pos = rng.uniform([0.0, 0.0, 0.0], [20.0, 10.0, 20.0], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm.astype(np.float32)
    tocam = rng.normal(size=(n, 3))
    tocam /= np.linalg.norm(tocam, axis=1, keepdims=True)
    tocam = tocam.astype(np.float32)
    out = probes.sample(*(torch.from_numpy(a).cuda() for a in (pos, nrm, tocam))).cpu().numpy()
    irr = probes.irradiance()[..., :3].astype(np.float64)
    dep = probes.depth().astype(np.float64)
    want = orc.sample(og, probes.states(), irr, dep, pos, nrm, tocam)
    err = np.abs(out - want) / np.maximum(np.abs(want), 1e-2)
    print(f"sample: max {err.max():.2e}, p99.9 {np.quantile(err, 0.999):.2e}, within 1e-4 {np.mean(err.max(1) <= 1e-4):.5f}")
    assert np.mean(err.max(1) <= 1e-4) >= 0.999
    probes.destroy()
