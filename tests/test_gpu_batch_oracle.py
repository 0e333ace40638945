"""The batched workload against the oracle: 24 probe volumes over 6 seeded scenes (780 to 262,144
triangles), grids from 4x4x4 to 32x16x32 and 64 to 512 rays per probe (scenes.batched_config),
plus a 25th volume of 100 rays per probe (not a multiple of the 64-lane wavefront) on the
262,144-triangle scene, all updated by ddgi_batch_update in one graph launch.

The reference updates one volume per IrradianceProbes::update (src/IrradianceProbes.cpp:426-594);
each batched volume must therefore equal that update on its own: for the compared volumes the
oracle starts from the GPU's atlases and states (teacher forcing) and the hit records (slot ->
instance / primitive) must be bit-exact, >= 99.99 % of the rays within 1e-4 and all within 1e-3,
every texel within 1e-5 relative (max(|oracle|, 1e-2) denominators), as for the large
single-volume update.  A full-workload run checks
the properties of all 25 volumes."""
import os

import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, ProbeBatch, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu

RAY_BAR, RAY_FRACTION, RAY_MAX, TEXEL_BAR = 1e-4, 0.9999, 1e-3, 1e-5
EXTRA = (5, (6, 5, 6), 100)  # scene 5 (262,144 triangles), 180 probes x 100 rays
# compared with the oracle: the smallest grid, both 262,144-triangle volumes of the mix, the
# 100-ray volume and the 32x16x32 x 128-ray volume on the room
COMPARED = (1, 5, 17, 24, 0)


@pytest.fixture(scope="module")
def workload():
    orc.set_threads(min(16, os.cpu_count() or 1))
    scene_list, vols = scenes.batched_config()
    lo, hi = scenes.inset_extent(scene_list[EXTRA[0]])
    vols.append((EXTRA[0], scenes.grid_params(lo, hi, EXTRA[1], EXTRA[2])))
    light = scenes.sun_light()
    handles = []
    for k, (si, grid) in enumerate(vols):
        p = IrradianceProbes(grid, max_rays_per_probe=int(grid["rays_per_probe"]))
        p.set_scene(scene_list[si])
        p.set_light(light)
        p.set_sky(scenes.SKY_RADIANCE)
        p.init_probes(scenes.orientations(200 + k, 1)[0])
        handles.append(p)
    batch = ProbeBatch(handles)
    oracles = {}
    yield scene_list, vols, handles, batch, oracles
    batch.destroy()
    for p in handles:
        p.destroy()


def _orientations(frame, count):
    return np.stack([scenes.orientations(7000 * frame + v, 1)[0] for v in range(count)])


def test_batched_volumes_against_oracle(workload):
    scene_list, vols, handles, batch, oracles = workload
    assert [int(g["rays_per_probe"]) for _, g in vols][-1] == 100
    # two warm-up frames so that shading samples non-zero atlases
    for f in range(2):
        batch.update(_orientations(f, len(handles)))
    for f in (2, 3):
        before = {}
        for v in COMPARED:
            p = handles[v]
            p.synchronize()
            before[v] = (p.irradiance()[..., :3].astype(np.float64), p.depth().astype(np.float64), p.states())
        Rs = _orientations(f, len(handles))
        batch.update(Rs)
        for v in COMPARED:
            si, grid = vols[v]
            p = handles[v]
            p.synchronize()
            if si not in oracles:
                oracles[si] = orc.OracleScene(scene_list[si])
            og = orc.Grid(grid)
            irr, dep, states = before[v]
            gray, gidx = p.last_update_rays()
            assert len(gidx) == og.count
            ghit = p.last_update_hit_records()
            res = oracles[si].update(og, scenes.sun_light(), scenes.SKY_RADIANCE.astype(np.float64), Rs[v], gidx,
                                     states, irr, dep, want_hits=True)
            bad = np.any(ghit != res["hits"], axis=-1)
            assert not bad.any(), f"volume {v} frame {f}: {bad.sum()} hit records differ"
            ray_err = np.abs(gray - res["rays"]) / np.maximum(np.abs(res["rays"]), 1e-2)
            ie = np.abs(p.irradiance()[..., :3] - irr) / np.maximum(np.abs(irr), 1e-2)
            de = np.abs(p.depth() - dep) / np.maximum(np.abs(dep), 1e-2)
            st = np.mean(p.states() == states)
            print(f"volume {v} ({oracles[si].tri_count} tris, {og.count} probes x {int(grid['rays_per_probe'])} rays) "
                  f"frame {f}: rays within {RAY_BAR:g} {np.mean(ray_err.max(-1) <= RAY_BAR):.5f}, irr max {ie.max():.2e}, "
                  f"ray max {ray_err.max():.2e}, depth max {de.max():.2e}, states eq {st:.4f}", flush=True)
            assert np.mean(ray_err.max(-1) <= RAY_BAR) >= RAY_FRACTION and ray_err.max() <= RAY_MAX
            assert ie.max() <= TEXEL_BAR and de.max() <= TEXEL_BAR and st >= 0.99


def test_batched_workload_properties(workload):
    """Ten more batched updates of all 25 volumes (BF16 blend on odd frames): states stay in
    {1..8} (every probe is updated, so none stays disabled), atlases finite and non-negative,
    depth means within the cell diagonal."""
    scene_list, vols, handles, batch, oracles = workload
    for f in range(10):
        batch.update(_orientations(100 + f, len(handles)), _lib.UPDATE_BLEND_BF16 if f % 2 else 0)
    for v, (p, (si, grid)) in enumerate(zip(handles, vols)):
        p.synchronize()
        st = p.states()
        irr, dep = p.irradiance(), p.depth()
        cell = (grid["extent_max"] - grid["extent_min"]) / (np.array(grid["resolution"]) - 1)
        assert st.min() >= 1 and st.max() <= 8, v
        assert np.isfinite(irr).all() and np.isfinite(dep).all(), v
        assert irr[..., :3].min() >= 0.0 and dep[..., 0].min() >= 0.0, v
        assert dep[..., 0].max() <= np.linalg.norm(cell) * (1 + 1e-5), v
    total, per = batch.timings()
    print(f"25 volumes, {sum(p.get_probe_count() for p in handles)} probes: batched update {total:.3f} ms, "
          f"per-volume completion p50 {np.median(per):.3f} ms")
