"""Dynamic scenes on the GPU: instances moved with ddgi_update_instances (re-transform + bottom-up
refit of the BVH, the TLAS update of src/Renderer.cpp:671-742) give exactly the hits of a scene
built from the new transforms -- against the FP64-checked oracle on the moved scene and against a
second handle built from it -- and so the same G-buffer, probe states and atlases; moving the
instances back gives the original results.  Room + boxes scene (open top), 8x4x8 probes x 128 rays;
the moves rotate, translate, scale and mirror boxes (a mirrored instance flips its facing)."""
import dataclasses

import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
from vulkanexp_amd import _lib

pytestmark = pytest.mark.gpu

EYE = (2.0, 7.0, 2.0)


def _moves(scene, ids, seed):
    """New 3x4 transforms for instances `ids`: rotate about the box centre (about y, and one
    mirrored in x, one scaled by 1.4), then shift by up to 2 m in x/z."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = scene.instances["triangle_count"].astype(np.int64)
    out = []
    for j, i in enumerate(ids):
        off = scene.offsets[scene.instances[i]["custom_index"]]
        tri = scene.indices[off["index_offset"]:off["index_offset"] + 3 * counts[i]]
        pos = scene.vertices["pos"][off["vertex_offset"] + tri].astype(np.float64)
        c = 0.5 * (pos.min(0) + pos.max(0))
        a = rng.uniform(0, 2 * np.pi)
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        if j % 5 == 1:
            R = np.diag([-1.0, 1.0, 1.0]) @ R
        if j % 5 == 2:
            R = 1.4 * R
        shift = np.array([rng.uniform(-2, 2), 0.0, rng.uniform(-2, 2)])
        M0 = scene.instances[i]["transform"].astype(np.float64).reshape(3, 4)
        M = np.zeros((3, 4))
        M[:, :3] = R @ M0[:, :3]
        M[:, 3] = R @ (M0[:, 3] - c) + c + shift
        out.append(M.astype(np.float32).reshape(12))
    return np.array(out)


def _moved(scene, ids, xf):
    inst = scene.instances.copy()
    for i, m in zip(ids, xf):
        inst[i]["transform"] = m
    return dataclasses.replace(scene, instances=inst)


def _rays(scene, k=20000, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = np.asarray(scene.bounds_min, np.float64), np.asarray(scene.bounds_max, np.float64)
    o = lo + rng.uniform(0.02, 0.98, (k, 3)) * (hi - lo)
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((k, 8), np.float32)
    r[:, :3], r[:, 4:7], r[:, 7] = o, d, 100.0
    cam = gbuffer.camera_rays(96, 54, eye=EYE, look_at=(12.0, 1.0, 12.0))
    return np.concatenate([r, cam])


@pytest.fixture(scope="module")
def setup():
    import torch
    scene, grid = scenes.config_small(seed=1, open_top=True)
    ids_a = list(range(3, 19))      # one call with a range
    ids_b = [40]                    # a second call with one instance
    xf_a, xf_b = _moves(scene, ids_a, 11), _moves(scene, [40, 40], 12)[1:2]  # index 1: mirrored
    moved = _moved(_moved(scene, ids_a, xf_a), ids_b, xf_b)

    def make(sc):
        p = IrradianceProbes(grid, max_rays_per_probe=128)
        p.set_scene(sc)
        p.set_light(scenes.sun_light())
        p.set_sky(scenes.SKY_RADIANCE)
        return p
    p, fresh = make(scene), make(moved)
    rays = _rays(scene)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    before = gbuffer.hits_from_device(p.trace_rays(t(rays)).cpu().numpy())
    p.update_instances(ids_a[0], xf_a)
    p.update_instances(ids_b[0], xf_b)
    p.synchronize()
    yield {"p": p, "fresh": fresh, "scene": scene, "moved": moved, "grid": grid, "rays": rays, "t": t,
           "before": before, "ids": (ids_a, ids_b)}
    p.destroy()
    fresh.destroy()


def test_moved_hits_match_oracle_and_rebuild(setup):
    s = setup
    got = gbuffer.hits_from_device(s["p"].trace_rays(s["t"](s["rays"])).cpu().numpy())
    want, _ = orc.OracleScene(s["moved"]).trace(s["rays"])
    assert np.array_equal(got, want), f"{np.sum(np.any(got != want, 1))} records differ"
    rebuilt = gbuffer.hits_from_device(s["fresh"].trace_rays(s["t"](s["rays"])).cpu().numpy())
    assert np.array_equal(got, rebuilt)
    # the moves are visible: a sizeable share of the rays sees different geometry
    assert np.mean(np.any(got != s["before"], 1)) > 0.05
    # back faces of the mirrored box are classified as for a rebuilt scene (flip of facing)
    for cull in (True,):
        a = gbuffer.hits_from_device(s["p"].trace_rays(s["t"](s["rays"]), cull_back=cull).cpu().numpy())
        b, _ = orc.OracleScene(s["moved"]).trace(s["rays"], cull_back=cull)
        assert np.array_equal(a, b)


def test_any_hit_and_mask_after_move(setup):
    s = setup
    rays = s["t"](s["rays"])
    for mask in (0xFF, scenes.MASK_STATIC):
        a = s["p"].trace_rays(rays, mask=mask, any_hit=True).cpu().numpy()
        b = s["fresh"].trace_rays(rays, mask=mask, any_hit=True).cpu().numpy()
        assert np.array_equal(a[:, 3] == -1, b[:, 3] == -1)


def test_gbuffer_after_move_matches_rebuild(setup):
    s = setup
    cam = s["t"](gbuffer.camera_rays(96, 54, eye=EYE, look_at=(12.0, 1.0, 12.0)))
    a = s["p"].gbuffer(cam).cpu().numpy()
    b = s["fresh"].gbuffer(cam).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_probe_update_after_move_matches_rebuild(setup):
    """The reference's "Update State" flow after a scene change: initProbes re-classifies the
    probes, then updates blend the moved scene's radiance."""
    s = setup
    for p in (s["p"], s["fresh"]):
        p.init_probes(scenes.orientations(3, 1)[0])
        for R in scenes.orientations(41, 2):
            p.update(R, _lib.UPDATE_ALL_PROBES)
        p.synchronize()
    assert np.array_equal(s["p"].states(), s["fresh"].states())
    assert np.array_equal(s["p"].irradiance().view(np.uint32), s["fresh"].irradiance().view(np.uint32))
    assert np.array_equal(s["p"].depth().view(np.uint32), s["fresh"].depth().view(np.uint32))


def test_move_back_restores_original_hits(setup):
    s = setup
    ids_a, ids_b = s["ids"]
    sc = s["scene"]
    p = s["p"]
    try:
        p.update_instances(ids_a[0], sc.instances["transform"][ids_a])
        p.update_instances(ids_b[0], sc.instances["transform"][ids_b])
        got = gbuffer.hits_from_device(p.trace_rays(s["t"](s["rays"])).cpu().numpy())
        assert np.array_equal(got, s["before"])
    finally:
        # leave the handle moved for any later test of the module
        p.update_instances(ids_a[0], s["moved"].instances["transform"][ids_a])
        p.update_instances(ids_b[0], s["moved"].instances["transform"][ids_b])
        p.synchronize()


def test_update_instances_argument_errors(setup):
    p = setup["p"]
    n = len(setup["scene"].instances)
    eye = np.eye(3, 4, dtype=np.float32).reshape(1, 12)
    with pytest.raises(_lib.DdgiError):
        p.update_instances(n, eye)                      # first past the end
    with pytest.raises(_lib.DdgiError):
        p.update_instances(n - 1, np.repeat(eye, 2, 0))  # range past the end
    bad = eye.copy()
    bad[0, 5] = np.nan
    with pytest.raises(_lib.DdgiError):
        p.update_instances(0, bad)
    with pytest.raises(_lib.DdgiError):
        p.update_instances(0, np.zeros((1, 12), np.float32))  # singular
    p.update_instances(0, np.zeros((0, 12), np.float32))     # empty move: nothing to do
    # a rejected call leaves the scene as it was
    got = gbuffer.hits_from_device(p.trace_rays(setup["t"](setup["rays"])).cpu().numpy())
    want, _ = orc.OracleScene(setup["moved"]).trace(setup["rays"])
    assert np.array_equal(got, want)
