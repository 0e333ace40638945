"""Pins the oracle's ray/triangle decisions: FP32 watertight test against FP64 geometry, facing,
interval rules, watertightness on shared edges, BVH traversal against brute force, and the
instance flattening of the scene (src/Renderer.cpp:532-551)."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import scenes


def _ray(o, d, tmin=0.0, tmax=1e30):
    return np.array([*o, tmin, *d, tmax], np.float32)


def _mt64(o, d, v0, v1, v2):
    """FP64 Moller-Trumbore: (t, b1, b2, det) without any culling or range test."""
    o, d, v0, v1, v2 = (np.asarray(x, np.float64) for x in (o, d, v0, v1, v2))
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    tv = o - v0
    q = np.cross(tv, e1)
    return (e2 @ q) / det, (tv @ p) / det, (d @ q) / det, det


def test_intersection_matches_fp64_geometry():
    rng = np.random.Generator(np.random.PCG64(11))
    checked = 0
    for _ in range(8000):
        tri = rng.uniform(-1, 1, (3, 3))
        o = rng.uniform(-3, 3, 3)
        target = tri.mean(0) + rng.normal(0, 0.6, 3)
        d = target - o
        d /= np.linalg.norm(d)
        t, b1, b2, det = _mt64(o, d, *tri)
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        cosang = abs(d @ n) / np.linalg.norm(n)
        margin = min(b1, b2, 1 - b1 - b2)
        inside = margin > 0
        if abs(margin) < 1e-4 or cosang < 1e-2 or t < 1e-3:   # ambiguous in FP32: skip
            continue
        ok, out = orc.tri_intersect32(_ray(o, d, 0.0, 1e30), tri.astype(np.float32))
        assert ok == (inside and t > 0)
        if ok:
            checked += 1
            assert out[0] == pytest.approx(t, rel=2e-5, abs=2e-6)
            assert out[1] == pytest.approx(b1, abs=2e-5) and out[2] == pytest.approx(b2, abs=2e-5)
            # facing: det > 0 <=> the ray meets the side cross(v1 - v0, v2 - v0) points to
            assert (out[3] > 0) == (d @ n < 0)
    assert checked > 600


def test_interval_is_closed():
    tri = np.array([[-1, -1, 1], [1, -1, 1], [0, 1, 1]], np.float32)
    ok, out = orc.tri_intersect32(_ray((0, 0, 0), (0, 0, 1), 0.0, 1.0), tri)
    assert ok and out[0] == 1.0                                       # t == tmax accepted
    ok, _ = orc.tri_intersect32(_ray((0, 0, 0), (0, 0, 1), 0.0, np.nextafter(np.float32(1), np.float32(0))), tri)
    assert not ok
    ok, out = orc.tri_intersect32(_ray((0, 0, 0), (0, 0, 1), 1.0, 2.0), tri)
    assert ok and out[0] == 1.0                                       # t == tmin accepted
    ok, _ = orc.tri_intersect32(_ray((0, 0, 0), (0, 0, 1), np.nextafter(np.float32(1), np.float32(2)), 2.0), tri)
    assert not ok


def test_watertight_on_shared_edges():
    # a quad split along its diagonal: rays aimed exactly at points of the shared edge hit at least
    # one of the two triangles (no cracks), for many origins
    rng = np.random.Generator(np.random.PCG64(12))
    for _ in range(3000):
        c = rng.uniform(-5, 5, 3).astype(np.float32)
        a = (c + rng.uniform(-1, 1, 3)).astype(np.float32)
        b = (c + rng.uniform(-1, 1, 3)).astype(np.float32)
        p = (c + rng.uniform(-1, 1, 3)).astype(np.float32)
        q = (c + rng.uniform(-1, 1, 3)).astype(np.float32)
        s = np.float32(rng.uniform(0.05, 0.95))
        on_edge = (a + s * (b - a)).astype(np.float32)
        o = (on_edge + rng.normal(0, 3, 3)).astype(np.float32)
        d = (on_edge - o).astype(np.float32)
        h1, _ = orc.tri_intersect32(_ray(o, d, 0.0, 10.0), np.array([a, b, p]))
        h2, _ = orc.tri_intersect32(_ray(o, d, 0.0, 10.0), np.array([b, a, q]))
        n1, n2 = np.cross(b - a, p - a), np.cross(a - b, q - b)
        # only meaningful when p and q lie on opposite sides (a proper quad) and the ray is not
        # parallel to either triangle
        side = np.cross(b - a, d)
        if (side @ (p - a)) * (side @ (q - a)) >= 0 or abs(d @ n1) < 1e-3 or abs(d @ n2) < 1e-3:
            continue
        assert h1 or h2


def _soup(seed, k=300):
    rng = np.random.Generator(np.random.PCG64(seed))
    sb = scenes.SceneBuilder()
    m = sb.add_material((0.5, 0.5, 0.5))
    centers = rng.uniform(-5, 5, (k, 1, 3))
    pos = (centers + rng.normal(0, 0.5, (k, 3, 3))).reshape(-1, 3).astype(np.float32)
    tris = np.arange(3 * k).reshape(k, 3)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (len(pos), 1))
    sb.add_instance(sb.add_mesh(pos, nrm, tris, m))
    return sb.build((-6, -6, -6), (6, 6, 6), "soup")


def _random_rays(rng, k, lo, hi, tmin=0.0, tmax=100.0):
    o = rng.uniform(lo, hi, (k, 3))
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((k, 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = o, tmin, d, tmax
    return r


@pytest.mark.parametrize("which", ["soup", "room_closed", "room_open"])
def test_bvh_traversal_equals_brute_force(which):
    scene = {"soup": lambda: _soup(1), "room_closed": lambda: scenes.room_boxes_scene(3),
             "room_open": lambda: scenes.room_boxes_scene(3, open_top=True)}[which]()
    osc = orc.OracleScene(scene)
    rng = np.random.Generator(np.random.PCG64(2))
    rays = _random_rays(rng, 20000, scene.bounds_min - 1, scene.bounds_max + 1)
    for any_hit in (False, True):
        a, sa = osc.trace(rays, any_hit=any_hit)
        b, sb = osc.trace(rays, any_hit=any_hit, brute=True)
        if any_hit:
            assert np.array_equal(a[:, 1] != 0xFFFFFFFF, b[:, 1] != 0xFFFFFFFF)
        else:
            assert np.array_equal(a, b)
            assert sa[1] < sb[1] / 4          # the BVH actually culls


def test_closest_hit_is_the_minimum_over_triangles():
    scene = _soup(5, 120)
    osc = orc.OracleScene(scene)
    tris = osc.world_triangles()
    rng = np.random.Generator(np.random.PCG64(7))
    rays = _random_rays(rng, 300, -6, 6)
    hits, _ = osc.trace(rays)
    for r, h in zip(rays, hits):
        best = None
        for gid, tri in enumerate(tris):
            ok, out = orc.tri_intersect32(r, tri)
            if ok and (best is None or out[0] < best[0]):
                best = (out[0], gid)
        if best is None:
            assert h[1] == 0xFFFFFFFF
        else:
            assert h[2] == best[1] and np.float32(best[0]).view(np.uint32) == h[0]


def test_instance_flattening_and_ids():
    scene = scenes.clustered_scene(seed=4, n_instances=8)
    osc = orc.OracleScene(scene)
    tris = osc.world_triangles()
    k = 0
    for inst in scene.instances:
        M = inst["transform"].astype(np.float64).reshape(3, 4)
        oe = scene.offsets[inst["custom_index"]]
        idx = scene.indices[oe["index_offset"]: oe["index_offset"] + 3 * inst["triangle_count"]].reshape(-1, 3)
        p = scene.vertices["pos"][oe["vertex_offset"] + idx].astype(np.float64)
        w = (p @ M[:, :3].T + M[:, 3]).astype(np.float32)
        assert np.allclose(tris[k: k + len(idx)], w, rtol=0, atol=1e-5)
        k += len(idx)
    assert k == osc.tri_count == scene.triangle_count
    # instance / primitive ids of hits point at the triangle that was hit
    rng = np.random.Generator(np.random.PCG64(8))
    rays = _random_rays(rng, 5000, scene.bounds_min, scene.bounds_max)
    hits, _ = osc.trace(rays)
    base = np.concatenate([[0], np.cumsum(scene.instances["triangle_count"].astype(np.int64))[:-1]])
    ok = hits[:, 1] != 0xFFFFFFFF
    assert ok.sum() > 100
    gid = base[hits[ok, 1]] + hits[ok, 2]
    t = hits[ok, 0].view(np.float32)
    for r, g, tt in zip(rays[ok][:200], gid[:200], t[:200]):
        good, out = orc.tri_intersect32(r, tris[g])
        assert good and out[0] == tt


def test_analytic_plane_distances():
    # one large floor quad at y = 0: downward rays from height h hit at t = h / -d.y
    sb = scenes.SceneBuilder()
    m = sb.add_material((0.5, 0.5, 0.5))
    pos = np.array([[-100, 0, -100], [100, 0, -100], [100, 0, 100], [-100, 0, 100]], np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    sb.add_instance(sb.add_mesh(pos, nrm, np.array([[0, 2, 1], [0, 3, 2]]), m))
    scene = sb.build((-100, 0, -100), (100, 0, 100))
    osc = orc.OracleScene(scene)
    rng = np.random.Generator(np.random.PCG64(9))
    k = 5000
    rays = _random_rays(rng, k, [-10, 0.5, -10], [10, 5, 10])
    rays[:, 5] = -np.abs(rays[:, 5]) - 0.05
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)
    hits, _ = osc.trace(rays)
    assert np.all(hits[:, 1] == 0)
    t = hits[:, 0].view(np.float32).astype(np.float64)
    want = rays[:, 1].astype(np.float64) / -rays[:, 5].astype(np.float64)
    # FP32 error of the sheared edge functions scales with the coordinates (|x| <= 100 here)
    assert np.allclose(t, want, rtol=3e-6, atol=3e-5)
    # the floor's CCW normal (+y) faces the rays: front hits
    assert np.all(hits[:, 3] == 0)
    # from below the floor the same plane is hit from the back
    rays[:, 1] = -rays[:, 1]
    rays[:, 5] = -rays[:, 5]
    hits, _ = osc.trace(rays)
    assert np.all(hits[:, 3] == 1)


def test_thread_count_does_not_change_results():
    scene = scenes.room_boxes_scene(2)
    osc = orc.OracleScene(scene)
    rng = np.random.Generator(np.random.PCG64(3))
    rays = _random_rays(rng, 30000, [0, 0, 0], [20, 10, 20])
    orc.set_threads(1)
    a, _ = osc.trace(rays)
    orc.set_threads(8)
    b, _ = osc.trace(rays)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("which", ["soup", "room_closed"])
def test_back_face_culling(which):
    """Culled traversal = the minimum over front-facing triangles only: BVH and brute force agree,
    no culled hit is a back face, and where the two-sided closest hit is front-facing it is also the
    culled hit."""
    scene = {"soup": lambda: _soup(4), "room_closed": lambda: scenes.room_boxes_scene(3)}[which]()
    osc = orc.OracleScene(scene)
    rng = np.random.Generator(np.random.PCG64(21))
    rays = _random_rays(rng, 20000, scene.bounds_min - 1, scene.bounds_max + 1)
    culled, _ = osc.trace(rays, cull_back=True)
    brute, _ = osc.trace(rays, cull_back=True, brute=True)
    assert np.array_equal(culled, brute)
    hit = culled[:, 1] != 0xFFFFFFFF
    assert np.all(culled[hit, 3] == 0)
    plain, _ = osc.trace(rays)
    front = (plain[:, 1] != 0xFFFFFFFF) & (plain[:, 3] == 0)
    assert np.array_equal(plain[front], culled[front])
    # where the two-sided hit is a back face, culling finds a farther surface or nothing
    back = (plain[:, 1] != 0xFFFFFFFF) & (plain[:, 3] == 1)
    assert back.sum() > 100
    t_back = plain[back, 0].view(np.float32)
    t_cull = culled[back, 0].view(np.float32)
    gone = culled[back, 1] == 0xFFFFFFFF
    assert np.all(t_cull[~gone] >= t_back[~gone])
