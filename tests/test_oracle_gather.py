"""The oracle's full-frame consumer (G-buffer + final gather) pinned by closed forms: a textbook
floor plane for the G-buffer fields, and the gather's terms on constant probe atlases, where
sampleProbes returns the atlas value exactly (pinned in test_oracle_probes)."""
import numpy as np

from oracle import py as orc
from vulkanexp_amd import scenes


def _floor_scene(color=(0.25, 0.5, 0.75), vcolor=(1.0, 0.5, 2.0), metal=0.0, rough=1.0, emissive=(0.0, 0.0, 0.0)):
    sb = scenes.SceneBuilder()
    m = sb.add_material(color, metallic=metal, roughness=rough, emissive=emissive)
    pos = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    sb.add_instance(sb.add_mesh(pos, nrm, np.array([[0, 2, 1], [0, 3, 2]]), m, color=np.tile(vcolor, (4, 1))))
    return sb.build((-50, 0, -50), (50, 10, 50))


def _down_rays(k, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.uniform([-10, 1, -10], [10, 6, 10], (k, 3))
    d = rng.normal(size=(k, 3))
    d[:, 1] = -np.abs(d[:, 1]) - 0.3
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((k, 8), np.float32)
    r[:, :3], r[:, 4:7], r[:, 7] = o, d, 1e4
    return r


def test_gbuffer_on_a_floor_plane():
    sc = _floor_scene()
    osc = orc.OracleScene(sc)
    rays = _down_rays(2000)
    hits, _ = osc.trace(rays, cull_back=True)
    gb = orc.gbuffer(sc, rays, hits)
    assert np.all(gb[:, 15] == 0)
    o, d = rays[:, :3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    t = -o[:, 1] / d[:, 1]
    assert np.allclose(gb[:, 0:3], o + t[:, None] * d, atol=3e-4)       # position on the plane
    assert np.allclose(gb[:, 1], 0.0, atol=3e-4)
    assert np.allclose(gb[:, 3], t, rtol=1e-5, atol=1e-5)                 # depth = distance to the origin
    assert np.allclose(gb[:, 4:7], [0.0, 1.0, 0.0], atol=1e-12)           # the face normal
    assert np.allclose(gb[:, 8:11], np.array([0.25, 0.5, 0.75]) * [1.0, 0.5, 2.0], rtol=1e-6)  # vertex colour * base
    assert np.allclose(gb[:, 11], 1.0) and np.allclose(gb[:, 7], 0.0)
    # from below the floor is a back face: culled, so the G-buffer misses
    up = rays.copy()
    up[:, 1] = -up[:, 1]
    up[:, 5] = -up[:, 5]
    hits, _ = osc.trace(up, cull_back=True)
    g2 = orc.gbuffer(sc, up, hits)
    assert np.all(g2[:, 3] == 0) and np.all(g2[:, 15] == -1)


def test_gather_closed_forms_on_constant_atlases():
    sc = _floor_scene(emissive=(0.1, 0.2, 0.3))
    osc = orc.OracleScene(sc)
    g = scenes.grid_params((-50, -1, -50), (50, 8, 50), (6, 3, 6), 64)  # contains every floor hit
    grid = orc.Grid(g)
    irr, dep = grid.zero_atlases()
    c = np.array([0.3, 0.6, 0.9])
    irr[:] = c
    dep[..., 0], dep[..., 1] = 100.0, 100.0 ** 2  # far: every probe fully visible
    states = np.ones(grid.count, np.uint32)
    rays = _down_rays(500)
    hits, _ = osc.trace(rays, cull_back=True)
    gb = orc.gbuffer(sc, rays, hits).astype(np.float32)
    gb[:, 15] = 0
    assert np.all(np.abs(gb[:, [0, 2]]) < 50)  # inside the probe grid
    light = scenes.sun_light()
    sky = np.array([0.5, 0.5, 0.5])
    albedo = gb[:, 8:11].astype(np.float64)
    # no sun, no reflection: indirect * diffuse + emissive, with sampleProbes == c
    out = orc.gather(grid, states, irr, dep, osc, light, sky, gb, (0, 5, 0), direct=np.zeros(len(gb)))
    want = c * albedo * (1 - 0.004) + gb[:, 12:15].astype(np.float64)  # the records are FP32
    assert np.allclose(out, want, rtol=1e-9, atol=1e-12)
    # reflection enters as mix(f0, albedo, metal) * reflection (metal 0: f0)
    refl = np.tile([1.0, 2.0, 3.0], (len(gb), 1))
    out2 = orc.gather(grid, states, irr, dep, osc, light, sky, gb, (0, 5, 0), direct=np.zeros(len(gb)), reflection=refl)
    assert np.allclose(out2 - out, 0.004 * refl, rtol=1e-9, atol=1e-12)
    # an open floor under the sun: the hard shadow ray escapes, visibility 1 -> the PBR term
    out3 = orc.gather(grid, states, irr, dep, osc, light, sky, gb, (0, 5, 0))
    pos = gb[:, 0:3].astype(np.float64)
    view = np.array([0.0, 5.0, 0.0]) - pos
    view /= np.linalg.norm(view, axis=1, keepdims=True)
    sun = np.array([orc.pbr([0, 1, 0], view[k], light["color"][:3], light["direction"][:3], albedo[k], 0.0, 1.0)
                    for k in range(len(gb))])
    assert np.allclose(out3 - out, sun, rtol=1e-9, atol=1e-12)
    # misses show the sky
    gm = gb.copy()
    gm[:7, 3] = 0
    out4 = orc.gather(grid, states, irr, dep, osc, light, sky, gm, (0, 5, 0))
    assert np.allclose(out4[:7], sky)
