"""Hand-evaluated cases that pin the oracle's sampleProbes weight chain
(src/shaders/irradiance.glsl:145-237), the probe ray's back-face record
(src/shaders/closesthit.glsl:137-141) and the terms of closest-hit shading
(closesthit.glsl:184-289).

Each case is built so that every factor takes a value written down by hand in the comments:
probes on the corners of a unit cell, constant (or linear) atlas tiles so that bilinear taps are
exact, a to-camera vector opposite the normal so that the shadow bias vanishes, and depth moments
that make the Chebyshev factor a chosen fraction.  A plausible slip in any factor (the trilinear
weights swapped, the back-face weight not squared, the Chebyshev factor squared instead of cubed,
another crush threshold, the fallback mix, the 0.8 back-face depth, the (1 - f0) of the diffuse
colour, the reflected direction of the specular lookup) changes one of the results below."""
import numpy as np
import pytest

from oracle import py as orc
from vulkanexp_amd import scenes

# unit cell: probes at the corners of [0,1]^3, linear index x + 2y + 4z
GRID = scenes.grid_params((0, 0, 0), (1, 1, 1), (2, 2, 2), 64, shadow_bias=0.3)
OG = orc.Grid(GRID)
UP = np.array([0.0, 1.0, 0.0])


def atlases(colors, moments):
    """Image-layout atlases whose probe tiles are constant: colors[p] (rgb), moments[p] (mean,
    mean of squares), border texels included, so every bilinear tap returns the constant."""
    irr, dep = OG.zero_atlases()
    for p in range(8):
        x, y, z = p % 2, (p // 2) % 2, p // 4
        irr[8 * z:8 * z + 8, 8 * (2 * y + x):8 * (2 * y + x) + 8] = colors[p]
        dep[16 * z:16 * z + 16, 16 * (2 * y + x):16 * (2 * y + x) + 16] = moments[p]
    return irr, dep


def sample(states, colors, moments, pos, nrm):
    irr, dep = atlases(colors, moments)
    n = np.asarray(nrm, np.float64)
    return orc.sample(OG, np.asarray(states, np.uint32), irr, dep, np.asarray(pos, np.float64)[None], n[None],
                      -n[None])[0]


S0 = np.array([0.5, 1.0, 0.4])   # sqrt of probe 0's colour
S1 = np.array([1.0, 0.5, 0.8])   # sqrt of probe 1's colour
COLORS = [S0 ** 2, S1 ** 2] + [np.zeros(3)] * 6
TWO_PROBES = [1, 1, 0, 0, 0, 0, 0, 0]  # probes (0,0,0) and (1,0,0) enabled
FAR = (100.0, 1e4)                       # mean 100: every probe is visible (distance <= mean)


def test_trilinear_weights():
    """x = (0.25, 0, 0), normal +y (perpendicular to both probe directions): back-face weight
    ((0 + 1) / 2)^2 + 0.2 = 0.45 for both, visible, not crushed; trilinear 0.75 for probe 0 and
    0.25 for probe 1, each + 0.001.  Total 0.45 * 1.002 > 1/8: the visibility-weighted blend."""
    got = sample(TWO_PROBES, COLORS, [FAR] * 8, (0.25, 0.0, 0.0), UP)
    want = ((0.751 * S0 + 0.251 * S1) / 1.002) ** 2
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_back_face_weight():
    """x = (0.5, 0, 0) (trilinear 0.5 each), normal (1/2, sqrt(3)/2, 0): the direction to probe 1
    is +x, dot 1/2, weight ((1/2 + 1) / 2)^2 + 0.2 = 0.7625; to probe 0 it is -x, dot -1/2,
    weight (1/4)^2 + 0.2 = 0.2625 (neither below the crush threshold)."""
    n = np.array([0.5, np.sqrt(3.0) / 2, 0.0])
    got = sample(TWO_PROBES, COLORS, [FAR] * 8, (0.5, 0.0, 0.0), n)
    want = ((0.2625 * S0 + 0.7625 * S1) / 1.025) ** 2
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_chebyshev_and_crush():
    """x = (0.5, 0, 0), normal +y: both probes 0.5 away, back-face weight 0.45.  Probe 0's moments:
    mean 0.25, mean of squares 0.125, so variance |0.25^2 - 0.125| = 0.0625 and distance - mean =
    0.25: Chebyshev 0.0625 / (0.0625 + 0.0625) = 1/2, cubed 1/8; weight 0.45 / 8 = 0.05625 < 0.2 is
    crushed to 0.05625^3 / 0.2^2.  Probe 1 is visible (mean 100): weight 0.45."""
    w0 = 0.05625 ** 3 / 0.04
    got = sample(TWO_PROBES, COLORS, [(0.25, 0.125)] + [FAR] * 7, (0.5, 0.0, 0.0), UP)
    want = ((w0 * S0 + 0.45 * S1) / (w0 + 0.45)) ** 2
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_fallback_mix():
    """Both probes partly occluded: probe 0 as above (crushed weight 0.05625^3 / 0.04); probe 1 with
    mean 0.25 and mean of squares 0.25 (variance 0.1875): Chebyshev 0.1875 / (0.1875 + 0.0625) =
    3/4, cubed 27/64, weight 0.45 * 27/64 = 0.18984375 < 0.2, crushed to 0.18984375^3 / 0.04.  Total
    visibility weight W = 0.501 (w0 + w1) ~ 0.088 < 1/8, so the result is
    mix(fallback^2, final^2, 8 W) with the fallback weighted by the back-face weights only (equal
    here: the mean of the two square roots)."""
    w0 = 0.05625 ** 3 / 0.04
    w1 = 0.18984375 ** 3 / 0.04
    W = 0.501 * (w0 + w1)
    assert 1e-3 < W < 1 / 8
    got = sample(TWO_PROBES, COLORS, [(0.25, 0.125), (0.25, 0.25)] + [FAR] * 6, (0.5, 0.0, 0.0), UP)
    final = ((w0 * S0 + w1 * S1) / (w0 + w1)) ** 2
    fallback = ((S0 + S1) / 2) ** 2
    want = fallback + (final - fallback) * 8 * W
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_fully_occluded_returns_the_fallback():
    """Chebyshev ~ 5e-4 (mean 0.05, variance 1e-4, 0.45 beyond the mean) for both probes: weights
    floor at 1e-6 and are crushed to 1e-18 / 0.04, so the total stays below 1e-3 (the sum is not
    normalised) and the mix factor 8 W ~ 2e-16: the fallback alone remains."""
    m = (0.05, 0.05 ** 2 + 1e-4)
    got = sample(TWO_PROBES, COLORS, [m, m] + [FAR] * 6, (0.5, 0.0, 0.0), UP)
    np.testing.assert_allclose(got, ((S0 + S1) / 2) ** 2, rtol=1e-12)


# ---- closest-hit records of probe rays -----------------------------------------------------------
A_COEF = np.array([0.2, 0.5, 0.1])   # irradiance tile texel (row, column c) = A + B c, every tile
B_COEF = np.array([0.1, -0.05, 0.02])
ALBEDO = np.array([0.6, 0.3, 0.9])
METAL = 0.25
EMISSIVE = np.array([0.05, 0.0, 0.1])


def _floor_scene():
    """A 100 m floor quad at y = 0 facing +y (material: ALBEDO, METAL, EMISSIVE) and a 100 m ceiling
    quad at y = 5 facing -y (grey), which shadows the floor from a sun above the horizon."""
    sb = scenes.SceneBuilder()
    mf = sb.add_material(ALBEDO, metallic=METAL, roughness=0.5, emissive=EMISSIVE)
    mc = sb.add_material((0.5, 0.5, 0.5))
    quad = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    up = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    sb.add_instance(sb.add_mesh(quad, up, np.array([[0, 2, 1], [0, 3, 2]]), mf))
    sb.add_instance(sb.add_mesh(quad + np.array([0, 5, 0], np.float32), -up, np.array([[0, 1, 2], [0, 2, 3]]), mc))
    return sb.build((-50, 0, -50), (50, 5, 50))


def _floor_frame(light):
    """One update of the 2x2x2 grid spanning [-1, 1]^3 around the floor, every probe enabled with
    the same linear irradiance tile and visible depth moments."""
    scene = _floor_scene()
    osc = orc.OracleScene(scene)
    grid = scenes.grid_params((-1, -1, -1), (1, 1, 1), (2, 2, 2), 512)
    og = orc.Grid(grid)
    irr, dep = og.zero_atlases()
    cols = np.arange(irr.shape[1]) % 8
    irr[:] = A_COEF + B_COEF * cols[None, :, None]
    dep[..., 0], dep[..., 1] = 100.0, 1e4
    states = np.ones(8, np.uint32)
    R = scenes.orientations(77, 1)[0]
    rays = osc.probe_rays(og, R, np.arange(8)).reshape(8, 512, 8)
    res = osc.update(og, light, scenes.SKY_RADIANCE.astype(np.float64), R, np.arange(8), states, irr, dep,
                     want_hits=True)
    return rays, res


def test_back_face_probe_ray_record():
    """Probes below the floor (y = -1) looking up see its back face at t = 1 / d_y: the ray record
    is black with depth 0.8 t (closesthit.glsl:137-141)."""
    rays, res = _floor_frame(scenes.sun_light())
    hits = res["hits"]
    back = (hits[..., 1] == 0) & (hits[..., 3] == 1)
    below = (np.arange(8) // 2) % 2 == 0  # probes x + 2y + 4z with y = 0 sit at world y = -1
    assert back[below].sum() > 50 and not back[~below].any()
    t = hits[..., 0].view(np.float32).astype(np.float64)[back]
    d_y = rays[..., 5][back].astype(np.float64)
    np.testing.assert_allclose(t, 1.0 / d_y, rtol=2e-5)  # FP32 watertight t over a 100 m triangle
    rec = res["rays"][back]
    assert np.all(rec[:, :3] == 0.0)
    # the record's depth is the FP64 hit distance (the plane y = 0 from y = -1: exactly 1 / d_y) x 0.8
    np.testing.assert_allclose(rec[:, 3], 0.8 / d_y, rtol=1e-12)


def _tile_value(direction):
    """Irradiance tile sampled along ``direction`` (z >= 0, no octahedral fold): u = 1/2 + x / (2
    (|x| + |y| + |z|)), tile-local texel coordinate 1 + 6 u, bilinear over a tile linear in the
    column: A + B (1 + 6 u - 1/2)."""
    d = np.asarray(direction, np.float64)
    u = 0.5 + 0.5 * d[..., 0] / np.abs(d).sum(-1)
    return A_COEF + B_COEF * (0.5 + 6 * u)[..., None]


def test_floor_shading_terms():
    """Front-face floor hits inside the cell, shadowed by the ceiling (sun above the horizon):
    radiance = emissive + mix(0.04, albedo, metal) * irradiance along reflect(d, n) +
    albedo * 0.96 * (1 - metal) * irradiance along n.  All eight probes carry the same tile, so
    sampleProbes returns the tile's value along its direction whatever the weights."""
    rays, res = _floor_frame(scenes.sun_light())
    hits = res["hits"]
    d = rays[..., 4:7].astype(np.float64)
    o = rays[..., 0:3].astype(np.float64)
    t = hits[..., 0].view(np.float32).astype(np.float64)
    p = o + t[..., None] * d
    refl = d * np.array([1.0, -1.0, 1.0])
    sel = (hits[..., 1] == 0) & (hits[..., 3] == 0) & (np.abs(p[..., 0]) < 0.95) & (np.abs(p[..., 2]) < 0.95) & \
          (refl[..., 2] >= 0)
    assert sel.sum() >= 20, sel.sum()
    spec = 0.04 * (1 - METAL) + ALBEDO * METAL
    diffuse = ALBEDO * 0.96 * (1 - METAL)
    want = EMISSIVE + spec * _tile_value(refl[sel]) + diffuse * _tile_value(np.array([0.0, 1.0, 0.0]))
    got = res["rays"][sel]
    np.testing.assert_allclose(got[:, :3], want, rtol=1e-6)  # tap positions are FP32, as on the GPU
    # depth: the FP64 hit distance, exactly -1 / d_y from y = 1 down to the plane y = 0
    np.testing.assert_allclose(got[:, 3], -1.0 / d[sel][:, 1], rtol=1e-12)


def test_sun_below_the_horizon_fades_the_whole_colour():
    """Sun 0.2 below the horizon: a floor point's shadow ray (toward the sun, downward) escapes, so
    the point is lit, and closesthit.glsl:283-285 multiplies the whole colour, emissive and probe
    terms included, by 1 - clamp(0.2, 0, 0.1) / 0.1 = 0."""
    light = scenes.sun_light(direction_of_travel=(0.3, 1.0, 0.2))
    assert light["direction"][1] < -0.1
    rays, res = _floor_frame(light)
    hits = res["hits"]
    front_floor = (hits[..., 1] == 0) & (hits[..., 3] == 0)
    assert front_floor.sum() > 50
    assert np.all(res["rays"][front_floor][:, :3] == 0.0)


# ---- intersection: the FP64 edge-function fallback and the closest-hit tie break ----
_EPS = float(np.finfo(np.float32).eps)
# Seen along +z from (0, 0, -1), the edge B C of this triangle passes through the ray's projected
# origin in FP32 arithmetic (Cx By and Cy Bx round to the same float) but not exactly: the exact
# edge function is -2^-46, so the ray misses, just outside the edge, while A's side is inside.
EDGE_TRI = np.array([[-1.0, 1.0, 0.0], [1.0, 1.0 + _EPS, 0.0], [-(1.0 + _EPS), -(1.0 + 2 * _EPS), 0.0]], np.float32)


def edge_and_tie_scene():
    """Instance 0: EDGE_TRI.  Instance 1: two coincident triangles (the same three vertices) at
    z = 5 around x = 10, so that a ray meets both at the same t.  Returns the scene and three rays:
    through the edge (a miss), just inside the edge triangle, and through the coincident pair."""
    sb = scenes.SceneBuilder()
    m = sb.add_material((0.5, 0.5, 0.5))
    up = np.tile(np.array([0, 0, -1], np.float32), (3, 1))
    sb.add_instance(sb.add_mesh(EDGE_TRI, up, np.array([[0, 1, 2]]), m))
    pair = np.array([[9.0, -1.0, 5.0], [11.0, -1.0, 5.0], [10.0, 1.0, 5.0]], np.float32)
    sb.add_instance(sb.add_mesh(pair, up, np.array([[0, 2, 1], [0, 2, 1]]), m))
    scene = sb.build((-2, -2, -2), (12, 2, 6))
    rays = np.array([[0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 1.0, 100.0],
                     [-1e-3, 1e-3, -1.0, 0.0, 0.0, 0.0, 1.0, 100.0],
                     [10.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 100.0]], np.float32)
    return scene, rays


def test_edge_function_fp64_fallback():
    a, b, c = EDGE_TRI.astype(np.float32)
    u32 = np.float32(c[0] * b[1]) - np.float32(c[1] * b[0])        # FP32 edge function of B C
    u64 = float(c[0]) * float(b[1]) - float(c[1]) * float(b[0])
    assert u32 == 0.0 and u64 < 0.0
    ok, _ = orc.tri_intersect32(np.array([0, 0, -1, 0, 0, 0, 1, 100], np.float32), EDGE_TRI)
    assert not ok, "the FP64 recomputation puts the ray outside the edge"
    ok, out = orc.tri_intersect32(np.array([-1e-3, 1e-3, -1, 0, 0, 0, 1, 100], np.float32), EDGE_TRI)
    assert ok and out[0] == 1.0


def test_closest_hit_ties_go_to_the_lower_triangle():
    scene, rays = edge_and_tie_scene()
    for brute in (False, True):
        hits, _ = orc.OracleScene(scene).trace(rays, brute=brute)
        assert hits[0, 1] == 0xFFFFFFFF                               # through the edge: miss
        assert hits[1, 1] == 0 and hits[1, 2] == 0                    # inside the edge triangle
        assert hits[2, 1] == 1 and hits[2, 2] == 0                    # the coincident pair: the lower id
        assert hits[2, 0].view(np.float32) == np.float32(5.0)


@pytest.mark.parametrize("state,change,want", [(2, 0.0125, 2), (4, 0.009, 4), (4, 0.0045, 5)])
def test_state_thresholds(state, change, want):
    """probesUpdate.glsl:110-118 between the thresholds: for s = 2 a change of 0.0125 lies between
    0.02 / s and 0.04 / s (no step); for s = 4, 0.009 lies between 0.005 and 0.01 (no step) and
    0.0045 below 0.005 (one step up)."""
    n = 64
    g = orc.Grid(scenes.grid_params((0, 0, 0), (4, 4, 4), (3, 3, 3), n, hysteresis=0.0))
    dirs = orc.frame_dirs(scenes.orientations(5, 1)[0], n)
    rays = np.zeros((1, n, 4))
    rays[0, :, :3] = 1.0 + change
    rays[0, :, 3] = 0.1
    irr, dep = g.zero_atlases()
    irr[:] = 1.0
    states = np.ones(g.count, np.uint32)
    states[13] = state
    mc, _ = orc.blend(g, dirs, rays, [13], states, irr, dep)
    assert mc[0] == pytest.approx(change, abs=1e-12)
    assert states[13] == want


@pytest.mark.parametrize("out_of_range,want", [(63, 5), (64, 8)])
def test_state_far_probe_needs_every_ray_out_of_range(out_of_range, want):
    """A probe goes to state 8 only when all of its rays are out of range (a miss, or a hit beyond
    the cell diagonal); 63 of 64 keep the normal state machine (probesUpdate.glsl:110-118)."""
    n = 64
    g = orc.Grid(scenes.grid_params((0, 0, 0), (4, 4, 4), (3, 3, 3), n, hysteresis=0.0))
    dirs = orc.frame_dirs(scenes.orientations(5, 1)[0], n)
    rays = np.zeros((1, n, 4))
    rays[0, :, :3] = 1.0
    rays[0, :, 3] = 0.1
    rays[0, :out_of_range, 3] = -1.0
    irr, dep = g.zero_atlases()
    irr[:] = 1.0
    states = np.ones(g.count, np.uint32)
    states[13] = 4
    _, oor = orc.blend(g, dirs, rays, [13], states, irr, dep)
    assert oor[0] == out_of_range
    assert states[13] == want
