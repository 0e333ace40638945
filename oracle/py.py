"""TEST INFRASTRUCTURE — ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Only tests/, the smoke test and bench.py's cpu_baseline leg import this module; the product
package never does.  Atlases here use the reference's 2D image layout in float64:
irradiance (H, W, 3) with W = 8*Rx*Ry, H = 8*Rz; depth (H, W, 2) with W = 16*Rx*Ry, H = 16*Rz.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

HIT_FIELDS = ("t_bits", "instance", "primitive", "flags", "u_bits", "v_bits")


def build(force: bool = False) -> None:
    args = ["make", "-C", _HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u32, u64 = C.c_uint32, C.c_uint64
        sig = {
            "orc_set_threads": (None, [u32]),
            "orc_get_threads": (u32, []),
            "orc_fib64": (None, [u32, u32, P]),
            "orc_fib_table32": (None, [u32, u32, P]),
            "orc_oct_decode": (None, [C.c_double, C.c_double, P]),
            "orc_oct_encode": (None, [P, P]),
            "orc_oct_uv": (None, [P, P]),
            "orc_texel_dirs": (None, [u32, P]),
            "orc_tri_intersect32": (C.c_int, [P, P, P]),
            "orc_scene_create": (P, [P, u64, P, u64, P, u64, P, u64, P, u64]),
            "orc_scene_destroy": (None, [P]),
            "orc_scene_tri_count": (u64, [P]),
            "orc_scene_world_tris": (None, [P, P]),
            "orc_trace": (None, [P, P, u64, u32, C.c_int, C.c_int, P, P]),
            "orc_probe_rays": (None, [P, P, P, u32, P]),
            "orc_update": (C.c_int, [P, P, P, P, P, P, u32, P, P, P, P, P, P, P, P, C.c_int, P, C.c_int]),
            "orc_sky": (None, [u64, P, P, P, P, C.c_double, C.c_int, P]),
            "orc_sky_scale": (C.c_double, [C.c_double]),
            "orc_sin_series": (C.c_double, [C.c_float]),
            "orc_cos_series": (C.c_double, [C.c_float]),
            "orc_direct_light": (None, [P, P, P, u64, P, P, P]),
            "orc_reflection": (None, [P, P, P, P, C.c_int, P, P, P, P, u64, P, P, P, P]),
            "orc_shade_aux_size": (u64, []),
            "orc_blend": (None, [P, P, P, P, u32, P, P, P, P, P, C.c_int]),
            "orc_sample": (None, [P, P, P, P, P, P, P, u64, P, C.c_int]),
            "orc_border_copy": (None, [P, P, u32, P, P]),
            "orc_diff": (C.c_double, [P, P, P, P, P, P, u32, P, P, u32, P, P, P, P, P, P, u64, P, P, P, P]),
            "orc_pbr": (None, [P, P, P, P, P, C.c_double, C.c_double, P]),
            "orc_init_probes": (None, [P, P, P, P, P]),
            "orc_select": (u32, [P, u32, u32, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype, f.argtypes = res, args
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> int:
    lib().orc_set_threads(int(n))
    return int(lib().orc_get_threads())


def fib64(i: int, n: int) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_fib64(i, n, _p(out))
    return out


def fib_table32(count: int, n: int) -> np.ndarray:
    out = np.zeros((count, 3), np.float32)
    lib().orc_fib_table32(count, n, _p(out))
    return out


def oct_decode(ox: float, oy: float) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_oct_decode(ox, oy, _p(out))
    return out


def oct_encode(v) -> np.ndarray:
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(2)
    lib().orc_oct_encode(_p(v), _p(out))
    return out


def oct_uv(d) -> np.ndarray:
    d = np.ascontiguousarray(d, np.float64)
    out = np.zeros(2)
    lib().orc_oct_uv(_p(d), _p(out))
    return out


def texel_dirs(kind: int) -> np.ndarray:
    r = 6 if kind == 0 else 14
    out = np.zeros((r, r, 3))
    lib().orc_texel_dirs(kind, _p(out))
    return out


def tri_intersect32(ray, tri):
    ray = np.ascontiguousarray(ray, np.float32)
    tri = np.ascontiguousarray(tri, np.float32)
    out = np.zeros(4, np.float32)
    ok = lib().orc_tri_intersect32(_p(ray), _p(tri), _p(out))
    return bool(ok), out


def pbr(n, v, light_color, light_dir, albedo, metal, rough):
    a = [np.ascontiguousarray(x, np.float64) for x in (n, v, light_color, light_dir, albedo)]
    out = np.zeros(3)
    lib().orc_pbr(*[_p(x) for x in a], float(metal), float(rough), _p(out))
    return out


def select(states: np.ndarray, cap: int, last_offset: int, loop_index: int):
    states = np.ascontiguousarray(states, np.uint32)
    out = np.zeros(len(states), np.uint32)
    lo = np.array([last_offset], np.uint32)
    li = np.array([loop_index], np.uint32)
    cnt = lib().orc_select(_p(states), len(states), cap, _p(lo), _p(li), _p(out))
    return out[:cnt].copy(), int(lo[0]), int(li[0])


def hysteresis_ramp(probe_count: int, counts, target: float = 0.98, hysteresis: float = 0.0,
                    uploaded: float = 0.0, updated: int = 0):
    """The hysteresis each of a run of updates blends with (the grid UBO's value when the update is
    recorded), src/IrradianceProbes.cpp:462-476: ``counts`` are the probes each update selects;
    ``hysteresis`` is GridParameters.hysteresis (0 after init, IrradianceProbes.hpp:53 and
    .cpp:358), ``uploaded`` the value last written to the UBO (updateUniforms, .cpp:106-108),
    ``updated`` the function-static counter (.cpp:463).  Before every update, once ``updated``
    reaches the probe count: while |h - target| > 0.05 (a float distance against the double 0.05)
    the UBO receives h BEFORE h moves a tenth of the way to the target (float arithmetic), so the
    kernels lag one step; then h snaps to the target and is uploaded at once.  Returns the list of
    values used and the final (hysteresis, uploaded, updated)."""
    f = np.float32
    h, up, t = f(hysteresis), f(uploaded), f(target)
    used = []
    for c in counts:
        if updated >= probe_count:
            if float(abs(f(h - t))) > 0.05:
                up = h
                h = f(h + f(f(0.1) * f(t - h)))
            elif h != t:
                h = t
                up = t
            updated -= probe_count
        updated += int(c)
        used.append(float(up))
    return used, (float(h), float(up), updated)


class Grid:
    def __init__(self, params: np.ndarray):
        self.params = np.array(params)  # 0-d structured GRID_DTYPE
        r = [int(x) for x in self.params["resolution"]]
        self.res = r
        self.count = r[0] * r[1] * r[2]
        self.irr_shape = (8 * r[2], 8 * r[0] * r[1], 3)
        self.dep_shape = (16 * r[2], 16 * r[0] * r[1], 2)

    def zero_atlases(self):
        return np.zeros(self.irr_shape), np.zeros(self.dep_shape)


class OracleScene:
    """The oracle's copy of a scene (flattened FP32 world triangles + its own BVH)."""

    def __init__(self, scene):
        self.scene = scene
        self._arrays = [np.ascontiguousarray(a) for a in (scene.vertices, scene.indices, scene.offsets,
                                                            scene.instances, scene.materials)]
        v, i, o, inst, m = self._arrays
        self.h = lib().orc_scene_create(_p(v), len(v), _p(i), len(i), _p(o), len(o), _p(inst), len(inst), _p(m), len(m))
        self.tri_count = int(lib().orc_scene_tri_count(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_destroy(self.h)
            self.h = None

    def world_triangles(self) -> np.ndarray:
        out = np.zeros((self.tri_count, 3, 3), np.float32)
        lib().orc_scene_world_tris(self.h, _p(out))
        return out

    def trace(self, rays: np.ndarray, mask: int = 0xFF, any_hit: bool = False, brute: bool = False,
              cull_back: bool = False):
        """rays: (n, 8) float32 (origin, tmin, direction, tmax) -> (n, 6) uint32 hit records, stats.
        cull_back: ignore back-facing triangles."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((len(rays), 6), np.uint32)
        stats = np.zeros(2, np.uint64)
        flags = (1 if any_hit else 0) | (2 if cull_back else 0)
        lib().orc_trace(self.h, _p(rays), len(rays), mask, flags, int(brute), _p(out), _p(stats))
        return out, stats

    def probe_rays(self, grid: Grid, R: np.ndarray, indices: np.ndarray) -> np.ndarray:
        indices = np.ascontiguousarray(indices, np.uint32)
        n = int(grid.params["rays_per_probe"])
        out = np.zeros((len(indices) * n, 8), np.float32)
        lib().orc_probe_rays(_p(grid.params), _p(np.ascontiguousarray(R, np.float32)), _p(indices), len(indices), _p(out))
        return out

    def update(self, grid: Grid, light: np.ndarray, sky, R, indices, states, irr, dep, fp32=False,
               want_hits=False, want_aux=False, sky_model=0):
        """One probe update frame.  ``states``, ``irr``, ``dep`` are updated in place.  Returns a dict
        with the ray buffer (count, n, 4), optional hit records, max_change, out_of_range, phase ms.
        ``sky_model`` 1: misses see the atmospheric sky (sky_atmosphere) instead of ``sky``."""
        indices = np.ascontiguousarray(indices, np.uint32)
        count = len(indices)
        n = int(grid.params["rays_per_probe"])
        assert states.dtype == np.uint32 and states.flags.c_contiguous
        assert irr.shape == grid.irr_shape and dep.shape == grid.dep_shape
        rays = np.zeros((count, n, 4))
        hits = np.zeros((count, n, 6), np.uint32) if want_hits else None
        aux = np.zeros(count * n * int(lib().orc_shade_aux_size()), np.uint8) if want_aux else None
        mc = np.zeros(count)
        oor = np.zeros(count, np.uint32)
        ms = np.zeros(3)
        sky = np.ascontiguousarray(sky, np.float64)
        lib().orc_update(self.h, _p(grid.params), _p(light), _p(sky), _p(np.ascontiguousarray(R, np.float32)),
                         _p(indices), count, _p(states), _p(irr), _p(dep), _p(rays), _p(hits), _p(aux),
                         _p(mc), _p(oor), int(fp32), _p(ms), int(sky_model))
        res = {"rays": rays, "max_change": mc, "out_of_range": oor,
               "trace_ms": ms[0], "blend_ms": ms[1], "border_ms": ms[2]}
        if want_hits:
            res["hits"] = hits
        if want_aux:
            res["aux"] = aux.view(AUX_DTYPE).reshape(count, n)
        return res

    def init_probes(self, grid: Grid, R):
        states = np.zeros(grid.count, np.uint32)
        back = np.zeros(grid.count, np.uint32)
        lib().orc_init_probes(self.h, _p(grid.params), _p(np.ascontiguousarray(R, np.float32)), _p(states), _p(back))
        return states, back


AUX_DTYPE = np.dtype([("c0", "<f8", 3), ("A", "<f8", 3), ("B", "<f8", 3), ("C", "<f8", 3),
                      ("lit", "<u4"), ("material", "<u4")])


class DiffProblem:
    """Inputs of the differentiable update (see orc_diff in ddgi_oracle.cpp): the per-ray factors
    of one update, the atlases/states it leaves fixed, and a G-buffer with targets."""

    def __init__(self, grid: Grid, states, irr_prev, dep, dirs, indices, aux, pos, nrm, tocam, gmat, n_materials):
        self.grid = grid
        self.states = np.ascontiguousarray(states, np.uint32)
        self.irr_prev = np.ascontiguousarray(irr_prev, np.float64)
        self.dep = np.ascontiguousarray(dep, np.float64)
        self.dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, np.uint32)
        self.aux = np.ascontiguousarray(aux)
        assert self.aux.dtype == AUX_DTYPE
        self.pos, self.nrm, self.tocam = (np.ascontiguousarray(a, np.float64).reshape(-1, 3) for a in (pos, nrm, tocam))
        self.gmat = np.ascontiguousarray(gmat, np.uint32)
        self.M = int(n_materials)
        self.N = len(self.gmat)

    def run(self, albedo, sun, target, grad=False):
        albedo = np.ascontiguousarray(albedo, np.float64).reshape(self.M, 3)
        sun = np.ascontiguousarray(sun, np.float64).reshape(3)
        target = np.ascontiguousarray(target, np.float64).reshape(self.N, 3)
        shade = np.zeros((self.N, 3))
        irr = np.zeros_like(self.irr_prev)
        ga = np.zeros((self.M, 3)) if grad else None
        gs = np.zeros(3) if grad else None
        loss = lib().orc_diff(_p(self.grid.params), _p(self.states), _p(self.irr_prev), _p(self.dep), _p(self.dirs),
                              _p(self.indices), len(self.indices), _p(self.aux), _p(albedo), self.M, _p(sun),
                              _p(self.pos), _p(self.nrm), _p(self.tocam), _p(self.gmat), _p(target), self.N,
                              _p(irr), _p(shade), _p(ga), _p(gs))
        return {"loss": loss, "shade": shade, "irradiance": irr, "g_albedo": ga, "g_sun": gs}


def blend(grid: Grid, dirs: np.ndarray, rays: np.ndarray, indices, states, irr, dep, fp32=False):
    indices = np.ascontiguousarray(indices, np.uint32)
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    rays = np.ascontiguousarray(rays, np.float64)
    mc = np.zeros(len(indices))
    oor = np.zeros(len(indices), np.uint32)
    lib().orc_blend(_p(grid.params), _p(dirs), _p(rays), _p(indices), len(indices), _p(states), _p(irr), _p(dep),
                    _p(mc), _p(oor), int(fp32))
    return mc, oor


def border_copy(grid: Grid, indices, irr, dep) -> None:
    indices = np.ascontiguousarray(indices, np.uint32)
    assert irr.flags.c_contiguous and dep.flags.c_contiguous
    lib().orc_border_copy(_p(grid.params), _p(indices), len(indices), _p(irr), _p(dep))


def sample(grid: Grid, states, irr, dep, pos, nrm, tocam, fp32=False) -> np.ndarray:
    pos, nrm, tocam = (np.ascontiguousarray(a, np.float64).reshape(-1, 3) for a in (pos, nrm, tocam))
    out = np.zeros_like(pos)
    lib().orc_sample(_p(grid.params), _p(np.ascontiguousarray(states, np.uint32)), _p(np.ascontiguousarray(irr)),
                     _p(np.ascontiguousarray(dep)), _p(pos), _p(nrm), _p(tocam), len(pos), _p(out), int(fp32))
    return out


def frame_dirs(R: np.ndarray, n: int) -> np.ndarray:
    """World ray directions of one frame, exactly as traced (FP32 R * fib32, no FMA)."""
    f = fib_table32(n, n)
    R = np.asarray(R, np.float32)
    out = np.zeros((n, 3), np.float32)
    for k in range(3):
        out[:, k] = (R[k, 0] * f[:, 0] + R[k, 1] * f[:, 1]) + R[k, 2] * f[:, 2]
    return out


# ---- full-frame consumer --------------------------------------------------------------------------
GBUFFER_COLUMNS = 16  # position(3) depth normal(3) metalness albedo(3) roughness emissive(3) material


def gbuffer(scene, rays: np.ndarray, hits: np.ndarray) -> np.ndarray:
    """G-buffer texels (N, 16) float64 of camera rays in FP64 (material id in column 15, -1 on a miss).

    Follows GBuffer.vert.glsl:35-48 and GBuffer.frag:32-66: position = o + t d of the nearest
    front-facing hit (``hits`` from OracleScene.trace(..., cull_back=True); back faces are culled,
    src/RenderPasses/GBufferPipeline.cpp:212), depth = |position - o|, normal = normalize(transpose(
    inverse(M)) * sum_i b_i n_i), albedo = sum_i b_i colour_i * baseColorFactor, metalness,
    roughness and emissive factors of the material.  A miss leaves depth 0 and carries the ray
    direction in the normal columns (for the gather's sky)."""
    rays = np.asarray(rays, np.float32).reshape(-1, 8)
    out = np.zeros((len(rays), GBUFFER_COLUMNS))
    out[:, 15] = -1.0
    out[:, 4:7] = rays[:, 4:7]
    hit = hits[:, 1] != 0xFFFFFFFF
    idx = np.nonzero(hit)[0]
    if len(idx) == 0:
        return out
    h = hits[idx]
    t = h[:, 0].view(np.float32).astype(np.float64)
    u = h[:, 4].view(np.float32).astype(np.float64)
    v = h[:, 5].view(np.float32).astype(np.float64)
    o = rays[idx, :3].astype(np.float64)
    d = rays[idx, 4:7].astype(np.float64)
    pos = o + t[:, None] * d
    inst = scene.instances[h[:, 1]]
    oe = scene.offsets[inst["custom_index"]]
    base = oe["index_offset"].astype(np.int64) + 3 * h[:, 2].astype(np.int64)
    tri = scene.indices[base[:, None] + np.arange(3)].astype(np.int64) + oe["vertex_offset"].astype(np.int64)[:, None]
    b = np.stack([1 - u - v, u, v], 1)[:, :, None]
    n_obj = (scene.vertices["normal"][tri].astype(np.float64) * b).sum(1)
    col = (scene.vertices["color"][tri].astype(np.float64) * b).sum(1)
    M = inst["transform"].astype(np.float64).reshape(-1, 3, 4)[:, :, :3]
    nmat = np.transpose(np.linalg.inv(M), (0, 2, 1))
    n = np.einsum("kij,kj->ki", nmat, n_obj)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    mat = scene.materials[oe["material_index"]]
    out[idx, 0:3] = pos
    out[idx, 3] = np.linalg.norm(pos - o, axis=1)
    out[idx, 4:7] = n
    out[idx, 7] = mat["metallic"]
    out[idx, 8:11] = col * mat["base_color"].astype(np.float64)
    out[idx, 11] = mat["roughness"]
    out[idx, 12:15] = mat["emissive"]
    out[idx, 15] = oe["material_index"]
    return out


def sky_atmosphere(origins, dirs, light: np.ndarray, brightness=None, show_sun=True) -> np.ndarray:
    """sky() of src/shaders/sky.glsl:60-127 in FP64 for rays (origins, dirs: (N, 3), taken as
    float32 values) under ``light`` (LIGHT_DTYPE record); the sun colour is scaled by
    ``brightness`` (default: the light colour's 4th component, as miss.rmiss:22 passes it)."""
    o = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(-1, 3))
    d = np.ascontiguousarray(np.asarray(dirs, np.float32).reshape(-1, 3))
    if len(o) == 1 and len(d) > 1:
        o = np.ascontiguousarray(np.repeat(o, len(d), 0))
    assert o.shape == d.shape
    sun = np.ascontiguousarray(light["direction"][:3], np.float32)
    col = np.ascontiguousarray(light["color"][:3], np.float32)
    b = float(light["color"][3]) if brightness is None else float(brightness)
    out = np.zeros((len(d), 3))
    lib().orc_sky(len(d), _p(o), _p(d), _p(sun), _p(col), b, int(bool(show_sun)), _p(out))
    return out


def sky_scale(c: float) -> float:
    """scale() of sky.glsl:44-48: O'Neil's polynomial fit of the optical depth along a ray leaving
    the ground at cos(zenith angle) c, in units of the atmosphere thickness."""
    return float(lib().orc_sky_scale(float(c)))


def sincos_series(x: float):
    """(sin, cos) of an FP32 angle as the camera-effect kernels compute them (Taylor series after a
    Cody-Waite reduction, all FP32)."""
    return float(lib().orc_sin_series(float(x))), float(lib().orc_cos_series(float(x)))


def direct_light(oscene: "OracleScene", light: np.ndarray, gb: np.ndarray, noise=None, previous=None) -> np.ndarray:
    """directLight.rgen:42-99 over G-buffer texels ``gb`` ((N, 16) float32): (N, 4) float32, exact
    (the FP32 ray directions and the FP32 intersection replica are the kernels' operations)."""
    gb = np.ascontiguousarray(np.asarray(gb, np.float32).reshape(-1, GBUFFER_COLUMNS))
    sd = np.ascontiguousarray(shadow_direction32(light["direction"]))
    nz = None if noise is None else np.ascontiguousarray(np.asarray(noise, np.float32).reshape(-1, 2))
    pv = None if previous is None else np.ascontiguousarray(np.asarray(previous, np.float32).reshape(-1, 4))
    out = np.zeros((len(gb), 4), np.float32)
    lib().orc_direct_light(oscene.h, _p(sd), _p(gb), len(gb), _p(nz), _p(pv), _p(out))
    return out


def reflection(grid: Grid, states, irr, dep, oscene: "OracleScene", light: np.ndarray, sky, gb: np.ndarray, origin,
               noise=None, sky_model=0, want_hits=False):
    """reflection.rgen:101-175 over G-buffer texels ``gb``: (N, 4) float64 (colorCompression(rgb),
    roughness); with ``want_hits`` also the reflected rays' hit records (N, 6) uint32.  Ray
    directions and intersections FP32 as the kernel forms them, shading in FP64."""
    gb = np.ascontiguousarray(np.asarray(gb, np.float32).reshape(-1, GBUFFER_COLUMNS))
    nz = None if noise is None else np.ascontiguousarray(np.asarray(noise, np.float32).reshape(-1, 2))
    o = np.ascontiguousarray(origin, np.float32)
    sky = np.ascontiguousarray(sky, np.float64)
    states = np.ascontiguousarray(states, np.uint32)
    irr = np.ascontiguousarray(irr, np.float64)
    dep = np.ascontiguousarray(dep, np.float64)
    out = np.zeros((len(gb), 4))
    hits = np.zeros((len(gb), 6), np.uint32) if want_hits else None
    lib().orc_reflection(oscene.h, _p(grid.params), _p(light), _p(sky), int(sky_model), _p(states), _p(irr), _p(dep),
                         _p(gb), len(gb), _p(o), _p(nz), _p(out), _p(hits))
    return (out, hits) if want_hits else out


def shadow_direction32(light_dir) -> np.ndarray:
    """normalize(light direction) in FP32, operation by operation as the library computes it."""
    L = np.asarray(light_dir, np.float32)[:3]
    s = np.float32(np.float32(L[0] * L[0]) + np.float32(L[1] * L[1]))
    s = np.float32(s + np.float32(L[2] * L[2]))
    return (L / np.sqrt(s)).astype(np.float32)


def gather(grid: Grid, states, irr, dep, oscene: OracleScene, light: np.ndarray, sky, gb: np.ndarray, origin,
           direct=None, reflection=None, sky_model=0) -> np.ndarray:
    """Final gather (FinalGather.frag:38-77) in FP64 over G-buffer texels ``gb`` ((N, 16) float32
    records as the library writes them).  Miss (depth <= 0): the constant sky, or with
    ``sky_model`` 1 the atmospheric sky from ``origin`` along the ray direction the texel carries in
    its normal, sun colour scale 1 (FinalGather.frag:52-54).  Hit:
    visibility * pbrMetallicRoughness(n, view, L colour, L, albedo, metal, rough) + mix(f0, albedo,
    metal) * reflection + sampleProbes(position, n, view) * albedo (1 - f0)(1 - metal) + emissive,
    f0 = 0.004, view = normalize(origin - position).  ``direct`` None: visibility from a hard shadow
    ray (directLight.rgen:66-90 without its jitter): the FP32 ray from the texel position along
    normalize(L) (tmin 0.01, tmax 10000, any hit) when dot(normalize(L), n) > 0 in FP32; 1 if it
    escapes (times 1 - clamp(-L.y, 0, 0.1) / 0.1 below the horizon), else 0."""
    gb = np.asarray(gb, np.float32).reshape(-1, GBUFFER_COLUMNS)
    N = len(gb)
    out = np.zeros((N, 3))
    sky = np.asarray(sky, np.float64)
    miss = gb[:, 3] <= 0
    out[miss] = sky
    if sky_model == 1 and miss.any():
        out[miss] = sky_atmosphere(np.asarray(origin, np.float32)[None, :], gb[miss, 4:7], light, brightness=1.0)
    idx = np.nonzero(~miss)[0]
    if len(idx) == 0:
        return out
    g = gb[idx]
    pos32 = g[:, 0:3]
    n32 = g[:, 4:7]
    # FP32 unit normal and visibility test, as the kernel computes them
    ln = np.sqrt((n32[:, 0] * n32[:, 0] + n32[:, 1] * n32[:, 1]) + n32[:, 2] * n32[:, 2])
    nn32 = (n32 / ln[:, None]).astype(np.float32)
    sd = shadow_direction32(light["direction"])
    if direct is None:
        facing = ((sd[0] * nn32[:, 0] + sd[1] * nn32[:, 1]) + sd[2] * nn32[:, 2]) > 0
        vis = np.zeros(len(idx))
        f = np.nonzero(facing)[0]
        if len(f):
            rays = np.zeros((len(f), 8), np.float32)
            rays[:, 0:3] = pos32[f]
            rays[:, 3] = 0.01
            rays[:, 4:7] = sd
            rays[:, 7] = 10000.0
            hits, _ = oscene.trace(rays, mask=0xFF, any_hit=True)
            lit = hits[:, 1] == 0xFFFFFFFF
            fade = 1.0 - min(max(-float(sd[1]), 0.0), 0.1) / 0.1 if sd[1] < 0 else 1.0
            vis[f[lit]] = fade
    else:
        vis = np.asarray(direct, np.float64).reshape(-1)[idx]
    pos = pos32.astype(np.float64)
    n = n32.astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    view = np.asarray(origin, np.float64)[None, :] - pos
    view /= np.linalg.norm(view, axis=1, keepdims=True)
    albedo = g[:, 8:11].astype(np.float64)
    metal = g[:, 7].astype(np.float64)
    rough = g[:, 11].astype(np.float64)
    lc = np.asarray(light["color"][:3], np.float64)
    ld = np.asarray(light["direction"][:3], np.float64)
    color = np.zeros((len(idx), 3))
    for k in range(len(idx)):
        if vis[k] != 0.0:
            color[k] = vis[k] * pbr(n[k], view[k], lc, ld, albedo[k], metal[k], rough[k])
    f0 = 0.004
    diffuse = albedo * (1 - f0) * (1 - metal)[:, None]
    if reflection is not None:
        spec = f0 * (1 - metal)[:, None] + albedo * metal[:, None]
        color += spec * np.asarray(reflection, np.float64).reshape(-1, 3)[idx]
    color += sample(grid, states, irr, dep, pos, n, view) * diffuse
    color += g[:, 12:15].astype(np.float64)
    out[idx] = color
    return out
