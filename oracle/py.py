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
            "orc_update": (C.c_int, [P, P, P, P, P, P, u32, P, P, P, P, P, P, P, P, C.c_int, P]),
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

    def trace(self, rays: np.ndarray, mask: int = 0xFF, any_hit: bool = False, brute: bool = False):
        """rays: (n, 8) float32 (origin, tmin, direction, tmax) -> (n, 6) uint32 hit records, stats."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((len(rays), 6), np.uint32)
        stats = np.zeros(2, np.uint64)
        lib().orc_trace(self.h, _p(rays), len(rays), mask, int(any_hit), int(brute), _p(out), _p(stats))
        return out, stats

    def probe_rays(self, grid: Grid, R: np.ndarray, indices: np.ndarray) -> np.ndarray:
        indices = np.ascontiguousarray(indices, np.uint32)
        n = int(grid.params["rays_per_probe"])
        out = np.zeros((len(indices) * n, 8), np.float32)
        lib().orc_probe_rays(_p(grid.params), _p(np.ascontiguousarray(R, np.float32)), _p(indices), len(indices), _p(out))
        return out

    def update(self, grid: Grid, light: np.ndarray, sky, R, indices, states, irr, dep, fp32=False,
               want_hits=False, want_aux=False):
        """One probe update frame.  ``states``, ``irr``, ``dep`` are updated in place.  Returns a dict
        with the ray buffer (count, n, 4), optional hit records, max_change, out_of_range, phase ms."""
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
                         _p(mc), _p(oor), int(fp32), _p(ms))
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
