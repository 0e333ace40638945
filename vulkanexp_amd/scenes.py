"""Seeded synthetic scenes and probe-grid presets for the DDGI workloads.

Record layouts are the reference's (and ``include/ddgi.h``'s):
  * vertex, 64 B ........ src/vulkan/Vertex.hpp:8-15
  * material, 48 B ...... src/vulkan/Material.hpp:16-25 (metallic must be set: the reference
                           default 1.0 has no diffuse term)
  * offset entry, 12 B .. src/Renderer.hpp:21-25
  * instance ............ src/Renderer.cpp:532-551 (3x4 row-major transform, custom index = mesh)
  * light, 32 B ......... src/Light.hpp:6-9 (direction points toward the light)
  * grid params, 64 B ... src/IrradianceProbes.hpp:49-60

Randomness: numpy's PCG64 bit generator, one ``Generator`` per (scene, purpose) seed.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

VERTEX_DTYPE = np.dtype([("pos", "<f4", 3), ("color", "<f4", 3), ("normal", "<f4", 3),
                         ("tangent", "<f4", 4), ("tex_coord", "<f4", 2), ("padding", "<u4")])
MATERIAL_DTYPE = np.dtype([("metallic", "<f4"), ("roughness", "<f4"), ("base_color", "<f4", 3),
                           ("emissive", "<f4", 3), ("albedo_texture", "<u4"), ("normal_texture", "<u4"),
                           ("metallic_roughness_texture", "<u4"), ("emissive_texture", "<u4")])
OFFSET_DTYPE = np.dtype([("material_index", "<u4"), ("vertex_offset", "<u4"), ("index_offset", "<u4")])
INSTANCE_DTYPE = np.dtype([("transform", "<f4", 12), ("custom_index", "<u4"), ("mask", "<u4"),
                           ("flags", "<u4"), ("triangle_count", "<u4")])
LIGHT_DTYPE = np.dtype([("direction", "<f4", 4), ("color", "<f4", 4)])
GRID_DTYPE = np.dtype([("extent_min", "<f4", 3), ("depth_sharpness", "<f4"), ("extent_max", "<f4", 3),
                       ("hysteresis", "<f4"), ("resolution", "<i4", 3), ("rays_per_probe", "<u4"),
                       ("color_res", "<u4"), ("depth_res", "<u4"), ("shadow_bias", "<f4"), ("padding", "<u4")])
assert VERTEX_DTYPE.itemsize == 64 and MATERIAL_DTYPE.itemsize == 48 and OFFSET_DTYPE.itemsize == 12
assert INSTANCE_DTYPE.itemsize == 64 and LIGHT_DTYPE.itemsize == 32 and GRID_DTYPE.itemsize == 64

NO_TEXTURE = 0xFFFFFFFF
MASK_STATIC = 0x1
IDENTITY_3x4 = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


@dataclasses.dataclass
class Scene:
    vertices: np.ndarray
    indices: np.ndarray
    offsets: np.ndarray
    instances: np.ndarray
    materials: np.ndarray
    bounds_min: np.ndarray
    bounds_max: np.ndarray
    name: str = ""

    @property
    def triangle_count(self) -> int:
        return int(self.instances["triangle_count"].sum())


class SceneBuilder:
    """Accumulates meshes (one offset-table entry each) and instances."""

    def __init__(self):
        self.vertices, self.indices, self.offsets, self.instances, self.materials = [], [], [], [], []
        self._nv = 0
        self._ni = 0

    def add_material(self, albedo, metallic=0.0, roughness=1.0, emissive=(0.0, 0.0, 0.0)) -> int:
        m = np.zeros((), MATERIAL_DTYPE)
        m["metallic"], m["roughness"] = metallic, roughness
        m["base_color"], m["emissive"] = albedo, emissive
        for k in ("albedo_texture", "normal_texture", "metallic_roughness_texture", "emissive_texture"):
            m[k] = NO_TEXTURE
        self.materials.append(m)
        return len(self.materials) - 1

    def add_mesh(self, pos: np.ndarray, nrm: np.ndarray, tris: np.ndarray, material: int, color=None) -> int:
        v = np.zeros(len(pos), VERTEX_DTYPE)
        v["pos"], v["normal"] = pos, nrm
        v["color"] = 1.0 if color is None else color
        v["tangent"] = (1.0, 0.0, 0.0, 1.0)
        self.vertices.append(v)
        self.indices.append(np.ascontiguousarray(tris, np.uint32).reshape(-1))
        e = np.zeros((), OFFSET_DTYPE)
        e["material_index"], e["vertex_offset"], e["index_offset"] = material, self._nv, self._ni
        self.offsets.append(e)
        self._nv += len(pos)
        self._ni += tris.size
        return len(self.offsets) - 1

    def add_instance(self, mesh: int, transform=IDENTITY_3x4, mask=MASK_STATIC, flags=0):
        ntri = (self.indices[mesh].size) // 3
        i = np.zeros((), INSTANCE_DTYPE)
        i["transform"] = np.asarray(transform, np.float32).reshape(12)
        i["custom_index"], i["mask"], i["flags"], i["triangle_count"] = mesh, mask, flags, ntri
        self.instances.append(i)

    def build(self, bmin, bmax, name="") -> Scene:
        return Scene(np.concatenate(self.vertices), np.concatenate(self.indices),
                     np.array(self.offsets, OFFSET_DTYPE), np.array(self.instances, INSTANCE_DTYPE),
                     np.array(self.materials, MATERIAL_DTYPE), np.asarray(bmin, np.float32),
                     np.asarray(bmax, np.float32), name)


def box_mesh(lo, hi, segments=(1, 1, 1), inward=False):
    """Axis-aligned box with flat per-face normals; face (axis a, sign s) tessellated into
    segments[b] x segments[c] quads (b, c = the other axes in cyclic order).  Outward faces wind
    counter-clockwise seen from outside (cross(e1, e2) points out); ``inward`` flips both."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    pos, nrm, tris = [], [], []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        nb, nc = segments[b], segments[c]
        for s in (-1.0, 1.0):
            base = sum(len(p) for p in pos)
            ub = np.linspace(lo[b], hi[b], nb + 1)
            uc = np.linspace(lo[c], hi[c], nc + 1)
            gb, gc = np.meshgrid(ub, uc, indexing="ij")
            p = np.zeros((nb + 1, nc + 1, 3))
            p[..., a] = hi[a] if s > 0 else lo[a]
            p[..., b], p[..., c] = gb, gc
            pos.append(p.reshape(-1, 3))
            out = s * (-1.0 if inward else 1.0)
            n = np.zeros(3)
            n[a] = out
            nrm.append(np.tile(n, ((nb + 1) * (nc + 1), 1)))
            ii, jj = np.meshgrid(np.arange(nb), np.arange(nc), indexing="ij")
            v00 = base + ii * (nc + 1) + jj
            v10, v11, v01 = v00 + (nc + 1), v00 + (nc + 1) + 1, v00 + 1
            if out > 0:   # cross(e_b, e_c) = +e_a
                t = np.stack([np.stack([v00, v10, v11], -1), np.stack([v00, v11, v01], -1)], -2)
            else:
                t = np.stack([np.stack([v00, v11, v10], -1), np.stack([v00, v01, v11], -1)], -2)
            tris.append(t.reshape(-1, 3))
    return np.concatenate(pos).astype(np.float32), np.concatenate(nrm).astype(np.float32), np.concatenate(tris)


def latlong_sphere(stacks=16, slices=32, radius_fn=None):
    """Unit lat-long sphere: (stacks+1)*(slices+1) vertices, stacks*slices*2 triangles (the pole rows
    hold degenerate triangles).  ``radius_fn(theta, phi)`` displaces radially; normals are the
    area-weighted mean of the incident faces."""
    th = np.linspace(0.0, math.pi, stacks + 1)
    ph = np.linspace(0.0, 2.0 * math.pi, slices + 1)
    T, PH = np.meshgrid(th, ph, indexing="ij")
    r = np.ones_like(T) if radius_fn is None else radius_fn(T, PH)
    pos = np.stack([r * np.sin(T) * np.cos(PH), r * np.cos(T), r * np.sin(T) * np.sin(PH)], -1).reshape(-1, 3)
    ii, jj = np.meshgrid(np.arange(stacks), np.arange(slices), indexing="ij")
    v00 = ii * (slices + 1) + jj
    v10, v11, v01 = v00 + slices + 1, v00 + slices + 2, v00 + 1
    # outward winding: seen from outside, (v00, v01, v11) is counter-clockwise
    tris = np.stack([np.stack([v00, v01, v11], -1), np.stack([v00, v11, v10], -1)], -2).reshape(-1, 3)
    p = pos[tris]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    nrm = np.zeros_like(pos)
    for k in range(3):
        np.add.at(nrm, tris[:, k], fn)
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.maximum(ln, 1e-30), pos / np.maximum(np.linalg.norm(pos, axis=1, keepdims=True), 1e-30))
    return pos.astype(np.float32), nrm.astype(np.float32), tris


def room_boxes_scene(seed: int = 1, n_boxes: int = 64, room=(20.0, 10.0, 20.0), room_albedo=0.5,
                     open_top: bool = False) -> Scene:
    """A closed room (12 inward triangles) plus ``n_boxes`` axis-aligned boxes of 12 outward
    triangles, sizes U[0.3, 2]^3 m, fully inside the room, albedo U[0.1, 0.9]^3, metallic 0,
    roughness 1.  64 boxes give 780 triangles.

    In the closed room every shadow ray toward the sun hits the ceiling, so the only direct light is
    the ceiling's own (its shadow ray starts past it, tmin 0.1, with NdotL clamped to 0.001).
    ``open_top`` drops the 2 ceiling triangles so that sun and sky reach the interior."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sb = SceneBuilder()
    room_mat = sb.add_material((room_albedo,) * 3)
    p, n, t = box_mesh((0, 0, 0), room, inward=True)
    if open_top:
        up = np.all(n == np.array([0, -1, 0], np.float32), axis=1)  # inward ceiling normal points down
        t = t[~np.all(up[t], axis=1)]
    sb.add_instance(sb.add_mesh(p, n, t, room_mat))
    room = np.asarray(room)
    for _ in range(n_boxes):
        size = rng.uniform(0.3, 2.0, 3)
        lo = rng.uniform(0.0, 1.0, 3) * (room - size)
        albedo = rng.uniform(0.1, 0.9, 3)
        m = sb.add_material(albedo)
        p, n, t = box_mesh(lo, lo + size)
        sb.add_instance(sb.add_mesh(p, n, t, m))
    return sb.build((0, 0, 0), room, f"room_boxes(seed={seed},boxes={n_boxes})")


def _transform(scale, angle, translate):
    c, s = math.cos(angle), math.sin(angle)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    m = np.zeros((3, 4))
    m[:, :3] = rot * scale
    m[:, 3] = translate
    return m.astype(np.float32).reshape(12)


def clustered_scene(seed: int = 2, n_instances: int = 256, extent=(60.0, 20.0, 40.0), n_materials=256,
                    base_scale: float = 1.5) -> Scene:
    """Instance 0: a 60x40 m floor grid of 1024 triangles; instances 1..: 1024-triangle meshes, half
    radially displaced lat-long spheres, half 8x8x12-segment boxes, each with its own mesh and
    material (albedo U[0.05, 0.95]^3), uniform scale base_scale x lognormal(sigma 0.5) (capped at
    5 m), a random yaw, positions clustered around 8 centres.  256 instances give 262,144 triangles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sb = SceneBuilder()
    mats = [sb.add_material(rng.uniform(0.05, 0.95, 3)) for _ in range(n_materials)]
    ex = np.asarray(extent)
    # floor: 32 x 16 quads in x, z at y = 0, normal +y
    fx, fz = np.meshgrid(np.linspace(0, ex[0], 33), np.linspace(0, ex[2], 17), indexing="ij")
    fpos = np.stack([fx, np.zeros_like(fx), fz], -1).reshape(-1, 3).astype(np.float32)
    fn = np.tile(np.array([0, 1, 0], np.float32), (len(fpos), 1))
    ii, jj = np.meshgrid(np.arange(32), np.arange(16), indexing="ij")
    v00 = ii * 17 + jj
    v10, v11, v01 = v00 + 17, v00 + 18, v00 + 1
    # +y normal: cross(e1, e2) must point up: (v00, v01, v11) with x along i and z along j
    ftris = np.stack([np.stack([v00, v01, v11], -1), np.stack([v00, v11, v10], -1)], -2).reshape(-1, 3)
    sb.add_instance(sb.add_mesh(fpos, fn, ftris, mats[0]))
    centres = np.stack([rng.uniform(6, ex[0] - 6, 8), rng.uniform(4, ex[2] - 4, 8)], -1)
    for k in range(1, n_instances):
        scale = base_scale * math.exp(rng.normal(0.0, 0.5))
        scale = min(scale, 5.0)
        cx, cz = centres[rng.integers(0, 8)] + rng.normal(0.0, 4.0, 2)
        cx, cz = float(np.clip(cx, 2, ex[0] - 2)), float(np.clip(cz, 2, ex[2] - 2))
        cy = scale + rng.uniform(0.0, 0.6 * ex[1] - 2 * scale) if 0.6 * ex[1] > 2 * scale else scale
        if k % 2 == 1:
            amp = rng.uniform(0.05, 0.25, 3)
            fr = rng.integers(1, 6, 3)
            phs = rng.uniform(0, 2 * math.pi, 3)

            def radius(T, PH, amp=amp, fr=fr, phs=phs):
                return 1.0 + np.sin(T) * sum(a * np.sin(f * T + q) * np.cos(f * PH + q) for a, f, q in zip(amp, fr, phs))

            p, n, t = latlong_sphere(16, 32, radius)
        else:
            half = rng.uniform(0.5, 1.0, 3)
            p, n, t = box_mesh(-half, half, segments=(8, 8, 12))
        mesh = sb.add_mesh(p, n, t, mats[k % n_materials])
        sb.add_instance(mesh, _transform(scale, rng.uniform(0, 2 * math.pi), (cx, cy, cz)))
    return sb.build((0, 0, 0), ex, f"clustered(seed={seed},instances={n_instances})")


def sun_light(direction_of_travel=(0.3, -1.0, 0.2), radiance=3.0) -> np.ndarray:
    """Directional light record: the reference's direction points TOWARD the light, so it is the
    negated direction of travel."""
    d = -np.asarray(direction_of_travel, np.float64)
    d /= np.linalg.norm(d)
    light = np.zeros((), LIGHT_DTYPE)
    light["direction"] = (*d, 0.0)
    light["color"] = (radiance, radiance, radiance, 1.0)
    return light


def grid_params(extent_min, extent_max, resolution, rays_per_probe, hysteresis=0.97, depth_sharpness=12.0,
                shadow_bias=0.3) -> np.ndarray:
    g = np.zeros((), GRID_DTYPE)
    g["extent_min"], g["extent_max"] = extent_min, extent_max
    g["resolution"], g["rays_per_probe"] = resolution, rays_per_probe
    g["hysteresis"], g["depth_sharpness"], g["shadow_bias"] = hysteresis, depth_sharpness, shadow_bias
    g["color_res"], g["depth_res"] = 8, 16
    return g


def inset_extent(scene: Scene, inset=0.5):
    """Probe-grid extent = scene bounds shrunk by ``inset`` metres on every side, so that edge probes
    are not placed exactly on the walls (the reference puts them on the bounds)."""
    return scene.bounds_min + inset, scene.bounds_max - inset


def gen_basis(z: np.ndarray):
    """genBasis (src/IrradianceProbes.cpp:347-355) in float64."""
    b1 = np.array([0.0, 1.0, 0.0]) if z[0] > 0.9 else np.array([1.0, 0.0, 0.0])
    b1 = b1 - z * np.dot(b1, z)
    b1 /= np.linalg.norm(b1)
    return b1, np.cross(z, b1)


def orientation(rng: np.random.Generator) -> np.ndarray:
    """Per-frame random rotation (src/IrradianceProbes.cpp:455-460): Z uniform on the sphere (the
    reference's unseeded sphericalRand replaced by a seeded generator), X, Y from genBasis; the
    result is row-major R with rows X, Y, Z, i.e. transpose(mat3(X, Y, Z)), as float32."""
    z = rng.normal(size=3)
    z /= np.linalg.norm(z)
    x, y = gen_basis(z)
    return np.stack([x, y, z]).astype(np.float32)


def orientations(seed: int, frames: int) -> list[np.ndarray]:
    rng = np.random.Generator(np.random.PCG64(seed))
    return [orientation(rng) for _ in range(frames)]


# ---- workload presets -------------------------------------------------------------------------
def config_small(seed=1, rays=128, open_top=False):
    """Plain DDGI, small: room + 64 boxes (780 tris), 8x4x8 probes, 128 rays, hysteresis 0.97."""
    sc = room_boxes_scene(seed, open_top=open_top)
    lo, hi = inset_extent(sc)
    return sc, grid_params(lo, hi, (8, 4, 8), rays)


def config_large(seed=2, rays=256):
    """Plain DDGI, large: 262,144 tris, 32x16x32 probes, 256 rays, hysteresis 0.97."""
    sc = clustered_scene(seed)
    lo, hi = inset_extent(sc)
    return sc, grid_params(lo, hi, (32, 16, 32), rays)


SKY_RADIANCE = np.array([0.5, 0.5, 0.5], np.float32)


BATCH_SCENE_INSTANCES = (None, 2, 8, 25, 80, 256)  # None: the 780-triangle room


def batched_scenes(seed=5):
    """Six scenes at log-spaced sizes: the room + 64 boxes (780 triangles), then clustered scenes of
    2, 8, 25, 80 and 256 instances of 1024 triangles (2,048 ... 262,144 triangles)."""
    out = []
    for k, n in enumerate(BATCH_SCENE_INSTANCES):
        out.append(room_boxes_scene(seed + k) if n is None else clustered_scene(seed + k, n_instances=n))
    return out


def batched_volumes(seed=5, count=24, n_scenes=6):
    """Volume specs of the batched workload: volume i uses scene i % n_scenes; its grid is
    (4*8^s, 4*4^s, 4*8^s) probes rounded, s ~ U[0, 1] (log-uniform between 4x4x4 and 32x16x32;
    volume 0 is the largest, volume 1 the smallest grid), rays per probe drawn from
    {64, 128, 256, 512}.  Returns [(scene index, resolution, rays)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = rng.uniform(0.0, 1.0, count)
    s[0], s[1] = 1.0, 0.0
    rays = rng.choice(np.array([64, 128, 256, 512]), count)
    out = []
    for i in range(count):
        res = (int(round(4 * 8 ** s[i])), int(round(4 * 4 ** s[i])), int(round(4 * 8 ** s[i])))
        out.append((i % n_scenes, res, int(rays[i])))
    return out


def batched_config(seed=5, count=24):
    """Batched volumes: ([scene], [(scene index, grid params)]) with each grid spanning its scene's
    bounds inset by 0.5 m."""
    scenes = batched_scenes(seed)
    vols = []
    for si, res, n in batched_volumes(seed, count, len(scenes)):
        lo, hi = inset_extent(scenes[si])
        vols.append((si, grid_params(lo, hi, res, n)))
    return scenes, vols


# ---- scene files (.scene / .gltf / .glb / .mat) through libddgi's loaders ------------------------
def _scene_from_handle(lib, h, name: str) -> Scene:
    import ctypes as C
    from . import _lib
    t = _lib.SceneTables()
    _lib.check(lib.ddgi_scene_get_tables(h, C.byref(t)))

    def arr(ptr, count, dtype):
        if count == 0:
            return np.zeros(0, dtype)
        buf = (C.c_char * (int(count) * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype).copy()
    return Scene(arr(t.vertices, t.vertex_count, VERTEX_DTYPE), arr(t.indices, t.index_count, np.uint32),
                 arr(t.offsets, t.offset_count, OFFSET_DTYPE), arr(t.instances, t.instance_count, INSTANCE_DTYPE),
                 arr(t.materials, t.material_count, MATERIAL_DTYPE), np.array(t.bounds_min[:], np.float32),
                 np.array(t.bounds_max[:], np.float32), name)


def load(path: str, materials=()) -> Scene:
    """Load a .scene, .gltf, .glb or .obj file (Scene::load, src/Scene.cpp:30-49), add the materials of the
    given .mat files, and flatten it into the tables ddgi_set_scene takes."""
    import ctypes as C
    import os
    from . import _lib
    lib = _lib.load()
    h = C.c_void_p()
    _lib.check(lib.ddgi_scene_load(os.fsencode(path), C.byref(h)))
    try:
        for m in materials:
            _lib.check(lib.ddgi_scene_load_material(h, os.fsencode(m)))
        return _scene_from_handle(lib, h, os.path.basename(path))
    finally:
        lib.ddgi_scene_destroy(h)


def save(scene: Scene, path: str) -> None:
    """Write ``scene`` (static instances) as a .scene container (Scene::save, src/Scene.cpp:710-816)."""
    import ctypes as C
    import os
    from . import _lib
    lib = _lib.load()
    arrs = [np.ascontiguousarray(a) for a in (scene.vertices, scene.indices, scene.offsets, scene.instances,
                                              scene.materials)]
    args = []
    for a in arrs:
        args += [a.ctypes.data_as(C.c_void_p), len(a)]
    h = C.c_void_p()
    _lib.check(lib.ddgi_scene_create(*args, C.byref(h)))
    try:
        _lib.check(lib.ddgi_scene_save(h, os.fsencode(path)))
    finally:
        lib.ddgi_scene_destroy(h)
