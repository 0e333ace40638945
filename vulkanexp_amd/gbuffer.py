"""Synthetic G-buffers for the differentiable-probes workload: a seeded pinhole camera, its primary
hits on visible (front-facing) surfaces, and per texel the position, the shading normal (the
interpolated vertex normal in world space, closesthit.glsl:158-159), the vector to the camera
and the material id (the input of FinalGather.frag:38-77)."""
from __future__ import annotations

import numpy as np

from .scenes import Scene


def camera_rays(width: int, height: int, eye, look_at, fov_y_deg: float = 60.0, tmax: float = 1e4) -> np.ndarray:
    """(width*height, 8) float32 rays (origin, tmin 0, direction, tmax), row-major pixels."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(look_at, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = np.tan(np.radians(fov_y_deg) / 2)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width * 2 - 1
    py = 1 - (np.arange(height) + 0.5) / height * 2
    X, Y = np.meshgrid(px * th * aspect, py * th)
    d = fwd + X[..., None] * right + Y[..., None] * up
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros((height * width, 8), np.float32)
    rays[:, :3] = eye
    rays[:, 4:7] = d.reshape(-1, 3)
    rays[:, 7] = tmax
    return rays


def hits_from_device(h: np.ndarray) -> np.ndarray:
    """ddgi_hit records (t, u, v, instance, primitive, flags) as uint32 -> (t, instance, primitive,
    flags, u, v)."""
    h = np.asarray(h).view(np.uint32)
    return np.stack([h[:, 0], h[:, 3], h[:, 4], h[:, 5], h[:, 1], h[:, 2]], 1)


def build(scene: Scene, rays: np.ndarray, hits: np.ndarray) -> dict:
    """G-buffer of the front-facing hits.  hits: (N, 6) uint32 (t, instance, primitive, flags, u, v)."""
    ok = (hits[:, 1] != 0xFFFFFFFF) & (hits[:, 3] == 0)
    idx = np.nonzero(ok)[0]
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
    nv = scene.vertices["normal"][tri].astype(np.float64)                  # (N, 3, 3)
    b = np.stack([1 - u - v, u, v], 1)
    n_obj = (nv * b[:, :, None]).sum(1)
    n_obj /= np.linalg.norm(n_obj, axis=1, keepdims=True)
    M = inst["transform"].astype(np.float64).reshape(-1, 3, 4)[:, :, :3]
    nmat = np.transpose(np.linalg.inv(M), (0, 2, 1))
    n = np.einsum("kij,kj->ki", nmat, n_obj)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    tocam = -d
    return {"pixel": idx.astype(np.int64), "position": pos.astype(np.float32), "normal": n.astype(np.float32),
            "to_camera": tocam.astype(np.float32), "material": oe["material_index"].astype(np.uint32)}


def perturbed_params(base_albedo: np.ndarray, sun, seed: int, albedo_scale=0.2, sun_scale=1.3, n_slots=256):
    """Target parameters: every material albedo scaled by U[1 - s, 1 + s] per channel, sun x 1.3.
    Albedo tables are padded to n_slots materials (unused slots stay 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = np.zeros((n_slots, 3))
    a[: len(base_albedo)] = base_albedo
    t = a * rng.uniform(1 - albedo_scale, 1 + albedo_scale, a.shape)
    return a, t, np.asarray(sun, np.float64), np.asarray(sun, np.float64) * sun_scale
