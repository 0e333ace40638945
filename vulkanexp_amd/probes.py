"""Host-side mirror of VulkanExp's ``IrradianceProbes`` (src/IrradianceProbes.hpp:16-129) over
libddgi.so.  Method names follow the reference (snake_case); every call goes through the C-ABI and
runs on the GPU.  PyTorch is used only for device buffers and streams.

    probes = IrradianceProbes(grid_params, max_rays_per_probe=256)     # init()
    probes.set_scene(scene)                                            # Renderer TLAS/BLAS + descriptors
    probes.set_light(light); probes.set_sky(rgb)                       # setLightBuffer / miss shader
    probes.init_probes(R)                                              # initProbes()
    probes.update(R, flags)                                            # update()
    probes.irradiance(), probes.depth(), probes.states()               # atlas / state getters
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, load
from .scenes import GRID_DTYPE, Scene


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _stream_ptr(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _tensor_ptr(t):
    return C.c_void_p(t.data_ptr())


class IrradianceProbes:
    """One probe volume (a libddgi handle)."""

    MaxRaysPerProbe = 512

    def __init__(self, grid_params: np.ndarray, max_rays_per_probe: int = 512, probes_per_update: int = 0,
                 target_hysteresis: float = 0.98, device: int = -1):
        self._lib = load()
        self.GridParameters = np.array(grid_params, GRID_DTYPE)
        opts = _lib.Options(max_rays_per_probe, probes_per_update, target_hysteresis, device, (C.c_uint32 * 4)())
        h = C.c_void_p()
        check(self._lib.ddgi_create(_p(self.GridParameters), C.byref(opts), C.byref(h)))
        self.handle = h
        self.max_rays_per_probe = max_rays_per_probe

    # -- lifetime -------------------------------------------------------------------------------
    def destroy(self) -> None:
        if getattr(self, "handle", None):
            check(self._lib.ddgi_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- configuration --------------------------------------------------------------------------
    def get_probe_count(self) -> int:
        r = self.GridParameters["resolution"]
        return int(r[0]) * int(r[1]) * int(r[2])

    def update_uniforms(self) -> None:
        check(self._lib.ddgi_set_params(self.handle, _p(self.GridParameters)), self.handle)

    def params(self) -> np.ndarray:
        out = np.zeros((), GRID_DTYPE)
        check(self._lib.ddgi_get_params(self.handle, _p(out)), self.handle)
        return out

    def set_light(self, light: np.ndarray) -> None:
        light = np.ascontiguousarray(light)
        check(self._lib.ddgi_set_light(self.handle, _p(light)), self.handle)

    def set_sky(self, rgb) -> None:
        rgb = np.ascontiguousarray(rgb, np.float32)
        check(self._lib.ddgi_set_sky(self.handle, _p(rgb)), self.handle)

    def set_sky_model(self, model: int) -> None:
        """_lib.SKY_CONSTANT (misses return set_sky's radiance) or _lib.SKY_ATMOSPHERE (sky.glsl)."""
        check(self._lib.ddgi_set_sky_model(self.handle, int(model)), self.handle)

    def set_scene(self, scene: Scene) -> None:
        arrs = [np.ascontiguousarray(a) for a in (scene.vertices, scene.indices, scene.offsets, scene.instances,
                                                   scene.materials)]
        v, i, o, inst, m = arrs
        check(self._lib.ddgi_set_scene(self.handle, _p(v), len(v), _p(i), len(i), _p(o), len(o), _p(inst), len(inst),
                                       _p(m), len(m)), self.handle)

    def update_instances(self, first: int, transforms: np.ndarray, stream=None) -> None:
        """Move instances first .. first + len(transforms) - 1 to new row-major 3x4 transforms
        ((count, 3, 4) or (count, 12) float32) and refit the BVH (src/Renderer.cpp:671-742)."""
        xf = np.ascontiguousarray(transforms, np.float32).reshape(-1, 12)
        check(self._lib.ddgi_update_instances(self.handle, int(first), len(xf), _p(xf), _stream_ptr(stream)),
              self.handle)

    def bvh_info(self) -> dict:
        """nodes, triangle slots ("triangles": a triangle cut by a spatial split has one slot per
        piece) and bytes of the uploaded BVH."""
        n, t, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(self._lib.ddgi_get_bvh_info(self.handle, C.byref(n), C.byref(t), C.byref(b)), self.handle)
        return {"nodes": n.value, "triangles": t.value, "bytes": b.value}

    # -- per-frame work -------------------------------------------------------------------------
    def init_probes(self, orientation: np.ndarray, stream=None) -> None:
        R = np.ascontiguousarray(orientation, np.float32).reshape(9)
        check(self._lib.ddgi_init_probes(self.handle, _p(R), _stream_ptr(stream)), self.handle)

    def update(self, orientation: np.ndarray, flags: int = 0, stream=None) -> None:
        R = np.ascontiguousarray(orientation, np.float32).reshape(9)
        check(self._lib.ddgi_update(self.handle, _p(R), int(flags), _stream_ptr(stream)), self.handle)

    def blend_rays(self, rays, probe_indices, orientation: np.ndarray, flags: int = 0, stream=None) -> None:
        """rays: CUDA float32 tensor (count, n, 4); probe_indices: CUDA int32/uint32 tensor (count,)."""
        R = np.ascontiguousarray(orientation, np.float32).reshape(9)
        check(self._lib.ddgi_blend_rays(self.handle, _tensor_ptr(rays), _tensor_ptr(probe_indices),
                                        int(probe_indices.numel()), _p(R), int(flags), _stream_ptr(stream)), self.handle)

    def gbuffer(self, rays, stream=None):
        """G-buffer of camera rays (a (N, 8) float32 CUDA tensor of ddgi_ray records) -> (N, 16) float32
        CUDA tensor of ddgi_gbuffer_texel records (position, depth, normal, metalness, albedo,
        roughness, emissive, material bits)."""
        import torch
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.shape[-1] == 8
        rays = rays.contiguous()
        out = torch.empty((rays.shape[0], 16), dtype=torch.float32, device=rays.device)
        check(self._lib.ddgi_gbuffer(self.handle, _tensor_ptr(rays), int(rays.shape[0]), _tensor_ptr(out),
                                     _stream_ptr(stream)), self.handle)
        return out

    def gather(self, gbuffer, view_origin, direct=None, reflection=None, stream=None):
        """Final gather (FinalGather.frag:38-77) of a G-buffer -> (N, 3) float32 CUDA tensor."""
        import torch
        gbuffer = gbuffer.contiguous()
        n = int(gbuffer.shape[0])
        out = torch.empty((n, 3), dtype=torch.float32, device=gbuffer.device)
        origin = np.ascontiguousarray(view_origin, np.float32).reshape(3)
        d = direct.contiguous() if direct is not None else None
        r = reflection.contiguous() if reflection is not None else None
        check(self._lib.ddgi_gather(self.handle, _tensor_ptr(gbuffer), n, _p(origin),
                                    _tensor_ptr(d) if d is not None else None,
                                    _tensor_ptr(r) if r is not None else None, _tensor_ptr(out),
                                    _stream_ptr(stream)), self.handle)
        return out

    def direct_light(self, gbuffer, noise=None, previous=None, stream=None):
        """Ray-traced shadow pass (directLight.rgen) of a G-buffer -> (N, 4) float32 CUDA tensor
        (visibility, previous g, previous b, 1), -1 on misses, 0 facing away.  noise: (N, 2) CUDA
        float32 blue-noise samples (None: no jitter); previous: (N, 4) CUDA float32 or None."""
        import torch
        gbuffer = gbuffer.contiguous()
        n = int(gbuffer.shape[0])
        out = torch.empty((n, 4), dtype=torch.float32, device=gbuffer.device)
        nz = noise.contiguous() if noise is not None else None
        pv = previous.contiguous() if previous is not None else None
        check(self._lib.ddgi_direct_light(self.handle, _tensor_ptr(gbuffer), n,
                                          _tensor_ptr(nz) if nz is not None else None,
                                          _tensor_ptr(pv) if pv is not None else None, _tensor_ptr(out),
                                          _stream_ptr(stream)), self.handle)
        return out

    def reflection(self, gbuffer, view_origin, noise=None, stream=None):
        """Reflection pass (reflection.rgen) of a G-buffer -> (N, 4) float32 CUDA tensor
        (compressed rgb, roughness).  noise: (N, 2) CUDA float32 samples or None (mirror rays)."""
        import torch
        gbuffer = gbuffer.contiguous()
        n = int(gbuffer.shape[0])
        out = torch.empty((n, 4), dtype=torch.float32, device=gbuffer.device)
        origin = np.ascontiguousarray(view_origin, np.float32).reshape(3)
        nz = noise.contiguous() if noise is not None else None
        check(self._lib.ddgi_reflection(self.handle, _tensor_ptr(gbuffer), n, _p(origin),
                                        _tensor_ptr(nz) if nz is not None else None, _tensor_ptr(out),
                                        _stream_ptr(stream)), self.handle)
        return out

    def reflection_noalloc(self, gbuffer, view_origin, out, noise=None, stream=None):
        """reflection() without device allocations (ddgi_reflection_noalloc): the caller passes the
        (N, 4) float32 CUDA output tensor; N <= 2,097,152 texels per call."""
        gbuffer = gbuffer.contiguous()
        n = int(gbuffer.shape[0])
        origin = np.ascontiguousarray(view_origin, np.float32).reshape(3)
        nz = noise.contiguous() if noise is not None else None
        check(self._lib.ddgi_reflection_noalloc(self.handle, _tensor_ptr(gbuffer), n, _p(origin),
                                                _tensor_ptr(nz) if nz is not None else None, _tensor_ptr(out),
                                                _stream_ptr(stream)), self.handle)
        return out

    def trace_rays(self, rays, mask: int = 0xFF, any_hit: bool = False, stream=None, cull_back: bool = False):
        """rays: CUDA float32 tensor (N, 8) = (origin, tmin, direction, tmax).  Returns a CUDA int32
        tensor (N, 6): t bits, u bits, v bits, instance, primitive, flags."""
        import torch
        hits = torch.empty((rays.shape[0], 6), dtype=torch.int32, device=rays.device)
        check(self._lib.ddgi_trace_rays(self.handle, _tensor_ptr(rays), _tensor_ptr(hits), int(rays.shape[0]), mask,
                                        (_lib.TRACE_ANY_HIT if any_hit else 0) |
                                        (_lib.TRACE_CULL_BACK_FACES if cull_back else 0), _stream_ptr(stream)), self.handle)
        return hits

    def sample(self, position, normal, to_camera, stream=None):
        """sampleProbes at CUDA float32 (N, 3) points; returns (N, 3)."""
        import torch
        out = torch.empty_like(position)
        check(self._lib.ddgi_sample(self.handle, _tensor_ptr(position), _tensor_ptr(normal), _tensor_ptr(to_camera),
                                    int(position.shape[0]), _tensor_ptr(out), _stream_ptr(stream)), self.handle)
        return out

    def synchronize(self) -> None:
        check(self._lib.ddgi_synchronize(self.handle), self.handle)

    # -- getters --------------------------------------------------------------------------------
    def _atlas(self, kind: int, layout: int) -> np.ndarray:
        r = [int(x) for x in self.GridParameters["resolution"]]
        res, ch = (8, 4) if kind == 0 else (16, 2)
        P = r[0] * r[1] * r[2]
        if layout == _lib.LAYOUT_TILED:
            out = np.zeros((P, res, res, ch), np.float32)
        else:
            out = np.zeros((res * r[2], res * r[0] * r[1], ch), np.float32)
        check(self._lib.ddgi_get_atlas(self.handle, kind, layout, _p(out), out.nbytes), self.handle)
        return out

    def irradiance(self, layout: int = _lib.LAYOUT_IMAGE) -> np.ndarray:
        return self._atlas(_lib.ATLAS_IRRADIANCE, layout)

    def depth(self, layout: int = _lib.LAYOUT_IMAGE) -> np.ndarray:
        return self._atlas(_lib.ATLAS_DEPTH, layout)

    def set_atlas(self, kind: int, data: np.ndarray, layout: int = _lib.LAYOUT_IMAGE) -> None:
        data = np.ascontiguousarray(data, np.float32)
        check(self._lib.ddgi_set_atlas(self.handle, kind, layout, _p(data), data.nbytes), self.handle)

    def states(self) -> np.ndarray:
        out = np.zeros(self.get_probe_count(), np.uint32)
        check(self._lib.ddgi_get_states(self.handle, _p(out), out.size), self.handle)
        return out

    def set_states(self, states: np.ndarray) -> None:
        states = np.ascontiguousarray(states, np.uint32)
        check(self._lib.ddgi_set_states(self.handle, _p(states), states.size), self.handle)

    def timings(self) -> dict:
        t = _lib.Timings()
        check(self._lib.ddgi_get_timings(self.handle, C.byref(t)), self.handle)
        return {k: getattr(t, k) for k, _ in _lib.Timings._fields_}

    def trace_stats(self) -> dict:
        """BVH fetch counts of the last update run with UPDATE_COUNT_VISITS."""
        out = np.zeros(5, np.uint64)
        check(self._lib.ddgi_get_trace_stats(self.handle, _p(out)), self.handle)
        return {"primary_nodes": int(out[0]), "primary_tris": int(out[1]), "shadow_nodes": int(out[2]),
                "shadow_tris": int(out[3]), "shadow_rays": int(out[4])}

    def device_view(self) -> _lib.DeviceView:
        v = _lib.DeviceView()
        check(self._lib.ddgi_get_device_view(self.handle, C.byref(v)), self.handle)
        return v

    def save_state(self) -> bytes:
        n = C.c_size_t(0)
        check(self._lib.ddgi_save_state(self.handle, None, C.byref(n)), self.handle)
        buf = np.zeros(n.value, np.uint8)
        check(self._lib.ddgi_save_state(self.handle, _p(buf), C.byref(n)), self.handle)
        return buf.tobytes()

    def load_state(self, blob: bytes) -> None:
        buf = np.frombuffer(blob, np.uint8).copy()
        check(self._lib.ddgi_load_state(self.handle, _p(buf), buf.size), self.handle)

    def _ray_order(self, v, n):
        import torch
        return _device_copy(v.ray_order, n, torch.int16).numpy().view(np.uint16).astype(np.int64)

    def last_update_rays(self):
        """(rays (count, n, 4) float32 in ray-index order, probe indices (count,)) of the last update."""
        import torch
        v = self.device_view()
        t = self.timings()
        count = t["probes_updated"]
        n = int(self.GridParameters["rays_per_probe"])
        stored = _device_copy(v.rays, count * n * 4, torch.float32).numpy().reshape(count, n, 4)
        rays = np.empty_like(stored)
        rays[:, self._ray_order(v, n)] = stored
        idx = _device_copy(v.probe_indices, count, torch.int32)
        return rays, idx.numpy().view(np.uint32)

    def last_update_hits(self):
        """Internal hit records (count, n, 4) as uint32 of the last update, in ray-index order."""
        import torch
        v = self.device_view()
        count = self.timings()["probes_updated"]
        n = int(self.GridParameters["rays_per_probe"])
        stored = _device_copy(v.hits, count * n * 4, torch.int32).numpy().view(np.uint32).reshape(count, n, 4)
        hits = np.empty_like(stored)
        hits[:, self._ray_order(v, n)] = stored
        return hits

    def slot_ids(self):
        """(instance, primitive) uint32 arrays indexed by BVH triangle slot."""
        tc = self.bvh_info()["triangles"]
        inst = np.zeros(tc, np.uint32)
        prim = np.zeros(tc, np.uint32)
        check(self._lib.ddgi_get_slot_ids(self.handle, _p(inst), _p(prim), tc), self.handle)
        return inst, prim

    def last_update_hit_records(self, slot_ids=None):
        """Hit records of the last update as (count, n, 6) uint32 (t bits, instance, primitive,
        back-face flag, u bits, v bits) in ray-index order, the layout of a ddgi_trace_rays record
        reordered as (t, instance, primitive, flags, u, v); misses: t = -1, ids 0xFFFFFFFF, u = v = 0."""
        h = self.last_update_hits()
        inst, prim = self.slot_ids() if slot_ids is None else slot_ids
        code = h[..., 1]
        miss = code == 0xFFFFFFFF
        slot = np.where(miss, 0, code & 0x7FFFFFFF)
        out = np.empty(h.shape[:-1] + (6,), np.uint32)
        out[..., 0] = h[..., 0]
        out[..., 1] = np.where(miss, 0xFFFFFFFF, inst[slot])
        out[..., 2] = np.where(miss, 0xFFFFFFFF, prim[slot])
        out[..., 3] = np.where(miss, 0, code >> 31)
        out[..., 4] = h[..., 2]
        out[..., 5] = h[..., 3]
        return out


def _device_copy(ptr: int, numel: int, dtype):
    """Copy `numel` elements at device address `ptr` to a host tensor (synchronous)."""
    import torch
    out = torch.empty(numel, dtype=dtype)
    if numel == 0:
        return out
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemcpy.restype = C.c_int
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    err = lib.hipMemcpy(C.c_void_p(out.data_ptr()), C.c_void_p(ptr), out.element_size() * numel, 2)
    if err != 0:
        raise RuntimeError(f"hipMemcpy failed: {err}")
    return out
