"""ctypes binding of libddgi.so (the C-ABI declared in include/ddgi.h).

The library is built in-tree (``vulkanexp_amd/libddgi.so``) by ``__graft_entry__.build()``.  There
is no fallback: if the shared object is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddgi.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ddgi.h")

DDGI_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "HIP_ERROR", 4: "UNSUPPORTED",
                5: "NO_SCENE", 6: "INVALID_STATE", 7: "NO_DEVICE"}

UPDATE_ALL_PROBES = 0x1
UPDATE_BLEND_BF16 = 0x2
UPDATE_BLEND_FP16 = 0x4
UPDATE_HYSTERESIS_RAMP = 0x8
UPDATE_NO_TIMING = 0x10
UPDATE_COUNT_VISITS = 0x20
NODE_BYTES, TRIANGLE_BYTES = 128, 48  # 4-wide BVH node, triangle record
ATLAS_IRRADIANCE, ATLAS_DEPTH = 0, 1
LAYOUT_TILED, LAYOUT_IMAGE = 0, 1
TRACE_ANY_HIT = 0x1
TRACE_CULL_BACK_FACES = 0x2
SKY_CONSTANT = 0
SKY_ATMOSPHERE = 1
GBUFFER_FIELDS = ("position", "depth", "normal", "metalness", "albedo", "roughness", "emissive", "material")


class DdgiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ddgi status {status} ({STATUS_NAMES.get(status, '?')}): {message}")
        self.status = status


class Timings(C.Structure):
    _fields_ = [("select_ms", C.c_float), ("trace_ms", C.c_float), ("blend_ms", C.c_float),
                ("border_ms", C.c_float), ("copy_ms", C.c_float), ("total_ms", C.c_float),
                ("probes_updated", C.c_uint32), ("rays_traced", C.c_uint32),
                ("primary_ms", C.c_float), ("shade_ms", C.c_float)]


class DeviceView(C.Structure):
    _fields_ = [("irradiance", C.c_void_p), ("depth", C.c_void_p), ("states", C.c_void_p), ("rays", C.c_void_p),
                ("hits", C.c_void_p), ("ray_order", C.c_void_p), ("probe_indices", C.c_void_p),
                ("ray_directions", C.c_void_p),
                ("probe_count", C.c_uint64), ("max_rays_per_probe", C.c_uint32), ("last_update_count", C.c_uint32)]


class Options(C.Structure):
    _fields_ = [("max_rays_per_probe", C.c_uint32), ("probes_per_update", C.c_uint32),
                ("target_hysteresis", C.c_float), ("device", C.c_int32), ("reserved", C.c_uint32 * 4)]


class BatchConfig(C.Structure):
    _fields_ = [("trace_block", C.c_uint32), ("trace_item_blocks", C.c_uint32), ("blend_probes", C.c_uint32),
                ("reserved", C.c_uint32 * 5)]

    def as_dict(self) -> dict:
        return {"trace_block": self.trace_block, "trace_item_blocks": self.trace_item_blocks,
                "blend_probes": self.blend_probes}


class BatchTuneRecord(C.Structure):
    _fields_ = [("config", BatchConfig), ("ms", C.c_float), ("phase", C.c_uint32)]


class SceneTables(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint64), ("indices", C.c_void_p),
                ("index_count", C.c_uint64), ("offsets", C.c_void_p), ("offset_count", C.c_uint64),
                ("instances", C.c_void_p), ("instance_count", C.c_uint64), ("materials", C.c_void_p),
                ("material_count", C.c_uint64), ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3),
                ("mesh_count", C.c_uint64), ("node_count", C.c_uint64), ("texture_count", C.c_uint64)]


_P = C.c_void_p
_U32, _U64, _SZ = C.c_uint32, C.c_uint64, C.c_size_t
_SIGNATURES = {
    "ddgi_create": (C.c_int, [_P, _P, C.POINTER(C.c_void_p)]),
    "ddgi_destroy": (C.c_int, [_P]),
    "ddgi_abi_version": (_U32, []),
    "ddgi_status_string": (C.c_char_p, [C.c_int]),
    "ddgi_last_error": (C.c_char_p, [_P]),
    "ddgi_set_params": (C.c_int, [_P, _P]),
    "ddgi_get_params": (C.c_int, [_P, _P]),
    "ddgi_set_light": (C.c_int, [_P, _P]),
    "ddgi_set_sky": (C.c_int, [_P, _P]),
    "ddgi_set_sky_model": (C.c_int, [_P, C.c_uint32]),
    "ddgi_set_scene": (C.c_int, [_P, _P, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _P, _SZ]),
    "ddgi_get_bvh_info": (C.c_int, [_P, _P, _P, _P]),
    "ddgi_get_slot_ids": (C.c_int, [_P, _P, _P, _U64]),
    "ddgi_update_instances": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "ddgi_scene_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "ddgi_scene_load_material": (C.c_int, [_P, C.c_char_p]),
    "ddgi_scene_create": (C.c_int, [_P, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, C.POINTER(C.c_void_p)]),
    "ddgi_scene_get_tables": (C.c_int, [_P, C.POINTER(SceneTables)]),
    "ddgi_scene_save": (C.c_int, [_P, C.c_char_p]),
    "ddgi_scene_destroy": (C.c_int, [_P]),
    "ddgi_init_probes": (C.c_int, [_P, _P, _P]),
    "ddgi_update": (C.c_int, [_P, _P, _U32, _P]),
    "ddgi_blend_rays": (C.c_int, [_P, _P, _P, _U32, _P, _U32, _P]),
    "ddgi_trace_rays": (C.c_int, [_P, _P, _P, _U64, _U32, _U32, _P]),
    "ddgi_sample": (C.c_int, [_P, _P, _P, _P, _U64, _P, _P]),
    "ddgi_get_atlas": (C.c_int, [_P, _U32, _U32, _P, _SZ]),
    "ddgi_set_atlas": (C.c_int, [_P, _U32, _U32, _P, _SZ]),
    "ddgi_get_states": (C.c_int, [_P, _P, _SZ]),
    "ddgi_set_states": (C.c_int, [_P, _P, _SZ]),
    "ddgi_get_timings": (C.c_int, [_P, _P]),
    "ddgi_get_trace_stats": (C.c_int, [_P, _P]),
    "ddgi_get_device_view": (C.c_int, [_P, _P]),
    "ddgi_save_state": (C.c_int, [_P, _P, C.POINTER(C.c_size_t)]),
    "ddgi_load_state": (C.c_int, [_P, _P, _SZ]),
    "ddgi_synchronize": (C.c_int, [_P]),
    "ddgi_diff_prepare": (C.c_int, [_P, _P, _P, _P, _P, _P, _U64, _U32, _P]),
    "ddgi_diff_forward": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "ddgi_diff_backward": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "ddgi_gbuffer": (C.c_int, [_P, _P, _U64, _P, _P]),
    "ddgi_gather": (C.c_int, [_P, _P, _U64, _P, _P, _P, _P, _P]),
    "ddgi_direct_light": (C.c_int, [_P, _P, _U64, _P, _P, _P, _P]),
    "ddgi_reflection": (C.c_int, [_P, _P, _U64, _P, _P, _P, _P]),
    "ddgi_reflection_noalloc": (C.c_int, [_P, _P, _U64, _P, _P, _P, _P]),
    "ddgi_batch_create": (C.c_int, [_P, _U32, _P, C.POINTER(C.c_void_p)]),
    "ddgi_batch_destroy": (C.c_int, [_P]),
    "ddgi_batch_set_config": (C.c_int, [_P, _P]),
    "ddgi_batch_get_config": (C.c_int, [_P, _P]),
    "ddgi_batch_update": (C.c_int, [_P, _P, _U32, _P]),
    "ddgi_batch_get_timings": (C.c_int, [_P, _P, _P, _U32]),
    "ddgi_batch_autotune": (C.c_int, [_P, _P, _U32, _U32, _P, _P, _P, _U32, _P]),
}

_lib = None


def header_functions() -> list[str]:
    """Names of every function declared in include/ddgi.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:ddgi_status|uint32_t|const char\*)\s+(ddgi_\w+)\s*\(", text, re.M)))


def load():
    """Load libddgi.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


def check(status: int, handle=None) -> None:
    if status != DDGI_OK:
        msg = load().ddgi_last_error(handle)
        raise DdgiError(status, msg.decode() if msg else "")
