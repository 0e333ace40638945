"""Many probe volumes updated by one graph launch (ddgi_batch_* of include/ddgi.h).

    batch = ProbeBatch([probes_a, probes_b, ...])          # IrradianceProbes handles, one device
    best, records = batch.autotune(orientations, flags)     # optional: pick tile shapes
    batch.update(orientations, flags)                       # orientations: (N, 3, 3), one per volume
    total_ms, per_volume_ms = batch.timings()

The update is bit-identical to ``probes.update(R, UPDATE_ALL_PROBES | flags)`` on each volume.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, load
from .probes import _stream_ptr


class ProbeBatch:
    def __init__(self, volumes, config: dict | None = None):
        self._lib = load()
        self.volumes = list(volumes)
        arr = (C.c_void_p * len(self.volumes))(*[v.handle.value for v in self.volumes])
        cfg = self._config(config) if config else None
        h = C.c_void_p()
        check(self._lib.ddgi_batch_create(C.cast(arr, C.c_void_p), len(self.volumes),
                                          C.byref(cfg) if cfg is not None else None, C.byref(h)))
        self.handle = h

    @staticmethod
    def _config(d: dict) -> _lib.BatchConfig:
        return _lib.BatchConfig(int(d["trace_block"]), int(d["trace_item_blocks"]), int(d["blend_probes"]),
                                (C.c_uint32 * 5)())

    def _orientations(self, orientations) -> np.ndarray:
        R = np.ascontiguousarray(orientations, np.float32).reshape(-1)
        if R.size != 9 * len(self.volumes):
            raise ValueError("one 3x3 orientation per volume is required")
        return R

    def destroy(self) -> None:
        if getattr(self, "handle", None):
            check(self._lib.ddgi_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def config(self) -> dict:
        c = _lib.BatchConfig()
        check(self._lib.ddgi_batch_get_config(self.handle, C.byref(c)))
        return c.as_dict()

    @config.setter
    def config(self, d: dict) -> None:
        c = self._config(d)
        check(self._lib.ddgi_batch_set_config(self.handle, C.byref(c)))

    def update(self, orientations, flags: int = _lib.UPDATE_ALL_PROBES, stream=None) -> None:
        R = self._orientations(orientations)
        check(self._lib.ddgi_batch_update(self.handle, R.ctypes.data_as(C.c_void_p), int(flags) | _lib.UPDATE_ALL_PROBES,
                                          _stream_ptr(stream)))

    def timings(self):
        """(total ms of the last update, per-volume completion ms from the update's start)."""
        total = C.c_float()
        vol = np.zeros(len(self.volumes), np.float32)
        check(self._lib.ddgi_batch_get_timings(self.handle, C.byref(total), vol.ctypes.data_as(C.c_void_p),
                                               len(self.volumes)))
        return float(total.value), vol

    def autotune(self, orientations, flags: int = _lib.UPDATE_ALL_PROBES, iterations: int = 3, stream=None):
        R = self._orientations(orientations)
        best = _lib.BatchConfig()
        recs = (_lib.BatchTuneRecord * 64)()
        n = C.c_uint32()
        check(self._lib.ddgi_batch_autotune(self.handle, R.ctypes.data_as(C.c_void_p),
                                            int(flags) | _lib.UPDATE_ALL_PROBES, int(iterations), _stream_ptr(stream),
                                            C.byref(best), C.cast(recs, C.c_void_p), 64, C.byref(n)))
        records = [dict(recs[i].config.as_dict(), ms=float(recs[i].ms), phase=int(recs[i].phase))
                   for i in range(n.value)]
        return best.as_dict(), records
