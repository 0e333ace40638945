"""MI355X (gfx950) DDGI irradiance-probe update: a HIP/C++ library (libddgi.so, C-ABI in
include/ddgi.h) behind a host mirror of VulkanExp's IrradianceProbes interface."""
from . import scenes
from ._lib import DdgiError, load
from .batch import ProbeBatch
from .probes import IrradianceProbes

__all__ = ["IrradianceProbes", "ProbeBatch", "DdgiError", "load", "scenes"]
