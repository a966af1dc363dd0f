"""sgipc -- MI355X-native Newton-step hot path of Stiff-GIPC (host-side Python).

The compute path is libsgipc.so (HIP kernels for gfx950 behind the C ABI in
include/sgipc.h); this package holds the ctypes binding, the scene generator
and a Newton-step driver used by tests and benchmarks.
"""
from . import capi, scene  # noqa: F401

__all__ = ["capi", "scene"]
