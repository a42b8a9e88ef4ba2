"""MI355X-native JAX-ALFA per-time-step hot path.

HIP kernels for gfx950 behind the C ABI in include/alfa.h (libalfa.so);
``ops`` mirrors the reference's operator names, ``core.LES`` runs the
composed step (src/Main.py:158-540) on device-resident state.
"""
from . import _lib, cases  # noqa: F401

__all__ = ["_lib", "cases", "core", "ops"]


def __getattr__(name):
    if name in ("core", "ops"):
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
