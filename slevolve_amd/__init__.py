"""slevolve_amd -- the SLEvolve sea-level time step on AMD MI355X (gfx950).

Drop-in for the hot path of the reference's Python package (python/slevolve/__init__.py:12-16):
the same class names and call signatures, with the transform, the modal state and the whole
inner fixed point running as HIP kernels in libslx.so (ctypes C-ABI, include/slx.h).
"""
from importlib import import_module

_EXPORTS = {
    "SeaLevelSolver": ".solver",
    "OptSeaLevelSolver": ".solver",
    "Planet": ".solver",
    "InnerLoopError": ".solver",
    "LoveNumbers": ".love",
    "SHT": ".sht",
    "make_sht": ".sht",
    "ExactMemory": ".memory",
}

__all__ = sorted(_EXPORTS)
__version__ = "0.1.0"


def __getattr__(name):
    if name in _EXPORTS:
        return getattr(import_module(_EXPORTS[name], __name__), name)
    raise AttributeError(name)
