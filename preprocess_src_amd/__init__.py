"""MI355X-native coupled D3Q19-LBM / IBM / Skalak-membrane timestep.

The compute path is liblbmibm.so (hand-written HIP for gfx950, C ABI in
include/lbmibm.h); this package holds its build script, ctypes bindings, the
host-side mirror of the reference driver (``Simulation``) and the mesh /
workload generators.
"""
from .native import LibraryMissing, LbmibmError, lib  # noqa: F401

__all__ = ["LibraryMissing", "LbmibmError", "lib", "Simulation"]


def __getattr__(name):
    if name == "Simulation":
        from .simulation import Simulation
        return Simulation
    raise AttributeError(name)
