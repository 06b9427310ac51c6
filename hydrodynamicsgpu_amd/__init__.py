"""MI355X-native finite-volume update path (Roe/HLL/HLLC, RK integrators).

The compute path is libhydro_hip.so (hand-written HIP for gfx950 behind the C
ABI of include/hydro_capi.h).  This package holds the build script, the ctypes
binding, the reference-style configuration front-end and the initial-condition
generators.
"""
from . import config, ics  # noqa: F401

__version__ = "0.1.0"


def Sim(*args, **kwargs):
    """Create a GPU simulation (see capi.Sim)."""
    from .capi import Sim as _Sim

    return _Sim(*args, **kwargs)
