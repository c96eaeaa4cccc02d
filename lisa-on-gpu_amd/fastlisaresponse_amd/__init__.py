"""lisa-on-gpu_amd: the fastlisaresponse time-domain LISA response on AMD MI355X (gfx950).

Public API (mirrors the reference package `fastlisaresponse`):
    pyResponseTDI, ResponseWrapper        -- response.py
    Orbits, EqualArmlengthOrbits          -- orbits.py
    OrbitsWrap, LISAResponseWrap          -- backend.py (the binding-level classes)
    AET, get_overlap                      -- utility.py
"""
from .orbits import CartwheelOrbits, EqualArmlengthOrbits, Orbits, StaticOrbits, YRSID_SI
from .utility import AET, get_overlap

__version__ = "0.1.0"


def __getattr__(name):
    # The GPU-facing classes import torch and load libflr_hip.so lazily.
    if name in ("pyResponseTDI", "ResponseWrapper"):
        from . import response
        return getattr(response, name)
    if name in ("OrbitsWrap", "LISAResponseWrap"):
        from . import backend
        return getattr(backend, name)
    raise AttributeError(name)


__all__ = ["pyResponseTDI", "ResponseWrapper", "Orbits", "EqualArmlengthOrbits", "CartwheelOrbits",
           "StaticOrbits",
           "OrbitsWrap", "LISAResponseWrap", "AET", "get_overlap", "YRSID_SI"]
