"""MI355X (gfx950) shoot-bounce-reflect trace-and-field step of SBRSolver.

The compute path is libsbr_amd.so (HIP kernels behind the C-ABI in include/sbr_capi.h); this package
holds its ctypes binding, the build recipe and the synthetic scene generators.
"""
from .capi import (Context, Material, Result, SbrError, Server, Stats, TraceParams, DEFAULT_MATERIALS,  # noqa: F401
                   LIB_PATH, SBR_PENDING, lib, make_params)

__all__ = ["Context", "Material", "Result", "SbrError", "Server", "Stats", "TraceParams", "DEFAULT_MATERIALS",
           "LIB_PATH", "SBR_PENDING", "lib", "make_params"]
