"""MI355X-native DFSPH time step and neighbourhood search.

The compute path is the HIP library splishsplash_amd/libsph_amd.so (gfx950),
reached through the C-ABI declared in include/sph_capi.h; `capi` binds it
with ctypes. There is no CPU fallback: loading fails loudly when the library
is missing, and compute calls fail when no gfx950 device is present.
"""
from .params import (BOUNDARY_AKINCI2012, BOUNDARY_BENDER2019, BOUNDARY_NONE, DFSPHCounters,  # noqa: F401
                     DFSPHParams, STAGES, VISCOSITY_NONE, VISCOSITY_WEILER2018, default_params)
