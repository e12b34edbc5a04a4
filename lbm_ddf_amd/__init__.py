"""MI355X-native D3Q27 double-distribution Rayleigh-Benard step (gfx950 HIP).

The compute path is the HIP library built from ``lbm_ddf_amd/csrc`` (see
``include/lbm_ddf.h`` for the C ABI).  ``lbm_ddf_amd.capi`` is the Python
binding; importing the package does not load the library, so that a
CPU-only environment can import it to check the build.
"""
from .capi import (Cavity, Ensemble, LbmError, serve, LIB_PATH, PREC, FIELDS, POPS, bytes_required, derive, load,
                   make_params)

__all__ = ["Cavity", "Ensemble", "LbmError", "serve", "LIB_PATH", "PREC", "FIELDS", "POPS", "bytes_required", "derive",
           "load", "make_params"]
