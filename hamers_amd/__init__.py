"""hamers_amd -- MI355X-native convective-flux RHS of HAMeRS.

The compute path is ``libhamers_amd.so`` (hand-written HIP kernels for gfx950
behind the C-ABI in ``include/hamers_convective.h``).  This package only loads
that library and wraps it for Python callers and tests; there is no CPU
fallback: if the library is missing, importing the bindings raises.
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhamers_amd.so")

from . import capi  # noqa: E402,F401  (ctypes bindings; raises if the library is missing)
from .reconstructor import (  # noqa: E402,F401
    ConvectiveFluxReconstructor,
    ConvectiveFluxReconstructorManager,
    DiffusiveFluxReconstructor,
    PatchData,
    PatchList,
)
from .integrator import (  # noqa: E402,F401
    SSP_RK3_ALPHA,
    SSP_RK3_BETA,
    SSP_RK3_GAMMA,
    Hierarchy,
    Level,
    RKPatchData,
    advanceSingleStepOnPatch,
)
