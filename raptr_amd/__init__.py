"""raptr_amd -- MI355X (gfx950) implementation of raptr's reconstruction hot
path: cylinder-culled workqueue, Monte Carlo ray-sampled system matrix,
COO/CSR compaction, forward/back SpMV and the MLEM update, behind a C ABI
(include/raptr_hip.h) served by hand-written HIP kernels.

The Python modules are thin ctypes bindings over that ABI; the C++ host
mirror of the reference's template interfaces lives in include/raptr/.
"""
from . import _capi  # noqa: F401
from ._capi import (OP_NON_TRANSPOSE, OP_TRANSPOSE, RAYGEN_RAND, RAYGEN_REGU, RAYGEN_RAND_ELEMENT, SM_KEEP_ZEROS,  # noqa: F401
                    RaptrError, lib)
from .geometry import (MeasurementSetup, VoxelGrid, centred_grid, config_full,  # noqa: F401
                       config_small, config_tiny, two_head_setup)

__all__ = ["lib", "MeasurementSetup", "VoxelGrid", "RaptrError"]
