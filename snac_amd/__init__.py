"""snac_amd -- SNAC's explicit time step on MI355X (gfx950).

The step runs in hand-written HIP kernels behind the C ABI of
libsnac_hip.so (include/snac_hip.h); this package is the host side: problem
description (problem.py), the ctypes binding (_abi.py) and the driver that
mirrors the reference's context lifecycle (context.py).
"""
from . import _abi
from .problem import (MaterialSpec, Problem, elastic_box_compression, maxwell_viscoelastic,
                      mohr_coulomb_extension, WORKLOADS, CM_PER_YEAR)
from .context import Context

__all__ = ["_abi", "MaterialSpec", "Problem", "Context", "elastic_box_compression",
           "mohr_coulomb_extension", "maxwell_viscoelastic", "WORKLOADS", "CM_PER_YEAR"]
