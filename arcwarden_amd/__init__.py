"""arcwarden_amd -- MI355X-native 2D3V particle-in-cell hot path.

The product is libarcwarden_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/aw.h).  This package holds its build recipe, the C++20
facade headers mirroring the reference's ``arc::`` API (``include/pic``), and
a thin Python harness (ctypes) used by the tests, the bench and the autograd
bridge.  Nothing here computes on the CPU: every operation calls the library,
and importing the harness without the library raises ``LibraryMissing``.
"""
from ._lib import (AW_DEPOSIT_AUTO, AW_DEPOSIT_BINNED, AW_DEPOSIT_GLOBAL,  # noqa: F401
                   AW_DEPOSIT_LDS_GRID, AW_MODEL_DARWIN, AW_MODEL_ES, AwError, LibraryMissing,
                   grid, lib, run_params, species)

__all__ = ["lib", "grid", "run_params", "species", "AwError", "LibraryMissing"]
