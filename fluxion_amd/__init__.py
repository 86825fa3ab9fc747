"""fluxion_amd — MI355X (gfx950) backend for fluxion's matrix-free structured-grid solve path.

The compute path is libfluxion_hip.so (HIP kernels + C ABI, include/fluxion_hip.h).
This package exposes the value types (always importable) and, lazily, the
device API in ``fluxion_amd.hip`` which requires the built library.
"""
from .core import (BcKind, BoundaryCondition, Boundaries, ConvergenceReason, CoreError, DType, DimensionMismatch,
                   Grid, InvalidGrid, InvalidStencil, PeriodicNotImplemented, SmootherKind, SolveStats, Stencil,
                   StencilEntry, Transfer)

__all__ = [
    "BcKind", "BoundaryCondition", "Boundaries", "ConvergenceReason", "CoreError", "DType", "DimensionMismatch",
    "Grid", "InvalidGrid", "InvalidStencil", "PeriodicNotImplemented", "SmootherKind", "SolveStats", "Stencil",
    "StencilEntry", "Transfer",
]
