"""Value types that cross the C ABI, mirroring ``fluxion_core``.

Grid        crates/core/src/grid.rs:4-62
Stencil     crates/core/src/stencil.rs:3-143
Boundaries  crates/core/src/boundary.rs:1-44
DType       crates/core/src/dtype.rs:1-16
SolveStats  crates/runtime/src/cg.rs:9-41

These are plain Python: no device is needed to build them.  Stencil weights
are computed with the same IEEE-754 double operations, in the same order, as
the reference constructors (Python does not contract a*b+c into an FMA).
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field


class CoreError(Exception):
    """Raised for invalid arguments (crates/core/src/error.rs:3-20)."""


class InvalidGrid(CoreError):
    pass


class InvalidStencil(CoreError):
    pass


class DimensionMismatch(CoreError):
    pass


class PeriodicNotImplemented(CoreError):
    pass


class BackendError(CoreError):
    pass


class IoError(CoreError, OSError):
    """std::io::Error of the raw field format (crates/io/src/raw.rs:13-77)."""


class DType(enum.IntEnum):
    F32 = 0
    F64 = 1


@dataclass(frozen=True)
class Grid:
    """Row-major grid, index = row*cols + col (grid.rs:4-62)."""

    rows: int
    cols: int
    dx: float
    dy: float

    def __post_init__(self):
        # grid.rs:15-36
        if self.rows < 3 or self.cols < 3:
            raise InvalidGrid(f"minimum 3x3, got {self.rows}x{self.cols}")
        if not (math.isfinite(self.dx) and math.isfinite(self.dy)) or self.dx <= 0 or self.dy <= 0:
            raise InvalidGrid(f"spacing must be finite and positive, got dx={self.dx}, dy={self.dy}")

    @staticmethod
    def square(n: int, h: float) -> "Grid":
        return Grid(n, n, h, h)

    def __len__(self) -> int:
        return self.rows * self.cols

    def idx(self, row: int, col: int) -> int:
        return row * self.cols + col


@dataclass(frozen=True)
class StencilEntry:
    dr: int
    dc: int
    weight: float


class Stencil:
    """Finite-difference stencil as (dr, dc, weight) entries (stencil.rs:3-143)."""

    def __init__(self, entries):
        entries = [e if isinstance(e, StencilEntry) else StencilEntry(*e) for e in entries]
        if not entries:
            raise InvalidStencil("stencil must have at least one entry")
        self._entries = tuple(entries)

    @property
    def entries(self):
        return self._entries

    def __len__(self):
        return len(self._entries)

    def __repr__(self):
        return f"Stencil({list(self._entries)})"

    @staticmethod
    def laplacian_2d_5pt(dx: float, dy: float) -> "Stencil":
        """stencil.rs:56-72, entry order (0,0), (-1,0), (1,0), (0,-1), (0,1)."""
        inv_dx2 = 1.0 / (dx * dx)
        inv_dy2 = 1.0 / (dy * dy)
        return Stencil([
            StencilEntry(0, 0, -2.0 * (inv_dx2 + inv_dy2)),
            StencilEntry(-1, 0, inv_dy2),
            StencilEntry(1, 0, inv_dy2),
            StencilEntry(0, -1, inv_dx2),
            StencilEntry(0, 1, inv_dx2),
        ])

    @staticmethod
    def laplacian_2d_9pt(dx: float, dy: float) -> "Stencil":
        """Compact (Collatz) 9-point Laplacian, stencil.rs:95-142."""
        inv_dx2 = 1.0 / (dx * dx)
        inv_dy2 = 1.0 / (dy * dy)
        dx2 = dx * dx
        dy2 = dy * dy
        k = (dx2 + dy2) / 12.0
        inv_dx2dy2 = inv_dx2 * inv_dy2
        w_c = -2.0 * (inv_dx2 + inv_dy2) + 4.0 * k * inv_dx2dy2
        w_n = inv_dy2 - 2.0 * k * inv_dx2dy2
        w_w = inv_dx2 - 2.0 * k * inv_dx2dy2
        w_diag = k * inv_dx2dy2
        return Stencil([
            StencilEntry(0, 0, w_c),
            StencilEntry(-1, 0, w_n),
            StencilEntry(1, 0, w_n),
            StencilEntry(0, -1, w_w),
            StencilEntry(0, 1, w_w),
            StencilEntry(-1, -1, w_diag),
            StencilEntry(-1, 1, w_diag),
            StencilEntry(1, -1, w_diag),
            StencilEntry(1, 1, w_diag),
        ])

    @staticmethod
    def poisson_minus_identity(dx: float, dy: float) -> "Stencil":
        """S = −Δ_h − I, so that the reference operator v + 1·S(v) is −Δ_h (pure Poisson)."""
        inv_dx2 = 1.0 / (dx * dx)
        inv_dy2 = 1.0 / (dy * dy)
        return Stencil([
            StencilEntry(0, 0, 2.0 * inv_dx2 + 2.0 * inv_dy2 - 1.0),
            StencilEntry(-1, 0, -inv_dy2),
            StencilEntry(1, 0, -inv_dy2),
            StencilEntry(0, -1, -inv_dx2),
            StencilEntry(0, 1, -inv_dx2),
        ])


class BcKind(enum.IntEnum):
    DIRICHLET = 0
    NEUMANN = 1
    ROBIN = 2
    PERIODIC = 3


@dataclass(frozen=True)
class BoundaryCondition:
    """Dirichlet(a) | Neumann(a) | Robin{alpha=a, beta=b, g} | Periodic (boundary.rs:1-15)."""

    kind: BcKind
    a: float = 0.0
    b: float = 0.0
    g: float = 0.0

    @staticmethod
    def dirichlet(value: float) -> "BoundaryCondition":
        return BoundaryCondition(BcKind.DIRICHLET, value)

    @staticmethod
    def neumann(flux: float) -> "BoundaryCondition":
        return BoundaryCondition(BcKind.NEUMANN, flux)

    @staticmethod
    def robin(alpha: float, beta: float, g: float) -> "BoundaryCondition":
        return BoundaryCondition(BcKind.ROBIN, alpha, beta, g)

    @staticmethod
    def periodic() -> "BoundaryCondition":
        return BoundaryCondition(BcKind.PERIODIC)


@dataclass(frozen=True)
class Boundaries:
    top: BoundaryCondition
    bottom: BoundaryCondition
    left: BoundaryCondition
    right: BoundaryCondition

    @staticmethod
    def uniform(bc: BoundaryCondition) -> "Boundaries":
        return Boundaries(bc, bc, bc, bc)

    @staticmethod
    def zero_dirichlet() -> "Boundaries":
        """boundary.rs:35-38."""
        return Boundaries.uniform(BoundaryCondition.dirichlet(0.0))

    def sides(self):
        return (self.top, self.bottom, self.left, self.right)


class ConvergenceReason(enum.IntEnum):
    """cg.rs:11-18."""

    CONVERGED = 0
    MAX_ITERATIONS = 1
    BREAKDOWN = 2


@dataclass
class SolveStats:
    """cg.rs:22-38; `true_residual` (‖b − A·x‖₂ in FP64) is an extension."""

    iterations: int
    residual: float
    reason: ConvergenceReason
    true_residual: float = float("nan")
    history: list = field(default_factory=list)

    def converged(self) -> bool:
        return self.reason == ConvergenceReason.CONVERGED


class SmootherKind(enum.IntEnum):
    """multigrid.rs:12-20."""

    JACOBI = 0
    CHEBYSHEV = 1


class TimeScheme(enum.IntEnum):
    """solver.rs:14-23."""

    EULER = 0
    RK2 = 1
    RK4 = 2


class MgKernels(enum.IntEnum):
    """Level kernels of the fused FP32 V-cycle (fx_mg_kernels); all give bit-identical results."""

    TILED = 0            # LDS tiles
    RING = 1             # LDS-ring row streams
    WAVE = 2             # independent wave-wide row streams (default)


class Transfer(enum.IntEnum):
    REFERENCE = 0        # 2x2 average restriction + injection prolongation (ops.rs:369-423)
    FW_BILINEAR = 1      # full-weighting restriction + bilinear prolongation
