"""Python mirror of the reference runtime API over the HIP C ABI.

HipBackend    ~ crates/backend-hip/src/backend.rs (the `Backend` trait, core/src/backend.rs:43-406)
Field         ~ crates/runtime/src/field.rs (buffer + grid with per-op methods)
CgSolver, PipelinedCgSolver, BiCgStabSolver, GmresSolver
              ~ crates/runtime/src/{cg,pipelined_cg,bicgstab,gmres}.rs
Multigrid, Identity
              ~ crates/runtime/src/{multigrid,preconditioner}.rs
CrankNicolsonSolver, ImexSolver, BdfSolver
              ~ crates/runtime/src/{implicit,imex,bdf}.rs
DiffusionSolver
              ~ crates/runtime/src/solver.rs
write_raw, read_raw
              ~ crates/io/src/raw.rs

Every method calls one C entry point of libfluxion_hip.so; nothing here computes
on the CPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .core import (Boundaries, ConvergenceReason, DType, DimensionMismatch, Grid, SmootherKind, SolveStats,
                   Stencil, TimeScheme, Transfer)


class HipBackend:
    """One device, one stream (backend-hip/src/backend.rs:61-264)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        L.check(self._lib.fx_backend_create(device, C.byref(h)), "fx_backend_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_backend_destroy(h)
            self.h = None

    @property
    def name(self) -> str:
        return self._lib.fx_backend_name(self.h).decode()

    def synchronize(self):
        L.check(self._lib.fx_synchronize(self.h), "synchronize")

    @property
    def stream(self) -> int:
        return self._lib.fx_backend_stream(self.h) or 0

    # ── kernel probe (fx_probe_set / fx_probe_read) ───────────────────
    PROBE_OFF, PROBE_FCG_UPDATE, PROBE_FCG_STEP, PROBE_VC_POST0, PROBE_VC_RESTRICT0, PROBE_STENCIL_AXPY, \
        PROBE_REDUCE = range(7)

    def probe(self, kind: int) -> None:
        """Time every launch of one kernel family with HIP events on this stream (0 = off)."""
        L.check(self._lib.fx_probe_set(self.h, int(kind)), "probe_set")

    def probe_read(self) -> tuple[float, int]:
        """(summed ms, launches) of the probed family since probe()."""
        ms, cnt = C.c_double(), C.c_uint64()
        L.check(self._lib.fx_probe_read(self.h, C.byref(ms), C.byref(cnt)), "probe_read")
        return ms.value, cnt.value

    HBM_KINDS = {"read": 8, "write": 8, "copy": 16, "mix52": 52}   # kind -> bytes per point

    def hbm_ceilings(self, n: int = 1 << 24, reps: int = 20) -> dict:
        """Attainable bandwidth of plain streaming kernels (fx_hbm_ceilings): {kind: (us, GB/s)}."""
        out = (C.c_double * 4)()
        L.check(self._lib.fx_hbm_ceilings(self.h, int(n), int(reps), out), "hbm_ceilings")
        return {k: (out[i], b * n / (out[i] * 1e-6) / 1e9) for i, (k, b) in enumerate(self.HBM_KINDS.items())}

    # ── memory ────────────────────────────────────────────────────────
    def allocate(self, n: int, dtype: DType = DType.F64) -> "Buffer":
        h = C.c_void_p()
        L.check(self._lib.fx_allocate(self.h, n, int(dtype), C.byref(h)), "allocate")
        return Buffer(self, h)

    def upload(self, data) -> "Buffer":
        a = np.ascontiguousarray(data)
        h = C.c_void_p()
        if a.dtype == np.float32:
            L.check(self._lib.fx_upload_f32(self.h, a.ctypes.data, a.size, C.byref(h)), "upload_f32")
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            L.check(self._lib.fx_upload_f64(self.h, a.ctypes.data, a.size, C.byref(h)), "upload_f64")
        return Buffer(self, h)


class Buffer:
    """Device buffer owned by the library (backend-hip/src/buffer.rs:16-123)."""

    def __init__(self, be: HipBackend, h: C.c_void_p):
        self.be = be
        self.h = h
        self._lib = be._lib

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_buffer_free(h)
            self.h = None

    def __len__(self):
        return int(self._lib.fx_buffer_len(self.h))

    @property
    def dtype(self) -> DType:
        return DType(self._lib.fx_buffer_dtype(self.h))

    @property
    def device_ptr(self) -> int:
        return self._lib.fx_buffer_device_ptr(self.h) or 0

    def to_numpy(self, dtype=np.float64) -> np.ndarray:
        n = len(self)
        out = np.empty(n, dtype=dtype)
        if dtype == np.float32:
            L.check(self._lib.fx_copy_to_host_f32(self.be.h, self.h, out.ctypes.data, n), "copy_to_host_f32")
        else:
            L.check(self._lib.fx_copy_to_host_f64(self.be.h, self.h, out.ctypes.data, n), "copy_to_host_f64")
        return out

    def write(self, data) -> None:
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        L.check(self._lib.fx_write_f64(self.be.h, self.h, a.ctypes.data, a.size), "write_f64")


def _backend(be: Optional[HipBackend]) -> HipBackend:
    global _DEFAULT
    if be is not None:
        return be
    if _DEFAULT is None:
        _DEFAULT = HipBackend(0)
    return _DEFAULT


_DEFAULT: Optional[HipBackend] = None


def default_backend() -> HipBackend:
    """Process-wide backend, like the reference's OnceLock singleton (storage.rs:120-131)."""
    return _backend(None)


class Field:
    """A buffer with its grid (crates/runtime/src/field.rs:15-1050)."""

    def __init__(self, grid: Grid, buf: Buffer):
        if len(buf) != len(grid):
            raise DimensionMismatch(f"expected {len(grid)}, got {len(buf)}")
        self.grid = grid
        self.buf = buf
        self.be = buf.be
        self._lib = buf._lib

    # constructors (field.rs:27-110)
    @staticmethod
    def zeros(grid: Grid, dtype: DType = DType.F64, be: Optional[HipBackend] = None) -> "Field":
        be = _backend(be)
        return Field(grid, be.allocate(len(grid), dtype))

    @staticmethod
    def from_f64(grid: Grid, data, be: Optional[HipBackend] = None) -> "Field":
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        if a.size != len(grid):
            raise DimensionMismatch(f"expected {len(grid)}, got {a.size}")
        return Field(grid, _backend(be).upload(a))

    def to_numpy(self) -> np.ndarray:
        return self.buf.to_numpy().reshape(self.grid.rows, self.grid.cols)

    def to_vec_f64(self) -> np.ndarray:
        return self.buf.to_numpy()

    def write(self, data) -> None:
        self.buf.write(data)

    @property
    def h(self):
        return self.buf.h

    def _bh(self):
        return self.be.h

    # ── ops (field.rs:147-1050) ───────────────────────────────────────
    def copy_from(self, other: "Field"):
        L.check(self._lib.fx_copy(self._bh(), other.h, self.h), "copy")

    def fill(self, value: float):
        L.check(self._lib.fx_fill(self._bh(), self.h, value), "fill")

    def scale(self, alpha: float):
        L.check(self._lib.fx_scale(self._bh(), self.h, alpha), "scale")

    def axpy(self, alpha: float, x: "Field"):
        L.check(self._lib.fx_axpy(self._bh(), alpha, x.h, self.h), "axpy")

    def waxpy(self, alpha: float, x: "Field", beta: float, y: "Field"):
        L.check(self._lib.fx_waxpy(self._bh(), alpha, x.h, beta, y.h, self.h), "waxpy")

    def aypx(self, alpha: float, x: "Field"):
        L.check(self._lib.fx_aypx(self._bh(), alpha, x.h, self.h), "aypx")

    def pointwise_mult(self, x: "Field", y: "Field"):
        L.check(self._lib.fx_pointwise_mult(self._bh(), x.h, y.h, self.h), "pointwise_mult")

    def pointwise_div(self, x: "Field", y: "Field"):
        L.check(self._lib.fx_pointwise_div(self._bh(), x.h, y.h, self.h), "pointwise_div")

    def pointwise_max(self, x: "Field", y: "Field"):
        L.check(self._lib.fx_pointwise_max(self._bh(), x.h, y.h, self.h), "pointwise_max")

    def pointwise_min(self, x: "Field", y: "Field"):
        L.check(self._lib.fx_pointwise_min(self._bh(), x.h, y.h, self.h), "pointwise_min")

    def reciprocal(self):
        L.check(self._lib.fx_reciprocal(self._bh(), self.h), "reciprocal")

    def abs_val(self):
        L.check(self._lib.fx_abs_val(self._bh(), self.h), "abs_val")

    def zero_boundaries(self):
        L.check(self._lib.fx_zero_boundaries(self._bh(), self.h, self.grid.rows, self.grid.cols), "zero_boundaries")

    def _red(self, fn, *args) -> float:
        out = C.c_double()
        L.check(fn(self._bh(), *args, C.byref(out)), fn.__name__)
        return out.value

    def norm_l2(self) -> float:
        return self._red(self._lib.fx_norm_l2, self.h)

    def sum(self) -> float:
        return self._red(self._lib.fx_reduce_sum, self.h)

    def max(self) -> float:
        return self._red(self._lib.fx_reduce_max, self.h)

    def min(self) -> float:
        return self._red(self._lib.fx_reduce_min, self.h)

    def dot(self, other: "Field") -> float:
        return self._red(self._lib.fx_dot, self.h, other.h)

    def dot2(self, a: "Field", b: "Field", c: "Field"):
        """(self·a, b·c) in one pass (field.rs:637-679)."""
        out = (C.c_double * 2)()
        L.check(self._lib.fx_dot2(self._bh(), self.h, a.h, b.h, c.h, out), "dot2")
        return out[0], out[1]

    def integral(self) -> float:
        """field.rs:686-690: sum·dx·dy."""
        return self.sum() * self.grid.dx * self.grid.dy

    def apply_stencil_into(self, stencil: Stencil, bc: Boundaries, output: "Field"):
        st = L.StencilC(stencil)
        g, b = L.grid_c(self.grid), L.bcs_c(bc)
        L.check(self._lib.fx_apply_stencil(self._bh(), self.h, output.h, C.byref(g), C.byref(st.c), C.byref(b)),
                "apply_stencil")

    def apply_stencil_var_into(self, stencil: Stencil, bc: Boundaries, coeff: "Field", output: "Field"):
        st = L.StencilC(stencil)
        g, b = L.grid_c(self.grid), L.bcs_c(bc)
        L.check(self._lib.fx_apply_stencil_var(self._bh(), self.h, output.h, coeff.h, C.byref(g), C.byref(st.c),
                                               C.byref(b)), "apply_stencil_var")

    def stencil_axpy(self, alpha: float, x: "Field", stencil: Stencil, bc: Boundaries):
        """self += alpha·S(x) on interior points (field.rs:309-351)."""
        st = L.StencilC(stencil)
        g, b = L.grid_c(self.grid), L.bcs_c(bc)
        L.check(self._lib.fx_stencil_axpy(self._bh(), alpha, x.h, self.h, C.byref(g), C.byref(st.c), C.byref(b)),
                "stencil_axpy")

    def weighted_jacobi(self, b: "Field", omega: float, stencil: Stencil, bc: Boundaries):
        st = L.StencilC(stencil)
        g, bb = L.grid_c(self.grid), L.bcs_c(bc)
        L.check(self._lib.fx_weighted_jacobi(self._bh(), self.h, b.h, omega, C.byref(g), C.byref(st.c),
                                             C.byref(bb)), "weighted_jacobi")

    def apply_conv_diff_into(self, kappa: "Field", vx: "Field", vy: "Field", bc: Boundaries, output: "Field"):
        g, b = L.grid_c(self.grid), L.bcs_c(bc)
        L.check(self._lib.fx_apply_conv_diff(self._bh(), self.h, output.h, kappa.h, vx.h, vy.h, C.byref(g),
                                             C.byref(b)), "apply_conv_diff")

    def conv_diff_axpy(self, alpha: float, u: "Field", kappa: "Field", vx: "Field", vy: "Field", bc: Boundaries):
        g, b = L.grid_c(self.grid), L.bcs_c(bc)
        L.check(self._lib.fx_conv_diff_axpy(self._bh(), alpha, u.h, self.h, kappa.h, vx.h, vy.h, C.byref(g),
                                            C.byref(b)), "conv_diff_axpy")


def restrict(fine: Field, coarse: Field):
    """Backend::restrict (core/src/backend.rs:240-246)."""
    fg, cg = L.grid_c(fine.grid), L.grid_c(coarse.grid)
    L.check(fine._lib.fx_restrict(fine._bh(), fine.h, coarse.h, C.byref(fg), C.byref(cg)), "restrict")


def prolong(coarse: Field, fine: Field):
    """Backend::prolong (core/src/backend.rs:254-260)."""
    fg, cg = L.grid_c(fine.grid), L.grid_c(coarse.grid)
    L.check(fine._lib.fx_prolong(fine._bh(), coarse.h, fine.h, C.byref(cg), C.byref(fg)), "prolong")


def pipelined_cg_fused(alpha, beta, z, t, p, x, r, w, q):
    """HipBackend::pipelined_cg_fused (backend-hip/src/backend.rs:469-504)."""
    L.check(z._lib.fx_pipelined_cg_fused(z._bh(), alpha, beta, z.h, t.h, p.h, x.h, r.h, w.h, q.h),
            "pipelined_cg_fused")


# ── operators (extension) ─────────────────────────────────────────────
class Operator:
    """A·v = sigma·v + coeff·L(v) on interior points, identity rows on the boundary.

    kind 0 reproduces the reference operator v + coeff·S(v) (cg.rs:104-111)."""

    STENCIL, CONV_DIFF, VARK = 0, 1, 2

    def __init__(self, kind, coeff, sigma=1.0, stencil=None, bc=None, kappa=None, kappa_scalar=0.0, vx=None,
                 vy=None, k=None):
        self.kind, self.coeff, self.sigma = kind, coeff, sigma
        self._st = L.StencilC(stencil) if stencil is not None else None
        self._bc = L.bcs_c(bc) if bc is not None else None
        self._keep = (kappa, vx, vy, k)
        self.c = L.FxOperator(
            kind, 0, coeff, sigma,
            C.pointer(self._st.c) if self._st else None,
            C.pointer(self._bc) if self._bc else None,
            kappa.h if kappa is not None else None, kappa_scalar,
            vx.h if vx is not None else None, vy.h if vy is not None else None,
            k.h if k is not None else None)

    @staticmethod
    def stencil(coeff: float, stencil: Stencil, bc: Boundaries) -> "Operator":
        return Operator(Operator.STENCIL, coeff, 1.0, stencil=stencil, bc=bc)

    @staticmethod
    def conv_diff(vx: Field, vy: Field, kappa: Optional[Field] = None, kappa_scalar: float = 0.0,
                  coeff: float = 1.0, sigma: float = 0.0) -> "Operator":
        """sigma·u + coeff·(κΔu + v·∇⁻u) with the reference upwind discretisation (ops.rs:739-791)."""
        return Operator(Operator.CONV_DIFF, coeff, sigma, kappa=kappa, kappa_scalar=kappa_scalar, vx=vx, vy=vy)

    @staticmethod
    def var_k(k: Field, coeff: float = 1.0, sigma: float = 0.0) -> "Operator":
        """sigma·u + coeff·(−∇·(k∇u)), harmonic face averages of nodal k."""
        return Operator(Operator.VARK, coeff, sigma, k=k)

    def apply(self, x: Field, y: Field):
        g = L.grid_c(x.grid)
        L.check(x._lib.fx_apply_operator(x._bh(), C.byref(self.c), x.h, y.h, C.byref(g)), "apply_operator")


# ── preconditioners ───────────────────────────────────────────────────
class Identity:
    """preconditioner.rs:29-37."""

    def _pc(self):
        return L.FxPrecond(0, 0, None)


class Multigrid:
    """Geometric multigrid V-cycle (multigrid.rs:12-608).

    Multigrid(grid, bc, n_pre, n_post, omega) is Multigrid::new; the keyword
    options select the extensions (transfer operators, coarsest size, FP32
    levels, level operator sigma·I + scale·Δ_h, or with k sigma·I + scale·(−∇·k∇),
    or with vx, vy the upwind convection-diffusion sigma·I + scale·(κΔ + v·∇⁻))."""

    def __init__(self, grid: Grid, bc: Boundaries, n_pre: int = 2, n_post: int = 2, omega: float = 2.0 / 3.0,
                 smoother: SmootherKind = SmootherKind.JACOBI, *, transfer: Transfer = Transfer.REFERENCE,
                 coarsest: int = 4, extra_coarse: int = 10, precision: DType = DType.F64, fused: bool = False,
                 sigma: float = 0.0, scale: float = 1.0, k: Optional["Field"] = None,
                 vx: Optional["Field"] = None, vy: Optional["Field"] = None, kappa: float = 0.0,
                 be: Optional[HipBackend] = None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grid = grid
        self._k = (k, vx, vy)
        opt = L.FxMgOptions(n_pre, n_post, omega, int(smoother), int(transfer), coarsest, extra_coarse,
                            int(precision), int(bool(fused)), sigma, scale, k.h if k is not None else None,
                            vx.h if vx is not None else None, vy.h if vy is not None else None, kappa)
        g, b = L.grid_c(grid), L.bcs_c(bc)
        h = C.c_void_p()
        L.check(self._lib.fx_multigrid_create_ex(self.be.h, C.byref(g), C.byref(b), C.byref(opt), C.byref(h)),
                "multigrid_create")
        self.h = h

    @staticmethod
    def build(grid, bc, n_pre, n_post, omega, smoother, be=None) -> "Multigrid":
        return Multigrid(grid, bc, n_pre, n_post, omega, smoother, be=be)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_multigrid_destroy(h)
            self.h = None

    def depth(self) -> int:
        return int(self._lib.fx_multigrid_depth(self.h))

    def v_cycle(self, x: Field, b: Field):
        L.check(self._lib.fx_multigrid_vcycle(self.h, x.h, b.h), "v_cycle")

    def set_kernels(self, kernels: int):
        """Level kernels of the fused FP32 cycle (MgKernels)."""
        L.check(self._lib.fx_multigrid_set_kernels(self.h, int(kernels)), "multigrid_set_kernels")

    def autotune(self, reps: int = 5) -> float:
        """Picks each level's wave segment heights by timing them (fx_multigrid_autotune); returns the
        summed per-launch time (ms) of the chosen shapes."""
        ms = C.c_double()
        L.check(self._lib.fx_multigrid_autotune(self.h, reps, C.byref(ms)), "multigrid_autotune")
        return ms.value

    def set_wave_rows(self, level: int, post_rows: int, restrict_rows: int):
        """Wave segment heights of one level (0: built-in choice; -1 on levels >= 1: LDS-tile kernel)."""
        L.check(self._lib.fx_multigrid_set_wave_rows(self.h, level, post_rows, restrict_rows),
                "multigrid_set_wave_rows")

    def apply(self, z: Field, r: Field):
        self.v_cycle(z, r)

    def _pc(self):
        return L.FxPrecond(1, 0, self.h)


# ── solvers ───────────────────────────────────────────────────────────
class _Solver:
    KIND = 0
    NAME = "?"

    def __init__(self, grid: Grid, tol: float, max_iter: int, restart: int = 0, be: Optional[HipBackend] = None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grid = grid
        self.tol = tol
        self.max_iter = max_iter
        g = L.grid_c(grid)
        h = C.c_void_p()
        L.check(self._lib.fx_solver_create(self.be.h, self.KIND, C.byref(g), restart, tol, max_iter, C.byref(h)),
                "solver_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_solver_destroy(h)
            self.h = None

    def name(self) -> str:
        return self._lib.fx_solver_name(self.h).decode()

    def _stats(self, st: L.FxSolveStats) -> SolveStats:
        n = self._lib.fx_solver_history(self.h, None, 0)
        hist = (C.c_double * max(n, 1))()
        self._lib.fx_solver_history(self.h, hist, n)
        return SolveStats(int(st.iterations), st.residual, ConvergenceReason(st.reason), st.true_residual,
                          list(hist[:n]))

    def solve(self, x: Field, b: Field, coeff: float, stencil: Stencil, bc: Boundaries) -> SolveStats:
        st = L.StencilC(stencil)
        bb = L.bcs_c(bc)
        out = L.FxSolveStats()
        L.check(self._lib.fx_solver_solve(self.h, x.h, b.h, coeff, C.byref(st.c), C.byref(bb), C.byref(out)),
                "solve")
        return self._stats(out)

    def solve_preconditioned(self, x: Field, b: Field, coeff: float, stencil: Stencil, bc: Boundaries,
                             pc) -> SolveStats:
        st = L.StencilC(stencil)
        bb = L.bcs_c(bc)
        p = pc._pc()
        out = L.FxSolveStats()
        L.check(self._lib.fx_solver_solve_preconditioned(self.h, x.h, b.h, coeff, C.byref(st.c), C.byref(bb),
                                                         C.byref(p), C.byref(out)), "solve_preconditioned")
        return self._stats(out)

    def solve_op(self, x: Field, b: Field, op: Operator, pc=None) -> SolveStats:
        p = pc._pc() if pc is not None else None
        out = L.FxSolveStats()
        L.check(self._lib.fx_solver_solve_op(self.h, x.h, b.h, C.byref(op.c), C.byref(p) if p is not None else None,
                                             C.byref(out)), "solve_op")
        return self._stats(out)


class CgSolver(_Solver):
    """cg.rs:52-256."""
    KIND = 0

    def __init__(self, grid, tol, max_iter, be=None):
        super().__init__(grid, tol, max_iter, 0, be)


class PipelinedCgSolver(_Solver):
    """pipelined_cg.rs:25-238 (+ preconditioned Ghysels–Vanroose variant)."""
    KIND = 1

    def __init__(self, grid, tol, max_iter, be=None):
        super().__init__(grid, tol, max_iter, 0, be)


class BiCgStabSolver(_Solver):
    """bicgstab.rs:19-346."""
    KIND = 2

    def __init__(self, grid, tol, max_iter, be=None):
        super().__init__(grid, tol, max_iter, 0, be)


class GmresSolver(_Solver):
    """gmres.rs:24-490: GmresSolver::new(grid, device, restart, tol, max_iter)."""
    KIND = 3

    def __init__(self, grid, restart, tol, max_iter, be=None):
        super().__init__(grid, tol, max_iter, restart, be)


class FgmresSolver(_Solver):
    """Flexible right-preconditioned GMRES(m) (extension; Saad 1993): the residual estimate is
    ‖b − A·x‖ and the preconditioner may vary; unpreconditioned it is GmresSolver."""
    KIND = 5

    def __init__(self, grid, restart, tol, max_iter, be=None):
        super().__init__(grid, tol, max_iter, restart, be)


class FcgSolver(_Solver):
    """Flexible CG with FP64 true-residual refinement (extension)."""
    KIND = 4

    def __init__(self, grid, tol, max_iter, be=None):
        super().__init__(grid, tol, max_iter, 0, be)


# ── batched multi-problem solver (fx_batch_*; no reference counterpart) ──
class BatchSolver:
    """Many −∇·(k∇u) = f problems (zero Dirichlet) on grids of uneven sizes, each by FP64 flexible CG
    around its own fused FP32 coefficient-aware V-cycle, all advancing together from a captured HIP
    graph.  Vectors are the problems' points concatenated in order (offsets()).  `k` is an FP64
    Buffer of the concatenated coefficients and must outlive the solver."""

    def __init__(self, grids, k: "Buffer", n_pre: int = 2, n_post: int = 2, omega: float = 2.0 / 3.0,
                 coarsest: int = 16, extra_coarse: int = 10, be: Optional[HipBackend] = None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grids = list(grids)
        self._k = k
        n = len(self.grids)
        arr = (L.FxGrid * n)(*[L.grid_c(g) for g in self.grids])
        opt = L.FxMgOptions(n_pre, n_post, omega, int(SmootherKind.JACOBI), int(Transfer.FW_BILINEAR), coarsest,
                            extra_coarse, int(DType.F32), 1, 0.0, 1.0, None)
        h = C.c_void_p()
        L.check(self._lib.fx_batch_create(self.be.h, n, C.cast(arr, C.c_void_p), k.h, C.byref(opt), C.byref(h)),
                "batch_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_batch_destroy(h)
            self.h = None

    def offsets(self) -> np.ndarray:
        sizes = np.array([g.rows * g.cols for g in self.grids], dtype=np.int64)
        return np.concatenate([[0], np.cumsum(sizes)])

    def info(self) -> dict:
        rpb, ppb, pts, gi = C.c_int32(), C.c_int32(), C.c_uint64(), C.c_int32()
        L.check(self._lib.fx_batch_info(self.h, C.byref(rpb), C.byref(ppb), C.byref(pts), C.byref(gi)), "batch_info")
        wv, ph, rh = C.c_int32(), C.c_int32(), C.c_int32()
        L.check(self._lib.fx_batch_level_kernels(self.h, C.byref(wv), C.byref(ph), C.byref(rh)), "batch_level_kernels")
        return {"rows_per_block": rpb.value, "points_per_block": ppb.value, "points": pts.value,
                "graph_iterations": gi.value, "wave_level_kernels": bool(wv.value), "wave_post_rows": ph.value,
                "wave_restrict_rows": rh.value}

    def solve(self, x: "Buffer", b: "Buffer", tol, max_iter: int) -> list:
        n = len(self.grids)
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, dtype=np.float64), (n,)))
        st = (L.FxSolveStats * n)()
        L.check(self._lib.fx_batch_solve(self.h, x.h, b.h, t.ctypes.data, int(max_iter), C.cast(st, C.c_void_p)),
                "batch_solve")
        return [SolveStats(s.iterations, s.residual, ConvergenceReason(s.reason), s.true_residual) for s in st]


# ── implicit time stepping (fx_cn_*) ──────────────────────────────────
def _mg_opts(n_pre, n_post, omega, smoother=SmootherKind.JACOBI, transfer=Transfer.REFERENCE, coarsest=4,
             extra_coarse=10, precision=DType.F64, fused=False, sigma=0.0, scale=1.0, k=None) -> L.FxMgOptions:
    return L.FxMgOptions(n_pre, n_post, omega, int(smoother), int(transfer), coarsest, extra_coarse, int(precision),
                         int(bool(fused)), sigma, scale, k.h if k is not None else None)


class CrankNicolsonSolver:
    """CrankNicolsonSolver (crates/runtime/src/implicit.rs:1-238), stepped on the device by fx_cn_*.

    new(grid, alpha, dt, bc, cg_tol, cg_max_iter)                          implicit.rs:41-75
    with_multigrid(..., mg_pre, mg_post, mg_omega)                         implicit.rs:77-107
    with_multigrid_smoother(..., mg_pre, mg_post, mg_omega, smoother)      implicit.rs:109-154
    var_k(grid, alpha, dt, k, cg_tol, cg_max_iter, mixed=True)             extension: ∂u/∂t = α∇·(k∇u)

    step(u, source=None) -> SolveStats, step_n(u, n), dt(), steps_done(), sim_time() as in implicit.rs:156-237.
    """

    def __init__(self, grid: Grid, alpha: float, dt: float, bc: Boundaries, cg_tol: float, cg_max_iter: int, *,
                 mg: Optional[L.FxMgOptions] = None, op: int = 0, solver: int = 0, k: Optional[Field] = None,
                 be: Optional[HipBackend] = None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grid = grid
        self._k = k
        self._mg = mg
        g, b = L.grid_c(grid), L.bcs_c(bc)
        opt = L.FxCnOptions(op, solver, k.h if k is not None else None,
                            C.pointer(mg) if mg is not None else None)
        h = C.c_void_p()
        L.check(self._lib.fx_cn_create(self.be.h, C.byref(g), alpha, dt, C.byref(b), cg_tol, cg_max_iter,
                                       C.byref(opt), C.byref(h)), "cn_create")
        self.h = h

    @staticmethod
    def new(grid, alpha, dt, bc, cg_tol, cg_max_iter, be=None) -> "CrankNicolsonSolver":
        return CrankNicolsonSolver(grid, alpha, dt, bc, cg_tol, cg_max_iter, be=be)

    @staticmethod
    def with_multigrid(grid, alpha, dt, bc, cg_tol, cg_max_iter, mg_pre, mg_post, mg_omega,
                       be=None) -> "CrankNicolsonSolver":
        return CrankNicolsonSolver.with_multigrid_smoother(grid, alpha, dt, bc, cg_tol, cg_max_iter, mg_pre, mg_post,
                                                           mg_omega, SmootherKind.JACOBI, be=be)

    @staticmethod
    def with_multigrid_smoother(grid, alpha, dt, bc, cg_tol, cg_max_iter, mg_pre, mg_post, mg_omega, smoother,
                                be=None) -> "CrankNicolsonSolver":
        return CrankNicolsonSolver(grid, alpha, dt, bc, cg_tol, cg_max_iter,
                                   mg=_mg_opts(mg_pre, mg_post, mg_omega, smoother), be=be)

    @staticmethod
    def var_k(grid, alpha, dt, k: Field, cg_tol, cg_max_iter, mixed: bool = True, n_pre=2, n_post=2,
              omega=2.0 / 3.0, coarsest=16, be=None) -> "CrankNicolsonSolver":
        """Extension: ∂u/∂t = α·∇·(k∇u) (zero Dirichlet) by Crank–Nicolson.  mixed: flexible CG around the
        fused FP32 V-cycle of the step operator I + (dt/2)·α·(−∇·k∇); otherwise CG around its FP64 V-cycle."""
        scale = 0.5 * dt * alpha
        mg = _mg_opts(n_pre, n_post, omega, SmootherKind.JACOBI, Transfer.FW_BILINEAR, coarsest, 10,
                      DType.F32 if mixed else DType.F64, mixed, 1.0, scale, k)
        return CrankNicolsonSolver(grid, alpha, dt, Boundaries.zero_dirichlet(), cg_tol, cg_max_iter, mg=mg,
                                   op=2, solver=4 if mixed else 0, k=k, be=be)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_cn_destroy(h)
            self.h = None

    def step(self, u: Field, source: Optional[Field] = None) -> SolveStats:
        out = L.FxSolveStats()
        L.check(self._lib.fx_cn_step(self.h, u.h, source.h if source is not None else None, C.byref(out)),
                "cn_step")
        return SolveStats(int(out.iterations), out.residual, ConvergenceReason(out.reason), out.true_residual)

    def step_n(self, u: Field, n: int) -> None:
        L.check(self._lib.fx_cn_step_n(self.h, u.h, int(n)), "cn_step_n")

    def dt(self) -> float:
        return float(self._lib.fx_cn_dt(self.h))

    def steps_done(self) -> int:
        return int(self._lib.fx_cn_steps_done(self.h))

    def sim_time(self) -> float:
        return float(self._lib.fx_cn_sim_time(self.h))

    def history(self) -> list:
        """Residual history of the last step's solve."""
        s = self._lib.fx_cn_solver(self.h)
        n = self._lib.fx_solver_history(s, None, 0)
        hist = (C.c_double * max(n, 1))()
        self._lib.fx_solver_history(s, hist, n)
        return list(hist[:n])

    def autotune(self, reps: int = 5) -> float:
        """Wave segment heights of the step's multigrid (fx_multigrid_autotune); ms per cycle of level kernels."""
        m = self._lib.fx_cn_multigrid(self.h)
        if not m:
            raise ValueError("the stepper has no multigrid")
        ms = C.c_double()
        L.check(self._lib.fx_multigrid_autotune(m, reps, C.byref(ms)), "multigrid_autotune")
        return ms.value

    def multigrid(self) -> "_MgView":
        """The stepper's own multigrid (owned by the stepper), e.g. to build a learned preconditioner on."""
        m = self._lib.fx_cn_multigrid(self.h)
        if not m:
            raise ValueError("the stepper has no multigrid")
        return _MgView(m, self)

    def set_preconditioner(self, pc) -> None:
        """Solve every step with `pc` (kept alive by the stepper object) instead of its multigrid; None
        restores the multigrid (fx_cn_set_preconditioner)."""
        self._pc_keep = pc
        if pc is None:
            L.check(self._lib.fx_cn_set_preconditioner(self.h, None), "cn_set_preconditioner")
        else:
            c = pc._pc()
            L.check(self._lib.fx_cn_set_preconditioner(self.h, C.byref(c)), "cn_set_preconditioner")


class _MgView:
    """A multigrid owned by another object (no destructor call); `owner` keeps that object alive."""

    def __init__(self, h, owner):
        self.h = h
        self._owner = owner

    def _pc(self):
        return L.FxPrecond(1, 0, self.h)


class ImexSolver:
    """ImexSolver (crates/runtime/src/imex.rs:1-274), stepped on the device by fx_imex_*.

    new(grid, kappa, vx, vy, dt, boundaries, cg_tol, cg_max_iter)                       imex.rs:46-93
    with_multigrid(..., mg_pre, mg_post, mg_omega)                                      imex.rs:95-137
    step(u) -> SolveStats, step_n(u, n), dt(), steps_done(), sim_time().
    """

    def __init__(self, grid, kappa, vx, vy, dt, boundaries, cg_tol, cg_max_iter, mg=None, be=None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grid = grid
        self._mg = mg
        g, b = L.grid_c(grid), L.bcs_c(boundaries)
        h = C.c_void_p()
        L.check(self._lib.fx_imex_create(self.be.h, C.byref(g), kappa, vx, vy, dt, C.byref(b), cg_tol,
                                         int(cg_max_iter), C.byref(mg) if mg is not None else None, C.byref(h)),
                "imex_create")
        self.h = h

    @staticmethod
    def new(grid, kappa, vx, vy, dt, boundaries, cg_tol, cg_max_iter, be=None) -> "ImexSolver":
        return ImexSolver(grid, kappa, vx, vy, dt, boundaries, cg_tol, cg_max_iter, be=be)

    @staticmethod
    def with_multigrid(grid, kappa, vx, vy, dt, boundaries, cg_tol, cg_max_iter, mg_pre, mg_post, mg_omega,
                       be=None) -> "ImexSolver":
        return ImexSolver(grid, kappa, vx, vy, dt, boundaries, cg_tol, cg_max_iter,
                          mg=_mg_opts(mg_pre, mg_post, mg_omega), be=be)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_imex_destroy(h)
            self.h = None

    def step(self, u: Field) -> SolveStats:
        out = L.FxSolveStats()
        L.check(self._lib.fx_imex_step(self.h, u.h, C.byref(out)), "imex_step")
        return SolveStats(int(out.iterations), out.residual, ConvergenceReason(out.reason), out.true_residual)

    def step_n(self, u: Field, n: int) -> None:
        L.check(self._lib.fx_imex_step_n(self.h, u.h, int(n)), "imex_step_n")

    def dt(self) -> float:
        return float(self._lib.fx_imex_dt(self.h))

    def steps_done(self) -> int:
        return int(self._lib.fx_imex_steps_done(self.h))

    def sim_time(self) -> float:
        return float(self._lib.fx_imex_sim_time(self.h))


class BdfSolver:
    """BdfSolver (crates/runtime/src/bdf.rs:1-256), stepped on the device by fx_bdf_*.

    new(grid, alpha, dt, order, boundaries, cg_tol, cg_max_iter)                      bdf.rs:58-108
    with_multigrid(..., mg_pre, mg_post, mg_omega)                                   bdf.rs:110-137
    step(u) -> SolveStats, step_n(u, n), dt(), steps_done(), sim_time(), order().
    """

    def __init__(self, grid: Grid, alpha: float, dt: float, order: int, boundaries: Boundaries, cg_tol: float,
                 cg_max_iter: int, mg: Optional[L.FxMgOptions] = None, be: Optional[HipBackend] = None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grid = grid
        self._mg = mg
        g, b = L.grid_c(grid), L.bcs_c(boundaries)
        h = C.c_void_p()
        L.check(self._lib.fx_bdf_create(self.be.h, C.byref(g), alpha, dt, int(order), C.byref(b), cg_tol,
                                        int(cg_max_iter), C.byref(mg) if mg is not None else None, C.byref(h)),
                "bdf_create")
        self.h = h

    @staticmethod
    def new(grid, alpha, dt, order, boundaries, cg_tol, cg_max_iter, be=None) -> "BdfSolver":
        return BdfSolver(grid, alpha, dt, order, boundaries, cg_tol, cg_max_iter, be=be)

    @staticmethod
    def with_multigrid(grid, alpha, dt, order, boundaries, cg_tol, cg_max_iter, mg_pre, mg_post, mg_omega,
                       be=None) -> "BdfSolver":
        return BdfSolver(grid, alpha, dt, order, boundaries, cg_tol, cg_max_iter,
                         mg=_mg_opts(mg_pre, mg_post, mg_omega), be=be)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_bdf_destroy(h)
            self.h = None

    def step(self, u: Field) -> SolveStats:
        out = L.FxSolveStats()
        L.check(self._lib.fx_bdf_step(self.h, u.h, C.byref(out)), "bdf_step")
        return SolveStats(int(out.iterations), out.residual, ConvergenceReason(out.reason), out.true_residual)

    def step_n(self, u: Field, n: int) -> None:
        L.check(self._lib.fx_bdf_step_n(self.h, u.h, int(n)), "bdf_step_n")

    def dt(self) -> float:
        return float(self._lib.fx_bdf_dt(self.h))

    def steps_done(self) -> int:
        return int(self._lib.fx_bdf_steps_done(self.h))

    def sim_time(self) -> float:
        return float(self._lib.fx_bdf_sim_time(self.h))

    def order(self) -> int:
        return int(self._lib.fx_bdf_order(self.h))


# ── explicit time stepping (fx_diffusion_*) ───────────────────────────
class DiffusionSolver:
    """DiffusionSolver (crates/runtime/src/solver.rs:25-494): ∂u/∂t = α·Δu + source, explicit.

    new(grid, alpha, dt)                                  solver.rs:51-69 (Euler, zero Dirichlet)
    with_boundaries(grid, alpha, dt, boundaries)          solver.rs:71-93
    build(grid, alpha, dt, boundaries, scheme)            solver.rs:96-148
    dt=None takes the scheme's CFL-safe step.  step, step_n, step_n_with_source, enable_hip_graph,
    dt(), steps_done(), sim_time(), scheme() as in solver.rs:149-337.
    """

    def __init__(self, grid: Grid, alpha: float, dt: Optional[float], boundaries: Boundaries,
                 scheme: TimeScheme = TimeScheme.EULER, be: Optional[HipBackend] = None):
        self.be = _backend(be)
        self._lib = self.be._lib
        self.grid = grid
        g, b = L.grid_c(grid), L.bcs_c(boundaries)
        d = C.c_double(dt) if dt is not None else None
        h = C.c_void_p()
        L.check(self._lib.fx_diffusion_create(self.be.h, C.byref(g), alpha, C.byref(d) if d is not None else None,
                                              C.byref(b), int(scheme), C.byref(h)), "diffusion_create")
        self.h = h

    @staticmethod
    def new(grid, alpha, dt=None, be=None) -> "DiffusionSolver":
        return DiffusionSolver(grid, alpha, dt, Boundaries.zero_dirichlet(), TimeScheme.EULER, be)

    @staticmethod
    def with_boundaries(grid, alpha, dt, boundaries, be=None) -> "DiffusionSolver":
        return DiffusionSolver(grid, alpha, dt, boundaries, TimeScheme.EULER, be)

    @staticmethod
    def build(grid, alpha, dt, boundaries, scheme, be=None) -> "DiffusionSolver":
        return DiffusionSolver(grid, alpha, dt, boundaries, scheme, be)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_diffusion_destroy(h)
            self.h = None

    def enable_hip_graph(self) -> None:
        L.check(self._lib.fx_diffusion_enable_graph(self.h), "diffusion_enable_graph")

    def step(self, u: Field, source: Optional[Field] = None) -> None:
        L.check(self._lib.fx_diffusion_step(self.h, u.h, source.h if source is not None else None),
                "diffusion_step")

    def step_n(self, u: Field, n: int) -> None:
        L.check(self._lib.fx_diffusion_step_n(self.h, u.h, int(n)), "diffusion_step_n")

    def step_n_with_source(self, u: Field, n: int, source: Field) -> None:
        L.check(self._lib.fx_diffusion_step_n_with_source(self.h, u.h, int(n), source.h),
                "diffusion_step_n_with_source")

    def dt(self) -> float:
        return float(self._lib.fx_diffusion_dt(self.h))

    def steps_done(self) -> int:
        return int(self._lib.fx_diffusion_steps_done(self.h))

    def sim_time(self) -> float:
        return float(self._lib.fx_diffusion_sim_time(self.h))

    def scheme(self) -> TimeScheme:
        return TimeScheme(self._lib.fx_diffusion_scheme(self.h))


# ── raw field format (fx_*_raw; crates/io/src/raw.rs) ──────────────────
def write_raw(path, field: Field) -> None:
    """write_raw (raw.rs:13-32): 32-byte little-endian header rows, cols (u64), dx, dy (f64), then f64 data."""
    g = L.grid_c(field.grid)
    L.check(field._lib.fx_write_raw(field.be.h, str(path).encode(), field.h, C.byref(g)), "write_raw")


def read_raw(path, be: Optional[HipBackend] = None) -> Field:
    """read_raw (raw.rs:38-77)."""
    be = _backend(be)
    g = L.FxGrid()
    h = C.c_void_p()
    L.check(be._lib.fx_read_raw(be.h, str(path).encode(), C.byref(g), C.byref(h)), "read_raw")
    return Field(Grid(int(g.rows), int(g.cols), g.dx, g.dy), Buffer(be, h))


def write_raw_host(path, data, grid: Grid) -> None:
    """The raw format from host memory (no device)."""
    a = np.ascontiguousarray(data, dtype=np.float64).ravel()
    if a.size != len(grid):
        raise DimensionMismatch(f"expected {len(grid)}, got {a.size}")
    g = L.grid_c(grid)
    L.check(L.lib().fx_write_raw_host(str(path).encode(), a.ctypes.data, C.byref(g)), "write_raw_host")


def read_raw_host(path):
    """(grid, data) of a raw file, read into host memory (no device)."""
    g = L.FxGrid()
    L.check(L.lib().fx_read_raw_host(str(path).encode(), C.byref(g), None, 0), "read_raw_host")
    grid = Grid(int(g.rows), int(g.cols), g.dx, g.dy)
    out = np.empty(len(grid), dtype=np.float64)
    L.check(L.lib().fx_read_raw_host(str(path).encode(), C.byref(g), out.ctypes.data, out.size), "read_raw_host")
    return grid, out
