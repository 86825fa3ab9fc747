"""ctypes wrapper of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the fluxion_amd product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from fluxion_amd.core import Boundaries, ConvergenceReason, SolveStats, Stencil

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE), "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


class OStencil(C.Structure):
    _fields_ = [("n", C.c_int32), ("dr", C.c_int32 * 32), ("dc", C.c_int32 * 32), ("w", C.c_double * 32)]


class OBcs(C.Structure):
    _fields_ = [("kind", C.c_int32 * 4), ("param", C.c_double * 12)]


class OOp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("coeff", C.c_double), ("sigma", C.c_double),
                ("st", OStencil), ("bc", OBcs), ("kappa", C.c_void_p), ("kappa_scalar", C.c_double),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("k", C.c_void_p)]


class OMg(C.Structure):
    _fields_ = [("n_pre", C.c_int32), ("n_post", C.c_int32), ("omega", C.c_double), ("smoother", C.c_int32),
                ("transfer", C.c_int32), ("coarsest", C.c_int32), ("extra_coarse", C.c_int32),
                ("precision", C.c_int32), ("pad", C.c_int32), ("sigma", C.c_double), ("scale", C.c_double),
                ("bc", OBcs), ("k", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p), ("kappa", C.c_double)]


class OStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("residual", C.c_double), ("reason", C.c_int32), ("pad", C.c_int32)]


# Smoother codes of the oracle's V-cycle.
SMOOTH_REF_GS = 0      # in-place lexicographic sweep (ops.rs:447-471)
SMOOTH_JACOBI = 1      # two-buffer Jacobi (the HIP library's smoother)
SMOOTH_CHEBYSHEV = 2   # multigrid.rs:317-473

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        for name in ("orc_dot", "orc_norm_l2", "orc_reduce_sum", "orc_reduce_max", "orc_reduce_min"):
            getattr(_lib, name).restype = C.c_double
        for name in ("orc_axpy", "orc_scale", "orc_fill"):
            pass
        _lib.orc_axpy.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_int64]
        _lib.orc_scale.argtypes = [C.c_void_p, C.c_double, C.c_int64]
        _lib.orc_fill.argtypes = [C.c_void_p, C.c_double, C.c_int64]
        _lib.orc_waxpy.argtypes = [C.c_double, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int64]
        _lib.orc_aypx.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_int64]
        _lib.orc_dot.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        for name in ("orc_norm_l2", "orc_reduce_sum", "orc_reduce_max", "orc_reduce_min"):
            getattr(_lib, name).argtypes = [C.c_void_p, C.c_int64]
        _lib.orc_conv_diff.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                       C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                       C.c_double]
        _lib.orc_vark_apply.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.c_int64, C.c_double, C.c_double]
        _lib.orc_weighted_jacobi.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int64, C.c_int64, C.c_double,
                                             C.c_double, C.POINTER(OStencil), C.c_int]
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous and a.dtype == np.float64
    return a.ctypes.data


def ostencil(st: Stencil) -> OStencil:
    s = OStencil()
    s.n = len(st.entries)
    for i, e in enumerate(st.entries):
        s.dr[i], s.dc[i], s.w[i] = e.dr, e.dc, e.weight
    return s


def obcs(bc: Boundaries) -> OBcs:
    b = OBcs()
    for i, s in enumerate(bc.sides()):
        b.kind[i] = int(s.kind)
        b.param[3 * i], b.param[3 * i + 1], b.param[3 * i + 2] = s.a, s.b, s.g
    return b


def mg_cfg(n_pre=2, n_post=2, omega=2.0 / 3.0, smoother=SMOOTH_JACOBI, transfer=0, coarsest=4, extra_coarse=10,
           precision=0, sigma=0.0, scale=1.0, bc: Boundaries = None, k=None, vx=None, vy=None, kappa=0.0) -> OMg:
    """Multigrid options; `k` (nodal coefficient, kept alive by the returned object) selects the
    level operator sigma·I + scale·(−∇·k∇); `vx`, `vy` (nodal velocity) select the upwind
    convection-diffusion level operator sigma·I + scale·(κΔ + v·∇⁻) with κ = kappa."""
    c = OMg()
    if k is not None:
        c._k = np.ascontiguousarray(k, dtype=np.float64)
        c.k = c._k.ctypes.data
    if vx is not None:
        c._vx = np.ascontiguousarray(vx, dtype=np.float64)
        c._vy = np.ascontiguousarray(vy, dtype=np.float64)
        c.vx, c.vy = c._vx.ctypes.data, c._vy.ctypes.data
    c.kappa = kappa
    c.n_pre, c.n_post, c.omega, c.smoother = n_pre, n_post, omega, smoother
    c.transfer, c.coarsest, c.extra_coarse, c.precision = transfer, coarsest, extra_coarse, precision
    c.sigma, c.scale = sigma, scale
    c.bc = obcs(bc or Boundaries.zero_dirichlet())
    return c


class Op:
    """Oracle operator (see orc_op in oracle.cpp)."""

    def __init__(self, kind=0, coeff=1.0, sigma=1.0, stencil=None, bc=None, kappa=None, kappa_scalar=0.0,
                 vx=None, vy=None, k=None):
        self._keep = [a for a in (kappa, vx, vy, k) if a is not None]
        self.c = OOp()
        self.c.kind, self.c.coeff, self.c.sigma = kind, coeff, sigma
        if stencil is not None:
            self.c.st = ostencil(stencil)
        self.c.bc = obcs(bc or Boundaries.zero_dirichlet())
        self.c.kappa = _p(kappa) if kappa is not None else None
        self.c.kappa_scalar = kappa_scalar
        self.c.vx = _p(vx) if vx is not None else None
        self.c.vy = _p(vy) if vy is not None else None
        self.c.k = _p(k) if k is not None else None

    @staticmethod
    def stencil(coeff, st, bc=None):
        return Op(0, coeff, 1.0, stencil=st, bc=bc)

    @staticmethod
    def conv_diff(vx, vy, kappa=None, kappa_scalar=0.0, coeff=1.0, sigma=0.0):
        return Op(1, coeff, sigma, kappa=kappa, kappa_scalar=kappa_scalar, vx=vx, vy=vy)

    @staticmethod
    def var_k(k, coeff=1.0, sigma=0.0):
        return Op(2, coeff, sigma, k=k)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


# ── ops ──────────────────────────────────────────────────────────────
def apply_stencil(x, grid, st, bc):
    out = np.zeros_like(x)
    s, b = ostencil(st), obcs(bc)
    _check(lib().orc_apply_stencil(C.c_void_p(_p(x)), C.c_void_p(_p(out)), C.c_int64(grid.rows),
                                   C.c_int64(grid.cols), C.c_double(grid.dx), C.c_double(grid.dy), C.byref(s),
                                   C.byref(b)), "apply_stencil")
    return out


def apply_stencil_var(x, coeff, grid, st, bc):
    out = np.zeros_like(x)
    s, b = ostencil(st), obcs(bc)
    _check(lib().orc_apply_stencil_var(C.c_void_p(_p(x)), C.c_void_p(_p(out)), C.c_void_p(_p(coeff)),
                                       C.c_int64(grid.rows), C.c_int64(grid.cols), C.c_double(grid.dx),
                                       C.c_double(grid.dy), C.byref(s), C.byref(b)), "apply_stencil_var")
    return out


def stencil_axpy(alpha, x, y, grid, st, bc):
    """Returns a new array y + alpha·S(x) (interior)."""
    y = np.array(y, dtype=np.float64, copy=True)
    s, b = ostencil(st), obcs(bc)
    _check(lib().orc_stencil_axpy(C.c_double(alpha), C.c_void_p(_p(x)), C.c_void_p(_p(y)), C.c_int64(grid.rows),
                                  C.c_int64(grid.cols), C.c_double(grid.dx), C.c_double(grid.dy), C.byref(s),
                                  C.byref(b)), "stencil_axpy")
    return y


def axpy(alpha, x, y):
    y = y.copy()
    lib().orc_axpy(alpha, _p(x), _p(y), x.size)
    return y


def scale(x, alpha):
    x = x.copy()
    lib().orc_scale(_p(x), alpha, x.size)
    return x


def waxpy(alpha, x, beta, y):
    w = np.empty_like(x)
    lib().orc_waxpy(alpha, _p(x), beta, _p(y), _p(w), x.size)
    return w


def aypx(alpha, x, y):
    y = y.copy()
    lib().orc_aypx(alpha, _p(x), _p(y), x.size)
    return y


def pointwise(op, x, y):
    z = np.empty_like(x)
    lib().orc_pointwise(C.c_int(op), C.c_void_p(_p(x)), C.c_void_p(_p(y)), C.c_void_p(_p(z)), C.c_int64(x.size))
    return z


def dot(x, y):
    return lib().orc_dot(_p(x), _p(y), x.size)


def norm_l2(x):
    return lib().orc_norm_l2(_p(x), x.size)


def reduce_sum(x):
    return lib().orc_reduce_sum(_p(x), x.size)


def reduce_max(x):
    return lib().orc_reduce_max(_p(x), x.size)


def reduce_min(x):
    return lib().orc_reduce_min(_p(x), x.size)


def restrict(f, fg, cg, mode=0):
    c = np.zeros(cg.rows * cg.cols)
    lib().orc_restrict(C.c_void_p(_p(f)), C.c_void_p(_p(c)), C.c_int64(fg.rows), C.c_int64(fg.cols),
                       C.c_int64(cg.rows), C.c_int64(cg.cols), C.c_int(mode))
    return c


def prolong(c, f, cg, fg, mode=0):
    f = f.copy()
    lib().orc_prolong(C.c_void_p(_p(c)), C.c_void_p(_p(f)), C.c_int64(cg.rows), C.c_int64(cg.cols),
                      C.c_int64(fg.rows), C.c_int64(fg.cols), C.c_int(mode))
    return f


def weighted_jacobi(x, b, omega, grid, st, mode=0):
    """mode 0: reference in-place lexicographic sweep; 1: two-buffer Jacobi."""
    x = x.copy()
    s = ostencil(st)
    lib().orc_weighted_jacobi(_p(x), _p(b), omega, grid.rows, grid.cols, grid.dx, grid.dy, C.byref(s), mode)
    return x


def conv_diff(mode, alpha, u, kappa, vx, vy, grid, out=None, kappa_scalar=0.0, sigma=1.0):
    out = np.zeros_like(u) if out is None else out.copy()
    lib().orc_conv_diff(mode, alpha, _p(u), _p(out), _p(kappa) if kappa is not None else None, kappa_scalar,
                        _p(vx), _p(vy), grid.rows, grid.cols, grid.dx, grid.dy, sigma)
    return out


def vark_apply(sigma, alpha, k, u, grid):
    out = np.zeros_like(u)
    lib().orc_vark_apply(sigma, alpha, _p(k), _p(u), _p(out), grid.rows, grid.cols, grid.dx, grid.dy)
    return out


def apply_op(op: Op, x, grid):
    y = np.zeros_like(x)
    _check(lib().orc_apply_op(C.byref(op.c), C.c_void_p(_p(x)), C.c_void_p(_p(y)), C.c_int64(grid.rows),
                              C.c_int64(grid.cols), C.c_double(grid.dx), C.c_double(grid.dy)), "apply_op")
    return y


def zero_boundaries(x, rows, cols):
    x = x.copy()
    lib().orc_zero_boundaries(C.c_void_p(_p(x)), C.c_int64(rows), C.c_int64(cols))
    return x


def mg_depth(grid, cfg: OMg) -> int:
    return lib().orc_mg_depth(C.c_int64(grid.rows), C.c_int64(grid.cols), C.c_double(grid.dx),
                              C.c_double(grid.dy), C.byref(cfg))


def vcycle(x, b, grid, cfg: OMg, n_cycles=1):
    x = np.array(x, dtype=np.float64, copy=True)
    lib().orc_vcycle(C.c_void_p(_p(x)), C.c_void_p(_p(b)), C.c_int64(grid.rows), C.c_int64(grid.cols),
                     C.c_double(grid.dx), C.c_double(grid.dy), C.byref(cfg), C.c_int(n_cycles))
    return x


# ── solvers ──────────────────────────────────────────────────────────
def _solve(fn, x, b, grid, op, tol, max_iter, extra=(), hist=True):
    x = np.array(x, dtype=np.float64, copy=True)
    st = OStats()
    h = np.zeros(max(1, max_iter + 2)) if hist else None
    args = [C.c_void_p(_p(x)), C.c_void_p(_p(b)), C.c_int64(grid.rows), C.c_int64(grid.cols),
            C.c_double(grid.dx), C.c_double(grid.dy), C.byref(op.c)]
    args += list(extra)
    args += [C.c_double(tol), C.c_int64(max_iter), C.byref(st), C.c_void_p(_p(h)) if h is not None else None]
    _check(fn(*args), fn.__name__)
    n = min(int(st.iterations) + 1, h.size) if h is not None else 0
    stats = SolveStats(int(st.iterations), st.residual, ConvergenceReason(st.reason),
                       history=list(h[:n]) if h is not None else [])
    return x, stats


def _pc(mg):
    return (C.c_int(1 if mg is not None else 0), C.byref(mg) if mg is not None else None)


def cg(x, b, grid, op, tol, max_iter):
    """CgSolver::solve (cg.rs:96-164)."""
    return _solve(lib().orc_cg, x, b, grid, op, tol, max_iter)


def pcg(x, b, grid, op, tol, max_iter, mg: OMg = None, identity=False):
    """CgSolver::solve_preconditioned (cg.rs:178-256); mg=None & identity → Identity."""
    kind = 0 if mg is None else 1
    return _solve(lib().orc_pcg, x, b, grid, op, tol, max_iter,
                  (C.c_int(kind), C.byref(mg) if mg is not None else None))


def fcg(x, b, grid, op, tol, max_iter, mg: OMg = None):
    return _solve(lib().orc_fcg, x, b, grid, op, tol, max_iter, _pc(mg))


def pipelined_cg(x, b, grid, op, tol, max_iter):
    return _solve(lib().orc_pipelined_cg, x, b, grid, op, tol, max_iter)


def ppcg(x, b, grid, op, tol, max_iter, mg: OMg = None):
    return _solve(lib().orc_ppcg, x, b, grid, op, tol, max_iter, _pc(mg))


def bicgstab(x, b, grid, op, tol, max_iter, mg: OMg = None, identity=False):
    if mg is None and identity:
        extra = (C.c_int(3), None)   # any non-zero kind without a config = identity preconditioner
    else:
        extra = _pc(mg)
    return _solve(lib().orc_bicgstab, x, b, grid, op, tol, max_iter, extra)


def gmres(x, b, grid, op, restart, tol, max_iter, mg: OMg = None, identity=False):
    if mg is None and identity:
        pc = (C.c_int(3), None)
    else:
        pc = _pc(mg)
    return _solve(lib().orc_gmres, x, b, grid, op, tol, max_iter, (C.c_int64(restart),) + tuple(pc))


def fgmres(x, b, grid, op, restart, tol, max_iter, mg: OMg = None):
    """Flexible right-preconditioned GMRES(m) (extension): z_k = M⁻¹v_k stored, x += Z·y."""
    return _solve(lib().orc_fgmres, x, b, grid, op, tol, max_iter, (C.c_int64(restart),) + tuple(_pc(mg)))


# ── CrankNicolsonSolver::step (runtime/src/implicit.rs:178-222) ────────
def cn_step(u, grid, alpha, dt, bc=None, source=None, tol=1e-8, max_iter=200, mg: OMg = None, k=None,
            flexible=False):
    """One Crank–Nicolson step from u; returns (u_new, stats).

    h = 0.5·dt·α evaluated left to right (implicit.rs:183).  RHS = u + h·S(u) on the interior and u on the
    boundary (copy, copy, stencil_axpy: implicit.rs:187-194), then RHS += dt·source (axpy, 195-197); the
    system u − h·S(u) = RHS (identity boundary rows, coeff = −h) is solved from u by CG, or by PCG with the
    multigrid `mg` (implicit.rs:199-218), S = Stencil::laplacian_2d_5pt(dx, dy).  k (extension): the same
    recurrence with the variable-coefficient operator −∇·(k∇·) in place of −S (explicit side
    u − h·(−∇·k∇u), system u + h·(−∇·k∇u)); flexible selects FCG (needs mg)."""
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    bc = bc or Boundaries.zero_dirichlet()
    h = 0.5 * dt * alpha
    if k is None:
        st = Stencil.laplacian_2d_5pt(grid.dx, grid.dy)
        rhs = stencil_axpy(h, u, u, grid, st, bc)
        system = Op.stencil(-h, st, bc)
    else:
        rhs = apply_op(Op.var_k(k, coeff=-h, sigma=1.0), u, grid)
        system = Op.var_k(k, coeff=h, sigma=1.0)
    if source is not None:
        rhs = axpy(dt, np.ascontiguousarray(source, dtype=np.float64).ravel(), rhs)
    if flexible:
        return fcg(u, rhs, grid, system, tol, max_iter, mg=mg)
    if mg is None:
        return cg(u, rhs, grid, system, tol, max_iter)
    return pcg(u, rhs, grid, system, tol, max_iter, mg=mg)


# ── ImexSolver::step (runtime/src/imex.rs:147-197) ─────────────────────
def imex_step(u, grid, kappa, vx, vy, dt, bc=None, tol=1e-8, max_iter=200, mg: OMg = None):
    """conv = v·∇u by apply_conv_diff with a zero κ field and uniform velocity fields (imex.rs:215-274), then the
    Crank–Nicolson step in κ with conv as its source (rhs = u + h·S(u) + dt·conv, system u − h·S(u))."""
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    n = u.size
    conv = conv_diff(0, 0.0, u, np.zeros(n), np.full(n, float(vx)), np.full(n, float(vy)), grid)
    return cn_step(u, grid, kappa, dt, bc=bc, source=conv, tol=tol, max_iter=max_iter, mg=mg)


# ── BdfSolver::step (runtime/src/bdf.rs:163-240) ───────────────────────
BDF_COEFFS = {1: ([-1.0], 1.0), 2: ([-2.0, 0.5], 1.5), 3: ([-3.0, 1.5, -1.0 / 3.0], 11.0 / 6.0),
              4: ([-4.0, 3.0, -4.0 / 3.0, 0.25], 25.0 / 12.0)}   # bdf.rs:41-56


def bdf_step(u, history, steps_done, order, grid, alpha, dt, bc=None, tol=1e-8, max_iter=200, mg: OMg = None):
    """One BDF step of the oracle at order min(order, steps_done + 1): rhs = copy(u)·(−a0), rhs −= a_i·h_{i−1},
    rhs ·= 1/lhs, history shifted (history[0] = u^n), then CG / PCG on u + coeff·S(u), coeff = −dt·α/lhs,
    from u.  Returns (u_new, history_new, stats); history is a list of arrays, newest first."""
    bc = bc or Boundaries.zero_dirichlet()
    st = Stencil.laplacian_2d_5pt(grid.dx, grid.dy)
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    eff = min(order, steps_done + 1)
    a, lhs = BDF_COEFFS[eff]
    rhs = scale(u.copy(), -a[0])
    for i in range(1, eff):
        rhs = axpy(-a[i], history[i - 1], rhs)
    coeff = -dt * alpha / lhs
    rhs = scale(rhs, 1.0 / lhs)
    history = [u.copy()] + list(history[:order - 1])
    if mg is None:
        x, s = cg(u, rhs, grid, Op.stencil(coeff, st, bc), tol, max_iter)
    else:
        x, s = pcg(u, rhs, grid, Op.stencil(coeff, st, bc), tol, max_iter, mg=mg)
    return x, history, s


# ── DiffusionSolver (runtime/src/solver.rs:96-494) ─────────────────────
def diffusion_dt(grid, alpha, scheme):
    """Default step safety·min(dx, dy)²/(4α), safety 0.4 / 0.8 / 1.4 for Euler / RK2 / RK4 (solver.rs:103-113)."""
    h_min = min(grid.dx, grid.dy)
    safety = (0.4, 0.8, 1.4)[int(scheme)]
    return safety * h_min * h_min / (4.0 * alpha)


def diffusion_step(u, grid, alpha, dt, bc=None, scheme=0, source=None):
    """One explicit step of the oracle, the reference's operations in its order:
    Euler (solver.rs:344-367), Heun (371-418), RK4 (421-494).  Returns the new u."""
    bc = bc or Boundaries.zero_dirichlet()
    st = Stencil.laplacian_2d_5pt(grid.dx, grid.dy)
    u = np.array(u, dtype=np.float64, copy=True).ravel()
    src = None if source is None else np.ascontiguousarray(source, dtype=np.float64).ravel()
    rows, cols = grid.rows, grid.cols
    if int(scheme) == 0:
        u = stencil_axpy(alpha * dt, u.copy(), u, grid, st, bc)
        if src is not None:
            u = axpy(dt, src, u)
        return u
    if int(scheme) == 1:
        coeff = dt * alpha
        u_old = u.copy()
        for _ in range(2):
            u = stencil_axpy(coeff, u.copy(), u, grid, st, bc)
            if src is not None:
                u = axpy(dt, src, u)
        u = axpy(1.0, u_old, u)
        return scale(u, 0.5)
    u_old = u.copy()
    k1 = apply_stencil(u, grid, st, bc)
    u = stencil_axpy(0.5 * dt * alpha, u_old, u_old.copy(), grid, st, bc)
    k2 = apply_stencil(u, grid, st, bc)
    u = zero_boundaries(axpy(0.5 * dt * alpha, k2, u_old.copy()), rows, cols)
    k3 = apply_stencil(u, grid, st, bc)
    u = zero_boundaries(axpy(dt * alpha, k3, u_old.copy()), rows, cols)
    k4 = apply_stencil(u, grid, st, bc)
    u = u_old.copy()
    c = dt * alpha / 6.0
    u = axpy(c, k1, u)
    u = axpy(2.0 * c, k2, u)
    u = axpy(2.0 * c, k3, u)
    u = axpy(c, k4, u)
    if src is not None:
        u = axpy(dt, src, u)
    return u


def num_threads() -> int:
    return lib().orc_num_threads()


# ── learned preconditioner network (unet.cpp) ─────────────────────────
def unet_nparams(cin=2) -> int:
    L = lib()
    L.orc_unet_nparams.restype = C.c_size_t
    return int(L.orc_unet_nparams(C.c_int(cin)))


def unet_layout(cin=2):
    woff = (C.c_int64 * 15)()
    boff = (C.c_int64 * 15)()
    lcin = (C.c_int * 15)()
    lcout = (C.c_int * 15)()
    lib().orc_unet_layout(C.c_int(cin), woff, boff, lcin, lcout)
    return list(woff), list(boff), list(lcin), list(lcout)


def unet_forward(params, x, bf16=False):
    """x: (B, H, W, CIN) float64 -> (B, H, W).  bf16=True rounds input, weights, activations and
    pooled maps to BF16 (biases to FP32) as the GPU stores them; sums stay FP64."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    b, h, w, cin = x.shape
    if h % 8 or w % 8:
        raise ValueError(f"the U-Net pools three times: image sides must be multiples of 8, got {h}x{w}")
    y = np.zeros((b, h, w))
    p = np.ascontiguousarray(params, dtype=np.float64)
    lib().orc_unet_forward(C.c_void_p(_p(p)), C.c_void_p(_p(x)), C.c_void_p(_p(y)), C.c_int(b), C.c_int(h),
                           C.c_int(w), C.c_int(cin), C.c_int(int(bool(bf16))))
    return y


def unet_loss_grad(params, x, t, energy=False, bf16=False):
    """Relative squared error (or, with energy=True, relative energy norm under the unit-spacing
    −∇·(k∇·) with k = exp(x[..., 1])) summed over the batch, and its gradient w.r.t. the parameters.
    bf16=True emulates the GPU's BF16 storage (see unet_forward) and rounds the ReLU-masked
    gradient to BF16 where it multiplies weights and activations."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    b, h, w, cin = x.shape
    if h % 8 or w % 8:
        raise ValueError(f"the U-Net pools three times: image sides must be multiples of 8, got {h}x{w}")
    p = np.ascontiguousarray(params, dtype=np.float64)
    g = np.zeros_like(p)
    L = lib()
    fn = L.orc_unet_loss_grad_energy if energy else L.orc_unet_loss_grad
    fn.restype = C.c_double
    loss = fn(C.c_void_p(_p(p)), C.c_void_p(_p(x)), C.c_void_p(_p(t)), C.c_void_p(_p(g)),
              C.c_int(b), C.c_int(h), C.c_int(w), C.c_int(cin), C.c_int(int(bool(bf16))))
    return loss, g


# ── learned preconditioner inside flexible CG (composition of the pieces above) ──
def learned_fcg(f, grid, op: Op, mg: OMg, params, logk, tol, max_iter, stall=0.9, every=1, max_apps=0, first=0):
    """Flexible CG (Notay's β = −z·Ap_old / p_old·Ap_old) with the learned preconditioner
    z = V(r) + α*·c, c = s·h²·N([r1/s, log k]) on interior points, r1 = r − A·V(r), s = rms(r1),
    α* = (c·r1)/(c·A·c) clipped at 0, the network applied after first, first + every, … iterations (at
    most max_apps times, 0 = no limit) until an iteration whose direction came from a network application
    fails to cut the FP64 true residual ‖b − A·x‖ by the factor `stall`.  The operation order of the device solver; network in BF16 storage mode.
    Returns (x, iterations, fell_back)."""
    n = grid.rows * grid.cols
    h2 = grid.dx * grid.dy
    inner = np.zeros((grid.rows, grid.cols), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.ravel()
    A = lambda v: apply_op(op, v, grid)
    st = {"iter": 0, "plain": False, "apps": 0, "on": False}

    def pc(r):
        z = vcycle(np.zeros(n), r, grid, mg)
        on = ((not st["plain"]) and st["iter"] >= first and (st["iter"] - first) % every == 0
              and (max_apps <= 0 or st["apps"] < max_apps))
        st["on"] = on
        if not on:
            return z
        st["apps"] += 1
        r1 = r - A(z)
        s = float(np.sqrt(np.dot(r1, r1) / n))
        if s == 0.0:
            s = 1.0
        x = np.stack([(r1 / s).astype(np.float32).astype(np.float64).reshape(grid.rows, grid.cols),
                      logk.astype(np.float32).astype(np.float64).reshape(grid.rows, grid.cols)], axis=-1)[None]
        y = unet_forward(params, x, bf16=True)[0].astype(np.float32).astype(np.float64).ravel()
        c = np.where(inner, s * h2 * y, 0.0)
        ac = A(c)
        cr, cac = float(c @ r1), float(c @ ac)
        a = cr / cac if (cac > 0.0 and cr > 0.0) else 0.0
        return z + a * c

    x = np.zeros(n)
    r = f - A(x)
    ltres = float(r @ r)
    z = pc(r)
    rs = float(r @ z)
    if np.sqrt(r @ r) < tol:
        return x, 0, False
    p = np.zeros(n)
    ap = np.zeros(n)
    beta, pap = 0.0, 1.0
    for it in range(max_iter):
        p = z + beta * p
        ap = A(p)
        pap = float(p @ ap)
        alpha = rs / pap
        x = x + alpha * p
        r = r - alpha * ap
        st["iter"] = it + 1
        if np.sqrt(r @ r) < tol:
            return x, it + 1, st["plain"]
        if not st["plain"]:
            tr = f - A(x)
            t2 = float(tr @ tr)
            if st["on"] and not t2 < stall * stall * ltres:
                st["plain"] = True
            ltres = t2
        z = pc(r)
        rs = float(r @ z)
        beta = -float(z @ ap) / pap
    return x, max_iter, st["plain"]
