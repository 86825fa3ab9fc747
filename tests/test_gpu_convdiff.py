"""Upwind convection-diffusion −κΔu + v·∇⁻u = f (zero Dirichlet) with the convection-diffusion
multigrid preconditioner: BiCGSTAB (right, bicgstab.rs:218-346), GMRES(30) (left, gmres.rs:294-454)
and flexible GMRES(30) (right, extension) against the oracle, and the 2048² configuration (GPU).

Operator: sigma = 0, coeff = 1, κ field −κ, i.e. −κΔu + v·∇⁻u with the reference's upwind differences
(ops.rs:707-731).  The multigrid rediscretises the same operator on every level (FP32 or FP64 Jacobi
V(2,2), linear transfers, coarsest 16²)."""
import numpy as np
import pytest

import oracle as O
from fluxion_amd.core import Boundaries, DType, Grid, Transfer
from fluxion_amd.inputs import gaussian_bumps, streamfunction_velocity
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

ZD = Boundaries.zero_dirichlet()
KAPPA = 1e-3


def setup(be, rows, cols=None, prec=DType.F32):
    from fluxion_amd import hip
    cols = rows if cols is None else cols
    g = Grid(rows, cols, 1.0 / (cols - 1), 1.0 / (rows - 1))
    if rows == cols:
        vx, vy, pe = streamfunction_velocity(rows, KAPPA)
    else:   # a rectangular grid takes the square field's velocity on its own nodes
        vx, vy, pe = streamfunction_velocity(max(rows, cols), KAPPA)
        vx, vy, pe = vx[:rows, :cols].copy(), vy[:rows, :cols].copy(), pe[:rows, :cols]
    vx, vy = vx.ravel(), vy.ravel()
    f = gaussian_bumps(rows, cols=cols).ravel()
    FX, FY = hip.Field.from_f64(g, vx, be), hip.Field.from_f64(g, vy, be)
    op = hip.Operator.conv_diff(FX, FY, kappa_scalar=-KAPPA, coeff=1.0, sigma=0.0)
    mg = hip.Multigrid(g, ZD, 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16, precision=prec,
                       fused=False, scale=1.0, vx=FX, vy=FY, kappa=-KAPPA, be=be)
    cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI, transfer=1, coarsest=16,
                   precision=1 if prec == DType.F32 else 0, vx=vx, vy=vy, kappa=-KAPPA)
    op_r = O.Op.conv_diff(vx, vy, kappa_scalar=-KAPPA, coeff=1.0, sigma=0.0)
    return hip, g, f, op, mg, op_r, cfg, (FX, FY), pe


@pytest.mark.parametrize("rows,cols,prec", [(129, 129, DType.F32), (96, 160, DType.F32), (128, 128, DType.F64)])
def test_cd_vcycle_bitwise(be, rows, cols, prec):
    hip, g, f, op, mg, op_r, cfg, keep, _ = setup(be, rows, cols, prec)
    r = np.random.default_rng(rows).standard_normal(rows * cols)
    r = O.zero_boundaries(r, rows, cols)
    X = hip.Field.zeros(g, be=be)
    mg.v_cycle(X, hip.Field.from_f64(g, r, be))
    ref = O.vcycle(np.zeros(rows * cols), r, g, cfg)
    got = X.to_vec_f64()
    print(f"cd V-cycle {rows}x{cols} {prec.name}: max |diff| {np.max(np.abs(got - ref)):.3e}")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("solver", ["bicgstab", "fgmres", "gmres"])
def test_cd_preconditioned_parity(be, solver):
    hip, g, f, op, mg, op_r, cfg, keep, _ = setup(be, 128)
    tol = 1e-8 * np.linalg.norm(f)
    B = hip.Field.from_f64(g, f, be)
    X = hip.Field.zeros(g, be=be)
    x0 = np.zeros_like(f)
    if solver == "bicgstab":
        s = hip.BiCgStabSolver(g, tol, 1000, be=be).solve_op(X, B, op, mg)
        xr, sr = O.bicgstab(x0, f, g, op_r, tol, 1000, mg=cfg)
    elif solver == "fgmres":
        s = hip.FgmresSolver(g, 30, tol, 1000, be=be).solve_op(X, B, op, mg)
        xr, sr = O.fgmres(x0, f, g, op_r, 30, tol, 1000, mg=cfg)
    else:
        s = hip.GmresSolver(g, 30, tol, 1000, be=be).solve_op(X, B, op, mg)
        xr, sr = O.gmres(x0, f, g, op_r, 30, tol, 1000, mg=cfg)
    d = rel_l2(X.to_vec_f64(), xr)
    print(f"{solver} 128²: gpu {s.iterations} its, oracle {sr.iterations} its, rel diff {d:.2e}, "
          f"true rel res {s.true_residual / np.linalg.norm(f):.2e}")
    assert s.converged() and sr.converged()
    assert abs(s.iterations - sr.iterations) <= 2
    assert d < 1e-8
    if solver != "gmres":   # right preconditioning: the stopping residual is the true one
        assert s.true_residual < 1.5 * tol


def test_cd_2048_bicgstab_fgmres(be):
    # The full configuration: 2048², κ = 1e-3, cell Péclet up to 50, FP32 V-cycle inside FP64
    # BiCGSTAB and flexible GMRES(30), to 1e-8 relative FP64 residual.
    hip, g, f, op, mg, op_r, cfg, keep, pe = setup(be, 2048)
    print(f"2048²: cell Péclet min {pe.min():.3g}, median {np.median(pe):.3g}, max {pe.max():.3g}; "
          f"multigrid depth {mg.depth()}")
    fn = np.linalg.norm(f)
    tol = 1e-8 * fn
    B = hip.Field.from_f64(g, f, be)
    for name, solver in (("BiCGSTAB", hip.BiCgStabSolver(g, tol, 2000, be=be)),
                         ("FGMRES(30)", hip.FgmresSolver(g, 30, tol, 3000, be=be))):
        X = hip.Field.zeros(g, be=be)
        s = solver.solve_op(X, B, op, mg)
        print(f"{name} 2048²: {s.iterations} its, {s.reason.name}, true rel res {s.true_residual / fn:.2e}", flush=True)
        assert s.converged()
        assert s.true_residual < 1.5 * tol
