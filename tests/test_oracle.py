"""Pins the CPU oracle against the reference's own known answers, closed forms,
invariants and brute-force solves (CPU only).

Each reference case cites the test it comes from (crates/... paths).
"""
import numpy as np
import pytest

import oracle as O
from fluxion_amd.core import (Boundaries, BoundaryCondition, ConvergenceReason, Grid, Stencil)
from tests.conftest import gaussian_ic, rel_l2

ZD = Boundaries.zero_dirichlet()


def lap(h):
    return Stencil.laplacian_2d_5pt(h, h)


def test_fill_and_norm():
    # backend-cpu/src/ops.rs:1202-1209
    assert abs(O.norm_l2(np.full(16, 3.0)) - 12.0) < 1e-12


def test_axpy_basic():
    # ops.rs:1212-1223
    y = O.axpy(2.0, np.array([1.0, 2.0, 3.0]), np.array([10.0, 20.0, 30.0]))
    assert np.allclose(y, [12.0, 24.0, 36.0], atol=1e-14)


def test_laplacian_on_quadratic_field():
    # ops.rs:1227-1266: Δ(x²+y²) = 4 on 32².
    n = 32
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    out = O.apply_stencil((X * X + Y * Y).ravel(), Grid.square(n, h), lap(h), ZD).reshape(n, n)
    assert np.max(np.abs(out[2:-2, 2:-2] - 4.0)) < 1e-10


def test_dealii_quadratic():
    # runtime/tests/manufactured.rs:71-107: Δ(1 + x² + 2y²) = 6 exactly for the 5-point stencil.
    n = 64
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    u = 1.0 + X * X + 2.0 * Y * Y
    out = O.apply_stencil(u.ravel(), Grid.square(n, h), lap(h), ZD).reshape(n, n)
    assert np.max(np.abs(out[1:-1, 1:-1] - 6.0)) < 1e-10


def test_dirichlet_nonzero_boundary():
    # ops.rs:1269-1305
    n = 16
    h = 1.0 / (n - 1)
    bc = Boundaries.uniform(BoundaryCondition.dirichlet(100.0))
    out = O.apply_stencil(np.full(n * n, 50.0), Grid.square(n, h), lap(h), bc).reshape(n, n)
    assert np.all(np.abs(out[0] - 100.0) < 1e-12) and np.all(np.abs(out[-1] - 100.0) < 1e-12)


def test_neumann_zero_flux():
    # ops.rs:1308-1338
    n = 16
    h = 1.0 / (n - 1)
    bc = Boundaries.uniform(BoundaryCondition.neumann(0.0))
    out = O.apply_stencil(np.full(n * n, 42.0), Grid.square(n, h), lap(h), bc).reshape(n, n)
    assert np.all(np.abs(out[0] - 42.0) < 1e-10)


@pytest.mark.parametrize("c", [492.4908687776044, -999.5, 0.0, 1.0])
def test_laplacian_of_constant(c):
    # runtime/tests/properties.rs:17-40 with the committed regression seed
    # (properties.proptest-regressions:7).
    n = 32
    h = 1.0 / (n - 1)
    out = O.apply_stencil(np.full(n * n, c), Grid.square(n, h), lap(h), ZD).reshape(n, n)
    assert np.max(np.abs(out[1:-1, 1:-1])) < 1e-10 * (1.0 + abs(c))


def test_stencil_weights():
    # core/src/stencil.rs:150-188 (5-pt sums to 0, centre −4/h²) and 214-258 (Collatz 9-pt).
    h = 0.1
    w5 = {(e.dr, e.dc): e.weight for e in Stencil.laplacian_2d_5pt(h, h).entries}
    assert abs(sum(w5.values())) < 1e-9 and abs(w5[(0, 0)] + 4.0 / (h * h)) < 1e-9
    w9 = {(e.dr, e.dc): e.weight * 6 * h * h for e in Stencil.laplacian_2d_9pt(h, h).entries}
    assert abs(w9[(0, 0)] + 20) < 1e-9
    for k in [(-1, 0), (1, 0), (0, -1), (0, 1)]:
        assert abs(w9[k] - 4) < 1e-9
    for k in [(-1, -1), (-1, 1), (1, -1), (1, 1)]:
        assert abs(w9[k] - 1) < 1e-9


def test_9pt_exact_on_quartic_harmonic_part():
    # The Collatz stencil is 4th-order: exact (to rounding) on u = x⁴ − 6x²y² + y⁴ (harmonic, Δu = 0)
    # while the 5-point stencil is not.
    n = 20
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    u = (X ** 4 - 6 * X * X * Y * Y + Y ** 4).ravel()
    g = Grid.square(n, h)
    o9 = O.apply_stencil(u, g, Stencil.laplacian_2d_9pt(h, h), ZD).reshape(n, n)[1:-1, 1:-1]
    o5 = O.apply_stencil(u, g, lap(h), ZD).reshape(n, n)[1:-1, 1:-1]
    assert np.max(np.abs(o9)) < 1e-8
    assert np.max(np.abs(o5)) > 1e-3


def test_dot2_closed_form():
    # runtime/tests/solvers.rs:1118-1147
    nn = 256
    out = (O.dot(np.full(nn, 1.0), np.full(nn, 2.0)), O.dot(np.full(nn, 3.0), np.full(nn, 4.0)))
    assert abs(out[0] - 2 * nn) < 1e-8 and abs(out[1] - 12 * nn) < 1e-8


def test_kahan_sum_exact_on_cancelling_series():
    # Compensated sum: 1e16 + (1, ..., 1) − 1e16 keeps the small terms (plain summation would lose them).
    x = np.concatenate([[1e16], np.ones(10000), [-1e16]])
    assert O.reduce_sum(x) == 10000.0


def test_blas_fma_sites():
    # ops.rs:203 axpy = alpha.mul_add(x, y): single rounding; waxpy ops.rs:551-555; aypx ops.rs:578-582.
    a = 1.0 + 2.0 ** -30
    x = np.array([1.0 + 2.0 ** -30])
    y = np.array([-1.0])
    exact = a * x[0] + y[0]  # not fused in Python: 2 roundings
    fused = float(np.longdouble(a) * np.longdouble(x[0]) + np.longdouble(y[0]))
    assert O.axpy(a, x, y)[0] == fused
    assert fused != exact   # the inputs are chosen so that one rounding and two differ
    assert O.aypx(a, y, x)[0] == float(np.longdouble(a) * np.longdouble(x[0]) + np.longdouble(y[0]))


def test_cg_zero_rhs():
    # solvers.rs:80-97
    n = 16
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    x, st = O.cg(np.zeros(n * n), np.zeros(n * n), g, O.Op.stencil(-0.01, lap(h)), 1e-10, 100)
    assert st.converged() and st.iterations == 0


def test_cg_max_iterations_reason():
    # solvers.rs:1153-1176
    n = 32
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    x, st = O.cg(np.zeros(n * n), gaussian_ic(n, h), g, O.Op.stencil(-0.01, lap(h)), 1e-15, 2)
    assert st.reason == ConvergenceReason.MAX_ITERATIONS and st.iterations == 2


def _manufactured(n, coeff):
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    u = gaussian_ic(n, h)
    b = O.stencil_axpy(coeff, u, u, g, lap(h), ZD)
    return g, h, u, b


def test_cg_converges_and_pcg_identity_matches():
    # solvers.rs:44-76 and 103-154
    g, h, u, b = _manufactured(32, -0.01)
    op = O.Op.stencil(-0.01, lap(h))
    x1, s1 = O.cg(np.zeros_like(b), b, g, op, 1e-8, 500)
    x2, s2 = O.pcg(np.zeros_like(b), b, g, op, 1e-8, 500)
    assert s1.converged() and s1.iterations < 200
    assert s2.converged() and abs(s1.iterations - s2.iterations) <= 2
    assert rel_l2(x1, u) < 1e-6


def _dense_matrix(op, g):
    n = g.rows * g.cols
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = O.apply_op(op, e, g)
    return A


@pytest.mark.parametrize("coeff", [-0.01, -0.5])
def test_brute_force_dense_solve_cg_pipelined_gmres_bicgstab(coeff):
    # Assemble A column by column from the oracle's operator and solve densely;
    # every Krylov solver must reach the dense solution.
    # Zero boundary values keep the iterates in the subspace where A is symmetric (the boundary
    # rows of A are identity rows, cg.rs:104-111).
    n = 12
    g, h, u, b = _manufactured(n, coeff)
    b = O.zero_boundaries(b, n, n)
    op = O.Op.stencil(coeff, lap(h))
    A = _dense_matrix(op, g)
    xd = np.linalg.solve(A, b)
    tol = 1e-11 * np.linalg.norm(b)
    for fn in (lambda: O.cg(np.zeros_like(b), b, g, op, tol, 2000),
               lambda: O.pipelined_cg(np.zeros_like(b), b, g, op, tol, 2000),
               lambda: O.bicgstab(np.zeros_like(b), b, g, op, tol, 2000),
               lambda: O.gmres(np.zeros_like(b), b, g, op, 30, tol, 2000)):
        x, st = fn()
        assert st.converged(), st
        assert rel_l2(x, xd) < 1e-9


def test_operator_symmetry_on_interior_subspace():
    # ⟨Au, v⟩ = ⟨u, Av⟩ for zero-boundary u, v (Laplacian and variable-k operators).
    n = 14
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((n, n)); u[0] = u[-1] = 0; u[:, 0] = u[:, -1] = 0
    v = rng.standard_normal((n, n)); v[0] = v[-1] = 0; v[:, 0] = v[:, -1] = 0
    u, v = u.ravel(), v.ravel()
    k = np.exp(rng.standard_normal(n * n))
    for op in (O.Op.stencil(-0.3, lap(h)), O.Op.var_k(k, coeff=1.0, sigma=0.0), O.Op.var_k(k, coeff=0.5, sigma=1.0)):
        Au = O.apply_op(op, u, g) * (v != 0)
        Av = O.apply_op(op, v, g) * (u != 0)
        assert abs(np.dot(Au, v) - np.dot(u, Av)) < 1e-9 * np.linalg.norm(Au) * np.linalg.norm(v)


def test_vark_unit_k_is_poisson():
    # −∇·(1·∇u) equals −Δ_h u (5-point) on the interior.
    n = 17
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(n * n)
    a = O.vark_apply(0.0, 1.0, np.ones(n * n), u, g).reshape(n, n)
    b = -O.apply_stencil(u, g, lap(h), ZD).reshape(n, n)
    assert np.max(np.abs(a[1:-1, 1:-1] - b[1:-1, 1:-1])) < 1e-9 * np.max(np.abs(b))


def test_conv_diff_zero_velocity_is_kappa_laplacian():
    # apply_conv_diff with v = 0 equals κ·Δ_h u on the interior, 0 on the boundary (ops.rs:674-736).
    n = 15
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(n * n)
    kap = np.exp(rng.standard_normal(n * n))
    z = np.zeros(n * n)
    out = O.conv_diff(0, 0.0, u, kap, z, z, g).reshape(n, n)
    ref = (kap * O.apply_stencil(u, g, lap(h), ZD)).reshape(n, n)
    assert np.max(np.abs(out[1:-1, 1:-1] - ref[1:-1, 1:-1])) < 1e-9 * np.max(np.abs(ref))
    assert np.all(out[0] == 0) and np.all(out[:, 0] == 0)


def test_conv_diff_upwind_on_linear_field():
    # Upwind differences are exact for linear u: κΔu = 0, v·∇u = vx·a + vy·b.
    n = 11
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    u = (3.0 * X - 2.0 * Y).ravel()
    rng = np.random.default_rng(4)
    vx, vy = rng.standard_normal(n * n), rng.standard_normal(n * n)
    out = O.conv_diff(0, 0.0, u, np.ones(n * n), vx, vy, g).reshape(n, n)
    ref = (3.0 * vx - 2.0 * vy).reshape(n, n)
    assert np.max(np.abs(out[1:-1, 1:-1] - ref[1:-1, 1:-1])) < 1e-9


def test_mg_depth_hierarchy():
    # runtime/tests/solvers.rs:814-840
    for n, d in ((64, 5), (32, 4), (8, 2)):
        assert O.mg_depth(Grid.square(n, 1.0 / (n - 1)), O.mg_cfg(coarsest=4)) == d
    # coarsest 16: 4096 -> 16 is 9 levels.
    assert O.mg_depth(Grid.square(4096, 1.0 / 4095), O.mg_cfg(coarsest=16)) == 9


@pytest.mark.parametrize("smoother", [O.SMOOTH_REF_GS, O.SMOOTH_JACOBI, O.SMOOTH_CHEBYSHEV])
def test_vcycle_reduces_residual(smoother):
    # runtime/tests/analytical.rs:449-521: residual strictly decreases for 10 V-cycles (3+3 sweeps)
    # on the sin·sin Poisson source at 64², and ends below 10.
    n = 64
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    b = (np.sin(np.pi * X) * np.sin(np.pi * Y)).ravel()
    cfg = O.mg_cfg(3, 3, 2.0 / 3.0, smoother=smoother)
    xs = np.zeros(n * n)
    prev = np.inf
    for _ in range(10):
        xs = O.vcycle(xs, b, g, cfg)
        r = b - O.apply_stencil(xs, g, lap(h), ZD)
        res = np.sqrt(np.dot(r, r))
        assert res < prev
        prev = res
    assert prev < 10.0


def test_transfer_operators_preserve_linear_functions():
    # Nested case (33 -> 17): full weighting and bilinear interpolation reproduce linear functions
    # exactly at coarse / fine interior points.
    fn, cn = 33, 17
    fg, cg = Grid.square(fn, 1.0 / (fn - 1)), Grid.square(cn, 1.0 / (cn - 1))
    xf = np.arange(fn) / (fn - 1)
    Xf, Yf = np.meshgrid(xf, xf)
    lin_f = (0.3 + 2.0 * Xf - 1.5 * Yf).ravel()
    c = O.restrict(lin_f, fg, cg, mode=1).reshape(cn, cn)
    xc = np.arange(cn) / (cn - 1)
    Xc, Yc = np.meshgrid(xc, xc)
    lin_c = 0.3 + 2.0 * Xc - 1.5 * Yc
    assert np.max(np.abs(c[1:-1, 1:-1] - lin_c[1:-1, 1:-1])) < 1e-12
    f = O.prolong(lin_c.ravel(), np.zeros(fn * fn), cg, fg, mode=1).reshape(fn, fn)
    assert np.max(np.abs(f[1:-1, 1:-1] - lin_f.reshape(fn, fn)[1:-1, 1:-1])) < 1e-12
    # Reference 2x2 average of a constant is the constant; injection adds the coarse value.
    cg2 = Grid.square(16, 2.0 / (fn - 1))
    assert np.allclose(O.restrict(np.full(fn * fn, 7.0), fg, cg2, mode=0), 7.0)
    assert np.allclose(O.prolong(np.full(16 * 16, 2.0), np.ones(fn * fn), cg2, fg, mode=0), 3.0)


@pytest.mark.parametrize("fn,cn", [(32, 16), (33, 17), (100, 50), (47, 24)])
def test_linear_transfer_adjoint_and_interpolation(fn, cn):
    # Non-nested levels (even sizes halve): prolongation reproduces linear functions at fine interior
    # points, and restriction is its transpose scaled by (h/H)²: <R f, c> = (h/H)² <f, P c> for f and c
    # vanishing on the boundary.
    fg, cg = Grid.square(fn, 1.0 / (fn - 1)), Grid.square(cn, 1.0 / (cn - 1))
    xc = np.arange(cn) / (cn - 1)
    Xc, Yc = np.meshgrid(xc, xc)
    lin_c = (0.3 + 2.0 * Xc - 1.5 * Yc).ravel()
    f = O.prolong(lin_c, np.zeros(fn * fn), cg, fg, mode=1).reshape(fn, fn)
    xf = np.arange(fn) / (fn - 1)
    Xf, Yf = np.meshgrid(xf, xf)
    assert np.max(np.abs(f[1:-1, 1:-1] - (0.3 + 2.0 * Xf - 1.5 * Yf)[1:-1, 1:-1])) < 1e-12
    rng = np.random.default_rng(9)
    fv = O.zero_boundaries(rng.standard_normal(fn * fn), fn, fn)
    cv = O.zero_boundaries(rng.standard_normal(cn * cn), cn, cn)
    lhs = np.dot(O.restrict(fv, fg, cg, mode=1), cv)
    rhs = ((cn - 1) / (fn - 1)) ** 2 * np.dot(fv, O.prolong(cv, np.zeros(fn * fn), cg, fg, mode=1))
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)


def test_reference_sweep_is_lexicographic_gauss_seidel():
    # ops.rs:447-471 writes x in place: the first interior point's update is visible to its east
    # neighbour in the same sweep, unlike two-buffer Jacobi.
    n = 6
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    b = np.zeros(n * n)
    b[n + 1] = 1.0
    x0 = np.zeros(n * n)
    gs = O.weighted_jacobi(x0, b, 1.0, g, lap(h), mode=0)
    jac = O.weighted_jacobi(x0, b, 1.0, g, lap(h), mode=1)
    assert gs[n + 2] != 0.0 and jac[n + 2] == 0.0
    assert gs[n + 1] == jac[n + 1]


def _cn_step_system(n, dt=0.1, alpha=1.0):
    """Crank–Nicolson step as in implicit.rs:178-222: rhs = u + (dt·α/2)·Δu, coeff = −dt·α/2,
    initial guess u."""
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    u = gaussian_ic(n, h)
    half = 0.5 * dt * alpha
    rhs = O.stencil_axpy(half, u, u, g, lap(h), ZD)
    return g, h, u, rhs, -half


def test_pcg_multigrid_fewer_iterations():
    # solvers.rs:158-207: the CN step with MG-PCG needs no more iterations than CG.
    g, h, u, b, coeff = _cn_step_system(64)
    op = O.Op.stencil(coeff, lap(h))
    _, s1 = O.cg(u, b, g, op, 1e-8, 500)
    assert s1.converged()
    for sm in (O.SMOOTH_REF_GS, O.SMOOTH_JACOBI):
        _, s2 = O.pcg(u, b, g, op, 1e-8, 100, mg=O.mg_cfg(2, 2, 2.0 / 3.0, smoother=sm))
        assert s2.converged() and s2.iterations <= s1.iterations


def cheb_T(k, z):
    return np.polynomial.chebyshev.chebval(z, [0.0] * k + [1.0])


@pytest.mark.parametrize("sweeps", [1, 2, 3, 4, 7])
def test_chebyshev_mode_factor_closed_form(sweeps):
    # multigrid.rs:317-473: with zero rhs, each discrete sine mode u_pq is an eigenvector of D⁻¹Δ_h
    # (D = −(2/dx² + 2/dy²)) with eigenvalue λ = (4/dy²·sin²(pπ/2(R−1)) + 4/dx²·sin²(qπ/2(C−1)))/(2/dx² + 2/dy²),
    # and n sweeps of the first-kind Chebyshev smoother with bounds [emin, emax] = [0.2, 2] scale it by
    # T_n((emax + emin − 2λ)/(emax − emin)) / T_n((emax + emin)/(emax − emin)).  A grid no larger than
    # the coarsest level is a one-level hierarchy whose V-cycle is n_pre + n_post + extra_coarse sweeps.
    R, Cc = 13, 9
    dx, dy = 1.0 / (Cc - 1), 0.7 / (R - 1)
    g = Grid(R, Cc, dx, dy)
    cfg = O.mg_cfg(0, 0, 2.0 / 3.0, smoother=O.SMOOTH_CHEBYSHEV, coarsest=64, extra_coarse=sweeps)
    assert O.mg_depth(g, cfg) == 1
    emax, emin = 2.0, 0.2
    rr, cc = np.meshgrid(np.arange(R), np.arange(Cc), indexing="ij")
    for p_, q_ in ((1, 1), (3, 2), (R - 2, Cc - 2), (6, 1), (2, 7)):
        u = (np.sin(p_ * np.pi * rr / (R - 1)) * np.sin(q_ * np.pi * cc / (Cc - 1))).ravel()
        lam = (4 / dy ** 2 * np.sin(p_ * np.pi / (2 * (R - 1))) ** 2 + 4 / dx ** 2 * np.sin(q_ * np.pi / (2 * (Cc - 1))) ** 2) \
            / (2 / dx ** 2 + 2 / dy ** 2)
        factor = cheb_T(sweeps, (emax + emin - 2 * lam) / (emax - emin)) / cheb_T(sweeps, (emax + emin) / (emax - emin))
        out = O.vcycle(u, np.zeros_like(u), g, cfg)
        assert np.max(np.abs(out - factor * u)) < 1e-13, (p_, q_, sweeps, factor)


def test_chebyshev_not_slower_than_jacobi():
    # solvers.rs:845-894
    g, h, u, b, coeff = _cn_step_system(64)
    op = O.Op.stencil(coeff, lap(h))
    _, sj = O.pcg(u, b, g, op, 1e-8, 100, mg=O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_REF_GS))
    _, sc = O.pcg(u, b, g, op, 1e-8, 100, mg=O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_CHEBYSHEV))
    assert sj.converged() and sc.converged() and sc.iterations <= sj.iterations


def test_pipelined_and_gmres_match_cg():
    # solvers.rs:324-377 and 1257-1304: max difference < 1e-6.
    g, h, u, b = _manufactured(32, -0.01)
    op = O.Op.stencil(-0.01, lap(h))
    x1, _ = O.cg(np.zeros_like(b), b, g, op, 1e-8, 500)
    x2, s2 = O.pipelined_cg(np.zeros_like(b), b, g, op, 1e-8, 500)
    x3, s3 = O.gmres(np.zeros_like(b), b, g, op, 50, 1e-8, 500)
    assert s2.converged() and s3.converged()
    assert np.max(np.abs(x1 - x2)) < 1e-6 and np.max(np.abs(x1 - x3)) < 1e-6


def test_gmres_restart_converges():
    # solvers.rs:1411-1443: GMRES(5) still converges.
    g, h, u, b = _manufactured(32, -0.01)
    x, st = O.gmres(np.zeros_like(b), b, g, O.Op.stencil(-0.01, lap(h)), 5, 1e-8, 2000)
    assert st.converged() and rel_l2(x, u) < 1e-5


def test_preconditioned_variants_converge():
    g, h, u, b = _manufactured(64, -0.5)
    b = O.zero_boundaries(b, 64, 64)
    op = O.Op.stencil(-0.5, lap(h))
    mg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI)
    xr, _ = O.cg(np.zeros_like(b), b, g, op, 1e-12, 2000)
    for fn in (O.bicgstab, O.ppcg, O.fcg):
        x, st = fn(np.zeros_like(b), b, g, op, 1e-9, 500, mg=mg)
        assert st.converged(), (fn.__name__, st)
        assert rel_l2(x, xr) < 1e-7
    x, st = O.gmres(np.zeros_like(b), b, g, op, 30, 1e-9, 500, mg=mg)
    assert st.converged()


def test_fp32_vcycle_close_to_fp64():
    # The FP32 V-cycle emulation differs from FP64 only by single-precision rounding.
    n = 64
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    rng = np.random.default_rng(5)
    r = rng.standard_normal((n, n)); r[0] = r[-1] = 0; r[:, 0] = r[:, -1] = 0
    r = r.ravel()
    for tr in (0, 1):
        z64 = O.vcycle(np.zeros(n * n), r, g, O.mg_cfg(2, 2, 2.0 / 3.0, transfer=tr, precision=0))
        z32 = O.vcycle(np.zeros(n * n), r, g, O.mg_cfg(2, 2, 2.0 / 3.0, transfer=tr, precision=1))
        assert rel_l2(z32, z64) < 1e-5


def test_fcg_fp32_reaches_fp64_true_residual():
    # Mixed precision: FP32 V-cycle inside FP64 flexible CG reaches 1e-10 relative true residual.
    n = 128
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    from fluxion_amd.inputs import gaussian_bumps
    f = gaussian_bumps(n).ravel()
    st5 = Stencil.poisson_minus_identity(h, h)
    op = O.Op.stencil(1.0, st5)
    tol = 1e-10 * np.linalg.norm(f)
    mg = O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16, precision=1, sigma=0.0, scale=-1.0)
    x, st = O.fcg(np.zeros_like(f), f, g, op, tol, 200, mg=mg)
    assert st.converged()
    r = f - O.apply_op(op, x, g)
    assert np.linalg.norm(r) < tol
    assert st.iterations <= 16


def test_refinement_stops_converged_at_the_fp64_floor():
    # A tolerance below FP64's attainable true residual (1e-14·‖f‖; the floor at 64² is about
    # 8e-14·‖f‖, growing as 1/h²): the recurrence still meets it, the refinement passes stop once a
    # restart no longer halves ‖f − A·x‖, and the solve counts as converged with that floor reported.
    # A budget the first pass cannot finish in stays MAX_ITERATIONS.
    n = 64
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    from fluxion_amd.inputs import gaussian_bumps
    f = gaussian_bumps(n).ravel()
    op = O.Op.stencil(1.0, Stencil.poisson_minus_identity(h, h))
    cfg = O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16, precision=1, scale=-1.0)
    tol = 1e-14 * np.linalg.norm(f)
    for solve in (O.fcg, O.ppcg):
        x, st = solve(np.zeros_like(f), f, g, op, tol, 400, mg=cfg)
        true = np.linalg.norm(f - O.apply_op(op, x, g))
        assert st.reason == ConvergenceReason.CONVERGED and st.iterations < 400, (solve, st.iterations)
        assert 2 * tol < true < 2e-13 * np.linalg.norm(f), (solve, true)
        _, st2 = solve(np.zeros_like(f), f, g, op, tol, 5, mg=cfg)
        assert st2.reason == ConvergenceReason.MAX_ITERATIONS and st2.iterations == 5


def test_planted_mistake_is_caught():
    # A transposed spacing (dx <-> dy) in the stencil on a rectangular grid fails the quadratic
    # closed form: the checks above discriminate such errors.
    rows, cols = 20, 30
    dx, dy = 1.0 / (cols - 1), 1.0 / (rows - 1)
    x = np.arange(cols) * dx
    y = np.arange(rows) * dy
    X, Y = np.meshgrid(x, y)
    u = (X * X + 3.0 * Y * Y).ravel()
    good = O.apply_stencil(u, Grid(rows, cols, dx, dy), Stencil.laplacian_2d_5pt(dx, dy), ZD)
    bad = O.apply_stencil(u, Grid(rows, cols, dx, dy), Stencil.laplacian_2d_5pt(dy, dx), ZD)
    good = good.reshape(rows, cols)[1:-1, 1:-1]
    bad = bad.reshape(rows, cols)[1:-1, 1:-1]
    assert np.max(np.abs(good - 8.0)) < 1e-9
    assert np.max(np.abs(bad - 8.0)) > 1.0


def test_errors():
    # runtime/tests/error_handling.rs: grid >= 3x3, periodic rejected, empty stencil.
    from fluxion_amd.core import InvalidGrid, InvalidStencil
    with pytest.raises(InvalidGrid):
        Grid(2, 10, 1.0, 1.0)
    with pytest.raises(InvalidGrid):
        Grid(10, 10, 0.0, 1.0)
    with pytest.raises(InvalidStencil):
        Stencil([])
    g = Grid.square(8, 1.0 / 7)
    with pytest.raises(RuntimeError):
        O.apply_stencil(np.zeros(64), g, lap(1.0 / 7), Boundaries.uniform(BoundaryCondition.periodic()))


def test_vark_multigrid_unit_k_matches_constant():
    # With k ≡ 1 the variable-coefficient level operator is −Δ_h: the V-cycle agrees with the
    # constant-coefficient one (scale −1) up to rounding.
    n = 65
    g = Grid.square(n, 1.0 / (n - 1))
    r = O.zero_boundaries(np.random.default_rng(12).standard_normal(n * n), n, n)
    a = O.vcycle(np.zeros(n * n), r, g, O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16, scale=-1.0))
    b = O.vcycle(np.zeros(n * n), r, g, O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16, scale=1.0,
                                                  k=np.ones(n * n)))
    assert rel_l2(a, b) < 1e-12


def test_vark_multigrid_pcg_converges():
    # Log-normal k (correlation 0.1, log-variance 1): MG-PCG with coefficient-aware levels converges
    # in far fewer iterations than CG and than MG built on the constant-coefficient Laplacian.
    from fluxion_amd.inputs import gaussian_bumps, grf_log_k
    n = 129
    g = Grid.square(n, 1.0 / (n - 1))
    k = grf_log_k(n).ravel()
    f = gaussian_bumps(n).ravel()
    op = O.Op.var_k(k)
    tol = 1e-10 * np.linalg.norm(f)
    _, s_cg = O.cg(np.zeros_like(f), f, g, op, tol, 5000)
    x, s_k = O.pcg(np.zeros_like(f), f, g, op, tol, 300, mg=O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16,
                                                                      scale=1.0, k=k))
    _, s_c = O.pcg(np.zeros_like(f), f, g, op, tol, 300, mg=O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16,
                                                                      scale=-1.0))
    assert s_k.converged() and s_k.iterations < s_c.iterations < s_cg.iterations
    assert np.linalg.norm(f - O.apply_op(op, x, g)) < tol


# ── convection-diffusion multigrid levels (extension) ───────────────────
def cd_problem(n, kappa=1e-3, seed=None):
    from fluxion_amd.inputs import gaussian_bumps, streamfunction_velocity
    g = Grid.square(n, 1.0 / (n - 1))
    vx, vy, _ = streamfunction_velocity(n, kappa) if seed is None else streamfunction_velocity(n, kappa, seed=seed)
    return g, vx.ravel(), vy.ravel(), gaussian_bumps(n).ravel()


def test_cd_level_jacobi_sweep_matches_operator():
    # A hierarchy of one level is a run of Jacobi sweeps x1 = x0 + (ω/d)(b − A·x0) with A the upwind
    # operator of ops.rs:707-731 (already pinned) and d its diagonal |vx|/dx + |vy|/dy + 2|κ|(1/dx² + 1/dy²).
    n, kappa = 12, 1e-3
    g, vx, vy, _ = cd_problem(n, kappa)
    rng = np.random.default_rng(7)
    x0 = rng.standard_normal(n * n)
    x0 = O.zero_boundaries(x0, n, n)
    b = rng.standard_normal(n * n)
    cfg = O.mg_cfg(0, 0, 0.6, transfer=1, coarsest=64, extra_coarse=1, vx=vx, vy=vy, kappa=-kappa)
    x1 = O.vcycle(x0, b, g, cfg)
    ax = O.apply_op(O.Op.conv_diff(vx, vy, kappa_scalar=-kappa, coeff=1.0, sigma=0.0), x0, g)
    ix = 1.0 / g.dx
    d = np.abs(vx) * ix + np.abs(vy) * ix + 2 * kappa * 2 * ix * ix
    want = x0 + 0.6 / d * (b - ax)
    inner = np.zeros((n, n), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.ravel()
    assert np.max(np.abs(x1[inner] - want[inner])) < 1e-12 * np.max(np.abs(want))
    assert np.array_equal(x1[~inner], x0[~inner])


def test_cd_levels_without_velocity_equal_scaled_laplacian():
    # v ≡ 0: the level operators are κΔ_h on every level, the plain multigrid with scale κ.
    n, kappa = 33, -2.5e-3
    g = Grid.square(n, 1.0 / (n - 1))
    b = np.random.default_rng(8).standard_normal(n * n)
    z = np.zeros(n * n)
    a = O.vcycle(z, b, g, O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=4, vx=z, vy=z, kappa=kappa))
    c = O.vcycle(z, b, g, O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=4, scale=kappa))
    assert np.max(np.abs(a - c)) < 1e-12 * np.max(np.abs(c))


@pytest.mark.parametrize("solver", ["bicgstab", "gmres"])
def test_cd_preconditioned_krylov_brute_force(solver):
    # BiCGSTAB (right) and GMRES(30) (left) preconditioned by the FP64 and FP32 convection-diffusion
    # V-cycles converge to the dense solution of −κΔu + v·∇⁻u = f on a 20² grid.
    n, kappa = 20, 1e-3
    g, vx, vy, f = cd_problem(n, kappa)
    op = O.Op.conv_diff(vx, vy, kappa_scalar=-kappa, coeff=1.0, sigma=0.0)
    A = np.stack([O.apply_op(op, e, g) for e in np.eye(n * n)], axis=1)
    xd = np.linalg.solve(A, f)
    for prec in (0, 1):
        cfg = O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=4, precision=prec, vx=vx, vy=vy, kappa=-kappa)
        if solver == "bicgstab":
            x, s = O.bicgstab(np.zeros_like(f), f, g, op, 1e-12 * np.linalg.norm(f), 200, mg=cfg)
        else:
            x, s = O.gmres(np.zeros_like(f), f, g, op, 30, 1e-13 * np.linalg.norm(f), 200, mg=cfg)
        assert s.converged(), (prec, s)
        # Left-preconditioned GMRES builds its Krylov basis from the rounded M⁻¹ applications and
        # judges ‖M⁻¹r‖, so with the FP32 cycle its answer is only as exact as FP32 (about 1e-7).
        bar = 1e-6 if (solver == "gmres" and prec == 1) else 1e-9
        assert rel_l2(x, xd) < bar, (prec, rel_l2(x, xd))


def test_fgmres_identity_is_gmres():
    # With the identity preconditioner z_k = v_k, so flexible GMRES is plain GMRES step for step.
    n, kappa = 24, 1e-3
    g, vx, vy, f = cd_problem(n, kappa)
    op = O.Op.conv_diff(vx, vy, kappa_scalar=-kappa, coeff=1.0, sigma=0.0)
    tol = 1e-9 * np.linalg.norm(f)
    xa, sa = O.gmres(np.zeros_like(f), f, g, op, 10, tol, 400)
    xb, sb = O.fgmres(np.zeros_like(f), f, g, op, 10, tol, 400)
    assert sa.iterations == sb.iterations and np.array_equal(xa, xb)
    assert sa.history == sb.history


@pytest.mark.parametrize("prec", [0, 1])
def test_fgmres_estimate_is_true_residual(prec):
    # Right preconditioning: the Givens estimate |g_{k+1}| is ‖b − A·x_k‖ (up to rounding) for the FP64
    # and the FP32 (not exactly linear) V-cycle, and the answer is the dense solution.
    n, kappa = 20, 1e-3
    g, vx, vy, f = cd_problem(n, kappa)
    op = O.Op.conv_diff(vx, vy, kappa_scalar=-kappa, coeff=1.0, sigma=0.0)
    A = np.stack([O.apply_op(op, e, g) for e in np.eye(n * n)], axis=1)
    xd = np.linalg.solve(A, f)
    cfg = O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=4, precision=prec, vx=vx, vy=vy, kappa=-kappa)
    tol = 1e-10 * np.linalg.norm(f)
    x, s = O.fgmres(np.zeros_like(f), f, g, op, 30, tol, 200, mg=cfg)
    assert s.converged() and s.iterations < 30
    true = np.linalg.norm(f - A @ x)
    assert true < 1.05 * tol and abs(true - s.residual) < 1e-3 * tol + 1e-6 * true
    assert rel_l2(x, xd) < 1e-8


def test_pipelined_flexible_cg_matches_fcg_and_dense():
    # With an FP32 multigrid the preconditioned pipelined CG runs its flexible single-reduction form;
    # in exact arithmetic it is flexible CG(1) step for step, so its iteration counts equal the
    # oracle's FCG with the same cycle (±1 for rounding), and a 16² solve is the dense solution.
    for n in (16, 64, 128):
        h = 1.0 / (n - 1)
        g = Grid.square(n, h)
        st = Stencil.poisson_minus_identity(h, h)
        from fluxion_amd.inputs import gaussian_bumps
        f = gaussian_bumps(n).ravel()
        op = O.Op.stencil(1.0, st)
        cfg = O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=4, precision=1, scale=-1.0)
        tol = 1e-11 * np.linalg.norm(f)
        xp, sp = O.ppcg(np.zeros_like(f), f, g, op, tol, 300, mg=cfg)
        xf, sf = O.fcg(np.zeros_like(f), f, g, op, tol, 300, mg=cfg)
        assert sp.converged() and sf.converged()
        assert abs(sp.iterations - sf.iterations) <= 1, (n, sp.iterations, sf.iterations)
        assert np.linalg.norm(f - O.apply_op(op, xp, g)) < tol
        if n == 16:
            A = np.stack([O.apply_op(op, e, g) for e in np.eye(n * n)], axis=1)
            assert rel_l2(xp, np.linalg.solve(A, f)) < 1e-10


# ── elementwise ops, extra reductions and the variable-coefficient stencil (ops.rs:129-154, 318-331,
#    476-585, 650-672) against known values ───────────────────────────────────────────────────
def test_scale_waxpy_pointwise_known_values():
    x = np.array([1.0, 2.0, 3.0])
    y = np.array([4.0, 5.0, 6.0])
    assert np.array_equal(O.scale(x, 2.5), [2.5, 5.0, 7.5])
    assert np.array_equal(O.waxpy(2.0, x, 3.0, y), [14.0, 19.0, 24.0])
    # waxpy is alpha.mul_add(x, beta·y): one rounding of α·x + (β·y).  With α = x = 1 + 2⁻³⁰ and
    # β·y = −(1 + 2⁻²⁹) the exact result is 2⁻⁶⁰; rounding α·x first would give 0.
    a = 1.0 + 2.0 ** -30
    w = O.waxpy(a, np.array([a]), 1.0, np.array([-(1.0 + 2.0 ** -29)]))
    assert w[0] == 2.0 ** -60
    assert np.array_equal(O.pointwise(0, np.array([2.0, 3.0]), np.array([4.0, 5.0])), [8.0, 15.0])
    assert np.array_equal(O.pointwise(1, np.array([1.0, 3.0]), np.array([4.0, 6.0])), [0.25, 0.5])
    # f64::max / f64::min ignore a NaN operand (ops.rs:650-672).
    nan = float("nan")
    assert np.array_equal(O.pointwise(2, np.array([nan, 1.0, -2.0]), np.array([1.0, nan, 3.0])), [1.0, 1.0, 3.0])
    assert np.array_equal(O.pointwise(3, np.array([nan, 1.0, -2.0]), np.array([1.0, nan, 3.0])), [1.0, 1.0, -2.0])


def test_reduce_max_min_known_values():
    v = np.array([3.0, -1.0, 7.0, 2.0])
    assert O.reduce_max(v) == 7.0 and O.reduce_min(v) == -1.0
    w = np.array([float("nan"), 2.0, -5.0])
    assert O.reduce_max(w) == 2.0 and O.reduce_min(w) == -5.0       # NaN ignored, as f64::max/min do
    e = np.zeros(0)
    assert O.reduce_max(e) == -np.inf and O.reduce_min(e) == np.inf  # identities of the folds


def test_apply_stencil_var_quadratic():
    # coeff_i·S(u)_i (ops.rs:129-154): Δ(x² + y²) = 4 exactly on the interior of a uniform grid, so the
    # result is 4·coeff there; with zero Dirichlet data the boundary values are coeff·0 = 0.
    n = 33
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    xs = np.arange(n) * h
    X, Y = np.meshgrid(xs, xs)
    u = (X ** 2 + Y ** 2).ravel()
    coeff = (1.0 + X + 2.0 * Y * Y).ravel()
    out = O.apply_stencil_var(u, coeff, g, lap(h), ZD).reshape(n, n)
    assert np.max(np.abs(out[1:-1, 1:-1] - 4.0 * coeff.reshape(n, n)[1:-1, 1:-1])) < 1e-9
    assert not out[0].any() and not out[-1].any() and not out[:, 0].any() and not out[:, -1].any()
    # A constant coefficient scales apply_stencil.
    assert np.array_equal(O.apply_stencil_var(u, np.full(n * n, 2.0), g, lap(h), ZD),
                          2.0 * O.apply_stencil(u, g, lap(h), ZD))


def _learned_problem(n=32):
    from fluxion_amd.inputs import gaussian_bumps, grf_log_k
    g = Grid.square(n, 1.0 / (n - 1))
    k = grf_log_k(n).ravel()
    f = gaussian_bumps(n).ravel()
    mg = O.mg_cfg(2, 2, 2.0 / 3.0, transfer=1, coarsest=16, precision=1, scale=1.0, k=k)
    return g, k, f, O.Op.var_k(k), mg


def test_learned_fcg_zero_network_is_fcg():
    # With all network parameters zero the correction c is 0, so α* = 0 and the learned preconditioner is
    # the V-cycle alone: the composition must then be flexible CG around the same V-cycle.
    g, k, f, op, mg = _learned_problem()
    tol = 1e-10 * np.linalg.norm(f)
    params = np.zeros(O.unet_nparams(2))
    xl, itl, fb = O.learned_fcg(f, g, op, mg, params, np.log(k), tol, 200)
    xf, sf = O.fcg(np.zeros_like(f), f, g, op, tol, 200, mg=mg)
    assert sf.converged() and abs(itl - sf.iterations) <= 1
    assert rel_l2(xl, xf) < 1e-8


def test_learned_fcg_schedule_first():
    # The network runs from iteration `first` on: with `first` beyond the iteration count it never runs,
    # and the trained-network solve is then flexible CG around the V-cycle step for step; with first = 0
    # it also corrects the preconditioner application made on the right-hand side itself.
    g, k, f, op, mg = _learned_problem()
    tol = 1e-10 * np.linalg.norm(f)
    params = np.random.default_rng(5).standard_normal(O.unet_nparams(2)) * 0.05
    xn, itn, _ = O.learned_fcg(f, g, op, mg, params, np.log(k), tol, 200, first=1000)
    xf, sf = O.fcg(np.zeros_like(f), f, g, op, tol, 200, mg=mg)
    assert abs(itn - sf.iterations) <= 1 and rel_l2(xn, xf) < 1e-8
    x0, it0, _ = O.learned_fcg(f, g, op, mg, params, np.log(k), tol, 200, first=0, stall=1e9)
    x2, it2, _ = O.learned_fcg(f, g, op, mg, params, np.log(k), tol, 200, first=2, stall=1e9)
    assert rel_l2(x0, x2) > 0 and it0 < 200 and it2 < 200


def test_learned_fcg_random_network_still_converges():
    # α* = (c·r1)/(c·A·c), clipped at 0, minimises the A-norm of the error left after the V-cycle along c,
    # so each application is at least as good as the V-cycle alone, whatever the network.  With an
    # untrained network (random weights) and the stall fallback disabled the solve therefore still
    # converges to the tolerance on the true residual.  It takes more iterations than the V-cycle alone
    # (30 against 15 here): the varying preconditioner costs flexible CG more than the per-application gain.
    g, k, f, op, mg = _learned_problem()
    tol = 1e-9 * np.linalg.norm(f)
    woff, boff, lcin, lcout = O.unet_layout(2)
    rng = np.random.default_rng(3)
    params = np.zeros(O.unet_nparams(2))
    for i in range(15):
        m = 9 * lcin[i] * lcout[i]
        params[woff[i]:woff[i] + m] = rng.standard_normal(m) * np.sqrt(2.0 / (9 * lcin[i]))
    xl, itl, fb = O.learned_fcg(f, g, op, mg, params, np.log(k), tol, 200, stall=1e9)
    _, sf = O.fcg(np.zeros_like(f), f, g, op, tol, 200, mg=mg)
    assert not fb and itl < 200
    assert np.linalg.norm(f - O.apply_op(op, xl, g)) < 2 * tol
    assert sf.iterations <= itl <= 3 * sf.iterations
