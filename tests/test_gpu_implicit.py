"""Crank–Nicolson stepping (runtime/src/implicit.rs) and the raw field format (io/src/raw.rs) on the GPU,
through the C ABI, against the oracle.

Parity bar of the FP64 paths, as for the solvers: iteration counts within ±2 and solution relative L2
difference ≤ 1e-10 against the oracle's step on the same input.  The mixed-precision variable-k path
(flexible CG around the fused FP32 cycle) is held to ≤ 1e-8, the bar of the fused FCG solver tests.
"""
import numpy as np
import pytest

import oracle as O
from fluxion_amd.core import Boundaries, BoundaryCondition, ConvergenceReason, DType, Grid, SmootherKind
from fluxion_amd.inputs import gaussian_bumps, grf_log_k
from tests.conftest import gaussian_ic, rel_l2

pytestmark = pytest.mark.gpu

ZD = Boundaries.zero_dirichlet()
IT_TOL = 2


def H():
    from fluxion_amd import hip
    return hip


def check(s, sr, x, xr, sol_tol=1e-10):
    assert s.reason == sr.reason, (s, sr)
    assert abs(s.iterations - sr.iterations) <= IT_TOL, (s.iterations, sr.iterations)
    assert rel_l2(x, xr) <= sol_tol, rel_l2(x, xr)


@pytest.mark.parametrize("n", [32, 64])
def test_cn_step_cg_parity(be, n):
    hip = H()
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    u0 = gaussian_ic(n, h)
    cn = hip.CrankNicolsonSolver.new(g, 1.0, 0.1, ZD, 1e-8, 500, be=be)
    u = hip.Field.from_f64(g, u0, be)
    s = cn.step(u)
    ur, sr = O.cn_step(u0, g, 1.0, 0.1, tol=1e-8, max_iter=500)
    check(s, sr, u.to_vec_f64(), ur)
    assert cn.steps_done() == 1 and cn.dt() == 0.1


@pytest.mark.parametrize("smoother", [SmootherKind.JACOBI, SmootherKind.CHEBYSHEV])
def test_cn_step_multigrid_parity(be, smoother):
    hip = H()
    n = 64
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    u0 = gaussian_ic(n, h)
    cn = hip.CrankNicolsonSolver.with_multigrid_smoother(g, 1.0, 0.1, ZD, 1e-8, 100, 2, 2, 2.0 / 3.0, smoother,
                                                         be=be)
    u = hip.Field.from_f64(g, u0, be)
    s = cn.step(u)
    osm = O.SMOOTH_JACOBI if smoother == SmootherKind.JACOBI else O.SMOOTH_CHEBYSHEV
    ur, sr = O.cn_step(u0, g, 1.0, 0.1, tol=1e-8, max_iter=100, mg=O.mg_cfg(2, 2, 2.0 / 3.0, smoother=osm))
    check(s, sr, u.to_vec_f64(), ur)


def test_cn_steps_with_source_and_dirichlet_data(be):
    # Rectangular grid, nonzero Dirichlet values, a source (zero on boundary nodes), three steps.
    hip = H()
    g = Grid(33, 47, 1.0 / 46, 1.0 / 32)
    bc = Boundaries(BoundaryCondition.dirichlet(1.0), BoundaryCondition.dirichlet(-0.5),
                    BoundaryCondition.dirichlet(0.25), BoundaryCondition.dirichlet(0.0))
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(len(g))
    f = rng.standard_normal((g.rows, g.cols))
    f[[0, -1], :] = 0.0
    f[:, [0, -1]] = 0.0
    f = f.ravel()
    cn = hip.CrankNicolsonSolver.new(g, 0.5, 0.02, bc, 1e-9, 1000, be=be)
    u = hip.Field.from_f64(g, u0, be)
    F = hip.Field.from_f64(g, f, be)
    ur = u0
    for i in range(3):
        s = cn.step(u, F)
        ur, sr = O.cn_step(ur, g, 0.5, 0.02, bc=bc, source=f, tol=1e-9, max_iter=1000)
        check(s, sr, u.to_vec_f64(), ur)
    assert cn.steps_done() == 3
    assert abs(cn.sim_time() - 0.06) < 1e-15


# ── the reference's own CN tests (runtime/tests/solvers.rs, analytical.rs, hip_gpu.rs) ──
def test_cn_unpreconditioned_converges(be):
    # solvers.rs:437-463
    hip = H()
    n = 32
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    cn = hip.CrankNicolsonSolver.new(g, 0.01, 0.01, ZD, 1e-8, 200, be=be)
    u = hip.Field.from_f64(g, gaussian_ic(n, h), be)
    s = cn.step(u)
    assert s.converged(), s
    assert cn.steps_done() == 1


def test_cn_stability_200_steps(be):
    # solvers.rs:465-510: peak decreases, stays finite and positive; sim_time = 200·0.01.
    hip = H()
    n = 32
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    data = gaussian_ic(n, h)
    cn = hip.CrankNicolsonSolver.with_multigrid(g, 0.01, 0.01, ZD, 1e-8, 200, 2, 2, 2.0 / 3.0, be=be)
    u = hip.Field.from_f64(g, data, be)
    cn.step_n(u, 200)
    peak = u.to_vec_f64().max()
    assert np.isfinite(peak) and 0.0 < peak < data.max()
    assert cn.steps_done() == 200
    assert abs(cn.sim_time() - 2.0) < 1e-10
    # The same 200 steps by the oracle.  Each step's solve stops anywhere below ‖r‖ = 1e-8, so the two
    # trajectories may drift apart by up to about that much per step: the bar is 200 steps' worth
    # (single steps are held to 1e-10 above).
    ur = data
    cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI)
    for _ in range(200):
        ur, _ = O.cn_step(ur, g, 0.01, 0.01, tol=1e-8, max_iter=200, mg=cfg)
    assert rel_l2(u.to_vec_f64(), ur) <= 1e-7


def test_cn_with_source(be):
    # solvers.rs:512-540: zero initial state, source 1 everywhere.
    hip = H()
    n = 32
    g = Grid.square(n, 1.0 / (n - 1))
    cn = hip.CrankNicolsonSolver.new(g, 0.01, 0.01, ZD, 1e-8, 200, be=be)
    u = hip.Field.zeros(g, be=be)
    src = hip.Field.zeros(g, be=be)
    src.fill(1.0)
    s = cn.step(u, src)
    assert s.converged()
    assert u.norm_l2() > 1e-10


def test_cn_pcg_fewer_iterations_than_cg(be):
    # solvers.rs:156-205 and hip_gpu.rs:285-340: n = 64, α = 1, dt = 0.1; then four more steps lower the peak.
    hip = H()
    n = 64
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    data = gaussian_ic(n, h)
    pcg = hip.CrankNicolsonSolver.with_multigrid(g, 1.0, 0.1, ZD, 1e-8, 100, 2, 2, 2.0 / 3.0, be=be)
    cg = hip.CrankNicolsonSolver.new(g, 1.0, 0.1, ZD, 1e-8, 500, be=be)
    u1 = hip.Field.from_f64(g, data, be)
    u2 = hip.Field.from_f64(g, data, be)
    s1, s2 = pcg.step(u1), cg.step(u2)
    assert s1.converged() and s2.converged()
    assert s1.iterations <= s2.iterations, (s1.iterations, s2.iterations)
    pcg.step_n(u1, 4)
    assert u1.to_vec_f64().max() < data.max()


def test_cn_pcg_chebyshev_converges(be):
    # analytical.rs:523-580: n = 64, α = dt = 0.01, tol 1e-10, Chebyshev (3, 3).
    hip = H()
    n = 64
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    cn = hip.CrankNicolsonSolver.with_multigrid_smoother(g, 0.01, 0.01, ZD, 1e-10, 200, 3, 3, 2.0 / 3.0,
                                                         SmootherKind.CHEBYSHEV, be=be)
    u = hip.Field.from_f64(g, gaussian_ic(n, h), be)
    s = cn.step(u)
    assert s.converged(), s


# ── variable-k Crank–Nicolson (extension) ─────────────────────────────
@pytest.mark.parametrize("n", [129, 256])
def test_cn_var_k_mixed_and_fp64_parity(be, n):
    hip = H()
    g = Grid.square(n, 1.0 / (n - 1))
    k = grf_log_k(n).ravel()
    u0 = gaussian_bumps(n, 20260101 + 7).ravel()
    alpha, dt = 1.0, 1e-3
    h = 0.5 * dt * alpha
    tol = 1e-10 * np.linalg.norm(u0)
    K = hip.Field.from_f64(g, k, be)
    for mixed in (True, False):
        cn = hip.CrankNicolsonSolver.var_k(g, alpha, dt, K, tol, 200, mixed=mixed, be=be)
        u = hip.Field.from_f64(g, u0, be)
        cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI, transfer=1, coarsest=16,
                       precision=1 if mixed else 0, sigma=1.0, scale=h, k=k)
        ur = u0
        for i in range(3):
            s = cn.step(u)
            ur, sr = O.cn_step(ur, g, alpha, dt, tol=tol, max_iter=200, mg=cfg, k=k, flexible=mixed)
            assert s.converged(), (mixed, i, s)
            check(s, sr, u.to_vec_f64(), ur, 1e-8 if mixed else 1e-10)
        if mixed:
            assert s.true_residual < tol


def test_cn_var_k_step_n_matches_steps(be):
    # step_n(n) is n steps without a source: same state as n calls of step.
    hip = H()
    n = 128
    g = Grid.square(n, 1.0 / (n - 1))
    K = hip.Field.from_f64(g, grf_log_k(n).ravel(), be)
    u0 = gaussian_bumps(n).ravel()
    a = hip.CrankNicolsonSolver.var_k(g, 1.0, 1e-3, K, 1e-9, 200, be=be)
    b = hip.CrankNicolsonSolver.var_k(g, 1.0, 1e-3, K, 1e-9, 200, be=be)
    ua, ub = hip.Field.from_f64(g, u0, be), hip.Field.from_f64(g, u0, be)
    a.step_n(ua, 5)
    for _ in range(5):
        b.step(ub)
    assert np.array_equal(ua.to_vec_f64(), ub.to_vec_f64())
    assert a.steps_done() == b.steps_done() == 5


def test_cn_create_errors(be):
    hip = H()
    from fluxion_amd.core import BackendError, DimensionMismatch
    g = Grid.square(16, 1.0 / 15)
    with pytest.raises(BackendError):
        hip.CrankNicolsonSolver(g, 1.0, 0.1, ZD, 1e-8, 10, op=1, be=be)          # conv-diff: not a CN operator
    with pytest.raises(BackendError):
        hip.CrankNicolsonSolver(g, 1.0, 0.1, ZD, 1e-8, 10, solver=4, be=be)      # FCG without a multigrid
    cn = hip.CrankNicolsonSolver.new(g, 1.0, 0.1, ZD, 1e-8, 10, be=be)
    with pytest.raises(DimensionMismatch):
        cn.step(hip.Field.zeros(Grid.square(17, 1.0 / 16), be=be))


# ── raw field format through device buffers ───────────────────────────
def test_raw_device_round_trip(be, tmp_path):
    hip = H()
    g = Grid(19, 23, 0.05, 0.07)
    data = np.random.default_rng(4).standard_normal(len(g))
    F = hip.Field.from_f64(g, data, be)
    p = tmp_path / "f.raw"
    hip.write_raw(p, F)
    g2, d2 = hip.read_raw_host(p)
    assert g2 == g and d2.tobytes() == data.tobytes()
    F2 = hip.read_raw(p, be)
    assert F2.grid == g and F2.to_vec_f64().tobytes() == data.tobytes()
    # an FP32 field is written widened to f64
    F32 = hip.Field(g, be.upload(data.astype(np.float32)))
    hip.write_raw(tmp_path / "s.raw", F32)
    _, d3 = hip.read_raw_host(tmp_path / "s.raw")
    assert np.array_equal(d3, data.astype(np.float32).astype(np.float64))
