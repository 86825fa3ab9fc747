"""Batched multi-problem solver (fx_batch_*) against the single-problem path and the oracle (GPU)."""
import numpy as np
import pytest

import oracle as O
from fluxion_amd.core import Boundaries, DType, Grid, Transfer
from fluxion_amd.inputs import gaussian_bumps, grf_log_k

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64), (100, 100), (129, 129), (160, 160), (96, 160), (200, 131), (256, 256), (97, 97), (333, 333),
          (65, 65), (128, 128), (80, 80)]


def problems(shapes, seed0=7000):
    grids, ks, fs = [], [], []
    for i, (r, c) in enumerate(shapes):
        g = Grid(r, c, 1.0 / (c - 1), 1.0 / (r - 1))
        if r == c:
            k = grf_log_k(r, seed=seed0 + i)
            f = gaussian_bumps(r, seed=seed0 + 100 + i)
        else:
            k = np.exp(0.5 * np.sin(np.linspace(0, 6, r))[:, None] * np.cos(np.linspace(0, 4, c))[None, :])
            f = gaussian_bumps(r, seed=seed0 + 100 + i, cols=c)
        grids.append(g)
        ks.append(k.ravel())
        fs.append(f.ravel())
    return grids, ks, fs


def test_batch_matches_single_problem_solves(be):
    from fluxion_amd import hip
    grids, ks, fs = problems(SHAPES)
    kc, fc = np.concatenate(ks), np.concatenate(fs)
    K = be.upload(kc)
    B = be.upload(fc)
    X = be.upload(np.zeros_like(fc))
    tols = np.array([1e-8 * np.linalg.norm(f) for f in fs])
    bs = hip.BatchSolver(grids, K, be=be)
    info = bs.info()
    print(f"batch of {len(grids)}: {info}")
    st = bs.solve(X, B, tols, 200)
    xb = X.to_numpy()
    off = bs.offsets()
    for i, g in enumerate(grids):
        assert st[i].converged() and st[i].true_residual < tols[i], (i, st[i])
        # The same problem on its own: FCG around the fused FP32 coefficient-aware V-cycle.
        Ki = hip.Field.from_f64(g, ks[i], be)
        mg = hip.Multigrid(g, Boundaries.zero_dirichlet(), 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR,
                           coarsest=16, precision=DType.F32, fused=True, scale=1.0, k=Ki, be=be)
        xi = hip.Field.zeros(g, be=be)
        si = hip.FcgSolver(g, tols[i], 200, be=be).solve_op(xi, hip.Field.from_f64(g, fs[i], be),
                                                            hip.Operator.var_k(Ki), mg)
        d = np.linalg.norm(xb[off[i]:off[i + 1]] - xi.to_vec_f64()) / np.linalg.norm(xi.to_vec_f64())
        print(f"problem {i} {g.rows}x{g.cols}: batch {st[i].iterations} its, single {si.iterations} its, rel diff {d:.2e}")
        assert abs(st[i].iterations - si.iterations) <= 1
        assert d < (1e-12 if st[i].iterations == si.iterations else 1e-6)
    # Two oracle problems (FP64 FCG around the oracle's FP32 V-cycle).
    for i in (0, 4):
        g = grids[i]
        cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI, transfer=1, coarsest=16, precision=1, scale=1.0,
                       k=ks[i])
        xr, sr = O.fcg(np.zeros_like(fs[i]), fs[i], g, O.Op.var_k(ks[i]), tols[i], 200, mg=cfg)
        d = np.linalg.norm(xb[off[i]:off[i + 1]] - xr) / np.linalg.norm(xr)
        assert abs(st[i].iterations - sr.iterations) <= 2 and d < 1e-6, (i, st[i].iterations, sr.iterations, d)
    # A second solve replays the captured graph and reproduces the first bit for bit.
    X2 = be.upload(np.zeros_like(fc))
    st2 = bs.solve(X2, B, tols, 200)
    assert [s.iterations for s in st2] == [s.iterations for s in st]
    assert np.array_equal(X2.to_numpy(), xb)


def test_batch_rejects_bad_input(be):
    from fluxion_amd import hip
    from fluxion_amd.core import CoreError
    grids, ks, fs = problems([(64, 64), (80, 80)])
    K = be.upload(np.concatenate(ks)[:-1])            # one point short
    with pytest.raises(CoreError):
        hip.BatchSolver(grids, K, be=be)


def config4_problems(count=512):
    """The batched configuration: square sides log-uniform in [64, 1024] (inputs.batch_sizes), k = exp of a
    unit-variance Gaussian field per problem, 8 Gaussian bumps per problem (seeds 20260101 + id)."""
    from fluxion_amd.inputs import SEED0, batch_sizes
    sizes = batch_sizes(count)
    grids = [Grid.square(n, 1.0 / (n - 1)) for n in sizes]
    ks = [grf_log_k(n, seed=SEED0 + i).ravel() for i, n in enumerate(sizes)]
    fs = [gaussian_bumps(n, seed=SEED0 + 1000 + i).ravel() for i, n in enumerate(sizes)]
    return sizes, grids, ks, fs


def test_batch_config4_full_size(be):
    # 512 problems, ≈96M points in one batch: every problem converges to 1e-8 of its own ‖f‖ in FP64
    # true residual; the largest and smallest problems agree with their single-problem solves.
    import time
    from fluxion_amd import hip
    sizes, grids, ks, fs = config4_problems()
    npts = sum(n * n for n in sizes)
    print(f"config 4: {len(sizes)} problems, sides {min(sizes)}..{max(sizes)}, {npts / 1e6:.1f}M points", flush=True)
    K = be.upload(np.concatenate(ks))
    B = be.upload(np.concatenate(fs))
    X = be.upload(np.zeros(npts))
    tols = np.array([1e-8 * np.linalg.norm(f) for f in fs])
    t0 = time.perf_counter()
    bs = hip.BatchSolver(grids, K, be=be)
    print(f"created in {time.perf_counter() - t0:.1f} s: {bs.info()}", flush=True)
    t0 = time.perf_counter()
    st = bs.solve(X, B, tols, 200)
    be.synchronize()
    print(f"first solve {time.perf_counter() - t0:.3f} s", flush=True)
    its = np.array([s.iterations for s in st])
    tr = np.array([s.true_residual for s in st])
    print(f"iterations min {its.min()} mean {its.mean():.1f} max {its.max()}; worst true residual / tol "
          f"{np.max(tr / tols):.3f}", flush=True)
    assert all(s.converged() for s in st)
    assert np.all(tr < tols)
    xb = X.to_numpy()
    off = bs.offsets()
    for i in (int(np.argmax(sizes)), int(np.argmin(sizes))):
        g = grids[i]
        Ki = hip.Field.from_f64(g, ks[i], be)
        mg = hip.Multigrid(g, Boundaries.zero_dirichlet(), 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR,
                           coarsest=16, precision=DType.F32, fused=True, scale=1.0, k=Ki, be=be)
        xi = hip.Field.zeros(g, be=be)
        si = hip.FcgSolver(g, tols[i], 200, be=be).solve_op(xi, hip.Field.from_f64(g, fs[i], be),
                                                            hip.Operator.var_k(Ki), mg)
        d = np.linalg.norm(xb[off[i]:off[i + 1]] - xi.to_vec_f64()) / np.linalg.norm(xi.to_vec_f64())
        print(f"problem {i} ({sizes[i]}²): batch {st[i].iterations} its, single {si.iterations} its, rel diff {d:.2e}")
        assert abs(st[i].iterations - si.iterations) <= 1
        assert d < (1e-12 if st[i].iterations == si.iterations else 1e-6)
