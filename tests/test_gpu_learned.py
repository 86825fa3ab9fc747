"""What the trained network gains over the V-cycle alone (GPU).

The network is trained on the GPU (fx_unet_* forward/backward kernels) on residual→error pairs the
solver generates on the log-normal-k family, then evaluated on held-out problems:

- gain per application: the error the fused FP32 V-cycle leaves, e1 = A⁻¹r1, measured in the energy
  norm before and after the network correction at its optimal step,
  min_α ‖e1 − α·N‖²_A / ‖e1‖²_A  (1.0 = no gain; the V-cycle alone has ratio 1 by definition);
- inside flexible CG: iterations and time with and without it (printed; the learned path keeps the
  V-cycle as its fallback and must still converge).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def energy_unit(v, logk):
    """vᵀA v for the unit-spacing −∇·(k∇·) on interior points (k = exp(logk), harmonic face means,
    values outside the interior taken as zero)."""
    n0, n1 = v.shape
    k = np.exp(logk)
    vi = np.zeros_like(v)
    vi[1:-1, 1:-1] = v[1:-1, 1:-1]
    e = 0.0
    # each interior–interior face once, each face from the interior to the boundary once
    hx = 2 * k[:, 1:] * k[:, :-1] / (k[:, 1:] + k[:, :-1])
    hy = 2 * k[1:, :] * k[:-1, :] / (k[1:, :] + k[:-1, :])
    e += np.sum(hx[1:-1, :] * (vi[1:-1, 1:] - vi[1:-1, :-1]) ** 2)
    e += np.sum(hy[:, 1:-1] * (vi[1:, 1:-1] - vi[:-1, 1:-1]) ** 2)
    return e


def energy_pair(t, y, logk):
    """(tᵀAt, min_α (t − αy)ᵀA(t − αy)) with A the unit-spacing operator."""
    tt = energy_unit(t, logk)
    yy = energy_unit(y, logk)
    ty = 0.25 * (energy_unit(t + y, logk) - energy_unit(t - y, logk))
    a = max(ty / yy, 0.0) if yy > 0 else 0.0
    return tt, tt - a * ty if a > 0 else tt


def test_learned_network_gain(be):
    from fluxion_amd import hip
    from fluxion_amd.core import Boundaries, DType, Grid, Transfer
    from fluxion_amd.inputs import gaussian_bumps, grf_log_k
    from fluxion_amd.learned import LearnedPreconditioner, PairSet, UNet, he_init, solver_pairs, train

    sets = []
    for n, seeds in ((128, range(8)), (256, range(3))):
        ps = [solver_pairs(be, n, seed=500 + 17 * n + s, max_its=8) for s in seeds]
        sets.append(PairSet(np.concatenate([p.x for p in ps]), np.concatenate([p.t for p in ps])))
    net = UNet(be, 2)
    net.set_params(he_init(2, 0))
    losses = train(net, sets, steps=400, batch=8, lr=1e-3, energy=True)
    print(f"training: energy loss {np.mean(losses[:10]):.3f} -> {np.mean(losses[-20:]):.3f}")

    # Held-out problems (seeds not used in training).
    ratios = []
    for n, seed in ((128, 9001), (256, 9002)):
        ps = solver_pairs(be, n, seed=seed, max_its=6, stride=2)
        y = net.forward(ps.x)
        for b in range(len(ps.x)):
            tt, rest = energy_pair(ps.t[b].astype(np.float64), y[b], ps.x[b, :, :, 1].astype(np.float64))
            ratios.append(rest / tt)
    ratios = np.array(ratios)
    print(f"held-out: error energy left after the network correction / before = "
          f"{ratios.mean():.3f} (min {ratios.min():.3f}, max {ratios.max():.3f}) over {len(ratios)} residuals")
    assert ratios.max() <= 1.0 + 1e-9        # the optimal step never increases the error
    assert ratios.mean() < 0.9               # the trained network removes a measurable part of it

    # Inside flexible CG at 256²: iterations and time with and without the network.
    n = 256
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    k = grf_log_k(n, seed=9003)
    f = gaussian_bumps(n, seed=9004).ravel()
    K = hip.Field.from_f64(g, k.ravel(), be)
    LK = hip.Field.from_f64(g, np.log(k).ravel(), be)
    op = hip.Operator.var_k(K)
    F = hip.Field.from_f64(g, f, be)
    mg = hip.Multigrid(g, Boundaries.zero_dirichlet(), 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16,
                       precision=DType.F32, fused=True, scale=1.0, k=K, be=be)
    tol = 1e-8 * np.linalg.norm(f)
    solver = hip.FcgSolver(g, tol, 200, be=be)
    X = hip.Field.zeros(g, be=be)
    s_mg = solver.solve_op(X, F, op, mg)
    lp = LearnedPreconditioner(net, mg, g, LK, stall=0.95)
    X.fill(0.0)
    s_l = solver.solve_op(X, F, op, lp)
    print(f"FCG 256² to 1e-8: V-cycle {s_mg.iterations} its; learned {s_l.iterations} its "
          f"(fell back to the V-cycle: {lp.fell_back()})")
    assert s_mg.converged() and s_l.converged()
    assert s_l.true_residual < tol
