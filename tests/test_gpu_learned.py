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


def _vark_problem(be, n, kseed=424242, fseed=434343):
    from fluxion_amd import hip
    from fluxion_amd.core import Boundaries, DType, Grid, Transfer
    from fluxion_amd.inputs import gaussian_bumps, grf_log_k
    g = Grid.square(n, 1.0 / (n - 1))
    k = grf_log_k(n, seed=kseed)
    f = gaussian_bumps(n, seed=fseed).ravel()
    K = hip.Field.from_f64(g, k.ravel(), be)
    LK = hip.Field.from_f64(g, np.log(k).ravel(), be)
    mg = hip.Multigrid(g, Boundaries.zero_dirichlet(), 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16,
                       precision=DType.F32, fused=True, scale=1.0, k=K, be=be)
    return hip, g, k, f, K, LK, mg


def test_learned_fcg_matches_oracle_composition(be):
    # The device's learned flexible CG (committed weights) against the oracle's composition of the same
    # pieces (FP32 V-cycle, BF16-storage U-Net, α* step, true-residual fallback) at 64².  The networks
    # differ by FP32 accumulation order, so the trajectories agree to the iteration count ±2, the same
    # fallback decision and solutions within the tolerance's reach.
    import oracle as O
    from fluxion_amd.learned import DEFAULT_WEIGHTS, LearnedPreconditioner, UNet, load_weights
    hip, g, k, f, K, LK, mg = _vark_problem(be, 64)
    params = load_weights(DEFAULT_WEIGHTS)
    net = UNet(be, 2)
    net.set_params(params)
    tol = 1e-8 * np.linalg.norm(f)
    lp = LearnedPreconditioner(net, mg, g, LK, stall=0.9)
    x = hip.Field.zeros(g, be=be)
    s = hip.FcgSolver(g, tol, 200, be=be).solve_op(x, hip.Field.from_f64(g, f, be), hip.Operator.var_k(K), lp)
    fb = lp.fell_back()
    cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI, transfer=1, coarsest=16, precision=1, scale=1.0,
                   k=k.ravel())
    xr, itr, fbr = O.learned_fcg(f, g, O.Op.var_k(k.ravel()), cfg, params.astype(np.float64), np.log(k).ravel(),
                                 tol, 200)
    d = np.linalg.norm(x.to_vec_f64() - xr) / np.linalg.norm(xr)
    print(f"learned FCG 64²: gpu {s.iterations} its (fell back {fb}), oracle {itr} its (fell back {fbr}), "
          f"rel diff {d:.2e}")
    assert s.converged() and s.true_residual < tol
    assert abs(s.iterations - itr) <= 2 and fb == fbr
    assert d < 1e-5


@pytest.mark.parametrize("n", [1024, 4096])
def test_learned_fcg_large(be, n):
    # The C5 evaluation sizes: the learned preconditioner inside FCG reaches 1e-8 in FP64 true residual;
    # iterations and time against the V-cycle alone are printed (the network's cost and gain are
    # weighed in the design notes).
    import time
    from fluxion_amd.learned import DEFAULT_WEIGHTS, LearnedPreconditioner, UNet, load_weights
    hip, g, k, f, K, LK, mg = _vark_problem(be, n)
    net = UNet(be, 2)
    net.set_params(load_weights(DEFAULT_WEIGHTS))
    tol = 1e-8 * np.linalg.norm(f)
    op = hip.Operator.var_k(K)
    F = hip.Field.from_f64(g, f, be)
    solver = hip.FcgSolver(g, tol, 300, be=be)
    out = {}
    for name, pc in (("V-cycle", mg), ("learned", LearnedPreconditioner(net, mg, g, LK, stall=0.9))):
        x = hip.Field.zeros(g, be=be)
        solver.solve_op(x, F, op, pc)
        x.fill(0.0)
        be.synchronize()
        t0 = time.perf_counter()
        s = solver.solve_op(x, F, op, pc)
        be.synchronize()
        out[name] = (s, time.perf_counter() - t0)
        assert s.converged() and s.true_residual < tol, (name, s)
    print(f"FCG {n}²: V-cycle {out['V-cycle'][0].iterations} its {1e3 * out['V-cycle'][1]:.2f} ms; learned "
          f"{out['learned'][0].iterations} its {1e3 * out['learned'][1]:.2f} ms")


def test_learned_crank_nicolson_steps(be):
    # Warm-started Crank–Nicolson steps of the 1024² variable-k diffusion with the learned preconditioner
    # (over the stepper's own multigrid) agree with the V-cycle stepper's states at every step's tolerance.
    from fluxion_amd.inputs import gaussian_bumps
    from fluxion_amd.learned import DEFAULT_WEIGHTS, LearnedPreconditioner, UNet, load_weights
    hip, g, k, f, K, LK, mg = _vark_problem(be, 1024)
    net = UNet(be, 2)
    net.set_params(load_weights(DEFAULT_WEIGHTS))
    u0 = gaussian_bumps(1024).ravel()
    tol = 1e-8 * np.linalg.norm(u0)
    finals, its = {}, {}
    for name in ("V-cycle", "learned"):
        stp = hip.CrankNicolsonSolver.var_k(g, 1.0, 1e-4, K, tol, 200, mixed=True, be=be)
        lp = None
        if name == "learned":
            lp = LearnedPreconditioner(net, stp.multigrid(), g, LK, stall=0.9)
            stp.set_preconditioner(lp)
        U = hip.Field.from_f64(g, u0, be)
        its[name] = [stp.step(U).iterations for _ in range(20)]
        finals[name] = U.to_vec_f64()
    d = np.linalg.norm(finals["learned"] - finals["V-cycle"]) / np.linalg.norm(finals["V-cycle"])
    print(f"CN 1024² x20: iterations V-cycle {sum(its['V-cycle'])}, learned {sum(its['learned'])}; final states differ "
          f"by {d:.1e}")
    assert d < 1e-7
