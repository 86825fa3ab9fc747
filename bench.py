"""Benchmark of the structured-grid solve path on one MI355X.

Headline workload (one step = one solve): −∇·(k∇u) = f on a 4096×4096 vertex
grid (h = 1/4095, zero Dirichlet), k = exp(g) with g a unit-variance Gaussian
random field (correlation length 0.1, 256 random Fourier features, seed
20260101), f = 8 seeded Gaussian bumps, solved from x = 0 to
‖f − A·x‖₂ ≤ 1e-8·‖f‖₂ judged on the FP64 true residual.

  value        solves/s of the mixed path: FP64 flexible CG around the fused
               FP32 coefficient-aware V-cycle (2+2 Jacobi ω = 2/3, linear
               transfers, coarsest 16²), or of learned+mixed (the same with the
               U-Net correction) when --learned names trained weights.
  plain        FP64 MG-preconditioned CG (CgSolver::solve_preconditioned) with
               the FP64 V-cycle composed of backend operations, same hierarchy.
  poisson      the same two paths on −Δu = f at 4096² in the reference form
               (coeff = 1, S = −Δ_h − I).
  roofline     the mixed solve's kernel with the largest total time, timed by
               the library's HIP-event probe on its own stream; algorithmic
               bytes per point are listed in kernel_bytes().
  cpu_baseline the CPU oracle (C++ restatement of the reference CPU backend and
               solvers) running one plain FP64 MG-PCG solve of the same problem.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]
                       [--learned weights.npy]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_FLOPS = 2.5e15   # MI355X dense BF16 MFMA
METRIC = "solves/s to 1e-8 rel residual: learned+mixed vs plain MG-PCG; stencil HBM GB/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(vk: bool, fused_update: bool) -> dict:
    """Algorithmic HBM bytes per grid point of each probed kernel (each array once at its dtype).
    fused_update: the FCG residual update rides in the level-0 wave restriction (even columns)."""
    return {
        # r, ap read (FP64); r written (FP64); FP32 copy of r written (the x update is deferred
        # to the next step kernel).
        "fcg_update": 2 * 8 + 8 + 4,
        # z (FP32), p_old and x read; p_new, A·p and x (the previous iteration's deferred
        # x += α·p_old) written (FP64); k read for the var-k operator.
        "fcg_step": 4 + 8 + 8 + 8 + 8 + 8 + (8 if vk else 0),
        # coarse correction (¼ point, FP32) read and z (FP32) written; r, A·p (FP64) read for the
        # dots (the rhs is float(r)); the level's FP32 face coefficients kx, ky and ω/d for var-k.
        "vc_post0": 1 + 4 + 8 + 8 + (12 if vk else 0),
        # level-0 rhs (FP32) -- or, with the fused update, r and A·p read and r_new written (FP64) --,
        # face coefficients and ω/d read, coarse rhs (¼ point, FP32) written.
        "vc_restrict0": (8 + 8 + 8 if fused_update else 4) + 1 + (12 if vk else 0),
        "stencil_axpy": 24,
    }


# HBM bytes per launch of the probed level-0 kernels on the 4096² var-k workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE passes over this bench, scaled by the same counters' reading of a copy of
# 2^24 doubles (fetch ×2.000, write ×1.000); summary and commands in
# profiles/vark4096_mixed_fcg_wave.md.  Only valid for --size 4096 with the wave level kernels, and
# only for the kernel sources the counters were read from: "sources_sha16" is the hash of
# TRAFFIC_SOURCES at that run, and a bench whose sources differ reports traffic null.  The level-0
# V-cycle kernels' fetch depends on the autotuned segment heights (lag rows): both passes ran the
# restriction at 8 coarse rows per wave and the post kernel at 32 fine rows, as the timed run did.
TRAFFIC_SOURCES = ("fluxion_amd/csrc/fx_device.hpp", "fluxion_amd/csrc/fx_mgfast.hpp", "fluxion_amd/csrc/fx_solvers.hip")
MEASURED_TRAFFIC_4096 = {
    "sources_sha16": "3d03464c8a56e57c",
    "vc_restrict0": (275562 * 2 + 151982) * 1024,
    "vc_post0": (292545 * 2 + 74924) * 1024,
    "fcg_step": (244647 * 2 + 393248) * 1024,
}


def poisson_residual_noise(x: np.ndarray, f: np.ndarray, n: int, h: float) -> tuple:
    """‖f − A·x‖ of the reference-form Poisson operator (A·v = v + (−Δ_h v − v) inside, identity rows
    on the boundary) evaluated in FP64 and in x87 extended precision, both relative to ‖f‖, and the
    relative size of their difference: the rounding noise of an FP64 true-residual evaluation, below
    which no FP64 solver can be judged converged."""
    out = []
    for dt in (np.float64, np.longdouble):
        X = x.reshape(n, n).astype(dt)
        c = dt(1.0) / (dt(h) * dt(h))
        lap = (X[:-2, 1:-1] + X[2:, 1:-1] + X[1:-1, :-2] + X[1:-1, 2:] - 4 * X[1:-1, 1:-1]) * c
        r = f.reshape(n, n)[1:-1, 1:-1].astype(dt) - (X[1:-1, 1:-1] + (-lap - X[1:-1, 1:-1]))
        out.append(r)
        del X, lap
    fn = float(np.linalg.norm(f))
    return (float(np.linalg.norm(out[0])) / fn, float(np.linalg.norm(out[1].astype(np.float64))) / fn,
            float(np.linalg.norm((out[0] - out[1]).astype(np.float64))) / fn)


def unet_forward_record(be, net, n: int, reps: int = 10) -> dict:
    """One network forward on an n×n two-channel input (15 convolutions, 61,920 FLOP per pixel), timed
    over `reps` forwards after two warm-ups; inputs and outputs stay on the device."""
    from fluxion_amd.core import DType
    XIN = be.allocate(n * n * 2, DType.F32)
    YOUT = be.allocate(n * n, DType.F32)
    for _ in range(2):
        net.forward_device(XIN, YOUT, 1, n, n)
    be.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward_device(XIN, YOUT, 1, n, n)
    be.synchronize()
    t_f = (time.perf_counter() - t0) / reps
    flop = 61920.0 * n * n
    del XIN, YOUT
    return {"grid": n, "ms_per_forward": 1e3 * t_f, "flop_per_forward": flop, "TFLOPs": flop / t_f / 1e12,
            "frac_bf16_dense_peak": flop / t_f / BF16_PEAK_FLOPS}


def unet_train_record(be, n: int = 256, batch: int = 8, reps: int = 10) -> dict:
    """One training step as learned.train_vark takes it -- forward, relative energy-norm loss, backward
    through all 15 layers (data and weight gradients), Adam update -- on a batch of n×n two-channel
    inputs resident on the device, committed weights as the start; timed over `reps` steps after two
    warm-ups (each step copies its loss to the host, as training does).  FLOP: the forward's 61,920 per
    pixel and the same again for each of the two backward products (data and weight gradients)."""
    from fluxion_amd.learned import DEFAULT_WEIGHTS, UNet, load_weights
    net = UNet(be, 2)
    net.set_params(load_weights(DEFAULT_WEIGHTS))
    rng = np.random.default_rng(7)
    XB = be.upload(rng.standard_normal(batch * n * n * 2).astype(np.float32))
    TB = be.upload(rng.standard_normal(batch * n * n).astype(np.float32))
    for _ in range(2):
        net.loss_grad_device(XB, TB, batch, n, n, energy=True)
        net.adam_step(1e-4)
    be.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.loss_grad_device(XB, TB, batch, n, n, energy=True)
        net.adam_step(1e-4)
    be.synchronize()
    t = (time.perf_counter() - t0) / reps
    flop = 3 * 61920.0 * batch * n * n
    del XB, TB, net
    return {"grid": n, "batch": batch, "ms_per_step": 1e3 * t, "pixels_per_s": batch * n * n / t,
            "TFLOPs_approx": flop / t / 1e12, "frac_bf16_dense_peak_approx": flop / t / BF16_PEAK_FLOPS}


# Which attainable ceiling each probed kernel is read against: the step kernel has the byte mix of the
# 52 B/pt streaming kernel; the level-0 V-cycle kernels read 28-29 B/pt and write 4-9 B/pt, so the
# read-only ceiling bounds them from above.
ATTAINABLE_KIND = {"fcg_step": "mix52", "vc_post0": "read", "vc_restrict0": "read", "fcg_update": "copy",
                   "stencil_axpy": "copy"}


def hbm_ceilings_record(be) -> dict:
    """Attainable HBM bandwidth of plain streaming kernels over 2^24 points (4 rotating copies of the
    arrays, past the 256 MB last-level cache): read, write, copy and the FCG step kernel's 52 B/pt mix."""
    rec = {k: {"us_per_launch": us, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "bytes_per_point": hip_bytes}
           for (k, (us, gbs)), hip_bytes in zip(be.hbm_ceilings(1 << 24, 20).items(), be.HBM_KINDS.values())}
    log("bench: attainable HBM " + ", ".join(f"{k} {v['GBs']:.0f} GB/s ({100 * v['frac']:.0f}%)" for k, v in rec.items()))
    return rec


def sources_sha16() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()[:16]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--c2-tol", type=float, default=1e-10, help="relative tolerance of the Poisson 4096^2 record")
    ap.add_argument("--no-convdiff", action="store_true", help="skip the convection-diffusion 2048^2 record")
    ap.add_argument("--no-reductions", action="store_true", help="skip the dot / dot2 / norm bandwidth record")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU baseline solves (after one warm-up)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-problem batched record")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--no-cn", action="store_true", help="skip the Crank-Nicolson record")
    ap.add_argument("--cn-steps", type=int, default=100)
    ap.add_argument("--no-explicit", action="store_true", help="skip the explicit diffusion record")
    ap.add_argument("--learned", type=str, default=None,
                    help="trained U-Net weights (.npz from fluxion_amd.learned; default: the committed weights)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner records")
    ap.add_argument("--mg-kernels", choices=("wave", "ring", "tiled"), default="wave",
                    help="level kernels of the fused FP32 V-cycle (all bit-identical)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in wave segment heights")
    ap.add_argument("--hbm-only", action="store_true",
                    help="only measure the attainable HBM bandwidth of plain streaming kernels and exit")
    ap.add_argument("--unet-only", action="store_true",
                    help="only time the U-Net forward at --size (committed weights) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run Backend::copy of 2^24 doubles 20 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the library has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    from fluxion_amd import _lib as L
    from fluxion_amd import hip
    from fluxion_amd.core import Boundaries, DType, Grid, Stencil, Transfer
    from fluxion_amd.inputs import gaussian_bumps, grf_log_k

    if args.hbm_only:
        rec = hbm_ceilings_record(hip.HipBackend(0))
        print(json.dumps({"hbm_ceilings": rec}), flush=True)
        return

    if args.unet_only:
        # The network forward alone (committed weights), for kernel traces of the convolutions.
        from fluxion_amd.learned import DEFAULT_WEIGHTS, UNet, load_weights
        be = hip.HipBackend(0)
        net = UNet(be, 2)
        net.set_params(load_weights(args.learned or DEFAULT_WEIGHTS))
        rec = unet_forward_record(be, net, args.size)
        log(f"bench: U-Net forward {args.size}^2 {rec['ms_per_forward']:.2f} ms, {rec['TFLOPs']:.0f} TFLOP/s")
        trec = unet_train_record(be)
        log(f"bench: U-Net training step 8x256^2 {trec['ms_per_step']:.2f} ms "
            f"({trec['pixels_per_s'] / 1e6:.0f} Mpixel/s, ~{trec['TFLOPs_approx']:.0f} TFLOP/s)")
        print(json.dumps({"unet_forward": rec, "unet_train_step": trec}), flush=True)
        return

    if args.copy_calibration:
        # A copy of known size: 2^24 doubles read and written per launch (128 MiB + 128 MiB).
        be = hip.HipBackend(0)
        src = hip.Field.zeros(Grid(4096, 4096, 1.0, 1.0), be=be)
        dst = hip.Field.zeros(Grid(4096, 4096, 1.0, 1.0), be=be)
        src.fill(1.0)
        for _ in range(20):
            dst.copy_from(src)
        be.synchronize()
        log("bench: copy calibration done (20 launches of 2^24 doubles)")
        return

    n = args.size
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    bc = Boundaries.zero_dirichlet()
    be = hip.HipBackend(0)
    log(f"bench: backend {be.name}, grid {n}x{n}")
    max_iter = 300

    def timed(fn, steps, warmup, probe=None):
        for _ in range(warmup):
            fn()
        be.synchronize()
        if probe is not None:
            be.probe(probe)
        t0 = time.perf_counter()
        for _ in range(steps):
            s = fn()
        be.synchronize()
        dt = (time.perf_counter() - t0) / steps
        pr = None
        if probe is not None:
            pr = be.probe_read()
            be.probe(be.PROBE_OFF)
        return dt, s, pr

    # ── variable-coefficient 4096² (headline) ─────────────────────────
    k = grf_log_k(n)
    f = gaussian_bumps(n).ravel()
    fnorm = float(np.linalg.norm(f))
    tol = args.tol * fnorm
    K = hip.Field.from_f64(g, k.ravel(), be)
    F = hip.Field.from_f64(g, f, be)
    X = hip.Field.zeros(g, be=be)
    op = hip.Operator.var_k(K)
    mgv64 = hip.Multigrid(g, bc, 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16, precision=DType.F64,
                          fused=False, scale=1.0, k=K, be=be)
    mgv32 = hip.Multigrid(g, bc, 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16, precision=DType.F32,
                          fused=True, scale=1.0, k=K, be=be)
    from fluxion_amd.core import MgKernels
    mgk = {"wave": MgKernels.WAVE, "ring": MgKernels.RING, "tiled": MgKernels.TILED}[args.mg_kernels]
    mgv32.set_kernels(mgk)
    if mgk == MgKernels.WAVE and not args.no_autotune:
        log(f"bench: autotuned wave segment heights, {mgv32.autotune(5):.3f} ms per cycle of level kernels")
    pcg = hip.CgSolver(g, tol, max_iter, be=be)
    fcg = hip.FcgSolver(g, tol, max_iter, be=be)

    def v_plain():
        X.fill(0.0)
        return pcg.solve_op(X, F, op, mgv64)

    def v_mixed(pc=mgv32):
        def run():
            X.fill(0.0)
            return fcg.solve_op(X, F, op, pc)
        return run

    t_plain, s_plain, _ = timed(v_plain, max(1, args.steps // 2), max(1, args.warmup // 2))
    x_plain = X.to_vec_f64()
    log(f"bench: var-k plain FP64 MG-PCG {s_plain.iterations} its, {1e3 * t_plain:.2f} ms/solve, "
        f"true rel res {s_plain.true_residual / fnorm:.2e}")
    t_mixed, s_mixed, pr_update = timed(v_mixed(), args.steps, args.warmup, be.PROBE_FCG_UPDATE)
    x_mixed = X.to_vec_f64()
    log(f"bench: var-k mixed FCG {s_mixed.iterations} its, {1e3 * t_mixed:.2f} ms/solve, "
        f"true rel res {s_mixed.true_residual / fnorm:.2e}")

    # ── learned + mixed: the same FCG with the U-Net correction (committed weights) ──
    learned, unet_rec, unet_train, net, LK = None, None, None, None, None
    t_value, s_value, label = t_mixed, s_mixed, "mixed"
    if not args.no_learned:
        from fluxion_amd.learned import DEFAULT_WEIGHTS, LearnedPreconditioner, UNet, load_weights
        wpath = args.learned or DEFAULT_WEIGHTS
        net = UNet(be, 2)
        net.set_params(load_weights(wpath))
        LK = hip.Field.from_f64(g, np.log(k).ravel(), be)
        lp = LearnedPreconditioner(net, mgv32, g, LK, stall=0.9)
        t_l, s_l, _ = timed(v_mixed(lp), max(1, args.steps // 4), 1)
        fb = lp.fell_back()
        log(f"bench: var-k learned+mixed FCG {s_l.iterations} its, {1e3 * t_l:.2f} ms/solve, "
            f"true rel res {s_l.true_residual / fnorm:.2e}, fell back to the V-cycle {fb}")
        learned = {"solves_per_s": 1.0 / t_l, "ms_per_solve": 1e3 * t_l, "iterations": s_l.iterations,
                   "true_rel_residual": s_l.true_residual / fnorm, "fell_back": fb,
                   "speedup_vs_plain": t_plain / t_l, "speedup_vs_mixed": t_mixed / t_l,
                   "weights": Path(wpath).name,
                   "note": "the network correction costs more iterations and time than it saves; the headline "
                           "value is the mixed path"}
        # The network alone: one forward at this size.
        unet_rec = unet_forward_record(be, net, n)
        t_f = 1e-3 * unet_rec["ms_per_forward"]
        unet_rec["mixed_fcg_iterations_per_forward"] = t_f / (t_mixed / max(1, s_mixed.iterations))
        log(f"bench: U-Net forward {n}^2 {1e3 * t_f:.2f} ms, {unet_rec['TFLOPs']:.0f} TFLOP/s "
            f"({100 * unet_rec['frac_bf16_dense_peak']:.1f}% of BF16 dense peak)")
        # One training step of the network at its training configuration (8 × 256², energy-norm loss).
        unet_train = unet_train_record(be)
        log(f"bench: U-Net training step 8x256^2 {unet_train['ms_per_step']:.2f} ms "
            f"({unet_train['pixels_per_s'] / 1e6:.0f} Mpixel/s, ~{unet_train['TFLOPs_approx']:.0f} TFLOP/s)")

    # Per-kernel probes of the mixed solve (the x/r update was probed over the timed region).
    kb = kernel_bytes(True, mgk == MgKernels.WAVE and n % 2 == 0)
    probes = {"fcg_update": pr_update}
    for name, kind in (("fcg_step", be.PROBE_FCG_STEP), ("vc_post0", be.PROBE_VC_POST0),
                       ("vc_restrict0", be.PROBE_VC_RESTRICT0)):
        _, _, probes[name] = timed(v_mixed(), 2, 0, kind)
    kernels = {}
    for name, (ms, cnt) in probes.items():
        if cnt == 0:
            continue
        us = 1e3 * ms / cnt
        bpp = kb[name]
        gbs = bpp * n * n / (us * 1e-6) / 1e9
        kernels[name] = {"us_per_launch": us, "launches": cnt, "bytes_per_point": bpp, "GBs": gbs,
                         "frac": gbs / HBM_PEAK_GBS, "ms_per_solve": ms / (args.steps if name == "fcg_update" else 2)}
        log(f"bench: kernel {name}: {us:.1f} us/launch, {gbs:.0f} GB/s algorithmic ({100 * gbs / HBM_PEAK_GBS:.0f}%)")
    dom = max(kernels, key=lambda m: kernels[m]["ms_per_solve"])
    ceilings = hbm_ceilings_record(be)
    for name, kd_ in kernels.items():
        c = ceilings[ATTAINABLE_KIND[name]]
        kd_["attainable_kind"] = ATTAINABLE_KIND[name]
        kd_["frac_of_attainable"] = kd_["GBs"] / c["GBs"]

    # Stencil HBM GB/s (Backend::stencil_axpy, 5-point) with the same probe.
    Y = hip.Field.zeros(g, be=be)
    lap = Stencil.laplacian_2d_5pt(h, h)
    for _ in range(5):
        Y.stencil_axpy(1e-9, X, lap, bc)
    be.synchronize()
    be.probe(be.PROBE_STENCIL_AXPY)
    for _ in range(50):
        Y.stencil_axpy(1e-9, X, lap, bc)
    ms, cnt = be.probe_read()
    be.probe(be.PROBE_OFF)
    st_us = 1e3 * ms / cnt
    st_gbs = 24.0 * n * n / (st_us * 1e-6) / 1e9
    log(f"bench: stencil_axpy {st_us:.1f} us/launch, {st_gbs:.0f} GB/s algorithmic")

    # ── Poisson 4096² (reference form) ────────────────────────────────
    poisson = None
    if not args.no_poisson:
        st = Stencil.poisson_minus_identity(h, h)
        fp = gaussian_bumps(n).ravel()
        fpn = float(np.linalg.norm(fp))
        B = hip.Field.from_f64(g, fp, be)

        def pmg(precision, fused):
            return hip.Multigrid(g, bc, 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16,
                                 precision=precision, fused=fused, scale=-1.0, be=be)
        mp64, mp32 = pmg(DType.F64, False), pmg(DType.F32, True)
        mp32.set_kernels(mgk)
        if mgk == MgKernels.WAVE and not args.no_autotune:
            mp32.autotune(5)
        ptol = args.c2_tol * fpn
        ppcg = hip.CgSolver(g, ptol, max_iter, be=be)
        pfcg = hip.FcgSolver(g, ptol, max_iter, be=be)
        ppipe = hip.PipelinedCgSolver(g, ptol, max_iter, be=be)

        def p_plain():
            X.fill(0.0)
            return ppcg.solve_preconditioned(X, B, 1.0, st, bc, mp64)

        def p_mixed():
            X.fill(0.0)
            return pfcg.solve_preconditioned(X, B, 1.0, st, bc, mp32)

        def p_pipe():
            X.fill(0.0)
            return ppipe.solve_preconditioned(X, B, 1.0, st, bc, mp32)
        tp, sp, _ = timed(p_plain, max(1, args.steps // 2), 1)
        tm, sm, _ = timed(p_mixed, max(1, args.steps // 2), 1)
        tq, sq, _ = timed(p_pipe, max(1, args.steps // 2), 1)
        r64, rext, noise = poisson_residual_noise(X.to_vec_f64(), fp, n, h)
        log(f"bench: poisson pipelined-CG solution: true rel res {r64:.2e} in FP64, {rext:.2e} in extended "
            f"precision; FP64 evaluation noise {noise:.2e}")
        log(f"bench: poisson to {args.c2_tol:g}: plain {sp.iterations} its {1e3 * tp:.2f} ms, mixed FCG {sm.iterations} its "
            f"{1e3 * tm:.2f} ms, pipelined CG + FP32 MG {sq.iterations} its {1e3 * tq:.2f} ms")

        def rec(t, s_):
            return {"solves_per_s": 1.0 / t, "ms_per_solve": 1e3 * t, "iterations": s_.iterations,
                    "reason": s_.reason.name, "true_rel_residual": s_.true_residual / fpn}
        poisson = {"tol_rel": args.c2_tol, "operator": "reference form coeff=1, S=-Lap_h-I (zero Dirichlet)",
                   "plain_fp64_mg_pcg": rec(tp, sp), "mixed_fcg": rec(tm, sm), "pipelined_cg_fp32_mg": rec(tq, sq),
                   "speedup_mixed_vs_plain": tp / tm, "speedup_pipelined_vs_plain": tp / tq,
                   "fp64_floor": {"pipelined_true_rel_residual_fp64": r64, "pipelined_true_rel_residual_extended": rext,
                                  "fp64_evaluation_noise_rel": noise,
                                  "note": "the tolerance 1e-10 is below the rounding noise of evaluating f - A.x in "
                                          "FP64 at this size; the refinement passes stop at that floor and count "
                                          "as converged, as the recurrence criterion of CgSolver does"}}
        del mp64, mp32

    # ── upwind convection-diffusion 2048² (κ = 1e-3, streamfunction velocity, cell Péclet ≤ 50) ──
    convdiff = None
    if not args.no_convdiff:
        from fluxion_amd.inputs import streamfunction_velocity
        nd, kap = 2048, 1e-3
        gd = Grid.square(nd, 1.0 / (nd - 1))
        vxd, vyd, pe = streamfunction_velocity(nd, kap)
        fd = gaussian_bumps(nd).ravel()
        fdn = float(np.linalg.norm(fd))
        VX, VY = hip.Field.from_f64(gd, vxd.ravel(), be), hip.Field.from_f64(gd, vyd.ravel(), be)
        FD, XD = hip.Field.from_f64(gd, fd, be), hip.Field.zeros(gd, be=be)
        opd = hip.Operator.conv_diff(VX, VY, kappa_scalar=-kap, coeff=1.0, sigma=0.0)
        mgd = hip.Multigrid(gd, bc, 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16, precision=DType.F32,
                            fused=False, scale=1.0, vx=VX, vy=VY, kappa=-kap, be=be)
        dtol = args.tol * fdn
        convdiff = {"grid": nd, "kappa": kap, "tol_rel": args.tol,
                    "operator": "-kappa*Lap_h u + v.grad^- u (reference upwind, ops.rs:707-731), zero Dirichlet",
                    "mg": "FP32 V(2,2) Jacobi w=2/3 on the same upwind operator per level, linear transfers, coarsest 16",
                    "peclet": {"min": float(pe.min()), "median": float(np.median(pe)), "max": float(pe.max()),
                               "p10": float(np.percentile(pe, 10)), "p90": float(np.percentile(pe, 90))}}
        for name, solver in (("bicgstab", hip.BiCgStabSolver(gd, dtol, 3000, be=be)),
                             ("fgmres30", hip.FgmresSolver(gd, 30, dtol, 3000, be=be)),
                             ("gmres30_left", hip.GmresSolver(gd, 30, dtol, 3000, be=be))):
            def run(sv=solver):
                XD.fill(0.0)
                return sv.solve_op(XD, FD, opd, mgd)
            t, s_, _ = timed(run, 2, 1)
            convdiff[name] = {"solves_per_s": 1.0 / t, "ms_per_solve": 1e3 * t, "iterations": s_.iterations,
                              "reason": s_.reason.name, "true_rel_residual": s_.true_residual / fdn}
            log(f"bench: conv-diff 2048^2 {name}: {s_.iterations} its, {1e3 * t:.1f} ms/solve, "
                f"true rel res {s_.true_residual / fdn:.2e}")
            del solver
        convdiff["gmres30_left"]["note"] = ("reference left preconditioning: stops on ||M^-1 r||, so the FP64 "
                                            "true residual is not held to the tolerance")
        del mgd, VX, VY, FD, XD

    # ── Backend reductions at 2^24 points: dot, dot2, norm_l2 (block-reduction kernel, HIP events) ──
    reductions = None
    if not args.no_reductions:
        reductions = {"points": n * n}
        Y2 = hip.Field.from_f64(g, np.linspace(0.0, 1.0, n * n), be)
        for name, fn_, bpp in (("dot", lambda: X.dot(F), 16), ("dot2", lambda: X.dot2(F, K, Y2), 32),
                               ("norm_l2", lambda: X.norm_l2(), 8)):
            for _ in range(3):
                fn_()
            be.probe(be.PROBE_REDUCE)
            for _ in range(20):
                fn_()
            ms, cnt = be.probe_read()
            be.probe(be.PROBE_OFF)
            us = 1e3 * ms / cnt
            gbs = bpp * n * n / (us * 1e-6) / 1e9
            reductions[name] = {"us_per_launch": us, "bytes_per_point": bpp, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
            log(f"bench: {name} {us:.1f} us/launch, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.0f}%)")
        del Y2

    # ── batched variable-coefficient diffusion: 512 problems, sides log-uniform in [64, 1024] ──
    batch = None
    if not args.no_batch:
        from fluxion_amd.inputs import SEED0, batch_sizes
        sizes = batch_sizes(512)
        bgrids = [Grid.square(m, 1.0 / (m - 1)) for m in sizes]
        bks = [grf_log_k(m, seed=SEED0 + i).ravel() for i, m in enumerate(sizes)]
        bfs = [gaussian_bumps(m, seed=SEED0 + 1000 + i).ravel() for i, m in enumerate(sizes)]
        npts = int(sum(m * m for m in sizes))
        BK, BB = be.upload(np.concatenate(bks)), be.upload(np.concatenate(bfs))
        BX = be.upload(np.zeros(npts))
        btols = np.array([args.tol * np.linalg.norm(v) for v in bfs])
        t0 = time.perf_counter()
        bsol = hip.BatchSolver(bgrids, BK, be=be)
        t_create = time.perf_counter() - t0

        # Every solve starts from x = 0, cleared on the device (a copy from a zero buffer).
        zero = be.upload(np.zeros(npts))

        def bsolve():
            L.check(be._lib.fx_copy(be.h, zero.h, BX.h), "copy")
            return bsol.solve(BX, BB, btols, max_iter)
        tb, sb, _ = timed(bsolve, 3, 1)
        bits = np.array([q.iterations for q in sb])
        btr = np.array([q.true_residual for q in sb])
        batch = {"problems": len(sizes), "points": npts, "side_min": int(min(sizes)), "side_max": int(max(sizes)),
                 "tol_rel": args.tol, "create_s": t_create, "ms_per_batch_solve": 1e3 * tb,
                 "batch_solves_per_s": 1.0 / tb, "problem_solves_per_s": len(sizes) / tb,
                 "iterations": {"min": int(bits.min()), "mean": float(bits.mean()), "max": int(bits.max())},
                 "all_converged": bool(all(q.converged() for q in sb)), "worst_true_res_over_tol": float(np.max(btr / btols)),
                 "point_iterations_per_s": float(np.sum(np.array(sizes, dtype=np.float64) ** 2 * bits)) / tb,
                 "info": bsol.info()}
        log(f"bench: batch of {len(sizes)} ({npts / 1e6:.1f}M points): {1e3 * tb:.1f} ms per batch solve, "
            f"its {bits.min()}..{bits.max()}, all converged {batch['all_converged']}")
        del bsol, BK, BB, BX, zero

    # ── config 1: Poisson 256², unpreconditioned CgSolver to 1e-8 (the reference's CPU-runnable case) ──
    c1 = None
    n1 = 256
    g1 = Grid.square(n1, 1.0 / (n1 - 1))
    st1 = Stencil.poisson_minus_identity(1.0 / (n1 - 1), 1.0 / (n1 - 1))
    f1 = gaussian_bumps(n1).ravel()
    tol1 = args.tol * float(np.linalg.norm(f1))
    B1, X1 = hip.Field.from_f64(g1, f1, be), hip.Field.zeros(g1, be=be)
    cg1 = hip.CgSolver(g1, tol1, 5000, be=be)

    def c1_solve():
        X1.fill(0.0)
        return cg1.solve(X1, B1, 1.0, st1, bc)
    t1, s1, _ = timed(c1_solve, 20, 3)
    c1 = {"grid": n1, "tol_rel": args.tol, "gpu_cg": {"solves_per_s": 1.0 / t1, "ms_per_solve": 1e3 * t1,
                                                      "iterations": s1.iterations}}
    log(f"bench: C1 Poisson 256^2 CG {s1.iterations} its, {1e3 * t1:.2f} ms/solve on the GPU")

    # ── Crank–Nicolson on variable k: 1024², warm-started steps (implicit.rs:178-222 recurrence) ──
    cn = None
    if not args.no_cn:
        nc, cn_dt = 1024, 1e-4
        gc = Grid.square(nc, 1.0 / (nc - 1))
        kc = hip.Field.from_f64(gc, grf_log_k(nc).ravel(), be)
        u0c = gaussian_bumps(nc).ravel()
        cn_tol = args.tol * float(np.linalg.norm(u0c))
        cn = {"grid": nc, "dt": cn_dt, "alpha": 1.0, "steps": args.cn_steps,
              "tol_abs": cn_tol, "tol_note": f"absolute ||r|| < {args.tol:g}*||u0|| per step"}
        finals = {}
        legs = [("mixed_fcg", True), ("plain_fp64_mg_pcg", False)]
        if net is not None:
            legs.append(("learned_mixed_fcg", True))
            LKc = hip.Field.from_f64(gc, np.log(grf_log_k(nc)).ravel(), be)
        for name, mixed in legs:
            stp = hip.CrankNicolsonSolver.var_k(gc, 1.0, cn_dt, kc, cn_tol, max_iter, mixed=mixed, be=be)
            if mixed and mgk == MgKernels.WAVE and not args.no_autotune:
                stp.autotune(5)
            if name == "learned_mixed_fcg":
                from fluxion_amd.learned import LearnedPreconditioner
                lpc = LearnedPreconditioner(net, stp.multigrid(), gc, LKc, stall=0.9)
                stp.set_preconditioner(lpc)
            U = hip.Field.from_f64(gc, u0c, be)
            for _ in range(2):
                stp.step(U)          # warm-up: first-use allocations, then reset
            U.write(u0c)
            be.synchronize()
            its = 0
            t0 = time.perf_counter()
            for _ in range(args.cn_steps):
                its += stp.step(U).iterations
            be.synchronize()
            t = (time.perf_counter() - t0) / args.cn_steps
            finals[name] = U.to_vec_f64()
            cn[name] = {"steps_per_s": 1.0 / t, "ms_per_step": 1e3 * t, "avg_iterations": its / args.cn_steps}
            del stp
        cn["speedup_mixed_vs_plain"] = cn["plain_fp64_mg_pcg"]["ms_per_step"] / cn["mixed_fcg"]["ms_per_step"]
        if "learned_mixed_fcg" in cn:
            cn["learned_final_state_rel_diff"] = float(np.linalg.norm(finals["learned_mixed_fcg"] - finals["mixed_fcg"])
                                                       / np.linalg.norm(finals["mixed_fcg"]))
            log(f"bench: CN learned+mixed {cn['learned_mixed_fcg']['ms_per_step']:.2f} ms/step "
                f"({cn['learned_mixed_fcg']['avg_iterations']:.1f} its)")
        cn["final_state_rel_diff"] = float(np.linalg.norm(finals["mixed_fcg"] - finals["plain_fp64_mg_pcg"])
                                           / np.linalg.norm(finals["plain_fp64_mg_pcg"]))
        log(f"bench: CN var-k {nc}^2, {args.cn_steps} steps dt={cn_dt:g}: mixed "
            f"{cn['mixed_fcg']['ms_per_step']:.2f} ms/step ({cn['mixed_fcg']['avg_iterations']:.1f} its), plain "
            f"{cn['plain_fp64_mg_pcg']['ms_per_step']:.2f} ms/step ({cn['plain_fp64_mg_pcg']['avg_iterations']:.1f} its), "
            f"final states differ by {cn['final_state_rel_diff']:.1e}")
        del kc

    # ── explicit diffusion (DiffusionSolver) on the 4096² grid, step_n from a captured graph ──
    explicit = None
    if not args.no_explicit:
        from fluxion_amd.core import TimeScheme
        ue0 = gaussian_bumps(n).ravel()
        UE = hip.Field.from_f64(g, ue0, be)
        explicit = {"grid": n, "alpha": 1.0, "dt": "scheme default (CFL-safe)", "steps": 200}
        for name, sch, bpp in (("euler", TimeScheme.EULER, 16), ("rk2", TimeScheme.RK2, 40),
                               ("rk4", TimeScheme.RK4, None)):
            dsol = hip.DiffusionSolver.build(g, 1.0, None, bc, sch, be=be)
            dsol.step_n(UE, 4)
            be.synchronize()
            t0 = time.perf_counter()
            dsol.step_n(UE, 200)
            be.synchronize()
            t = (time.perf_counter() - t0) / 200
            rec = {"us_per_step": 1e6 * t, "steps_per_s": 1.0 / t}
            if bpp is not None:
                rec["bytes_per_point"] = bpp
                rec["GBs_algorithmic"] = bpp * n * n / t / 1e9
                rec["frac"] = rec["GBs_algorithmic"] / HBM_PEAK_GBS
            explicit[name] = rec
            del dsol
        # Euler from the captured graph against direct launches of the same kernels, here and on a small
        # grid, where launch latency rather than bandwidth sets the step time.  The solver's graph holds 64
        # Euler steps: the warm-up replays it twice (the first launches of a new graph executable carry a
        # one-time cost), and the timed runs are 256 steps, whole replays, best of 3 in both modes.
        for gn, tag in ((n, ""), (256, "_256")):
            gg = g if gn == n else Grid.square(gn, 1.0 / (gn - 1))
            UU = UE if gn == n else hip.Field.from_f64(gg, gaussian_bumps(gn).ravel(), be)
            for use_graph in (True, False):
                dsol = hip.DiffusionSolver.new(gg, 1.0, be=be)
                if use_graph:
                    dsol.enable_hip_graph()
                dsol.step_n(UU, 128)
                be.synchronize()
                t = 1e9
                for _ in range(3):
                    t0 = time.perf_counter()
                    dsol.step_n(UU, 256)
                    be.synchronize()
                    t = min(t, (time.perf_counter() - t0) / 256)
                explicit["euler" + tag + ("_graph" if use_graph else "_no_graph")] = {"us_per_step": 1e6 * t}
                del dsol
        # The reference's Euler step as its own op sequence (copy u → s, then stencil_axpy s → u).
        SE = hip.Field.zeros(g, be=be)
        lap5 = Stencil.laplacian_2d_5pt(h, h)
        dt_e = 0.4 * h * h / 4.0
        for _ in range(4):
            SE.copy_from(UE)
            UE.stencil_axpy(dt_e, SE, lap5, bc)
        be.synchronize()
        t0 = time.perf_counter()
        for _ in range(200):
            SE.copy_from(UE)
            UE.stencil_axpy(dt_e, SE, lap5, bc)
        be.synchronize()
        t = (time.perf_counter() - t0) / 200
        explicit["euler_reference_op_sequence"] = {"us_per_step": 1e6 * t, "steps_per_s": 1.0 / t,
                                                   "bytes_per_point": 40}
        explicit["euler_speedup_vs_op_sequence"] = t / (1e-6 * explicit["euler"]["us_per_step"])
        log(f"bench: explicit diffusion {n}^2: Euler {explicit['euler']['us_per_step']:.1f} us/step "
            f"({explicit['euler']['GBs_algorithmic']:.0f} GB/s at 16 B/pt), reference op sequence "
            f"{1e6 * t:.1f} us/step, graph {explicit['euler_graph']['us_per_step']:.1f}; 256^2 graph "
            f"{explicit['euler_256_graph']['us_per_step']:.1f} vs {explicit['euler_256_no_graph']['us_per_step']:.1f} us/step; "
            f"RK2 {explicit['rk2']['us_per_step']:.1f}, RK4 {explicit['rk4']['us_per_step']:.1f} us/step")
        del UE, SE

    # ── CPU baseline: the oracle's plain FP64 MG-PCG on the same var-k problem ──
    cpu = None
    if not args.no_cpu:
        import oracle as O
        kk = k.ravel()
        cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI, transfer=1, coarsest=16, scale=1.0, k=kk)

        def cpu_solve():
            return O.pcg(np.zeros_like(f), f, g, O.Op.var_k(kk), tol, max_iter, mg=cfg)
        xr, sr = cpu_solve()   # warm-up
        times = []
        for _ in range(max(1, args.cpu_reps)):
            t0 = time.perf_counter()
            xr, sr = cpu_solve()
            times.append(time.perf_counter() - t0)
        t_cpu = float(np.median(times))
        cores = O.num_threads()
        model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        nr = float(np.linalg.norm(xr))
        dp = float(np.linalg.norm(x_plain - xr) / nr)
        dm = float(np.linalg.norm(x_mixed - xr) / nr)
        log(f"bench: CPU oracle plain MG-PCG {sr.iterations} its, median {t_cpu:.2f} s of {len(times)} on {cores} "
            f"threads ({model}); GPU plain vs CPU rel diff {dp:.2e} ({s_plain.iterations} vs {sr.iterations} its), "
            f"mixed {dm:.2e}")
        # Config 1 on the CPU: the reference's own CPU-runnable case.
        cfn = lambda: O.cg(np.zeros_like(f1), f1, g1, O.Op.stencil(1.0, st1), tol1, 5000)
        xr1, sr1 = cfn()
        t1s = []
        for _ in range(5):
            t0 = time.perf_counter()
            xr1, sr1 = cfn()
            t1s.append(time.perf_counter() - t0)
        t1c = float(np.median(t1s))
        c1["cpu_cg"] = {"solves_per_s": 1.0 / t1c, "ms_per_solve": 1e3 * t1c, "iterations": sr1.iterations,
                        "cores": cores, "kind": "port", "sample": "median of 5 solves after one warm-up"}
        c1["gpu_vs_cpu_rel_diff"] = float(np.linalg.norm(X1.to_vec_f64() - xr1) / np.linalg.norm(xr1))
        log(f"bench: C1 CPU oracle CG {sr1.iterations} its, {1e3 * t1c:.1f} ms/solve; GPU vs CPU rel diff "
            f"{c1['gpu_vs_cpu_rel_diff']:.2e}")
        cpu = {"value": 1.0 / t_cpu, "unit": "solves/s", "cores": cores, "kind": "port", "cpu_model": model,
               "sample": f"plain FP64 MG-PCG of the same {n}x{n} variable-coefficient problem ({sr.iterations} its): "
                         f"median of {len(times)} solves after one warm-up",
               "seconds": times, "iterations": sr.iterations, "gpu_plain_rel_diff": dp, "gpu_mixed_rel_diff": dm}

    kd = kernels[dom]
    sha = sources_sha16()
    traffic, traffic_note = None, None
    if n != 4096 or args.mg_kernels != "wave":
        traffic_note = "counter figures exist for the 4096^2 wave-kernel workload only"
    elif MEASURED_TRAFFIC_4096["sources_sha16"] != sha:
        traffic_note = f"kernel sources ({sha}) differ from those of the counter run: traffic not restated"
    else:
        traffic = MEASURED_TRAFFIC_4096.get(dom)
        traffic_note = f"rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated, sources {sha}"
    out = {
        "metric": METRIC,
        "value": 1.0 / t_value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_value,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 V-cycle",
        "data": "synthetic (seeded log-normal k, 8 Gaussian bumps)",
        "config": {"workload": f"variable-coefficient diffusion -div(k grad u)=f, {n}x{n}, log-normal k "
                               f"(corr 0.1, log-var 1), FCG to {args.tol:g} FP64 true relative residual",
                   "path": label, "grid": n, "tol_rel": args.tol,
                   "mg": "V(2,2) Jacobi w=2/3, linear transfers, coarsest 16, coefficient-aware levels",
                   "parallelism": "single GPU"},
        "vark_4096": {
            "mixed_fcg": {"solves_per_s": 1.0 / t_mixed, "iterations": s_mixed.iterations,
                          "true_rel_residual": s_mixed.true_residual / fnorm},
            "plain_fp64_mg_pcg": {"solves_per_s": 1.0 / t_plain, "iterations": s_plain.iterations,
                                  "true_rel_residual": s_plain.true_residual / fnorm},
            "learned_mixed_fcg": learned,
            "speedup_vs_plain": t_plain / t_value,
            "speedup_mixed_vs_plain": t_plain / t_mixed,
        },
        "poisson_4096": poisson,
        "convdiff_2048": convdiff,
        "reductions_4096": reductions,
        "c1_poisson_256": c1,
        "batch_512": batch,
        "cn_vark_1024": cn,
        "explicit_diffusion_4096": explicit,
        "unet_forward": unet_rec,
        "unet_train_step": unet_train,
        "kernels": kernels,
        "hbm_ceilings": ceilings,
        "stencil_axpy_GBs": st_gbs,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": kd["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kd["frac"],
                     "attainable": ceilings[ATTAINABLE_KIND[dom]]["GBs"],
                     "attainable_note": f"measured {ATTAINABLE_KIND[dom]} streaming kernel (fx_hbm_ceilings), same byte mix",
                     "frac_of_attainable": kd["frac_of_attainable"],
                     "traffic": traffic,
                     "traffic_note": traffic_note,
                     "bytes_per_point": kd["bytes_per_point"],
                     "points_per_launch": n * n, "us_per_launch": kd["us_per_launch"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
