"""Benchmark of the structured-grid solve path on one MI355X.

Workload (one step = one solve): 2D Poisson −Δu = f on a 4096×4096 vertex grid
(h = 1/4095, zero Dirichlet), f = 8 seeded Gaussian bumps, solved from x = 0
to ‖b − A·x‖₂ ≤ 1e-10·‖f‖₂ (true FP64 residual).  Operator in the reference
form A·v = v + 1·S(v), S = −Δ_h − I.  Multigrid V-cycle: 2+2 weighted Jacobi
(ω = 2/3), linear transfers between levels spanning the domain, coarsest 16×16.

  value       = solves/s of the mixed-precision path: FP32 V-cycle inside FP64
                flexible CG with FP64 true-residual refinement.
  plain       = FP64 MG-preconditioned CG (CgSolver::solve_preconditioned).
  roofline    = the FP64 5-point fused A·p kernel of the CG step, timed with HIP
                events on the library's stream; algorithmic bytes 32 B/pt.
  cpu_baseline= the CPU oracle (C++ restatement of the reference CPU backend and
                solvers) running the plain FP64 MG-PCG solve of the same problem
                on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


class HipEvents:
    """Minimal hipEvent timing on a given stream (libamdhip64)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    from fluxion_amd import hip
    from fluxion_amd.core import Boundaries, DType, Grid, Stencil, Transfer
    from fluxion_amd.inputs import gaussian_bumps

    n = args.size
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    bc = Boundaries.zero_dirichlet()
    st = Stencil.poisson_minus_identity(h, h)
    coeff = 1.0
    f = gaussian_bumps(n).ravel()
    fnorm = float(np.linalg.norm(f))
    tol = 1e-10 * fnorm

    be = hip.HipBackend(0)
    log(f"bench: backend {be.name}, grid {n}x{n}")
    B = hip.Field.from_f64(g, f, be)
    X = hip.Field.zeros(g, be=be)

    def mg(precision, fused):
        return hip.Multigrid(g, bc, 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16,
                             precision=precision, fused=fused, scale=-1.0, be=be)

    mg64, mg32 = mg(DType.F64, False), mg(DType.F32, True)
    max_iter = 200
    pcg = hip.CgSolver(g, tol, max_iter, be=be)
    fcg = hip.FcgSolver(g, tol, max_iter, be=be)

    def solve_plain():
        X.fill(0.0)
        return pcg.solve_preconditioned(X, B, coeff, st, bc, mg64)

    def solve_mixed():
        X.fill(0.0)
        return fcg.solve_preconditioned(X, B, coeff, st, bc, mg32)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            s = fn()
        be.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s = fn()
        be.synchronize()
        return (time.perf_counter() - t0) / steps, s

    t_plain, s_plain = timed(solve_plain, max(1, args.steps // 2), max(1, args.warmup // 2))
    x_plain = X.to_vec_f64()
    log(f"bench: plain FP64 MG-PCG {s_plain.iterations} its, {1e3 * t_plain:.2f} ms/solve, "
        f"true rel res {s_plain.true_residual / fnorm:.2e}")
    t_mixed, s_mixed = timed(solve_mixed, args.steps, args.warmup)
    x_mixed = X.to_vec_f64()
    log(f"bench: mixed FCG {s_mixed.iterations} its, {1e3 * t_mixed:.2f} ms/solve, "
        f"true rel res {s_mixed.true_residual / fnorm:.2e}")
    assert s_plain.converged() and s_mixed.converged()
    assert s_mixed.true_residual <= tol and s_plain.true_residual <= tol * 1.0001

    # Roofline of the dominant FP64 kernel: the fused A·p step is the same traversal as
    # stencil_axpy; time stencil_axpy launches with HIP events on the library stream.
    ev = HipEvents()
    Y = hip.Field.zeros(g, be=be)
    lap = Stencil.laplacian_2d_5pt(h, h)
    reps = 50
    for _ in range(5):
        Y.stencil_axpy(1e-9, X, lap, bc)
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, be.stream)
    for _ in range(reps):
        Y.stencil_axpy(1e-9, X, lap, bc)
    ev.record(e1, be.stream)
    ms = ev.elapsed_ms(e0, e1) / reps
    alg_bytes = 24.0 * n * n
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    log(f"bench: stencil_axpy {ms * 1e3:.1f} us/launch, {achieved:.0f} GB/s algorithmic")

    cpu = None
    if not args.no_cpu:
        import oracle as O
        cfg = O.mg_cfg(2, 2, 2.0 / 3.0, smoother=O.SMOOTH_JACOBI, transfer=1, coarsest=16, scale=-1.0)
        t0 = time.perf_counter()
        xr, sr = O.pcg(np.zeros_like(f), f, g, O.Op.stencil(coeff, st), tol, max_iter, mg=cfg)
        t_cpu = time.perf_counter() - t0
        cores = O.num_threads()
        diff = float(np.linalg.norm(x_mixed - xr) / np.linalg.norm(xr))
        diff_plain = float(np.linalg.norm(x_plain - xr) / np.linalg.norm(xr))
        log(f"bench: CPU oracle plain MG-PCG {sr.iterations} its, {t_cpu:.1f} s on {cores} threads; "
            f"GPU mixed vs CPU rel diff {diff:.2e}, plain {diff_plain:.2e}")
        cpu = {"value": 1.0 / t_cpu, "unit": "solves/s", "cores": cores, "kind": "port",
               "sample": f"one full plain FP64 MG-PCG solve of the same {n}x{n} problem ({sr.iterations} its)",
               "iterations": sr.iterations, "gpu_mixed_rel_diff": diff, "gpu_plain_rel_diff": diff_plain}

    out = {
        "metric": "solves/s to 1e-10 rel residual (mixed FP32-MG/FP64 FCG), 4096^2 Poisson; stencil HBM GB/s",
        "value": 1.0 / t_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 V-cycle",
        "data": "synthetic (8 seeded Gaussian bumps)",
        "config": {"workload": f"2D Poisson {n}x{n} MG-preconditioned CG to 1e-10 rel residual",
                   "grid": n, "tol_rel": 1e-10, "mg": "V(2,2) Jacobi w=2/3, linear transfers, coarsest 16",
                   "parallelism": "single GPU"},
        "mixed_fcg": {"solves_per_s": 1.0 / t_mixed, "iterations": s_mixed.iterations,
                      "true_rel_residual": s_mixed.true_residual / fnorm},
        "plain_fp64_mg_pcg": {"solves_per_s": 1.0 / t_plain, "iterations": s_plain.iterations,
                              "true_rel_residual": s_plain.true_residual / fnorm},
        "speedup_mixed_vs_plain": t_plain / t_mixed,
        "stencil_axpy_GBs": achieved,
        "roofline": {"kernel": "k_stream5 (stencil_axpy, FP64 5-point)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": alg_bytes, "us_per_launch": ms * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
