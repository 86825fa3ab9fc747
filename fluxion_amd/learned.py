"""Learned preconditioner: the U-Net on the GPU, its training on solver-generated
residual→error pairs, and its use inside flexible CG (extension; no reference
counterpart).

All tensor work runs in libfluxion_hip.so (fx_unet_* and the solvers); this
module only moves small host arrays and sequences the calls.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib as L
from .core import DType


class UNet:
    """4-level U-Net, 16/32/48/64 channels, BF16/MFMA (fx_unet_* in include/fluxion_hip.h)."""

    def __init__(self, be, cin: int = 2):
        self.be = be
        self._lib = be._lib
        self.cin = cin
        h = C.c_void_p()
        L.check(self._lib.fx_unet_create(be.h, cin, C.byref(h)), "unet_create")
        self.h = h
        self.nparams = int(self._lib.fx_unet_nparams(h))
        self._bufs = {}

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_unet_destroy(h)
            self.h = None

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.nparams
        L.check(self._lib.fx_unet_set_params(self.h, a.ctypes.data, a.size), "unet_set_params")

    def params(self) -> np.ndarray:
        a = np.empty(self.nparams, dtype=np.float32)
        L.check(self._lib.fx_unet_get_params(self.h, a.ctypes.data, a.size), "unet_get_params")
        return a

    def grad(self) -> np.ndarray:
        a = np.empty(self.nparams, dtype=np.float32)
        L.check(self._lib.fx_unet_get_grad(self.h, a.ctypes.data, a.size), "unet_get_grad")
        return a

    def _buf(self, key, n):
        b = self._bufs.get(key)
        if b is None or len(b) != n:
            b = self.be.allocate(n, DType.F32)
            self._bufs[key] = b
        return b

    def _upload(self, key, a):
        a = np.ascontiguousarray(a, dtype=np.float32).ravel()
        b = self.be.upload(a)
        self._bufs[key] = b
        return b

    def forward(self, x) -> np.ndarray:
        """x: (B, H, W, cin) -> (B, H, W)."""
        B, H, W, cin = x.shape
        bx = self._upload("x", x)
        by = self._buf("y", B * H * W)
        L.check(self._lib.fx_unet_forward(self.h, bx.h, by.h, B, H, W), "unet_forward")
        return by.to_numpy(np.float32).astype(np.float64).reshape(B, H, W)

    def forward_device(self, bx, by, B, H, W):
        L.check(self._lib.fx_unet_forward(self.h, bx.h, by.h, B, H, W), "unet_forward")

    def loss_grad(self, x, t, energy: bool = False):
        """Σ_b ‖y_b − t_b‖²/‖t_b‖² (or the relative energy norm, fx_unet_loss_grad_energy) and its
        gradient (host copy)."""
        B, H, W, cin = x.shape
        bx = self._upload("x", x)
        bt = self._upload("t", t)
        return self.loss_grad_device(bx, bt, B, H, W, energy), self.grad().astype(np.float64)

    def loss_grad_device(self, bx, bt, B, H, W, energy: bool = False) -> float:
        out = C.c_double()
        fn = self._lib.fx_unet_loss_grad_energy if energy else self._lib.fx_unet_loss_grad
        L.check(fn(self.h, bx.h, bt.h, B, H, W, C.byref(out)), "unet_loss_grad")
        return out.value

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L.check(self._lib.fx_unet_adam_step(self.h, lr, b1, b2, eps), "unet_adam_step")


def he_init(cin: int = 2, seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, in the fx_unet parameter layout."""
    from . import _unet_layout
    woff, boff, lcin, lcout = _unet_layout.layout(cin)
    n = boff[-1] + lcout[-1]
    p = np.zeros(n)
    rng = np.random.default_rng(seed)
    for i in range(15):
        k = 9 * lcin[i] * lcout[i]
        p[woff[i]:woff[i] + k] = rng.standard_normal(k) * np.sqrt(2.0 / (9 * lcin[i]))
    # Small output layer so the initial correction is small.
    p[woff[14]:woff[14] + 9 * lcin[14]] *= 0.1
    return p


class LearnedPreconditioner:
    """FX_PC_LEARNED: z = V(r) + α*·s·h²·N([r1/s, log k]) with r1 = r − A·V(r), s = rms(r1), α* the
    A-norm-optimal step; falls back to V(r) once the FP64 true residual stalls.

    `mg` is the V-cycle (also the fallback), `logk` an FP64 field of log k (None for k ≡ 1)."""

    def __init__(self, net: UNet, mg, grid, logk=None, stall: float = 0.9):
        self.net, self.mg, self.logk = net, mg, logk
        self.be = net.be
        self._lib = self.be._lib
        g = L.grid_c(grid)
        h = C.c_void_p()
        L.check(self._lib.fx_learned_create(self.be.h, net.h, mg.h, logk.h if logk is not None else None,
                                            C.byref(g), stall, C.byref(h)), "learned_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h and L.live():
            self._lib.fx_learned_destroy(h)
            self.h = None

    def fell_back(self) -> bool:
        return bool(self._lib.fx_learned_fell_back(self.h))

    def set_schedule(self, every: int = 1, max_apps: int = 0, first: int = 0) -> None:
        """Network from iteration `first` on, at every `every`-th iteration, at most `max_apps` times
        per solve (0: no limit)."""
        L.check(self._lib.fx_learned_set_schedule(self.h, every, max_apps, first), "learned_set_schedule")

    def _pc(self):
        return L.FxPrecond(2, 0, self.h)


@dataclass
class PairSet:
    x: np.ndarray   # (P, H, W, 2) float32: (r1/s, log k)
    t: np.ndarray   # (P, H, W) float32: e1/(s·h²)


def solver_pairs(be, n: int, seed: int, max_its: int = 12, stride: int = 1) -> PairSet:
    """Residual→error pairs from the solver's own iterates on one −∇·(k∇u) = f problem.

    x* is solved to 1e-12 with FCG and the coefficient-aware FP32 V-cycle; the iterate x_i of the
    same solve after i steps gives r = f − A·x_i and e = x* − x_i.  After one V-cycle z = V(r) the
    pair is (r1 = r − A·z, e1 = e − z), normalised by s = rms(r1) and h²."""
    from .core import Boundaries, DType, Grid, Transfer
    from .hip import FcgSolver, Field, Multigrid, Operator
    from .inputs import gaussian_bumps, grf_log_k

    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    k = grf_log_k(n, seed=seed)
    f = gaussian_bumps(n, seed=seed + 7919).ravel()
    K = Field.from_f64(g, k.ravel(), be)
    op = Operator.var_k(K)
    mg = Multigrid(g, Boundaries.zero_dirichlet(), 2, 2, 2.0 / 3.0, transfer=Transfer.FW_BILINEAR, coarsest=16,
                   precision=DType.F32, fused=True, scale=1.0, k=K, be=be)
    F = Field.from_f64(g, f, be)
    fn = float(np.linalg.norm(f))
    X = Field.zeros(g, be=be)
    FcgSolver(g, 1e-12 * fn, 400, be=be).solve_op(X, F, op, mg)
    xstar = X.to_vec_f64()
    logk = np.log(k).astype(np.float32)
    Ax, Z, R = Field.zeros(g, be=be), Field.zeros(g, be=be), Field.zeros(g, be=be)
    xs, ts = [], []
    for it in range(1, max_its + 1, stride):
        X.fill(0.0)
        FcgSolver(g, 1e-30, it, be=be).solve_op(X, F, op, mg)
        op.apply(X, Ax)
        R.copy_from(F)
        R.axpy(-1.0, Ax)
        Z.fill(0.0)
        mg.v_cycle(Z, R)
        op.apply(Z, Ax)
        r1 = R.to_vec_f64() - Ax.to_vec_f64()
        e1 = xstar - X.to_vec_f64() - Z.to_vec_f64()
        s = float(np.sqrt(np.mean(r1 * r1)))
        if not np.isfinite(s) or s == 0.0:
            continue
        xs.append(np.stack([(r1 / s).reshape(n, n).astype(np.float32), logk], axis=-1))
        ts.append((e1 / (s * h * h)).reshape(n, n).astype(np.float32))
    return PairSet(np.stack(xs), np.stack(ts))


def train(net: UNet, sets, steps: int = 400, batch: int = 8, lr: float = 1e-3, seed: int = 0, log=None,
          energy: bool = False):
    """Adam on the relative squared error (or relative energy norm) over minibatches drawn from
    same-size pair sets, with a cosine learning-rate schedule."""
    rng = np.random.default_rng(seed)
    losses = []
    for step in range(steps):
        ps = sets[rng.integers(len(sets))]
        idx = rng.choice(len(ps.x), size=min(batch, len(ps.x)), replace=False)
        x, t = ps.x[idx], ps.t[idx]
        lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps))
        B, H, W, _ = x.shape
        loss = net.loss_grad_device(net._upload("x", x), net._upload("t", t), B, H, W, energy)
        net.adam_step(lr_t)
        losses.append(loss / len(idx))
        if log and (step % 50 == 0 or step == steps - 1):
            log(f"train step {step}: loss {losses[-1]:.4f}")
    return losses


# Weights trained by train_vark() on the log-normal-k family (written by `python -m fluxion_amd.learned`).
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_vark.npz"


def train_vark(be, steps: int = 1500, sizes=((256, 16), (512, 6)), max_its: int = 10, batch: int = 8,
               lr: float = 1e-3, seed0: int = 31000, log=None):
    """Trains the network on residual→error pairs that the mixed FCG solver generates on the log-normal-k
    diffusion family: for each (n, count) in `sizes`, `count` problems at n² with seeds seed0 + 1000·n + i,
    pairs from the first `max_its` FCG iterates (solver_pairs); relative energy-norm loss, Adam with a
    cosine schedule.  Returns (net, losses)."""
    sets = []
    for n, count in sizes:
        ps = [solver_pairs(be, n, seed=seed0 + 1000 * n + i, max_its=max_its) for i in range(count)]
        sets.append(PairSet(np.concatenate([p.x for p in ps]), np.concatenate([p.t for p in ps])))
        if log:
            log(f"pairs at {n}^2: {len(sets[-1].x)}")
    net = UNet(be, 2)
    net.set_params(he_init(2, 0))
    losses = train(net, sets, steps=steps, batch=batch, lr=lr, energy=True, log=log)
    return net, losses


def save_weights(path, params, meta: dict) -> None:
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(p, params=np.asarray(params, dtype=np.float32),
                        meta=np.array(repr(sorted(meta.items()))))


def load_weights(path=DEFAULT_WEIGHTS) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["params"], dtype=np.float32)


if __name__ == "__main__":
    import argparse
    import sys
    import time

    from .hip import HipBackend

    ap = argparse.ArgumentParser(description="train the learned preconditioner's network on the GPU")
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    ap.add_argument("--steps", type=int, default=1500)
    a = ap.parse_args()
    be = HipBackend(0)
    t0 = time.time()
    net, losses = train_vark(be, steps=a.steps, log=lambda m: print(m, file=sys.stderr, flush=True))
    meta = {"steps": a.steps, "sizes": "256^2 x16, 512^2 x6", "loss": "relative energy norm",
            "final_loss": float(np.mean(losses[-50:])), "seconds": time.time() - t0}
    save_weights(a.out, net.params(), meta)
    print(f"saved {a.out}: {meta}", flush=True)
