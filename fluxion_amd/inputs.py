"""Seeded synthetic inputs for the solve-path workloads.

The reference ships no input generators; these are this project's own and are
fixed here so that the GPU runs, the CPU oracle and the committed fixtures all
see the same data.

PRNG: SplitMix64 (Steele, Lea & Flood 2014), u = (x >> 11)·2⁻⁵³; normals by
Box–Muller.  Instance seeds are 20260101 + instance id.
"""
from __future__ import annotations

import math

import numpy as np

SEED0 = 20260101
_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def normal(self) -> float:
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def node_coords(n_rows: int, n_cols: int):
    """Vertex-centred unit square: x = col·h, y = row·h with h = 1/(n−1)."""
    x = np.arange(n_cols, dtype=np.float64) / (n_cols - 1)
    y = np.arange(n_rows, dtype=np.float64) / (n_rows - 1)
    return x, y


def zero_boundary(f: np.ndarray) -> np.ndarray:
    f[0, :] = 0.0
    f[-1, :] = 0.0
    f[:, 0] = 0.0
    f[:, -1] = 0.0
    return f


def gaussian_bumps(n: int, seed: int = SEED0, nbumps: int = 8, cols: int | None = None) -> np.ndarray:
    """f = Σ_k a_k·exp(−((x−cx_k)² + (y−cy_k)²)/(2w_k²)), a~U(−1,1), c~U(0,1), w~U(0.05,0.2),
    drawn in the order (a, cx, cy, w) per bump; zero on boundary nodes.  Returns (n, cols)."""
    cols = n if cols is None else cols
    rng = SplitMix64(seed)
    x, y = node_coords(n, cols)
    f = np.zeros((n, cols))
    for _ in range(nbumps):
        a = rng.uniform(-1.0, 1.0)
        cx = rng.uniform()
        cy = rng.uniform()
        w = rng.uniform(0.05, 0.2)
        gx = np.exp(-((x - cx) ** 2) / (2.0 * w * w))
        gy = np.exp(-((y - cy) ** 2) / (2.0 * w * w))
        f += a * np.outer(gy, gx)
    return zero_boundary(f)


def streamfunction_velocity(n: int, kappa: float, seed: int = SEED0, modes: int = 16, pe_max: float = 50.0):
    """Divergence-free velocity v = (∂ψ/∂y, −∂ψ/∂x) from
    ψ = Σ_m A_m sin(π k_m x + φ_m) sin(π l_m y + θ_m), k,l∈{1..8}, A~N(0,1), phases U(0,2π),
    scaled so that the largest cell Péclet number |v|·h/κ equals pe_max.  Returns (vx, vy, pe)."""
    rng = SplitMix64(seed)
    x, y = node_coords(n, n)
    X, Y = np.meshgrid(x, y)
    vx = np.zeros((n, n))
    vy = np.zeros((n, n))
    for _ in range(modes):
        k = rng.randint(1, 8)
        l = rng.randint(1, 8)
        A = rng.normal()
        ph = rng.uniform(0.0, 2.0 * math.pi)
        th = rng.uniform(0.0, 2.0 * math.pi)
        sx, cx = np.sin(math.pi * k * X + ph), np.cos(math.pi * k * X + ph)
        sy, cy = np.sin(math.pi * l * Y + th), np.cos(math.pi * l * Y + th)
        vx += A * sx * (math.pi * l) * cy
        vy -= A * (math.pi * k) * cx * sy
    h = 1.0 / (n - 1)
    vmax = float(np.max(np.hypot(vx, vy)))
    s = pe_max * kappa / (h * vmax)
    vx *= s
    vy *= s
    pe = np.hypot(vx, vy) * h / kappa
    return vx, vy, pe


def grf_log_k(n: int, seed: int = SEED0, corr: float = 0.1, modes: int = 256) -> np.ndarray:
    """k = exp(g), g a unit-variance Gaussian field with correlation length `corr`, by random
    Fourier features (Rahimi & Recht 2007): ω ~ N(0, 1/corr²), φ ~ U(0, 2π),
    g = √(2/M)·Σ cos(ω·x + φ)."""
    rng = SplitMix64(seed)
    x, y = node_coords(n, n)
    w = np.empty((modes, 3))
    for m in range(modes):
        w[m, 0] = rng.normal() / corr
        w[m, 1] = rng.normal() / corr
        w[m, 2] = rng.uniform(0.0, 2.0 * math.pi)
    # cos(wx·x + wy·y + ph) = cos(a)cos(b) − sin(a)sin(b), a = wx·x + ph, b = wy·y; the sum over
    # modes is two (n×M)·(M×n) products.
    a = w[:, 0:1] * x[None, :] + w[:, 2:3]
    b = w[:, 1:2] * y[None, :]
    g = np.cos(b).T @ np.cos(a) - np.sin(b).T @ np.sin(a)
    g *= math.sqrt(2.0 / modes)
    return np.exp(g)


def batch_sizes(count: int = 512, lo: int = 64, hi: int = 1024, seed: int = SEED0) -> list[int]:
    """Square grid sides log-uniform in [lo, hi]: n = round(lo·(hi/lo)^U)."""
    rng = SplitMix64(seed)
    return [int(round(lo * (hi / lo) ** rng.uniform())) for _ in range(count)]
