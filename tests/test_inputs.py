"""Properties of the seeded input generators (fluxion_amd/inputs.py).

The reference ships no generators, so these pin the stated recipes by what
they must satisfy: the published SplitMix64 outputs, a divergence-free
velocity whose largest cell Péclet number is the requested one, a Gaussian
random field with unit variance and the Gaussian correlation of length 0.1,
and log-uniform grid sides.
"""
import math

import numpy as np

from fluxion_amd.inputs import (SEED0, SplitMix64, batch_sizes, gaussian_bumps, grf_log_k,
                                streamfunction_velocity)


def test_splitmix64_published_outputs():
    # First outputs of SplitMix64 from state 0 (Steele, Lea & Flood 2014; the reference generator of
    # java.util.SplittableRandom and of the xoshiro seeding code).
    r = SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = SplitMix64(1).uniform()
    assert 0.0 <= u < 1.0 and u * 2 ** 53 == int(u * 2 ** 53)   # 53-bit grid


def test_gaussian_bumps_zero_boundary_and_deterministic():
    f = gaussian_bumps(65)
    assert f.shape == (65, 65)
    assert not f[0].any() and not f[-1].any() and not f[:, 0].any() and not f[:, -1].any()
    assert np.array_equal(f, gaussian_bumps(65)) and not np.array_equal(f, gaussian_bumps(65, seed=SEED0 + 1))
    # Each bump is at most |a| ≤ 1 high, so |f| ≤ 8.
    assert np.abs(f).max() <= 8.0 and np.abs(f).max() > 0.1


def _divergence_rel(n):
    vx, vy, _ = streamfunction_velocity(n, 1e-3)
    h = 1.0 / (n - 1)
    dvx = (vx[1:-1, 2:] - vx[1:-1, :-2]) / (2 * h)
    dvy = (vy[2:, 1:-1] - vy[:-2, 1:-1]) / (2 * h)
    return np.abs(dvx + dvy).max() / np.abs(dvx).max()


def test_streamfunction_velocity_is_divergence_free():
    # v = (∂ψ/∂y, −∂ψ/∂x) has zero divergence; its central-difference divergence is the O(h²) truncation
    # error alone: small, and four times smaller on a grid twice as fine.
    a, b = _divergence_rel(129), _divergence_rel(257)
    assert a < 5e-3 and b < 1.5e-3
    assert 3.5 < a / b < 4.5


def test_streamfunction_peclet_range():
    n, kappa = 257, 1e-3
    vx, vy, pe = streamfunction_velocity(n, kappa)
    h = 1.0 / (n - 1)
    assert np.allclose(pe, np.hypot(vx, vy) * h / kappa)
    assert abs(pe.max() - 50.0) < 1e-12
    # Spread: almost every node above Pe 0.1, the bulk convection-dominated.
    assert (pe >= 0.1).mean() > 0.99
    assert 5.0 < np.median(pe) < 40.0
    # The scale is set by kappa: a ten times larger kappa gives the same Péclet field.
    _, _, pe10 = streamfunction_velocity(n, 10 * kappa)
    assert np.allclose(pe10, pe)


def test_grf_unit_variance_and_correlation_length():
    # g = log k over 200 seeds: ensemble mean 0, variance 1, and the correlation of random Fourier
    # features with ω ~ N(0, 1/0.1²), exp(−|d|²/(2·0.1²)): 0.607 at |d| = 0.1, 0.368 at the diagonal lag.
    n, lag = 41, 4   # h = 1/40, so 4 cells = 0.1
    G = np.stack([np.log(grf_log_k(n, seed=SEED0 + s)) for s in range(200)])
    var = (G ** 2).mean()
    assert abs(G.mean()) < 0.05 and abs(var - 1.0) < 0.06
    c1 = (G[:, :, :-lag] * G[:, :, lag:]).mean() / var
    c2 = (G[:, :-lag, :-lag] * G[:, lag:, lag:]).mean() / var
    assert abs(c1 - math.exp(-0.5)) < 0.03
    assert abs(c2 - math.exp(-1.0)) < 0.03
    assert np.all(grf_log_k(n) > 0)


def test_batch_sizes_log_uniform():
    s = np.array(batch_sizes(512))
    assert s.min() >= 64 and s.max() <= 1024
    u = np.log(s / 64.0) / np.log(16.0)
    # U(0, 1): mean 1/2, variance 1/12 (standard errors 0.013 and 0.005 for 512 draws).
    assert abs(u.mean() - 0.5) < 0.04 and abs(u.var() - 1.0 / 12.0) < 0.015
    assert batch_sizes(512) == list(s)
