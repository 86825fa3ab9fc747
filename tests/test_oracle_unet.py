"""The FP64 U-Net oracle: shapes, linearity pieces and gradients by central finite differences (CPU)."""
import numpy as np

import oracle as O


def init_params(cin=2, seed=0):
    woff, boff, lcin, lcout = O.unet_layout(cin)
    p = np.zeros(O.unet_nparams(cin))
    rng = np.random.default_rng(seed)
    for i in range(15):
        fan_in = 9 * lcin[i]
        n = 9 * lcin[i] * lcout[i]
        p[woff[i]:woff[i] + n] = rng.standard_normal(n) * np.sqrt(2.0 / fan_in)
        p[boff[i]:boff[i] + lcout[i]] = 0.01 * rng.standard_normal(lcout[i])
    return p


def test_param_count():
    # Σ 9·cin·cout + cout over the 15 layers of the 16/32/48/64 U-Net with 2 input channels.
    layers = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64),
              (112, 48), (48, 48), (80, 32), (32, 32), (48, 16), (16, 16), (16, 1)]
    assert O.unet_nparams(2) == sum(9 * a * b + b for a, b in layers)


def test_conv_of_delta_is_kernel():
    # With all weights zero except the output layer reading channel 0 of a 1-channel identity path,
    # the response to a delta input is the 3×3 kernel (flipped by the correlation convention).
    cin = 2
    woff, boff, lcin, lcout = O.unet_layout(cin)
    p = np.zeros(O.unet_nparams(cin))
    x = np.zeros((1, 8, 8, cin))
    y0 = O.unet_forward(p, x)
    assert np.all(y0 == 0)
    p[boff[14]] = 0.5
    assert np.allclose(O.unet_forward(p, x), 0.5)


def test_gradient_central_differences():
    cin = 2
    p = init_params(cin, seed=1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 8, 8, cin))
    t = rng.standard_normal((2, 8, 8))
    loss, g = O.unet_loss_grad(p, x, t)
    woff, boff, lcin, lcout = O.unet_layout(cin)
    idx = []
    for i in range(15):
        idx += list(rng.integers(woff[i], woff[i] + 9 * lcin[i] * lcout[i], 2))
        idx.append(boff[i])
    eps = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += eps
        pm[j] -= eps
        lp, _ = O.unet_loss_grad(pp, x, t)
        lm, _ = O.unet_loss_grad(pm, x, t)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[j]) <= 1e-6 * (abs(fd) + abs(g[j])) + 1e-9, (j, fd, g[j])


def test_energy_gradient_central_differences():
    # Relative energy-norm loss under the unit-spacing −∇·(k∇·) with k = exp(x[..., 1]).
    cin = 2
    p = init_params(cin, seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 8, 8, cin))
    t = rng.standard_normal((2, 8, 8))
    loss, g = O.unet_loss_grad(p, x, t, energy=True)
    woff, boff, lcin, lcout = O.unet_layout(cin)
    idx = []
    for i in range(15):
        idx += list(rng.integers(woff[i], woff[i] + 9 * lcin[i] * lcout[i], 2))
        idx.append(boff[i])
    eps = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += eps
        pm[j] -= eps
        lp, _ = O.unet_loss_grad(pp, x, t, energy=True)
        lm, _ = O.unet_loss_grad(pm, x, t, energy=True)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[j]) <= 1e-6 * (abs(fd) + abs(g[j])) + 1e-9, (j, fd, g[j])


def test_energy_loss_closed_form():
    # Zero network output (all-zero parameters): dᵀAd = tᵀAt, so the loss is exactly the batch size.
    cin = 2
    p = np.zeros(O.unet_nparams(cin))
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 8, 16, cin))
    t = rng.standard_normal((3, 8, 16))
    loss, _ = O.unet_loss_grad(p, x, t, energy=True)
    assert abs(loss - 3.0) < 1e-12
    # y ≡ 1 (output bias only), k ≡ 1, t a unit impulse at interior point (3, 3) of an 8×8 image:
    # d = 1 − t on the 6×6 interior.  dᵀAd = 4 faces around the impulse with Δd = 1, plus the 24
    # faces from the interior ring to the boundary with d = 1, = 28; tᵀAt = 4.  Loss 7.
    woff, boff, lcin, lcout = O.unet_layout(cin)
    p[boff[14]] = 1.0
    x = np.zeros((1, 8, 8, cin))
    t = np.zeros((1, 8, 8))
    t[0, 3, 3] = 1.0
    loss, _ = O.unet_loss_grad(p, x, t, energy=True)
    assert abs(loss - 7.0) < 1e-12


def test_sizes_not_multiple_of_8_rejected():
    import pytest
    p = np.zeros(O.unet_nparams(2))
    with pytest.raises(ValueError):
        O.unet_forward(p, np.zeros((1, 12, 16, 2)))
