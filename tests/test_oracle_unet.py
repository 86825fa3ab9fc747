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
