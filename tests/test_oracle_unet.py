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


# ── forward pins ──────────────────────────────────────────────────────
# Decoder layer reading the concatenation (up2(h), skip_l) at level l, and the channel index at which
# skip_l starts in it: h first, then the skip (layers 8, 10, 12).
DEC_FIRST = {2: (8, 64), 1: (10, 48), 0: (12, 32)}


def path_params(cin=2, path=None, k0=None, b0=0.0, b14=0.0):
    """Parameters under which every layer copies channel 0 of its input to channel 0 of its output
    (centre tap 1, bias 0), except: layer 0 applies the 3×3 kernel k0 (taps in row-major (dr, dc)
    order) to input channel 0 with bias b0; the first decoder layer of level l reads channel 0 of
    up2(h) when path[l] == "up" and channel 0 of skip_l when path[l] == "skip"; layer 14 has bias b14.
    With a nonnegative input every ReLU passes its value, so the network is a composition of the
    chosen conv, pools, upsamplings and channel selections that a test can write by hand."""
    path = path or {2: "skip", 1: "skip", 0: "skip"}
    woff, boff, lcin, lcout = O.unet_layout(cin)
    p = np.zeros(O.unet_nparams(cin))

    def w(li, tap, o, i, v):
        p[woff[li] + (tap * lcout[li] + o) * lcin[li] + i] = v

    for li in range(15):
        w(li, 4, 0, 0, 1.0)
    if k0 is not None:
        for t in range(9):
            w(0, t, 0, 0, float(np.ravel(k0)[t]))
    p[boff[0]] = b0
    p[boff[14]] = b14
    for lev, (li, ca) in DEC_FIRST.items():
        if path[lev] == "skip":
            w(li, 4, 0, 0, 0.0)
            w(li, 4, 0, ca, 1.0)
    return p


def img(ch0, cin=2, seed=0):
    """(1, H, W, cin) input with channel 0 = ch0 and the other channels random (ignored by
    path_params networks)."""
    h, w_ = ch0.shape
    x = np.random.default_rng(seed).standard_normal((1, h, w_, cin))
    x[0, :, :, 0] = ch0
    return x


def test_zero_parameters_give_zero_and_output_bias_is_constant():
    cin = 2
    woff, boff, lcin, lcout = O.unet_layout(cin)
    p = np.zeros(O.unet_nparams(cin))
    x = np.random.default_rng(1).standard_normal((1, 8, 8, cin))
    assert np.all(O.unet_forward(p, x) == 0)
    p[boff[14]] = 0.5
    assert np.all(O.unet_forward(p, x) == 0.5)


def test_conv_of_delta_is_flipped_kernel():
    # conv3 is a correlation, y[r][c] = Σ_t K[t]·x[r + dr_t][c + dc_t]: the response to a unit delta at
    # (r0, c0) holds K[(dr, dc)] at (r0 − dr, c0 − dc), the kernel flipped in both axes.
    K = np.arange(1.0, 10.0).reshape(3, 3)
    p = path_params(k0=K)
    for r0, c0 in ((3, 5), (0, 0), (7, 15)):
        d = np.zeros((8, 16))
        d[r0, c0] = 1.0
        y = O.unet_forward(p, img(d))[0]
        want = np.zeros((8, 16))
        for dr in (-1, 0, 1):
            for dc in (-1, 0, 1):
                r, c = r0 - dr, c0 - dc
                if 0 <= r < 8 and 0 <= c < 16:
                    want[r, c] = K[dr + 1, dc + 1]
        assert np.array_equal(y, want), (r0, c0)


def test_conv_zero_padding():
    # A constant input sees the taps that fall inside the image only: corners 4 taps, edges 6,
    # interior 9 (replicate or wrap-around padding would give the full kernel sum everywhere).
    K = np.array([[1.0, 2.0, 4.0], [8.0, 16.0, 32.0], [64.0, 128.0, 256.0]])
    y = O.unet_forward(path_params(k0=K), img(np.ones((8, 16))))[0]
    for r in range(8):
        for c in range(16):
            want = sum(K[dr + 1, dc + 1] for dr in (-1, 0, 1) for dc in (-1, 0, 1)
                       if 0 <= r + dr < 8 and 0 <= c + dc < 16)
            assert y[r, c] == want, (r, c)
    assert y[0, 0] == 16 + 32 + 128 + 256 and y[3, 7] == 511


def test_laplacian_conv_equals_apply_stencil():
    # A one-channel conv with the 5-point Laplacian taps equals the backend's apply_stencil on interior
    # points (unit spacing).  A bias of 100 keeps every ReLU in its linear range; layer 14 removes it.
    from fluxion_amd.core import Boundaries, Grid, Stencil
    H, W = 16, 24
    x0 = np.random.default_rng(2).uniform(0.0, 1.0, (H, W))
    K = np.array([[0.0, 1.0, 0.0], [1.0, -4.0, 1.0], [0.0, 1.0, 0.0]])
    y = O.unet_forward(path_params(k0=K, b0=100.0, b14=-100.0), img(x0))[0]
    ref = O.apply_stencil(x0.ravel(), Grid(H, W, 1.0, 1.0), Stencil.laplacian_2d_5pt(1.0, 1.0),
                          Boundaries.zero_dirichlet()).reshape(H, W)
    assert np.max(np.abs(y[1:-1, 1:-1] - ref[1:-1, 1:-1])) < 1e-12


def block_mean(x, b):
    h, w = x.shape
    m = x.reshape(h // b, b, w // b, b).mean(axis=(1, 3))
    return np.repeat(np.repeat(m, b, axis=0), b, axis=1)


def test_pool_then_upsample_is_block_mean():
    # Level 0 reads up2(h) and level 1 reads skip_1 = pool2(x): y = up2(pool2(x)), the mean of each
    # aligned 2×2 block repeated over it (nearest upsampling r -> r // 2).  A concatenation in the
    # other order (skip first) would read skip_0 = x at channel 0 and give y = x.
    path = {2: "skip", 1: "skip", 0: "up"}
    p = path_params(path=path)
    x0 = np.random.default_rng(3).uniform(0.0, 1.0, (16, 24))
    y = O.unet_forward(p, img(x0))[0]
    assert np.max(np.abs(y - block_mean(x0, 2))) < 1e-15
    assert np.max(np.abs(y - x0)) > 0.05
    assert np.all(O.unet_forward(p, img(np.full((8, 8), 0.75)))[0] == 0.75)
    checker = 1.0 + 2.0 * ((np.arange(8)[:, None] + np.arange(8)[None, :]) % 2)
    assert np.all(O.unet_forward(p, img(checker))[0] == 2.0)


def test_full_depth_path_is_8x8_block_mean():
    # Every decoder level reads up2(h): three pools to the bottom level and three upsamplings back,
    # y = up2³(pool2³(x)), the 8×8 block means.
    p = path_params(path={2: "up", 1: "up", 0: "up"})
    x0 = np.random.default_rng(4).uniform(0.0, 1.0, (16, 32))
    assert np.max(np.abs(O.unet_forward(p, img(x0))[0] - block_mean(x0, 8))) < 1e-14
    # Level 1 reading skip_1 = pool2(x) and level 0 reading its upsampling: the 2×2 block means; level
    # 2 reading up2 of the bottom level, then level 1 its skip... the middle choice is what selects
    # the block size.
    p = path_params(path={2: "up", 1: "skip", 0: "up"})
    assert np.max(np.abs(O.unet_forward(p, img(x0))[0] - block_mean(x0, 2))) < 1e-15
    p = path_params(path={2: "skip", 1: "up", 0: "up"})
    assert np.max(np.abs(O.unet_forward(p, img(x0))[0] - block_mean(x0, 4))) < 1e-15


def test_bf16_mode_rounds_to_nearest_even():
    # BF16 keeps 8 significant bits: 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7 and rounds to the even
    # neighbour 1; 1 + 3·2^-8 rounds up to 1 + 2^-6.  The identity path stores the input and every
    # activation in BF16, so the output shows the rounding; FP64 mode passes the values through.
    p = path_params()
    for v, want in ((1.0 + 2.0 ** -8, 1.0), (1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -6), (1.0 + 2.0 ** -7, 1.0 + 2.0 ** -7)):
        x = img(np.full((8, 8), v))
        assert np.all(O.unet_forward(p, x, bf16=True)[0] == want), v
        assert np.all(O.unet_forward(p, x)[0] == v), v
    # Pooled maps are rounded too: the 2×2 mean of {1, 1, 1, 1 + 2^-6} is 1 + 2^-8 -> 1.
    x0 = np.ones((8, 8))
    x0[0::2, 0::2] = 1.0 + 2.0 ** -6
    y = O.unet_forward(path_params(path={2: "skip", 1: "skip", 0: "up"}), img(x0), bf16=True)[0]
    assert np.all(y == 1.0)


def test_bf16_mode_close_to_fp64():
    # Random network: BF16 storage moves the output by about the BF16 precision, not more.
    p = init_params(2, seed=11)
    x = np.random.default_rng(12).standard_normal((2, 16, 16, 2))
    y64 = O.unet_forward(p, x)
    y16 = O.unet_forward(p, x, bf16=True)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < rel < 5e-2, rel


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
