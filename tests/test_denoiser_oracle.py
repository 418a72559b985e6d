"""CPU: pins the denoiser oracle (oracle/unet.py) before it checks the GPU.

* convolution: against a brute-force loop (correlation, zero padding, layout);
* pooling / upsampling / convolution backward: adjoint identities;
* whole-network backward: central finite differences in float64;
* BF16 rounding: against exhaustive properties (round to nearest even);
* input encoding: octahedral decode of the encoded normal.
"""
import numpy as np
import pytest

from oracle import unet


def brute_conv(x, w, b):
    N, H, W, C = x.shape
    out = np.zeros((N, H, W, w.shape[0]))
    for n in range(N):
        for i in range(H):
            for j in range(W):
                for ky in range(3):
                    for kx in range(3):
                        y, xx = i + ky - 1, j + kx - 1
                        if 0 <= y < H and 0 <= xx < W:
                            out[n, i, j] += w[:, ky, kx, :] @ x[n, y, xx]
                out[n, i, j] += b
    return out


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    np.testing.assert_allclose(unet.conv3x3(x, w, b), brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_backward_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 6, 4, 5))
    w = rng.normal(size=(3, 3, 3, 5))
    g = rng.normal(size=(1, 6, 4, 3))
    dx, dw, db = unet.conv3x3_backward(x, w, g)
    lin = unet.conv3x3(x, w, np.zeros(3))
    # <A x, g> = <x, A^T g> and the weight gradient is <A_w, g> per weight
    assert np.isclose((lin * g).sum(), (x * dx).sum())
    assert np.isclose((lin * g).sum(), (w * dw).sum())
    np.testing.assert_allclose(db, g.sum(axis=(0, 1, 2)))


def test_pool_and_upsample_adjoint():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 8, 6, 3))
    g = rng.normal(size=(2, 4, 3, 3))
    assert np.isclose((unet.pool2(x) * g).sum(), (x * unet.pool2_backward(g)).sum())
    y = rng.normal(size=(2, 4, 3, 3))
    h = rng.normal(size=(2, 8, 6, 3))
    assert np.isclose((unet.up2(y) * h).sum(), (y * unet.up2_backward(h)).sum())


def test_parameter_count_and_macs():
    # 0.48 M parameters; 84,624 multiply-adds per full-resolution pixel
    assert unet.num_params() == 480675
    macs = 0.0
    scale = {"0": 1.0, "1": 0.25, "2": 1.0 / 16}
    for name, ws, _ in unet.param_shapes():
        if name == "head":
            macs += np.prod(ws)
            continue
        lvl = 1.0 / 64 if name.startswith("mid") else scale[name[3]]
        macs += np.prod(ws) * lvl
    assert round(macs) == 84624


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(4)
    params = unet.he_init(7)
    for name in params:   # non-zero biases so every ReLU pattern is generic
        w, b = params[name]
        params[name] = (w, rng.normal(0.0, 0.05, b.shape))
    x = rng.normal(size=(1, 8, 8, unet.C_IN))
    t = rng.normal(size=(1, 8, 8, 3))
    out, cache = unet.forward(params, x)
    loss, dout = unet.loss_mse(out, t)
    grads = unet.backward(params, cache, dout)
    flat = unet.flatten(params)
    gflat = unet.flatten(grads)
    idx = rng.choice(flat.size, 40, replace=False)
    # include the head and one bias of every layer
    k = 0
    for name, ws, bs in unet.param_shapes():
        k += int(np.prod(ws))
        idx = np.append(idx, k)
        k += int(np.prod(bs))
    eps = 1e-6
    for i in idx:
        fp = flat.copy(); fp[i] += eps
        fm = flat.copy(); fm[i] -= eps
        lp, _ = unet.loss_mse(unet.forward(unet.unflatten(fp), x)[0], t)
        lm, _ = unet.loss_mse(unet.forward(unet.unflatten(fm), x)[0], t)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - gflat[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, gflat[i])


def test_bf16_rounding():
    # exactly representable values are unchanged
    v = np.array([0.0, 1.0, -2.5, 0.15625, 3.0e38, 1.0e-38])
    np.testing.assert_array_equal(unet.bf16(unet.bf16(v)), unet.bf16(v))
    assert unet.bf16(1.0) == 1.0
    # halfway cases round to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert unet.bf16(1.0 + 2.0 ** -8) == 1.0
    assert unet.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    # unit roundoff of an 8-bit significand: 2^-8
    x = np.random.default_rng(5).normal(size=10000)
    assert np.max(np.abs(unet.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_octahedral_round_trip():
    rng = np.random.default_rng(6)
    n = rng.normal(size=(1000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    e = unet.octahedral(n)
    assert np.abs(e).max() <= 1.0 + 1e-12
    # decode (textbook octahedral mapping)
    z = 1.0 - np.abs(e[:, 0]) - np.abs(e[:, 1])
    x, y = e[:, 0].copy(), e[:, 1].copy()
    neg = z < 0
    x[neg] = (1.0 - np.abs(e[neg, 1])) * np.where(e[neg, 0] >= 0, 1.0, -1.0)
    y[neg] = (1.0 - np.abs(e[neg, 0])) * np.where(e[neg, 1] >= 0, 1.0, -1.0)
    d = np.stack([x, y, z], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    np.testing.assert_allclose(d, n, atol=1e-12)


def test_pack_inputs_channels():
    f = np.zeros((1, 1, 2, unet.N_FEAT))
    f[0, 0, 0, :3] = [0.5, 1.0, 2.0]
    f[0, 0, 0, 3:6] = [0.1, 0.2, 0.3]
    f[0, 0, 0, 6] = 0.25
    f[0, 0, 0, 7:10] = [0.0, 0.0, -2.0]
    f[0, 0, 0, 10] = 6.0
    f[0, 0, 0, 11] = 4.0
    p = unet.pack_inputs(f)
    np.testing.assert_allclose(p[0, 0, 0, :3], np.log1p([0.5, 1.0, 2.0]))
    np.testing.assert_allclose(p[0, 0, 0, 3:6], [0.1, 0.2, 0.3])
    np.testing.assert_allclose(p[0, 0, 0, 6:8], [1.0, 1.0])      # -z maps to the corner
    assert p[0, 0, 0, 8] == 0.25
    assert np.isclose(p[0, 0, 0, 9], np.sqrt(6.0 / 12.0))
    assert p[0, 0, 1, 9] == 0.0                                  # n < 2: no variance


@pytest.mark.parametrize("emulate", [False, True])
def test_forward_shapes_and_residual(emulate):
    params = unet.he_init(1)
    for name in params:   # zero network: output = input log radiance
        w, b = params[name]
        params[name] = (np.zeros_like(w), np.zeros_like(b))
    x = np.random.default_rng(8).uniform(size=(2, 16, 8, unet.C_IN))
    out, _ = unet.forward(params, x, emulate_bf16=emulate)
    assert out.shape == (2, 16, 8, 3)
    np.testing.assert_allclose(out, x[..., :3], rtol=1e-6)


def test_relmse_loss_gradient_finite_differences():
    rng = np.random.default_rng(9)
    out = rng.normal(0.0, 0.5, size=(1, 2, 3, 3))
    t = rng.gamma(1.0, 0.3, size=(1, 2, 3, 3))
    _, g = unet.loss_relmse(out, t)
    eps = 1e-6
    for i in range(out.size):
        o1, o2 = out.copy().ravel(), out.copy().ravel()
        o1[i] += eps
        o2[i] -= eps
        fd = (unet.loss_relmse(o1.reshape(out.shape), t)[0] - unet.loss_relmse(o2.reshape(out.shape), t)[0]) / (2 * eps)
        assert abs(fd - g.ravel()[i]) <= 1e-7 + 1e-5 * abs(fd)
    # the value is the bench's relMSE of the linear prediction
    y = np.expm1(out)
    assert np.isclose(unet.loss_relmse(out, t)[0], np.mean((y - t) ** 2 / (t * t + 0.01)))
