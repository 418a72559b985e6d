"""GPU: the denoiser's BF16 MFMA kernels against the numpy oracle.

The oracle (oracle/unet.py, pinned by finite differences on the CPU) emulates
BF16 rounding where the GPU stores BF16, so the remaining differences are FP32
accumulation order and the rare BF16 rounding flip it causes. Bars:
  * forward: relative L2 error of the log-space output <= 5e-3 (BF16's unit
    roundoff is 2^-8 = 3.9e-3: a different FP32 accumulation order can flip
    an intermediate rounding, which later layers carry; a wrong tap, channel
    or layout gives errors of order 1);
  * gradients: per parameter tensor, relative L2 error <= 3e-2 and cosine
    similarity >= 0.999 (14 convolutions + head, through every ReLU, pool,
    upsample and skip);
  * padding: images whose sides are not multiples of 8 equal the oracle on
    the zero-padded image;
  * hh_denoise on a rendered surface equals the host-array path on the same
    features; training lowers the loss.
"""
import numpy as np
import pytest

from helios_amd import scenegen
from oracle import unet

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def random_feats(rng, n, h, w):
    f = np.zeros((n, h, w, unet.N_FEAT), np.float32)
    f[..., 0:3] = rng.gamma(1.0, 0.4, size=(n, h, w, 3))
    f[..., 3:6] = rng.uniform(size=(n, h, w, 3))
    f[..., 6] = rng.uniform(0.0, 1.0, size=(n, h, w))
    nrm = rng.normal(size=(n, h, w, 3))
    f[..., 7:10] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    f[..., 10] = rng.gamma(1.0, 0.5, size=(n, h, w))
    f[..., 11] = 4.0
    return f


def perturbed_params(seed):
    rng = np.random.default_rng(seed)
    p = unet.he_init(seed)
    for name in p:      # non-zero biases: every ReLU pattern is generic
        w, b = p[name]
        p[name] = (w, rng.normal(0.0, 0.05, b.shape))
    return p


@pytest.fixture(scope="module")
def dn(gpu_ctx):
    from helios_amd.denoise import Denoiser
    d = Denoiser(gpu_ctx, seed=1)
    yield d
    d.close()


def test_param_count(dn):
    assert dn.n == unet.num_params()


def test_forward_matches_oracle(dn):
    rng = np.random.default_rng(10)
    params = perturbed_params(3)
    dn.set_params(unet.flatten(params).astype(np.float32))
    feats = random_feats(rng, 2, 16, 24)
    got = dn.forward(feats)
    ref, _ = unet.forward(params, unet.pack_inputs(feats), emulate_bf16=True)
    assert rel(got, ref) <= 5e-3, rel(got, ref)
    # and the float64 network without BF16 is still close (BF16 is the only loss)
    ref64, _ = unet.forward(params, unet.pack_inputs(feats), emulate_bf16=False)
    assert rel(got, ref64) <= 5e-2


def test_forward_padding(dn):
    rng = np.random.default_rng(11)
    params = perturbed_params(4)
    dn.set_params(unet.flatten(params).astype(np.float32))
    feats = random_feats(rng, 1, 13, 21)
    got = dn.forward(feats)
    padded = np.zeros((1, 16, 24, unet.N_FEAT), np.float32)
    padded[:, :13, :21] = feats
    ref, _ = unet.forward(params, unet.pack_inputs(padded), emulate_bf16=True)
    assert rel(got, ref[:, :13, :21]) <= 5e-3


@pytest.mark.parametrize("kind", ["log_mse", "relmse"])
def test_gradients_match_oracle(dn, kind):
    rng = np.random.default_rng(12)
    params = perturbed_params(5)
    dn.set_params(unet.flatten(params).astype(np.float32))
    dn.set_loss(kind)
    feats = random_feats(rng, 2, 16, 16)
    target = rng.gamma(1.0, 0.4, size=(2, 16, 16, 3)).astype(np.float32)
    loss = dn.loss_grad(feats, target)
    dn.set_loss("log_mse")
    g = unet.unflatten(dn.grads())
    out, cache = unet.forward(params, unet.pack_inputs(feats), emulate_bf16=True)
    if kind == "log_mse":
        ref_loss, dout = unet.loss_mse(out, np.log1p(target.astype(np.float64)))
    else:
        ref_loss, dout = unet.loss_relmse(out, target.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    ref = unet.backward(params, cache, dout)
    for name, _, _ in unet.param_shapes():
        for k, what in ((0, "W"), (1, "b")):
            a, b = g[name][k], ref[name][k]
            cos = float(np.dot(a.ravel(), b.ravel()) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
            assert rel(a, b) <= 3e-2 and cos >= 0.999, (name, what, rel(a, b), cos)


def test_training_lowers_loss(gpu_ctx):
    """A fixed batch of noisy constant patches: 60 Adam steps cut the loss."""
    from helios_amd.denoise import Denoiser
    rng = np.random.default_rng(13)
    n, h, w = 4, 32, 32
    clean = np.broadcast_to(rng.uniform(0.1, 0.8, size=(n, 1, 1, 3)), (n, h, w, 3)).astype(np.float32)
    feats = random_feats(rng, n, h, w)
    feats[..., 0:3] = np.maximum(clean + rng.normal(0.0, 0.15, size=clean.shape), 0.0)
    d = Denoiser(gpu_ctx, seed=2, params=unet.flatten(unet.he_init(2)).astype(np.float32))
    first = d.train_step(feats, clean, 1e-3)
    for _ in range(59):
        last = d.train_step(feats, clean, 1e-3)
    d.close()
    assert np.isfinite(last) and last < 0.5 * first, (first, last)


def test_denoise_surface_equals_host_path(gpu_ctx, dn):
    from helios_amd import runtime, scenegen
    from helios_amd.denoise import raw_records
    d = scenegen.cornell_box()
    w, h = 40, 32          # 40 x 32: no padding; the host path pads nothing either
    scene = runtime.Scene(gpu_ctx, d)
    surf = runtime.Surface(gpu_ctx, w, h)
    surf.enable_features()
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(5)
    it.render(scene, d.camera, surf, frames=5)
    dn.set_params(unet.flatten(perturbed_params(6)).astype(np.float32))
    out = runtime.Surface(gpu_ctx, w, h)
    dn.denoise(surf, 5, out)
    got = out.read()
    rec = raw_records(surf, 5)
    ref = np.maximum(np.expm1(dn.forward(rec)[0].astype(np.float64)), 0.0)
    np.testing.assert_allclose(got[..., :3], ref, rtol=1e-5, atol=1e-6)
    assert np.all(got[..., 3] == 1.0)


def test_trained_network_gain_on_held_out_view(gpu_ctx):
    """What the learned part buys: on a view of the held-out benchmark scene
    (seed 0; the network was trained on seeds 1-16), after 3 frames (2
    samples), the trained network's relMSE against a 512-frame plain image is
    far below both the plain image's and the untrained network's."""
    from pathlib import Path
    from helios_amd import runtime
    from helios_amd.denoise import Denoiser
    weights = Path(__file__).resolve().parents[1] / "helios_amd" / "data" / "denoiser_unet.bin"
    d = scenegen.icosphere_field(seed=0, width=960, height=540, max_ray_bounces=8)
    scene = runtime.Scene(gpu_ctx, d)
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(8)
    it.set_max_samples(1 << 20)
    ref_s = runtime.Surface(gpu_ctx, d.width, d.height)
    it.render(scene, d.camera, ref_s, frames=512)
    ref = ref_s.read()
    it.restart_bake()
    surf = runtime.Surface(gpu_ctx, d.width, d.height)
    surf.enable_features()
    it.render(scene, d.camera, surf, frames=3)
    out = runtime.Surface(gpu_ctx, d.width, d.height)

    def err(img):
        x, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
        return float(np.mean((x - r) ** 2 / (r * r + 0.01)))

    plain = err(surf.read())
    untrained = Denoiser(gpu_ctx, seed=0)
    untrained.denoise(surf, 3, out)
    e_untrained = err(out.read())
    untrained.close()
    trained = Denoiser(gpu_ctx, params=np.fromfile(weights, np.float32))
    trained.denoise(surf, 3, out)
    e_trained = err(out.read())
    trained.close()
    assert e_trained < 0.25 * plain, (e_trained, plain)
    assert e_trained < 0.25 * e_untrained, (e_trained, e_untrained)
