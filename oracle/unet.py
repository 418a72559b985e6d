"""CPU oracle of the learned denoiser (TEST INFRASTRUCTURE ONLY).

The reference has no denoiser: this network is the build's own, specified by
the task (a U-Net with 4 levels of 32/48/64/96 channels, 3x3 convolutions and
ReLU, BF16 storage with FP32 accumulation). This module restates it in numpy
(float64) as the checker of the HIP kernels:

  * forward / backward of every layer, pinned by central finite differences
    in float64 (tests/test_denoiser_oracle.py);
  * optional BF16 emulation at exactly the points where the GPU stores BF16
    (inputs, weights, activations, activation gradients), so GPU-vs-oracle
    differences reduce to FP32 accumulation order.

Only tests/ and bench.py's baseline leg import this module.

Layout: images are NHWC arrays. Convolution weights are (Cout, 3, 3, Cin)
(ky, kx = 0..2 for offsets -1..+1), zero padding, stride 1. The flat
parameter vector holds, layer by layer, W (row-major) then b.
"""
import numpy as np

C_IN = 10          # network input channels
C_IN_PAD = 16      # input tensor channels on the GPU (zeros beyond C_IN)
N_FEAT = 12        # raw feature record per pixel (see pack_inputs)

# name, (source0 name, upsample), (source1 name or None), cout
LAYERS = [
    ("enc0a", ("in", 1), None, 32),
    ("enc0b", ("enc0a", 1), None, 32),
    ("enc1a", ("pool0", 1), None, 48),
    ("enc1b", ("enc1a", 1), None, 48),
    ("enc2a", ("pool1", 1), None, 64),
    ("enc2b", ("enc2a", 1), None, 64),
    ("mida", ("pool2", 1), None, 96),
    ("midb", ("mida", 1), None, 96),
    ("dec2a", ("midb", 2), "enc2b", 64),
    ("dec2b", ("dec2a", 1), None, 64),
    ("dec1a", ("dec2b", 2), "enc1b", 48),
    ("dec1b", ("dec1a", 1), None, 48),
    ("dec0a", ("dec1b", 2), "enc0b", 32),
    ("dec0b", ("dec0a", 1), None, 32),
]
POOLS = [("pool0", "enc0b"), ("pool1", "enc1b"), ("pool2", "enc2b")]
HEAD_IN = "dec0b"
HEAD_OUT = 3


def channels():
    ch = {"in": C_IN}
    for name, (s0, _), s1, cout in LAYERS:
        ch[name] = cout
        for p, src in POOLS:
            if src == name:
                ch[p] = cout
    return ch


def layer_cin(name):
    ch = channels()
    for n, (s0, _), s1, _ in LAYERS:
        if n == name:
            return ch[s0] + (ch[s1] if s1 else 0)
    raise KeyError(name)


def param_shapes():
    """[(name, W shape, b shape)] in flat order."""
    out = []
    for name, _, _, cout in LAYERS:
        out.append((name, (cout, 3, 3, layer_cin(name)), (cout,)))
    out.append(("head", (HEAD_OUT, channels()[HEAD_IN]), (HEAD_OUT,)))
    return out


def num_params():
    return sum(int(np.prod(w)) + int(np.prod(b)) for _, w, b in param_shapes())


def unflatten(flat):
    flat = np.asarray(flat, np.float64)
    out, k = {}, 0
    for name, ws, bs in param_shapes():
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        out[name] = (flat[k:k + nw].reshape(ws), flat[k + nw:k + nw + nb].reshape(bs))
        k += nw + nb
    assert k == flat.size
    return out


def flatten(params):
    return np.concatenate([np.concatenate([params[n][0].ravel(), params[n][1].ravel()]) for n, _, _ in param_shapes()])


def he_init(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, ws, bs in param_shapes():
        fan_in = int(np.prod(ws[1:]))
        p[name] = (rng.normal(0.0, np.sqrt(2.0 / fan_in), ws), np.zeros(bs))
    return p


# ---------------------------------------------------------------------------
# BF16 round-to-nearest-even of a float32 value (NaN-free inputs)
# ---------------------------------------------------------------------------
def bf16(x):
    u = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


# ---------------------------------------------------------------------------
# Input encoding
# ---------------------------------------------------------------------------
def octahedral(n):
    """Unit vector -> [-1, 1]^2 (zero vector -> 0)."""
    l1 = np.abs(n).sum(axis=-1, keepdims=True)
    p = np.where(l1 > 1e-12, n / np.maximum(l1, 1e-30), 0.0)
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    sx = np.where(x >= 0.0, 1.0, -1.0)
    sy = np.where(y >= 0.0, 1.0, -1.0)
    ox = np.where(z < 0.0, (1.0 - np.abs(y)) * sx, x)
    oy = np.where(z < 0.0, (1.0 - np.abs(x)) * sy, y)
    return np.stack([ox, oy], axis=-1)


def pack_inputs(feat):
    """feat (N, H, W, 12) = radiance rgb, albedo rgb, 1/(1+t), normal xyz, M2,
    samples held by the running mean -> network input (N, H, W, 10):
    log1p(radiance), albedo, octahedral(normal / |normal|), depth code, and the
    standard deviation of the mean luminance sqrt(M2 / (n (n - 1))) (0 for n < 2)."""
    feat = np.asarray(feat, np.float64)
    rad = np.maximum(feat[..., 0:3], 0.0)
    nrm = feat[..., 7:10]
    ln = np.linalg.norm(nrm, axis=-1, keepdims=True)
    nrm = np.where(ln > 1e-12, nrm / np.maximum(ln, 1e-30), 0.0)
    n = feat[..., 11:12]
    var = np.where(n >= 2.0, np.maximum(feat[..., 10:11], 0.0) / np.maximum(n * (n - 1.0), 1.0), 0.0)
    return np.concatenate([np.log1p(rad), feat[..., 3:6], octahedral(nrm), feat[..., 6:7], np.sqrt(var)], axis=-1)


# ---------------------------------------------------------------------------
# Layers (forward and backward)
# ---------------------------------------------------------------------------
OFFS = [(ky - 1, kx - 1) for ky in range(3) for kx in range(3)]


def _shift(x, dy, dx):
    """y[:, i, j] = x[:, i + dy, j + dx] with zeros outside."""
    N, H, W, C = x.shape
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    y[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv3x3(x, w, b):
    cols = np.stack([_shift(x, dy, dx) for dy, dx in OFFS], axis=3)        # N H W 9 C
    N, H, W_, _, C = cols.shape
    out = cols.reshape(-1, 9 * C) @ w.reshape(w.shape[0], -1).T
    return out.reshape(N, H, W_, -1) + b


def conv3x3_backward(x, w, dy_):
    cols = np.stack([_shift(x, dy, dx) for dy, dx in OFFS], axis=3)
    N, H, W_, _, C = cols.shape
    g = dy_.reshape(-1, w.shape[0])
    dw = (g.T @ cols.reshape(-1, 9 * C)).reshape(w.shape)
    db = g.sum(axis=0)
    dcols = (g @ w.reshape(w.shape[0], -1)).reshape(N, H, W_, 9, C)
    dx = np.zeros_like(x)
    for t, (dy, dx_) in enumerate(OFFS):
        dx += _shift(dcols[:, :, :, t], -dy, -dx_)
    return dx, dw, db


def pool2(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool2_backward(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_backward(g):
    N, H, W, C = g.shape
    return g.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


# ---------------------------------------------------------------------------
# Network
# ---------------------------------------------------------------------------
def forward(params, x, emulate_bf16=False):
    """x: packed input (N, H, W, 10), H and W multiples of 8. Returns
    (out (N, H, W, 3) in log1p space, cache). The head adds the input's
    log1p radiance (residual)."""
    q = bf16 if emulate_bf16 else (lambda a: a)
    acts = {"in": q(np.asarray(x, np.float64))}
    pre = {}
    for name, (s0, up), s1, cout in LAYERS:
        w, b = params[name]
        src = acts[s0]
        if up == 2:
            src = up2(src)
        if s1:
            src = np.concatenate([src, acts[s1]], axis=-1)
        z = conv3x3(src, q(w), f32(b) if emulate_bf16 else b)
        pre[name] = z
        acts[name] = q(np.maximum(z, 0.0))
        for p, srcn in POOLS:
            if srcn == name:
                acts[p] = q(pool2(acts[name]))
    wh, bh = params["head"]
    h = acts[HEAD_IN]
    out = h @ q(wh).T + (f32(bh) if emulate_bf16 else bh)
    resid = f32(x[..., 0:3]) if emulate_bf16 else np.asarray(x[..., 0:3], np.float64)
    out = out + resid
    return out, dict(acts=acts, pre=pre, x=x, emulate=emulate_bf16)


def loss_mse(out, target_log):
    d = out - target_log
    return float(np.mean(d * d)), 2.0 * d / d.size


def loss_relmse(out, target, eps=0.01):
    """relMSE of the linear prediction y = expm1(out) against a linear target:
    mean of (y - t)^2 / (t^2 + eps); gradient with respect to out."""
    y = np.expm1(out)
    t = np.maximum(np.asarray(target, np.float64), 0.0)
    wgt = 1.0 / (t * t + eps)
    e = y - t
    return float(np.mean(e * e * wgt)), 2.0 * e * wgt * (y + 1.0) / e.size


def backward(params, cache, dout):
    """Gradients of a scalar loss w.r.t. every parameter given dL/dout."""
    q = bf16 if cache["emulate"] else (lambda a: a)
    acts = cache["acts"]
    grads = {}
    wh, bh = params["head"]
    h = acts[HEAD_IN]
    grads["head"] = (np.einsum("nhwo,nhwc->oc", dout, h), dout.sum(axis=(0, 1, 2)))
    gact = {HEAD_IN: dout @ q(wh)}        # gradient w.r.t. the stored activation
    for name, (s0, up), s1, cout in reversed(LAYERS):
        w, b = params[name]
        g = gact.pop(name)
        # pooled copies of this activation feed later layers: add their grads
        for p, srcn in POOLS:
            if srcn == name and p in gact:
                g = g + pool2_backward(gact.pop(p))
        g = q(g * (acts[name] > 0.0))     # through the ReLU; stored BF16 on the GPU
        src = acts[s0]
        if up == 2:
            src = up2(src)
        if s1:
            src = np.concatenate([src, acts[s1]], axis=-1)
        dx, dw, db = conv3x3_backward(src, q(w), g)
        grads[name] = (dw, db)
        c0 = acts[s0].shape[-1]
        d0 = dx[..., :c0]
        if up == 2:
            d0 = up2_backward(d0)
        if s0 != "in":
            gact[s0] = gact.get(s0, 0.0) + d0
        if s1:
            gact[s1] = gact.get(s1, 0.0) + dx[..., c0:]
    return grads


def denoise(params, feat, emulate_bf16=False):
    """Raw features (N, H, W, 12) -> denoised radiance (N, H, W, 3)."""
    out, _ = forward(params, pack_inputs(feat), emulate_bf16)
    return np.maximum(np.expm1(out), 0.0)
