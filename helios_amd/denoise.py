"""The learned denoiser through the C ABI (hh_denoiser_*).

Denoiser wraps one network on a context: parameters in and out as one FP32
vector (480,675 floats, layer order of include/helios_hip.h), inference on a
surface with feature planes, and training steps on host batches of raw
feature records (n x h x w x 12 floats) against target radiance
(n x h x w x 3).
"""
import ctypes as C

import numpy as np

from .abi import ptr
from .runtime import check, lib

N_FEAT = 12

# (input channels, output channels, pyramid level) of the 14 convolutions
CONVS = [(10, 32, 0), (32, 32, 0), (32, 48, 1), (48, 48, 1), (48, 64, 2), (64, 64, 2), (64, 96, 3), (96, 96, 3),
         (160, 64, 2), (64, 64, 2), (112, 48, 1), (48, 48, 1), (80, 32, 0), (32, 32, 0)]


def forward_flops(h, w):
    """Algorithmic FLOPs (2 per multiply-add) of one forward pass on an h x w
    image padded to multiples of 8: 84,624 multiply-adds per pixel."""
    hp, wp = (h + 7) // 8 * 8, (w + 7) // 8 * 8
    macs = sum(9 * ci * co * (hp >> lv) * (wp >> lv) for ci, co, lv in CONVS) + 32 * 3 * hp * wp
    return 2 * macs

# Stored tensors each convolution reads and writes: ([(channels, level)] in, (channels, level) out).
# Inputs are BF16 (the first layer reads the 16-channel packed input); decoder
# convolutions read the coarser level once (the nearest upsample is in the
# loader) and the skip at their own level.
_CONV_IO = [([(16, 0)], (32, 0)), ([(32, 0)], (32, 0)), ([(32, 1)], (48, 1)), ([(48, 1)], (48, 1)),
            ([(48, 2)], (64, 2)), ([(64, 2)], (64, 2)), ([(64, 3)], (96, 3)), ([(96, 3)], (96, 3)),
            ([(96, 3), (64, 2)], (64, 2)), ([(64, 2)], (64, 2)), ([(64, 2), (48, 1)], (48, 1)), ([(48, 1)], (48, 1)),
            ([(48, 1), (32, 0)], (32, 0)), ([(32, 0)], (32, 0))]
_POOLS = [(32, 0), (48, 1), (64, 2)]


def forward_bytes(h, w):
    """Algorithmic HBM bytes of one hh_denoise pass on an h x w surface (padded
    to multiples of 8): every stored tensor read once and written once --
    the packing (radiance + two feature planes in, 16 BF16 channels + the
    FP32 residual out), the 14 convolutions, the three 2x2 poolings and the
    head (32 channels + residual in, network output + surface out)."""
    hp, wp = (h + 7) // 8 * 8, (w + 7) // 8 * 8

    def px(lv):
        return (hp >> lv) * (wp >> lv)

    total = px(0) * (48 + 32 + 16)                                       # pack
    for ins, (co, lo) in _CONV_IO:
        total += sum(2 * c * px(lv) for c, lv in ins) + 2 * co * px(lo)
    total += sum(2 * c * px(lv) + 2 * c * px(lv + 1) for c, lv in _POOLS)
    total += px(0) * (2 * 32 + 16 + 16 + 16)                            # head
    return total


_SIGS = {
    "hh_denoiser_create": [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)],
    "hh_denoiser_destroy": [C.c_void_p],
    "hh_denoiser_num_params": [C.POINTER(C.c_uint64)],
    "hh_denoiser_set_params": [C.c_void_p, C.c_void_p, C.c_uint64],
    "hh_denoiser_get_params": [C.c_void_p, C.c_void_p, C.c_uint64],
    "hh_denoiser_get_grads": [C.c_void_p, C.c_void_p, C.c_uint64],
    "hh_denoise": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "hh_denoiser_forward": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p],
    "hh_denoiser_loss_grad": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                              C.POINTER(C.c_float)],
    "hh_denoiser_set_loss": [C.c_void_p, C.c_uint32],
    "hh_denoiser_train_step": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                               C.POINTER(C.c_float)],
}
_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _bound = True
    return L


def num_params():
    n = C.c_uint64()
    check(_lib().hh_denoiser_num_params(C.byref(n)))
    return n.value


class Denoiser:
    def __init__(self, ctx, seed=0, params=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(_lib().hh_denoiser_create(ctx.h, seed, C.byref(self.h)))
        self.n = num_params()
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h:
            _lib().hh_denoiser_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_loss(self, kind):
        """"log_mse" (default) or "relmse"."""
        check(_lib().hh_denoiser_set_loss(self.h, {"log_mse": 0, "relmse": 1}[kind]))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32).ravel()
        check(_lib().hh_denoiser_set_params(self.h, ptr(p), p.size))

    def params(self):
        out = np.zeros(self.n, np.float32)
        check(_lib().hh_denoiser_get_params(self.h, ptr(out), self.n))
        return out

    def grads(self):
        out = np.zeros(self.n, np.float32)
        check(_lib().hh_denoiser_get_grads(self.h, ptr(out), self.n))
        return out

    def forward(self, feats):
        """Raw records (n, h, w, 12) -> network output in log1p space (n, h, w, 3)."""
        f = np.ascontiguousarray(feats, np.float32)
        n, h, w, _ = f.shape
        out = np.zeros((n, h, w, 3), np.float32)
        check(_lib().hh_denoiser_forward(self.h, ptr(f), n, h, w, ptr(out)))
        return out

    def loss_grad(self, feats, target):
        f = np.ascontiguousarray(feats, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        n, h, w, _ = f.shape
        loss = C.c_float()
        check(_lib().hh_denoiser_loss_grad(self.h, ptr(f), ptr(t), n, h, w, C.byref(loss)))
        return loss.value

    def train_step(self, feats, target, lr):
        f = np.ascontiguousarray(feats, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        n, h, w, _ = f.shape
        loss = C.c_float()
        check(_lib().hh_denoiser_train_step(self.h, ptr(f), ptr(t), n, h, w, lr, C.byref(loss)))
        return loss.value

    def denoise(self, surface_in, num_frames, surface_out):
        check(_lib().hh_denoise(self.h, surface_in.h, num_frames, surface_out.h))


def raw_records(surface, num_frames):
    """Raw feature records (1, h, w, 12) of a surface with feature planes."""
    rad = surface.read()
    a0, a1 = surface.read_features()
    held = num_frames - 1 if num_frames >= 2 else 1
    rec = np.concatenate([rad[..., :3], a0, a1, np.full(rad.shape[:2] + (1,), held, np.float32)], axis=-1)
    return rec[None].astype(np.float32)
