"""Training data from the tracer itself, and the training loop.

Pairs come from progressive renders of seeded icosphere-field scenes at
1920x1080 (the same generator, resolution and bounce count as the benchmark
workload; held-out seeds, random cameras): one run per view is snapshotted at
several frame counts (the noisy inputs, with their feature planes) and
continued to `ref_frames`. The target is the mean of the samples the inputs
do not hold, recovered from the two running means, so input and target noise
are independent:
    target = ((R - 1) m_R - (k - 1) m_k) / (R - k)
for the last snapshot k (the running mean after f frames holds samples
1..f-1). Random 128x128 crops of every snapshot form the dataset.

    python -m helios_amd.denoise_train --out helios_amd/data/denoiser_unet.bin
"""
import argparse
import json
import sys
import time

import numpy as np

from . import runtime, scenegen
from .denoise import Denoiser, raw_records

SNAPSHOTS = (2, 3, 5, 9, 17, 33)     # running means of 1, 2, 4, 8, 16, 32 samples


def random_view(rng, seed, width, height, bounces):
    d = scenegen.icosphere_field(seed=seed, n_spheres=48, level=4, n_leaves=1500, width=width, height=height,
                                 max_ray_bounces=bounces)
    ang = rng.uniform(0, 2 * np.pi)
    dist = rng.uniform(50.0, 140.0)
    eye = (dist * np.cos(ang), rng.uniform(8.0, 50.0), dist * np.sin(ang))
    tgt = (rng.uniform(-40, 40), rng.uniform(0.0, 6.0), rng.uniform(-40, 40))
    d.camera = scenegen.look_at_camera(eye, tgt, fov_y=float(rng.uniform(40.0, 70.0)), aperture=0.0)
    return d


def render_view(ctx, desc, ref_frames, crops, crop, rng, log=print):
    """Snapshots + independent target for one view -> (inputs, target) crops:
    inputs (len(SNAPSHOTS) * crops, crop, crop, 12), targets (crops, crop, crop, 3)
    and the index of each input's target."""
    scene = runtime.Scene(ctx, desc)
    w, h = desc.width, desc.height
    surf = runtime.Surface(ctx, w, h)
    surf.enable_features()
    it = runtime.PathIntegrator(ctx)
    it.set_max_ray_bounces(desc.max_ray_bounces)
    it.set_max_samples(1 << 30)
    ys = rng.integers(0, h - crop + 1, crops)
    xs = rng.integers(0, w - crop + 1, crops)
    inputs = []
    done = 0
    for k in SNAPSHOTS:
        it.render(scene, desc.camera, surf, frames=k - done)
        done = k
        rec = raw_records(surf, k)[0]
        for y, x in zip(ys, xs):
            inputs.append(rec[y:y + crop, x:x + crop])
    m_k = surf.read()[..., :3].astype(np.float64)
    it.render(scene, desc.camera, surf, frames=ref_frames - done)
    m_r = surf.read()[..., :3].astype(np.float64)
    k, R = SNAPSHOTS[-1], ref_frames
    target = ((R - 1) * m_r - (k - 1) * m_k) / (R - k)
    target = np.maximum(target, 0.0).astype(np.float32)
    tgt = np.stack([target[y:y + crop, x:x + crop] for y, x in zip(ys, xs)])
    tidx = np.tile(np.arange(crops), len(SNAPSHOTS))
    return np.stack(inputs), tgt, tidx


def build_dataset(ctx, seeds, views_per_seed, width, height, bounces, ref_frames, crops, crop, seed=0, log=print):
    rng = np.random.default_rng(seed)
    X, T, TI = [], [], []
    base = 0
    for s in seeds:
        for v in range(views_per_seed):
            t0 = time.time()
            desc = random_view(rng, s, width, height, bounces)
            x, t, ti = render_view(ctx, desc, ref_frames, crops, crop, rng)
            X.append(x)
            T.append(t)
            TI.append(ti + base)
            base += len(t)
            log(json.dumps({"view": f"seed {s} #{v}", "pairs": int(len(x)), "seconds": round(time.time() - t0, 2)}))
            sys.stdout.flush()
    return np.concatenate(X), np.concatenate(T), np.concatenate(TI)


def augment(rng, x, t):
    """Random flips / transposes (the network is not rotation-equivariant)."""
    if rng.random() < 0.5:
        x, t = x[:, ::-1], t[:, ::-1]
    if rng.random() < 0.5:
        x, t = x[:, :, ::-1], t[:, :, ::-1]
    if rng.random() < 0.5:
        x, t = x.transpose(0, 2, 1, 3), t.transpose(0, 2, 1, 3)
    return np.ascontiguousarray(x), np.ascontiguousarray(t)


def train(dn, X, T, TI, steps, batch, lr, seed=0, log=print, log_every=100):
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.time()
    for s in range(steps):
        idx = rng.integers(0, len(X), batch)
        x, t = augment(rng, X[idx], T[TI[idx]])
        cur_lr = lr * (0.5 * (1 + np.cos(np.pi * s / steps)))    # cosine schedule
        losses.append(dn.train_step(x, t, float(cur_lr)))
        if (s + 1) % log_every == 0 or s == 0:
            log(json.dumps({"step": s + 1, "loss": float(np.mean(losses[-log_every:])), "lr": float(cur_lr),
                            "seconds": round(time.time() - t0, 1)}))
            sys.stdout.flush()
    return losses


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--out", required=True)
    p.add_argument("--seeds", type=int, nargs="+", default=list(range(1, 9)))
    p.add_argument("--views-per-seed", type=int, default=2)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--bounces", type=int, default=8)
    p.add_argument("--ref-frames", type=int, default=1024)
    p.add_argument("--crops", type=int, default=24)
    p.add_argument("--crop", type=int, default=128)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--batch", type=int, default=8)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--loss", default="relmse", choices=["relmse", "log_mse"])
    p.add_argument("--init", default=None, help="start from these parameters (float32 file)")
    a = p.parse_args(argv)
    ctx = runtime.Context(0)
    ctx.set_tuning(frames_in_flight=8, extend_block=64, persistent_blocks_per_cu=32)
    X, T, TI = build_dataset(ctx, a.seeds, a.views_per_seed, a.width, a.height, a.bounces, a.ref_frames, a.crops,
                             a.crop)
    print(json.dumps({"dataset": int(len(X)), "targets": int(len(T))}))
    dn = Denoiser(ctx, seed=0)
    dn.set_loss(a.loss)
    if a.init:
        dn.set_params(np.fromfile(a.init, np.float32))
    losses = train(dn, X, T, TI, a.steps, a.batch, a.lr)
    dn.params().astype(np.float32).tofile(a.out)
    print(json.dumps({"saved": a.out, "final_loss": float(np.mean(losses[-100:]))}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
