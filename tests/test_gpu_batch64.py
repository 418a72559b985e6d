"""GPU: the batched workload at its generated sizes.

Nine views of scenegen.batched_workload (64 icosphere-field scenes, seeds
0..63, square sides log-uniform in 128..1024, triangle counts log-uniform in
10 k..2 M; 8 bounces, 16 spp each): the eight that sit at the 0, 1/7, ..., 7/7
quantiles of the triangle count (11.7 k .. 1.99 M) plus the largest view
(894 x 894). The reference traces each view with its own
vkCmdTraceRaysKHR (path_integrator.cpp:125-200); here they share wavefront
passes, so:

  * every view of one hh_render_batch is bit-identical to rendering it alone,
    with the default tuning and with the one the autotuner picks;
  * no ray was cut short by a traversal bound;
  * on reduced-resolution copies (same triangles and cameras, 48 x 48) of the
    smallest, the largest and the widest scene, primary IDs, t, u, v equal
    the oracle's and radiance is within relMSE 1e-4 of the FP32 transcription.
"""
import numpy as np
import pytest

from helios_amd import scenegen

pytestmark = pytest.mark.gpu


def _selection():
    P = scenegen.batched_workload_params()
    tris = np.array([p[2] for p in P])
    side = np.array([p[1] for p in P])
    order = np.argsort(tris)
    sel = [int(order[k]) for k in np.linspace(0, len(P) - 1, 8).round().astype(int)]
    widest = int(side.argmax())
    if widest not in sel:
        sel.append(widest)
    return sorted(sel), widest


def relmse(x, r, eps=0.01):
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(r, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


@pytest.fixture(scope="module")
def workload(gpu_ctx):
    from helios_amd import runtime
    sel, widest = _selection()
    descs = scenegen.batched_workload(select=set(sel))
    scenes = [runtime.Scene(gpu_ctx, d) for d in descs]
    return sel, widest, descs, scenes


def _batch(ctx, scenes, descs):
    from helios_amd import runtime
    surfs = [runtime.Surface(ctx, d.width, d.height) for d in descs]
    runtime.render_batch(ctx, [dict(scene=s, surface=f, camera=d.camera, spp=d.spp, max_ray_bounces=8)
                               for s, f, d in zip(scenes, surfs, descs)])
    imgs = [f.read() for f in surfs]
    for f in surfs:
        f.close()
    return imgs


def test_workload_spans_its_stated_ranges(workload):
    _, _, descs, _ = workload
    tris = [d.num_triangles for d in descs]
    sides = [d.width for d in descs]
    assert min(tris) < 15_000 and max(tris) > 1_900_000
    assert min(sides) < 150 and max(sides) > 850
    assert all(d.spp == 16 and d.max_ray_bounces == 8 for d in descs)


def test_batched_views_equal_single_renders(gpu_ctx, workload):
    _, _, descs, scenes = workload
    gpu_ctx.set_tuning()
    batch = _batch(gpu_ctx, scenes, descs)
    for s, d, img in zip(scenes, descs, batch):
        alone = _batch(gpu_ctx, [s], [d])[0]
        assert np.isfinite(img).all() and img[..., :3].any()
        assert np.array_equal(img, alone), d.num_triangles
    for s in scenes:
        assert s.traversal_faults() == dict(lane_stack=0, packet_stack=0, trips=0)


def test_autotuned_batch_equals_default(gpu_ctx, workload):
    from helios_amd import runtime
    _, _, descs, scenes = workload
    gpu_ctx.set_tuning()
    ref = _batch(gpu_ctx, scenes, descs)
    scratch = [runtime.Surface(gpu_ctx, d.width, d.height) for d in descs]
    views = [dict(scene=s, surface=f, camera=d.camera, max_ray_bounces=8) for s, f, d in zip(scenes, scratch, descs)]
    try:
        tuning, ms = runtime.autotune(gpu_ctx, views, frames=2)
        assert ms > 0.0
        tuned = _batch(gpu_ctx, scenes, descs)
    finally:
        gpu_ctx.set_tuning()
        for f in scratch:
            f.close()
    for a, b in zip(ref, tuned):
        assert np.array_equal(a, b)


def test_reduced_copies_match_oracle(gpu_ctx, oracle_mod, workload):
    from helios_amd import runtime
    sel, widest, descs, scenes = workload
    tris = [d.num_triangles for d in descs]
    picks = sorted({int(np.argmin(tris)), int(np.argmax(tris)), sel.index(widest)})
    for k in picks:
        d = descs[k]
        small = scenegen.batched_workload(select={sel[k]})[0]
        small.width = small.height = 48
        o = oracle_mod.OracleScene(small)
        hg = scenes[k].primary_hits(small.camera, 48, 48, 0)
        ho = o.primary_hits(frame=0, brute=False)
        for f in ("instance", "geometry", "primitive", "t", "u", "v"):
            assert np.array_equal(hg[f], ho[f]), (d.num_triangles, f)
        surf = runtime.Surface(gpu_ctx, 48, 48)
        runtime.render_batch(gpu_ctx, [dict(scene=scenes[k], surface=surf, camera=small.camera, spp=4,
                                            max_ray_bounces=8)])
        img = surf.read()
        surf.close()
        ref = o.render(frames=4)
        assert relmse(img, ref) < 1e-4, (d.num_triangles, relmse(img, ref))
