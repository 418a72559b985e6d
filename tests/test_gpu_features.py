"""GPU: first-hit feature planes (the denoiser's auxiliary inputs).

The planes have no counterpart in the reference; what pins them is algebra:
they must follow the radiance's own progressive recurrence (path_trace_rgen.glsl
accumulation: frame 0 stores, frame f > 0 adds (x - prev) / f), and M2 must be
Welford's sum of squared luminance deviations over exactly the samples that
the running mean holds (samples 1..N-1 after N frames). Per-sample values are
recovered independently: a batch render of one sample with first_frame = f
into a cleared surface stores x / f (x at f = 0).
"""
import numpy as np
import pytest

from helios_amd import scenegen

pytestmark = pytest.mark.gpu

LUMA = np.array([0.2126, 0.7152, 0.0722])


def _per_sample(ctx, scene, cam, w, h, frames, bounces):
    from helios_amd import runtime
    out = []
    for f in range(frames):
        s = runtime.Surface(ctx, w, h)
        s.enable_features()
        runtime.render_batch(ctx, [dict(scene=scene, surface=s, camera=cam, spp=1, max_ray_bounces=bounces,
                                        first_frame=f)])
        k = float(max(f, 1))
        a0, a1 = s.read_features()
        out.append((s.read()[..., :3].astype(np.float64) * k, a0.astype(np.float64) * k,
                    a1[..., :3].astype(np.float64) * k))
    return out


@pytest.mark.parametrize("frames", [2, 7])
def test_feature_planes_follow_the_radiance_recurrence(gpu_ctx, frames):
    from helios_amd import runtime
    d = scenegen.cornell_box()
    w, h, bounces = 48, 40, 5
    scene = runtime.Scene(gpu_ctx, d)
    per = _per_sample(gpu_ctx, scene, d.camera, w, h, frames, bounces)

    surf = runtime.Surface(gpu_ctx, w, h)
    surf.enable_features()
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(bounces)
    it.render(scene, d.camera, surf, frames=frames)
    rad = surf.read()[..., :3].astype(np.float64)
    a0, a1 = (p.astype(np.float64) for p in surf.read_features())

    # the running means hold samples 1..N-1 (sample 0 is overwritten by frame 1)
    held = per[1:] if frames > 1 else per
    mean_rad = np.mean([np.minimum(p[0], 1.0) for p in held], axis=0)
    mean_a0 = np.mean([p[1] for p in held], axis=0)
    mean_n = np.mean([p[2] for p in held], axis=0)
    np.testing.assert_allclose(rad, mean_rad, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(a0, mean_a0, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(a1[..., :3], mean_n, rtol=1e-4, atol=1e-5)

    lum = np.array([np.minimum(p[0], 1.0) @ LUMA for p in held])       # (n, h, w)
    m2 = ((lum - lum.mean(axis=0)) ** 2).sum(axis=0)
    np.testing.assert_allclose(a1[..., 3], m2, rtol=2e-3, atol=2e-5)

    # ranges: albedo and depth code in [0, 1], normals at most unit length
    assert a0[..., :3].min() >= 0.0 and a0[..., :3].max() <= 1.0 + 1e-6
    assert a0[..., 3].min() >= 0.0 and a0[..., 3].max() <= 1.0
    assert np.linalg.norm(a1[..., :3], axis=-1).max() <= 1.0 + 1e-5
    assert a1[..., 3].min() >= 0.0


def test_feature_planes_cleared_with_the_surface(gpu_ctx):
    from helios_amd import runtime
    d = scenegen.cornell_box()
    scene = runtime.Scene(gpu_ctx, d)
    surf = runtime.Surface(gpu_ctx, 32, 32)
    surf.enable_features()
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(3)
    it.render(scene, d.camera, surf, frames=3)
    a0, _ = surf.read_features()
    assert np.abs(a0).max() > 0.0
    surf.clear()
    a0, a1 = surf.read_features()
    assert not a0.any() and not a1.any()


def test_megakernel_writes_the_same_feature_planes(gpu_ctx):
    from helios_amd import runtime
    from helios_amd.abi import HH_FLAG_MEGAKERNEL
    d = scenegen.icosphere_field(seed=3, n_spheres=6, level=3, n_leaves=120, width=96, height=64, max_ray_bounces=6)
    scene = runtime.Scene(gpu_ctx, d)
    planes = []
    for flags in (0, HH_FLAG_MEGAKERNEL):
        surf = runtime.Surface(gpu_ctx, d.width, d.height)
        surf.enable_features()
        it = runtime.PathIntegrator(gpu_ctx)
        it.set_max_ray_bounces(6)
        it.set_flags(flags)
        it.render(scene, d.camera, surf, frames=4)
        planes.append(surf.read_features())
    for a, b in zip(*planes):
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6)


def test_surface_validate_counts_nonfinite(gpu_ctx):
    from helios_amd import runtime
    from helios_amd.runtime import HeliosError
    d = scenegen.cornell_box()
    scene = runtime.Scene(gpu_ctx, d)
    surf = runtime.Surface(gpu_ctx, 32, 24)
    it = runtime.PathIntegrator(gpu_ctx)
    it.render(scene, d.camera, surf, frames=2)
    assert surf.validate() == 0
    img = surf.read()
    img[3, 5, 1] = np.nan
    img[7, 2, 0] = np.inf
    img[0, 0, 3] = -np.inf
    surf.write(img)
    with pytest.raises(HeliosError) as e:
        surf.validate()
    assert e.value.status == 5 and "3 non-finite" in str(e.value)
