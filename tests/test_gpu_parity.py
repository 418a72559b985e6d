"""GPU parity: the HIP path against the CPU oracle on the same inputs.

Bars (from the path's contract):
  * primary-hit instance / geometry / primitive IDs bit-exact, except edge-tie
    pixels at no more than 0.01 % of the image;
  * traversal of arbitrary rays: same IDs and bit-identical t, u, v;
  * radiance: relMSE (mean of (x - r)^2 / (r^2 + 0.01)) within 1e-3 of the
    FP32 transcription and of the FP64 oracle (same per-pixel RNG streams);
  * integer/byte outputs (light records, tonemapped RGBA8) exact.
"""
import numpy as np
import pytest

from helios_amd import scenegen
from helios_amd.abi import HH_TRACE_OPAQUE, HH_TRACE_TERMINATE_ON_FIRST, HH_FLAG_MEGAKERNEL

pytestmark = pytest.mark.gpu

EDGE_TIE_FRACTION = 1e-4   # 0.01 % of pixels
RELMSE_BAR = 1e-3


def relmse(x, r, eps=0.01):
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(r, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def id_mismatch(a, b):
    return (a["instance"] != b["instance"]) | (a["geometry"] != b["geometry"]) | (a["primitive"] != b["primitive"])


@pytest.fixture(scope="module")
def cornell():
    return scenegen.cornell_box()


@pytest.fixture(scope="module")
def small_field():
    return scenegen.icosphere_field(seed=3, n_spheres=6, level=3, n_leaves=120, width=160, height=90,
                                    max_ray_bounces=8)


@pytest.fixture(scope="module")
def cornell_pair(gpu_ctx, oracle_mod, cornell):
    from helios_amd import runtime
    return runtime.Scene(gpu_ctx, cornell), oracle_mod.OracleScene(cornell)


@pytest.fixture(scope="module")
def field_pair(gpu_ctx, oracle_mod, small_field):
    from helios_amd import runtime
    return runtime.Scene(gpu_ctx, small_field), oracle_mod.OracleScene(small_field)


@pytest.mark.parametrize("frame", [0, 1, 7, 63])
def test_primary_ids_cornell_bit_exact(cornell, cornell_pair, frame):
    g, o = cornell_pair
    hg = g.primary_hits(cornell.camera, cornell.width, cornell.height, frame)
    ho = o.primary_hits(frame=frame, brute=True)
    bad = id_mismatch(hg, ho)
    assert bad.mean() <= EDGE_TIE_FRACTION, f"{bad.sum()} mismatching pixels"
    both = ~bad & (ho["instance"] != 0xFFFFFFFF)
    assert np.array_equal(hg["t"][both], ho["t"][both])
    assert np.array_equal(hg["u"][both], ho["u"][both])
    assert np.array_equal(hg["v"][both], ho["v"][both])


def test_primary_ids_field_bit_exact(small_field, field_pair):
    g, o = field_pair
    d = small_field
    for frame in (0, 5):
        hg = g.primary_hits(d.camera, d.width, d.height, frame)
        ho = o.primary_hits(frame=frame, brute=False)
        bad = id_mismatch(hg, ho)
        assert bad.mean() <= EDGE_TIE_FRACTION, f"frame {frame}: {bad.sum()} mismatching pixels"
        both = ~bad & (ho["instance"] != 0xFFFFFFFF)
        assert np.array_equal(hg["t"][both], ho["t"][both])


@pytest.mark.parametrize("flags", [0, HH_TRACE_OPAQUE, HH_TRACE_OPAQUE | HH_TRACE_TERMINATE_ON_FIRST])
def test_random_rays_cornell(cornell_pair, flags):
    g, o = cornell_pair
    rays = scenegen.random_rays(20000, (-0.99, 0.01, -0.99), (0.99, 1.98, 2.0), seed=11, tmin=1e-4)
    hg = g.trace_rays(rays, flags)
    ho = o.trace_rays(rays, flags, brute=True)
    if flags & HH_TRACE_TERMINATE_ON_FIRST:
        # any-hit: only the hit/miss decision is defined
        assert np.array_equal(hg["instance"] == 0xFFFFFFFF, ho["instance"] == 0xFFFFFFFF)
        return
    assert not id_mismatch(hg, ho).any()
    assert np.array_equal(hg["t"], ho["t"])
    assert np.array_equal(hg["u"], ho["u"]) and np.array_equal(hg["v"], ho["v"])


@pytest.mark.parametrize("flags", [0, HH_TRACE_OPAQUE])
def test_random_rays_field_transforms_alpha(field_pair, flags):
    g, o = field_pair
    rays = scenegen.random_rays(30000, (-90, 0.5, -90), (90, 25, 90), seed=5, tmin=1e-4)
    hg = g.trace_rays(rays, flags)
    ho = o.trace_rays(rays, flags, brute=False)
    bad = id_mismatch(hg, ho)
    assert bad.mean() <= EDGE_TIE_FRACTION, f"{bad.sum()} mismatching rays"
    ok = ~bad & (ho["instance"] != 0xFFFFFFFF)
    assert np.array_equal(hg["t"][ok], ho["t"][ok])


def test_scene_packing_matches(cornell_pair, field_pair):
    for g, o in (cornell_pair, field_pair):
        assert g.num_lights() == o.num_lights
        assert np.array_equal(g.lights(), o.lights())


def test_sky_bake_matches_oracle(field_pair):
    g, o = field_pair
    assert np.array_equal(g.sky_coefficients(), o.sky_coefficients())
    eg, eo = g.env(), o.env()
    assert eg.shape == eo.shape
    rel = np.abs(eg[..., :3] - eo[..., :3]) / np.maximum(np.abs(eo[..., :3]), 1e-6)
    assert rel.max() < 2e-5, rel.max()


def _render_gpu(ctx, scene, desc, frames, first=0, flags=0, width=None, height=None, bounces=None, camera=None):
    from helios_amd import runtime
    w, h = width or desc.width, height or desc.height
    surf = runtime.Surface(ctx, w, h)
    it = runtime.PathIntegrator(ctx)
    it.set_max_ray_bounces(bounces or desc.max_ray_bounces)
    it.set_max_samples(100000)
    it.set_flags(flags)
    it.render(scene, camera or desc.camera, surf, frames=frames)
    img = surf.read()
    stats = it.ray_stats()
    it.close()
    surf.close()
    return img, stats


def test_radiance_cornell_matches_fp32_transcription(gpu_ctx, cornell, cornell_pair):
    g, o = cornell_pair
    img_g, _ = _render_gpu(gpu_ctx, g, cornell, frames=16)
    img_o = o.render(frames=16)
    assert relmse(img_g, img_o) < 1e-4
    assert np.isfinite(img_g).all()


def test_radiance_cornell_4096spp_crop(gpu_ctx, cornell, cornell_pair, oracle_mod):
    """4096 progressive samples: GPU vs FP32 transcription and FP64 oracle on a 24x24 crop."""
    g, o = cornell_pair
    img_g, _ = _render_gpu(gpu_ctx, g, cornell, frames=4096)
    crop = (116, 70, 140, 94)
    x0, y0, x1, y1 = crop
    img32 = o.render(frames=4096, crop=crop)
    img64 = o.render(frames=4096, crop=crop, fp64=True)
    cg = img_g[y0:y1, x0:x1]
    assert relmse(cg, img32[y0:y1, x0:x1]) < RELMSE_BAR
    assert relmse(cg, img64[y0:y1, x0:x1]) < RELMSE_BAR


def test_radiance_field_matches_oracle(gpu_ctx, small_field, field_pair):
    g, o = field_pair
    img_g, _ = _render_gpu(gpu_ctx, g, small_field, frames=4)
    img_o = o.render(frames=4)
    assert relmse(img_g, img_o) < RELMSE_BAR


@pytest.mark.parametrize("bounces", [1, 2, 8])
def test_megakernel_equals_wavefront(gpu_ctx, small_field, field_pair, bounces):
    """The single-frame form (camera rays' bounce as wavefront launches, the
    rest of every path in one megakernel launch; with one bounce no
    megakernel runs) renders the wavefront's image bit for bit."""
    g, _ = field_pair
    a, sa = _render_gpu(gpu_ctx, g, small_field, frames=3, bounces=bounces)
    b, sb = _render_gpu(gpu_ctx, g, small_field, frames=3, flags=HH_FLAG_MEGAKERNEL, bounces=bounces)
    assert np.array_equal(a, b)
    assert sa["extension_rays"] == sb["extension_rays"] and sa["shadow_rays"] == sb["shadow_rays"]


@pytest.mark.parametrize("cont", [2, 3, 7])
def test_continuation_depth_equals_wavefront(gpu_ctx, small_field, field_pair, cont):
    """Wavefront launches for the first `cont` bounces, then one continuation
    launch for the rest of every path: the same image, bit for bit, and the
    same ray counts, with several frames per pass."""
    g, _ = field_pair
    gpu_ctx.set_tuning(frames_in_flight=3)
    try:
        a, sa = _render_gpu(gpu_ctx, g, small_field, frames=5)
        gpu_ctx.set_tuning(frames_in_flight=3, continuation_depth=cont)
        b, sb = _render_gpu(gpu_ctx, g, small_field, frames=5)
    finally:
        gpu_ctx.set_tuning()
    assert np.array_equal(a, b)
    assert sa == sb


def test_tonemap_matches_oracle(gpu_ctx, oracle_mod):
    from helios_amd import runtime
    rng = np.random.default_rng(0)
    img = (rng.random((64, 96, 4)) * np.array([4, 2, 1, 1])).astype(np.float32)
    surf = runtime.Surface(gpu_ctx, 96, 64)
    surf.write(img)
    for op, exposure in ((0, 1.0), (1, 0.7), (2, 1.3)):
        a = surf.tone_map(exposure, op)
        b = oracle_mod.tonemap(img, exposure, op)
        diff = np.abs(a.astype(int) - b.astype(int))
        assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    surf.close()


def test_progressive_protocol(gpu_ctx, cornell, cornell_pair):
    """Frames rendered in separate calls accumulate exactly like one call."""
    from helios_amd import runtime
    g, _ = cornell_pair
    a, _ = _render_gpu(gpu_ctx, g, cornell, frames=5)
    surf = runtime.Surface(gpu_ctx, cornell.width, cornell.height)
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(cornell.max_ray_bounces)
    for _ in range(5):
        it.render(g, cornell.camera, surf)
    assert it.num_accumulated_samples() == 5
    b = surf.read()
    assert np.array_equal(a, b)
    it.restart_bake()
    it.render(g, cornell.camera, surf)  # frame 0 overwrites
    c = surf.read()
    one, _ = _render_gpu(gpu_ctx, g, cornell, frames=1)
    assert np.array_equal(c, one)


@pytest.mark.parametrize("layout", [dict(bvh_width=8), dict(bvh_width=4, bvh_compressed=True),
                                    dict(bvh_builder=0), dict(bvh_builder=0, bvh_width=8),
                                    dict(bvh_builder=2), dict(bvh_builder=2, bvh_width=8),
                                    dict(bvh_builder=2, bvh_compressed=True),
                                    dict(bvh_builder=3), dict(bvh_builder=3, bvh_width=8)])
def test_bvh_layouts_bit_identical(gpu_ctx, small_field, field_pair, oracle_mod, layout):
    """BVH8, the quantized BVH4, the LBVH, the clustered (PLOC) and the
    GPU-built binned-SAH trees return the same hits as the oracle and the same
    images as the host-built binned-SAH float BVH4 (traversal results are
    independent of the tree's shape and visit order)."""
    from helios_amd import runtime
    g4, o = field_pair
    ref_img, _ = _render_gpu(gpu_ctx, g4, small_field, frames=3)
    gpu_ctx.set_tuning(**layout)
    try:
        g = runtime.Scene(gpu_ctx, small_field)
        rays = scenegen.random_rays(30000, (-90, 0.5, -90), (90, 25, 90), seed=8, tmin=1e-4)
        for flags in (0, HH_TRACE_OPAQUE):
            hg = g.trace_rays(rays, flags)
            ho = o.trace_rays(rays, flags)
            assert not id_mismatch(hg, ho).any()
            assert np.array_equal(hg["t"], ho["t"])
        img, _ = _render_gpu(gpu_ctx, g, small_field, frames=3)
        assert np.array_equal(img, ref_img)
        g.close()
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("level,width", [(3, 4), (5, 4), (5, 8)])
def test_gpu_sah_builder_builds_the_host_tree(gpu_ctx, level, width):
    """The GPU binned-SAH builder (bvh_builder 3) makes the host builder's
    splits: same node counts, and the counting traversal fetches exactly the
    same nodes, triangles and instances for every ray of three frames, at
    mesh sizes that use the block-wide level passes (level 5: 20,480
    triangles per sphere) and the one-wave subtree builds."""
    from helios_amd import runtime
    d = scenegen.icosphere_field(seed=4, n_spheres=5, level=level, n_leaves=300, width=96, height=64,
                                 max_ray_bounces=4)
    out = {}
    try:
        for b in (1, 3):
            gpu_ctx.set_tuning(bvh_builder=b, bvh_width=width)
            sc = runtime.Scene(gpu_ctx, d)
            gpu_ctx.set_profiling(timing=True, counting=True)
            gpu_ctx.reset_profile()
            img, stats = _render_gpu(gpu_ctx, sc, d, frames=3)
            counts = gpu_ctx.read_profile()["counts"]
            gpu_ctx.set_profiling()
            out[b] = (sc.bvh_stats(), counts, img, stats)
            sc.close()
    finally:
        gpu_ctx.set_profiling()
        gpu_ctx.set_tuning()
    assert out[1][0] == out[3][0]
    assert out[1][1] == out[3][1] and out[1][1]["ext_nodes"] > 0
    assert np.array_equal(out[1][2], out[3][2])
    assert out[1][3] == out[3][3]


@pytest.mark.parametrize("fif", [4, 8, 64])
def test_camera_ray_packets_bit_identical(gpu_ctx, small_field, field_pair, fif):
    """Camera rays are traced in wave-uniform packets (from 4 samples per
    pixel in flight) except in the counting variant, which traces every ray on
    its own: the images and the ray counts agree exactly (4 frames in flight:
    packets of 16 neighbouring pixels; 64: the 64 samples of one pixel)."""
    g, _ = field_pair
    gpu_ctx.set_tuning(frames_in_flight=fif)
    try:
        packed, s0 = _render_gpu(gpu_ctx, g, small_field, frames=fif)
        gpu_ctx.set_profiling(timing=True, counting=True)
        lanes, s1 = _render_gpu(gpu_ctx, g, small_field, frames=fif)
    finally:
        gpu_ctx.set_profiling()
        gpu_ctx.set_tuning()
    assert np.array_equal(packed, lanes)
    assert s0 == s1


@pytest.mark.parametrize("tune", [dict(persistent_blocks_per_cu=1), dict(persistent_blocks_per_cu=1, extend_block=256),
                                  dict(persistent_blocks_per_cu=3, extend_block=128, frames_in_flight=7)])
def test_launch_shapes_bit_identical(gpu_ctx, small_field, field_pair, tune):
    """The ray queues are split into segments sized from the launch shape
    (a producer wave writes at most one extension and one shadow ray per entry
    it reads); small persistent grids put many entries on each wave and fill
    the segments to their bound. Every shape renders the same image."""
    g, _ = field_pair
    gpu_ctx.set_tuning(frames_in_flight=tune.get("frames_in_flight", 4))
    ref, s0 = _render_gpu(gpu_ctx, g, small_field, frames=7)
    gpu_ctx.set_tuning(**tune)
    try:
        img, s1 = _render_gpu(gpu_ctx, g, small_field, frames=7)
    finally:
        gpu_ctx.set_tuning()
    assert np.array_equal(img, ref)
    assert s0 == s1


def test_frames_in_flight_bit_identical(gpu_ctx, small_field, field_pair):
    """Tracing several progressive frames per wavefront pass changes nothing."""
    g, _ = field_pair
    imgs = []
    for fif in (1, 3, 8):
        gpu_ctx.set_tuning(frames_in_flight=fif)
        imgs.append(_render_gpu(gpu_ctx, g, small_field, frames=7)[0])
    gpu_ctx.set_tuning()
    assert np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])


def test_tiled_mode_matches_untiled(gpu_ctx, oracle_mod):
    """Tiled progressive mode (128x128 tiles, max_samples each) ends at the same image."""
    from helios_amd import runtime
    d = scenegen.cornell_box(width=200, height=150)
    g = runtime.Scene(gpu_ctx, d)
    surf = runtime.Surface(gpu_ctx, 200, 150)
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(5)
    it.set_max_samples(3)
    it.set_tiled(True)
    it.render(g, d.camera, surf, frames=3 * 4)
    assert it.tile_idx() == 4
    tiled = surf.read()
    untiled, _ = _render_gpu(gpu_ctx, g, d, frames=3)
    assert np.array_equal(tiled, untiled)


def test_render_batch_matches_single_views(gpu_ctx, cornell, small_field, cornell_pair, field_pair):
    from helios_amd import runtime
    g1, _ = cornell_pair
    g2, _ = field_pair
    s1 = runtime.Surface(gpu_ctx, cornell.width, cornell.height)
    s2 = runtime.Surface(gpu_ctx, small_field.width, small_field.height)
    runtime.render_batch(gpu_ctx, [
        dict(scene=g1, surface=s1, camera=cornell.camera, spp=3, max_ray_bounces=5),
        dict(scene=g2, surface=s2, camera=small_field.camera, spp=2, max_ray_bounces=8),
    ])
    a1, _ = _render_gpu(gpu_ctx, g1, cornell, frames=3)
    a2, _ = _render_gpu(gpu_ctx, g2, small_field, frames=2)
    assert np.array_equal(s1.read(), a1)
    assert np.array_equal(s2.read(), a2)


def test_rr_start_depth_matches_oracle(gpu_ctx, cornell, cornell_pair):
    """Russian roulette from bounce 3 (the small workload's setting): same
    estimator on both sides, same RNG streams (the RR draw is always taken)."""
    from helios_amd import runtime
    g, o = cornell_pair
    d = cornell
    surf = runtime.Surface(gpu_ctx, d.width, d.height)
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(d.max_ray_bounces)
    it.set_rr_start_depth(3)
    it.render(g, d.camera, surf, frames=16)
    ref = o.render(frames=16, rr_start_depth=3)
    assert relmse(surf.read(), ref) < 1e-4
    # and it is a different estimator from RR at every bounce
    base = o.render(frames=16)
    assert relmse(base, ref) > 1e-5


@pytest.fixture(scope="module")
def combined_pair(gpu_ctx, oracle_mod):
    """The combined workload's scene (1,000 instances of 40 meshes, emissive
    panels, sky + sun, 16 bounces) at a reduced image size."""
    from helios_amd import runtime
    d, cams = scenegen.combined_workload(width=480, height=270)
    return d, cams, runtime.Scene(gpu_ctx, d), oracle_mod.OracleScene(d)


def test_combined_workload_primary_ids_bit_exact(combined_pair):
    d, cams, g, o = combined_pair
    for cam in cams[:2]:
        hg = g.primary_hits(cam, d.width, d.height, 0)
        ho = o.primary_hits(camera=cam, frame=0, brute=False)
        bad = id_mismatch(hg, ho)
        assert bad.mean() <= EDGE_TIE_FRACTION, f"{bad.sum()} mismatching pixels"
        both = ~bad & (ho["instance"] != 0xFFFFFFFF)
        assert np.array_equal(hg["t"][both], ho["t"][both])


def test_combined_workload_radiance_matches_oracle(gpu_ctx, combined_pair):
    """Instanced geometry, area lights from the emissive panels, 16 bounces:
    GPU vs the FP32 transcription on a crop, same RNG streams."""
    d, cams, g, o = combined_pair
    img_g, _ = _render_gpu(gpu_ctx, g, d, frames=16, camera=cams[1])
    crop = (200, 100, 248, 148)
    x0, y0, x1, y1 = crop
    img_o = o.render(camera=cams[1], frames=16, crop=crop)
    assert relmse(img_g[y0:y1, x0:x1], img_o[y0:y1, x0:x1]) < 1e-4


@pytest.fixture(scope="module")
def combined_ext_pair(gpu_ctx, oracle_mod):
    """The combined workload with its two extensions: an equirectangular
    RGBA16F sky and power-proportional emitter sampling."""
    from helios_amd import runtime
    d, cams = scenegen.combined_workload(width=480, height=270, equirect_width=2048)
    return d, cams, runtime.Scene(gpu_ctx, d), oracle_mod.OracleScene(d)


def test_equirect_sky_bake_matches_oracle(combined_ext_pair):
    _, _, g, o = combined_ext_pair
    eg, eo = g.env_equirect(), o.env_equirect()
    assert eg is not None and eg.shape == (1024, 2048, 4)
    # both round the same float evaluation to half; the float sky terms may
    # differ in the last bits (transcendentals), so allow one half ulp
    ulp = np.abs(eg.view(np.int16).astype(np.int32) - eo.view(np.int16).astype(np.int32))
    assert ulp.max() <= 1 and (ulp == 0).mean() > 0.99


def test_light_power_and_equirect_match_oracle(gpu_ctx, combined_ext_pair):
    from helios_amd.abi import HH_FLAG_LIGHT_POWER
    d, cams, g, o = combined_ext_pair
    img_g, _ = _render_gpu(gpu_ctx, g, d, frames=16, camera=cams[1], flags=HH_FLAG_LIGHT_POWER)
    crop = (200, 100, 248, 148)
    x0, y0, x1, y1 = crop
    img_o = o.render(camera=cams[1], frames=16, crop=crop, flags=HH_FLAG_LIGHT_POWER)
    assert relmse(img_g[y0:y1, x0:x1], img_o[y0:y1, x0:x1]) < 1e-4
    # a different estimator from the reference's area lights
    img_r = o.render(camera=cams[1], frames=16, crop=crop)
    assert relmse(img_r[y0:y1, x0:x1], img_o[y0:y1, x0:x1]) > 1e-6
