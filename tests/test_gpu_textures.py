"""Textured materials and spot lights: the GPU path against the oracle.

The showcase scene (scenegen.textured_showcase) puts every texture branch of
populate_surface_properties (path_trace_rchit.glsl:190-252) on screen: an
sRGB albedo texture, a tangent-space normal map, roughness read from the G
channel and metallic from the B channel of two separate textures, an
emissive RGBA32F texture (on an area-light mesh), plus a spot light with a
smoothstep cone (path_trace_rchit.glsl:319-341), a point light and the sun.

Bars: primary-hit IDs, t, u, v bit-exact; radiance within relMSE 1e-4 of the
FP32 transcription at 16 spp (the same per-pixel RNG streams); the
single-frame form equals the wavefront bit for bit.
"""
import numpy as np
import pytest

from helios_amd import scenegen
from helios_amd.abi import HH_FLAG_MEGAKERNEL


def relmse(x, r, eps=0.01):
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(r, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


@pytest.fixture(scope="module")
def showcase():
    return scenegen.textured_showcase()


def test_showcase_packs_every_texture_branch(oracle_mod, showcase):
    """CPU: the packed material records carry the normal, roughness (G),
    metallic (B) and emissive texture slots, and the spot light is packed."""
    o = oracle_mod.OracleScene(showcase)
    recs = np.zeros((8, 20), np.int32)
    n = oracle_mod.lib().orc_scene_read_materials(o.h, oracle_mod.ptr(recs), 8)
    assert n == 3
    tex0, tex1 = recs[:n, 0:4], recs[:n, 4:8]
    ground = 0
    assert tex0[ground, 0] >= 0 and tex0[ground, 1] >= 0            # albedo, normal map
    assert tex0[ground, 2] >= 0 and tex1[ground, 2] == 1             # roughness from G
    assert tex0[ground, 3] >= 0 and tex1[ground, 3] == 2             # metallic from B
    assert (tex1[:, 0] >= 0).sum() == 1                              # one emissive texture
    kinds = sorted(int(k) for k in o.lights()[:, 0])
    assert kinds == [0, 1, 2, 3, 4]                                  # dir, spot, point, env, area


@pytest.fixture(scope="module")
def showcase_pair(gpu_ctx, oracle_mod, showcase):
    from helios_amd import runtime
    return runtime.Scene(gpu_ctx, showcase), oracle_mod.OracleScene(showcase)


def _render(ctx, scene, desc, frames, flags=0):
    from helios_amd import runtime
    surf = runtime.Surface(ctx, desc.width, desc.height)
    it = runtime.PathIntegrator(ctx)
    it.set_max_ray_bounces(desc.max_ray_bounces)
    it.set_max_samples(100000)
    it.set_flags(flags)
    it.render(scene, desc.camera, surf, frames=frames)
    img = surf.read()
    it.close()
    surf.close()
    return img


@pytest.mark.gpu
def test_showcase_primary_ids_bit_exact(showcase, showcase_pair):
    g, o = showcase_pair
    d = showcase
    for frame in (0, 3):
        hg = g.primary_hits(d.camera, d.width, d.height, frame)
        ho = o.primary_hits(frame=frame, brute=True)
        for k in ("instance", "geometry", "primitive", "t", "u", "v"):
            assert np.array_equal(hg[k], ho[k]), (frame, k)


@pytest.mark.gpu
def test_showcase_radiance_matches_oracle(gpu_ctx, showcase, showcase_pair):
    g, o = showcase_pair
    img_g = _render(gpu_ctx, g, showcase, 16)
    img_o = o.render(frames=16)
    assert np.isfinite(img_g).all()
    e = relmse(img_g, img_o)
    assert e < 1e-4, e
    # the textured branches matter: the image is not that of constant materials
    assert img_o[..., :3].std() > 0.05


@pytest.mark.gpu
def test_showcase_megakernel_equals_wavefront(gpu_ctx, showcase, showcase_pair):
    g, _ = showcase_pair
    a = _render(gpu_ctx, g, showcase, 3)
    b = _render(gpu_ctx, g, showcase, 3, flags=HH_FLAG_MEGAKERNEL)
    assert np.array_equal(a, b)
