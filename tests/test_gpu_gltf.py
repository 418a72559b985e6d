"""GPU: a scene read with hh_scene_load_gltf renders like the scene it was
exported from (scenegen.write_gltf): identical primary hits and radiance
equal to float rounding (the albedo factor passes through v^(1/2.2) and
back), on the wavefront path and in the output-buffer views; for the
alpha-tested field and for the textured showcase (normal map, roughness and
metallic channels, float emission texture, spot light)."""
import numpy as np
import pytest

from helios_amd import scenegen
from helios_amd.abi import OUTPUT_BUFFER_ALBEDO, OUTPUT_BUFFER_NORMALS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,suffix", [("field", ".gltf"), ("field", ".glb"), ("textured", ".gltf")])
def test_loaded_scene_renders_like_the_original(gpu_ctx, tmp_path, which, suffix):
    from helios_amd import runtime
    if which == "field":
        d = scenegen.icosphere_field(seed=5, n_spheres=10, level=3, n_leaves=300, width=160, height=96,
                                     max_ray_bounces=6)
    else:
        d = scenegen.textured_showcase(width=128, height=96, max_ray_bounces=6)
    for l in d.lights:
        l["radius"] = 0.0                      # glTF lights are punctual
    path = scenegen.write_gltf(d, tmp_path / (which + suffix))
    a = runtime.Scene(gpu_ctx, d)
    b = runtime.Scene(gpu_ctx, None, commit=False)
    info = b.load_gltf(path)
    b.load(scenegen.SceneDesc(sky=d.sky))      # the sky is not part of the file
    b.commit()
    assert info["num_instances"] == len(d.instances) and info["num_lights"] == len(d.lights)
    assert info["num_triangles"] == d.num_triangles
    cam = d.camera
    ha, hb = a.primary_hits(cam, d.width, d.height, 3), b.primary_hits(cam, d.width, d.height, 3)
    for k in ("instance", "geometry", "primitive"):
        assert np.array_equal(ha[k], hb[k])
    imgs = []
    for s in (a, b):
        surf = runtime.Surface(gpu_ctx, d.width, d.height)
        it = runtime.PathIntegrator(gpu_ctx)
        it.set_max_ray_bounces(6)
        it.render(s, cam, surf, frames=4)
        views = []
        for which in (OUTPUT_BUFFER_ALBEDO, OUTPUT_BUFFER_NORMALS):
            v = runtime.Surface(gpu_ctx, d.width, d.height)
            runtime.render_output_buffer(s, cam, which, v)
            views.append(v.read())
        imgs.append((surf.read(), views))
    np.testing.assert_allclose(imgs[0][0], imgs[1][0], atol=5e-5)
    for va, vb in zip(imgs[0][1], imgs[1][1]):
        np.testing.assert_allclose(va, vb, atol=2e-6)


def test_ibl_from_hdr_file(gpu_ctx, oracle_mod, tmp_path):
    """hh_scene_load_ibl_hdr: the GPU scene lit by a .hdr environment renders
    like the oracle given the faces hh_ibl_from_hdr produces."""
    from helios_amd import runtime
    from test_hdr_ibl import latlong_dirs, write_hdr
    w, h = 256, 128
    write_hdr(tmp_path / "env.hdr", 0.4 + 0.3 * latlong_dirs(w, h) + np.array([0.2, 0.0, -0.1]))
    faces = runtime.ibl_from_hdr(tmp_path / "env.hdr", 32)
    d = scenegen.cornell_box(width=40, height=32, max_ray_bounces=3)
    d.ibl = (32, faces)
    scene = runtime.Scene(gpu_ctx, None, commit=False)
    blank = scenegen.cornell_box(width=40, height=32, max_ray_bounces=3)
    scene.load(blank)
    scene.load_ibl_hdr(tmp_path / "env.hdr", 32)
    scene.commit()
    assert np.array_equal(scene.env()[..., :4].reshape(faces.shape), faces)
    surf = runtime.Surface(gpu_ctx, d.width, d.height)
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(3)
    it.render(scene, d.camera, surf, frames=8)
    ref = oracle_mod.OracleScene(d).render(frames=8, max_ray_bounces=3)
    x, r = surf.read()[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    assert float(np.mean((x - r) ** 2 / (r * r + 0.01))) < 1e-4
