"""GPU: edge cases of the path against the oracle.

  * an empty scene (no instances): every ray misses; black without a sky,
    the baked sky when a directional light brings one (Q16);
  * a zero-area triangle and a mesh without triangles (the build gives it
    one inert triangle) change nothing;
  * 1x1 and odd-sized views: primary IDs bit-exact, radiance equal to the
    FP32 transcription (relMSE < 1e-4, the bar of the main parity tests);
  * a one-bounce integrator (max_ray_bounces 1: emission + NEE only);
  * the reference's instance limit (scene.h:14, 1024): 1024 instances commit,
    the 1025th is refused with HH_ERR_LIMIT_EXCEEDED;
  * a non-finite vertex position is refused with HH_ERR_INVALID_ARG.
"""
import numpy as np
import pytest

from helios_amd import scenegen
from helios_amd.abi import LIGHT_DIRECTIONAL, LIGHT_POINT

pytestmark = pytest.mark.gpu


def relmse(x, r, eps=0.01):
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(r, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def render_gpu(ctx, desc, frames, width=None, height=None, bounces=None):
    from helios_amd import runtime
    w, h = width or desc.width, height or desc.height
    scene = runtime.Scene(ctx, desc)
    surf = runtime.Surface(ctx, w, h)
    it = runtime.PathIntegrator(ctx)
    it.set_max_ray_bounces(bounces or desc.max_ray_bounces)
    it.render(scene, desc.camera, surf, frames=frames)
    img = surf.read()
    hits = scene.primary_hits(desc.camera, w, h, 1)
    return img, hits


def same_hits(a, b):
    """IDs everywhere, t / u / v bit-exact where a surface was hit (u, v of a
    miss are unspecified)."""
    for k in ("instance", "geometry", "primitive"):
        if not np.array_equal(a[k], b[k]):
            return False
    hit = a["instance"] != 0xFFFFFFFF
    return all(np.array_equal(a[k][hit], b[k][hit]) for k in ("t", "u", "v"))


def sun(elev=30.0, az=40.0):
    return dict(type=LIGHT_DIRECTIONAL, color=(1.0, 0.95, 0.9), intensity=2.0, radius=0.1, position=(0, 0, 0),
                forward=tuple(scenegen.sun_forward(elev, az)), inner_cone_deg=40.0, outer_cone_deg=50.0)


@pytest.mark.parametrize("with_sun", [False, True])
def test_empty_scene(gpu_ctx, oracle_mod, with_sun):
    d = scenegen.SceneDesc(width=24, height=16, max_ray_bounces=4)
    d.camera = scenegen.look_at_camera((0.0, 1.0, 3.0), (0.0, 1.2, 0.0), fov_y=70.0)
    if with_sun:
        d.lights.append(sun())
    img, hits = render_gpu(gpu_ctx, d, 3)
    assert np.all(hits["instance"] == 0xFFFFFFFF) and np.all(hits["t"] < 0)
    ref = oracle_mod.OracleScene(d).render(frames=3)
    if with_sun:
        assert img[..., :3].max() > 0.05
        np.testing.assert_allclose(img, ref, rtol=1e-4, atol=1e-6)
    else:
        assert not img[..., :3].any()
        assert np.all(img[..., 3] == 1.0)


def test_degenerate_triangle_and_empty_mesh_change_nothing(gpu_ctx, oracle_mod):
    base = scenegen.cornell_box(width=32, height=24)
    d = scenegen.cornell_box(width=32, height=24)
    d.materials.append(scenegen.material())
    V = np.array([[0.0, 1.0, 0.0]] * 3)
    d.meshes.append(scenegen._mesh_from_parts([(V, np.tile([0.0, 0.0, 1.0], (3, 1)), None,
                                                np.array([0, 1, 2], np.uint32), 0)], [len(d.materials) - 1]))
    d.instances.append(scenegen.identity_instance(len(d.meshes) - 1))
    d.meshes.append(scenegen.MeshDesc(scenegen.make_vertices(np.zeros((1, 3)), np.array([[0.0, 0.0, 1.0]])),
                                      np.zeros(0, np.uint32), [(0, 0, 1, 0, 0)], [0]))
    d.instances.append(scenegen.identity_instance(len(d.meshes) - 1))
    img, hits = render_gpu(gpu_ctx, d, 4)
    img0, hits0 = render_gpu(gpu_ctx, base, 4)
    assert same_hits(hits, hits0)
    assert np.array_equal(img, img0)
    assert same_hits(hits, oracle_mod.OracleScene(d).primary_hits(frame=1, brute=True))


@pytest.mark.parametrize("size", [(1, 1), (7, 3), (129, 65)])
def test_odd_view_sizes(gpu_ctx, oracle_mod, size):
    w, h = size
    d = scenegen.cornell_box(width=w, height=h)
    img, hits = render_gpu(gpu_ctx, d, 16)
    o = oracle_mod.OracleScene(d)
    assert same_hits(hits, o.primary_hits(frame=1, brute=False))
    ref = o.render(frames=16)
    assert relmse(img, ref) < 1e-4, relmse(img, ref)


def test_one_bounce(gpu_ctx, oracle_mod):
    d = scenegen.cornell_box(width=48, height=40, max_ray_bounces=1)
    img, _ = render_gpu(gpu_ctx, d, 16, bounces=1)
    ref = oracle_mod.OracleScene(d).render(frames=16, max_ray_bounces=1)
    assert relmse(img, ref) < 1e-4
    # no indirect light: the ceiling outside the light sees nothing but NEE
    assert img[..., :3].max() > 0.0


def test_instance_limit(gpu_ctx):
    from helios_amd import runtime
    from helios_amd.runtime import HeliosError
    d = scenegen.SceneDesc(width=8, height=8)
    d.materials.append(scenegen.material())
    V = np.array([[0.0, 0.0, 0.0], [1.4, 0.0, 0.0], [0.0, 1.4, 0.0]])
    d.meshes.append(scenegen._mesh_from_parts([(V, np.tile([0.0, 0.0, 1.0], (3, 1)), None,
                                                np.array([0, 1, 2], np.uint32), 0)], [0]))
    for i in range(1024):
        m, n = scenegen.trs(translate=(float(i % 32) * 1.5, float(i // 32) * 1.5, 0.0))
        d.instances.append((m, n, 0, -1))
    d.camera = scenegen.look_at_camera((24.0, 24.0, 50.0), (24.0, 24.0, 0.0), fov_y=60.0)
    scene = runtime.Scene(gpu_ctx, d)
    hits = scene.primary_hits(d.camera, 256, 256, 0)
    assert len(np.unique(hits["instance"][hits["t"] > 0])) == 1024
    with pytest.raises(HeliosError) as e:
        extra = runtime.Scene(gpu_ctx, None, commit=False)
        d.instances.append(d.instances[0])
        extra.load(d)
    assert e.value.status == 3


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_vertex_position_is_refused(gpu_ctx, bad):
    """A NaN or infinite vertex position is refused at hh_scene_add_mesh
    (HH_ERR_INVALID_ARG) before any BVH build sees it."""
    from helios_amd import runtime
    from helios_amd.runtime import HeliosError
    d = scenegen.SceneDesc(width=8, height=8)
    d.materials.append(scenegen.material())
    V = np.array([[0.0, 0.0, 0.0], [1.0, bad, 0.0], [0.0, 1.0, 0.0]])
    d.meshes.append(scenegen._mesh_from_parts([(V, np.tile([0.0, 0.0, 1.0], (3, 1)), None,
                                                np.array([0, 1, 2], np.uint32), 0)], [0]))
    d.instances.append(scenegen.identity_instance(0))
    with pytest.raises(HeliosError) as e:
        runtime.Scene(gpu_ctx, d)
    assert e.value.status == 1


def _quad(x0, x1, y0, y1, z):
    V = np.array([[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]], np.float64)
    N = np.tile([0.0, 0.0, 1.0], (4, 1))
    U = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64)
    return V, N, U, np.array([0, 1, 2, 0, 2, 3], np.uint32)


def test_alpha_sources(gpu_ctx, oracle_mod):
    """The any-hit alpha test reads one alpha source per instance when all its
    non-opaque geometry shares it (a texture slot or a constant alpha) and
    looks the material up per geometry otherwise; a material override of the
    instance supplies the alpha while the mesh's own material decides which
    geometry is tested. Every case returns the oracle's hits."""
    from helios_amd import runtime
    d = scenegen.SceneDesc(width=128, height=64, max_ray_bounces=2)
    tex = scenegen.leaf_texture(size=32)
    d.textures.append((tex.shape[1], tex.shape[0], scenegen.TEXTURE_RGBA8_SRGB, tex))
    d.materials += [scenegen.material((0.5, 0.5, 0.5, 1.0)),                                # 0 opaque
                    scenegen.material((1, 1, 1, 1), albedo_texture=0, alpha_test=True),     # 1 textured alpha
                    scenegen.material((1, 1, 1, 0.05), alpha_test=True),                   # 2 constant, cut away
                    scenegen.material((1, 1, 1, 1.0), alpha_test=True)]                    # 3 constant, solid
    # mesh 0: backdrop; mesh 1: two alpha geometries with different sources (mixed);
    # mesh 2: one constant-alpha quad, instanced with and without an override
    d.meshes.append(scenegen._mesh_from_parts([(*_quad(-8, 8, -4, 4, -1.0), 0)], [0]))
    d.meshes.append(scenegen._mesh_from_parts([(*_quad(-4.5, -2.5, -1, 1, 0.0), 0),
                                               (*_quad(-2.3, -0.3, -1, 1, 0.0), 1)], [1, 2]))
    d.meshes.append(scenegen._mesh_from_parts([(*_quad(-1, 1, -1, 1, 0.0), 0)], [3]))
    d.instances.append(scenegen.identity_instance(0))
    d.instances.append(scenegen.identity_instance(1))
    for x, ov in ((1.1, 1), (3.3, -1)):
        m, n = scenegen.trs(translate=(x, 0.0, 0.0))
        d.instances.append((m, n, 2, ov))
    # a point light in front: the backdrop's shadow rays cross every quad
    d.lights.append(dict(type=LIGHT_POINT, color=(1.0, 1.0, 1.0), intensity=20.0, radius=0.2, position=(0.0, 0.3, 2.5),
                         forward=(0.0, 0.0, -1.0), inner_cone_deg=40.0, outer_cone_deg=50.0))
    d.camera = scenegen.look_at_camera((0.0, 0.0, 6.0), (0.0, 0.0, 0.0), fov_y=40.0)
    scene = runtime.Scene(gpu_ctx, d)
    o = oracle_mod.OracleScene(d)
    for frame in (0, 1):
        g = scene.primary_hits(d.camera, d.width, d.height, frame)
        r = o.primary_hits(frame=frame, brute=True)
        assert same_hits(g, r)
        inst, geom = g["instance"], g["geometry"]
        assert np.any((inst == 1) & (geom == 0))             # textured geometry of the mixed mesh
        assert not np.any((inst == 1) & (geom == 1))         # constant alpha 0.05: never accepted
        assert np.any(inst == 2) and np.any(inst == 0)       # override texture: hits and see-through
        assert np.any(inst == 3)                             # constant alpha 1.0: solid
    img, _ = render_gpu(gpu_ctx, d, 8)
    ref = o.render(frames=8)
    assert relmse(img, ref) < 1e-4
