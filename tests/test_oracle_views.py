"""CPU: the oracle's output-buffer views and ray-debug variant.

The reference's views are a raster pass (debug_visualization.vert/.frag,
renderer.cpp:491-547, pipeline state :1178-1300) and the debug variant is
path_trace_debug.rgen with the RAY_DEBUG_VIEW branches of path_trace_rgen.glsl,
path_trace_rchit.glsl and path_trace_rmiss.glsl. Neither can run here; what
pins the restatement is independent construction:
  * pixel-centre rays built in float64 from the camera's own definition
    (quaternion, vertical fov, aspect; row 0 = NDC y -1) and traced by brute
    force give the surface each pixel must show, whose material, normals and
    transforms come straight from the scene description;
  * the clip range is the Vulkan depth range [0, 1] of glm's GL-style
    projection, i.e. view depth in [2fn/(f+n), f]: a surface at depth 1.5 with
    n = 1 is clipped, and kept with n = 0.5;
  * the debug paths never meet Russian roulette: inside a closed box every
    path has exactly max_ray_bounces - 1 recorded segments, chained end to
    origin; colours are the first three draws of RNG(i, 0, frame); the first
    origin is the primary hit of the camera ray that follows those draws.
"""
import math

import numpy as np
import pytest

from helios_amd import scenegen
from helios_amd.abi import (HH_TRACE_OPAQUE, OUTPUT_BUFFER_ALBEDO, OUTPUT_BUFFER_EMISSIVE, OUTPUT_BUFFER_METALLIC,
                            OUTPUT_BUFFER_NORMALS, OUTPUT_BUFFER_ROUGHNESS, RAY_DTYPE, TEXTURE_RGBA8_SRGB)

W, H = 48, 40


def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def centre_rays(cam, w, h):
    """Pixel-centre rays in float64 clipped to view depth [2fn/(f+n), f]."""
    R = quat_matrix(cam["orientation"])
    right, up, fwd = R[:, 0], R[:, 1], -R[:, 2]
    th = math.tan(math.radians(cam["fov_y_deg"]) / 2)
    ys, xs = np.mgrid[0:h, 0:w]
    nx = (xs + 0.5) / w * 2 - 1
    ny = (ys + 0.5) / h * 2 - 1
    d = fwd + (nx * th * w / h)[..., None] * right + (ny * th)[..., None] * up
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    cs = d @ fwd
    n, f = cam["near_plane"], cam["far_plane"]
    rays = np.zeros(w * h, RAY_DTYPE)
    rays["origin"] = cam["position"]
    rays["direction"] = d.reshape(-1, 3)
    rays["tmin"] = (2 * f * n / (f + n) / cs).ravel()
    rays["tmax"] = (f / cs).ravel()
    return rays, d


def hits(o, rays, w, h):
    return o.trace_rays(rays, flags=HH_TRACE_OPAQUE, brute=True, fp64=True).reshape(h, w)


def hit_material(desc, inst, geom):
    mesh = desc.meshes[desc.instances[inst][2]]
    return desc.materials[mesh.materials[mesh.submeshes[geom][0]]]


def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


@pytest.fixture(scope="module")
def box(oracle_mod):
    d = scenegen.cornell_box(width=W, height=H)
    d.materials[0]["roughness_value"] = 0.05       # below the path's 0.1 minimum
    d.materials[1]["metallic_value"] = 0.7
    # turn the two small boxes by 30 degrees about their own vertical axis
    model, normal = scenegen.trs(rot=rot_y(30.0))
    d.instances[1] = (model, normal, d.instances[1][2], -1)
    return d, oracle_mod.OracleScene(d)


def test_views_show_the_centre_ray_surface(box):
    d, o = box
    rays, _ = centre_rays(d.camera, W, H)
    hh = hits(o, rays, W, H)
    views = [o.output_buffer(k) for k in range(5)]
    agree = 0
    for y in range(H):
        for x in range(W):
            h = hh[y, x]
            if h["t"] < 0:
                for v in views:
                    assert np.array_equal(v[y, x], [0, 0, 0, 1])
                continue
            m = hit_material(d, int(h["instance"]), int(h["geometry"]))
            alb = np.asarray(m["albedo_value"][:3]) ** 2.2        # scene.cpp:1017 packing
            got = views[OUTPUT_BUFFER_ALBEDO][y, x, :3]
            if not np.allclose(got, alb, atol=2e-6):
                continue                                          # a silhouette pixel seen differently in fp32
            agree += 1
            np.testing.assert_allclose(views[OUTPUT_BUFFER_ROUGHNESS][y, x, :3], m["roughness_value"], atol=1e-7)
            np.testing.assert_allclose(views[OUTPUT_BUFFER_METALLIC][y, x, :3], m["metallic_value"], atol=1e-7)
            np.testing.assert_allclose(views[OUTPUT_BUFFER_EMISSIVE][y, x, :3], m["emissive_value"][:3], atol=1e-7)
            # flat faces: the three vertex normals agree; the view shows R n * 0.5 + 0.5
            mesh = d.meshes[d.instances[int(h["instance"])][2]]
            sm = mesh.submeshes[int(h["geometry"])]
            vi = mesh.indices[sm[4] + 3 * int(h["primitive"])]
            n_obj = mesh.vertices["normal"][vi, :3].astype(np.float64)
            nm = np.asarray(d.instances[int(h["instance"])][1], np.float64).reshape(4, 4).T[:3, :3]
            n = nm @ n_obj
            n /= np.linalg.norm(n)
            np.testing.assert_allclose(views[OUTPUT_BUFFER_NORMALS][y, x, :3], n * 0.5 + 0.5, atol=2e-6)
            assert views[OUTPUT_BUFFER_NORMALS][y, x, 3] == 1.0
    assert agree >= 0.97 * np.count_nonzero(hh["t"] >= 0)
    # the raw roughness (0.05) and the light's emission are both on screen
    assert np.any(np.isclose(views[OUTPUT_BUFFER_ROUGHNESS][..., 0], 0.05))
    assert np.any(np.all(np.isclose(views[OUTPUT_BUFFER_EMISSIVE][..., :3], [17, 12, 4]), axis=-1))
    # the turned boxes show normals that are not axis-aligned
    nrm = views[OUTPUT_BUFFER_NORMALS][..., :3] * 2 - 1
    assert np.any(np.abs(nrm[..., 0] - math.sin(math.radians(30))) < 1e-5)


def test_interpolated_normals_are_not_renormalised(oracle_mod):
    """A unit icosphere with radial vertex normals: inside a face the
    interpolated normal is shorter than 1 (the vertex stage normalises, the
    fragment stage does not) and points nearly radially."""
    d = scenegen.SceneDesc(width=W, height=H)
    d.materials.append(scenegen.material())
    V, F = scenegen.icosphere(2)
    d.meshes.append(scenegen._mesh_from_parts([(V, V, None, F.astype(np.uint32).ravel(), 0)], [0]))
    d.instances.append(scenegen.identity_instance(0))
    d.camera = scenegen.look_at_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov_y=50.0, near=0.1)
    o = oracle_mod.OracleScene(d)
    v = o.output_buffer(OUTPUT_BUFFER_NORMALS)
    rays, dirs = centre_rays(d.camera, W, H)
    hh = hits(o, rays, W, H)
    m = hh["t"] > 0
    n = v[..., :3][m].astype(np.float64) * 2 - 1
    p = (np.asarray(d.camera["position"]) + dirs * hh["t"][..., None])[m]
    length = np.linalg.norm(n, axis=-1)
    assert length.max() <= 1.0 + 1e-5
    assert np.mean(length < 1.0 - 1e-3) > 0.5
    cosang = np.sum(n * p, axis=-1) / (length * np.linalg.norm(p, axis=-1))
    assert cosang.min() > math.cos(math.radians(12.0))


def _clip_scene(near):
    d = scenegen.SceneDesc(width=W, height=H)
    d.materials.append(scenegen.material(scenegen.linear_albedo((0.2, 0.4, 0.6))))
    d.materials.append(scenegen.material(scenegen.linear_albedo((0.9, 0.5, 0.1))))
    wall = scenegen._quad((-20, -20, -10), (20, -20, -10), (20, 20, -10), (-20, 20, -10), (0, 0, 1))
    card = scenegen._quad((-0.2, -0.2, -1.5), (0.2, -0.2, -1.5), (0.2, 0.2, -1.5), (-0.2, 0.2, -1.5), (0, 0, 1))
    d.meshes.append(scenegen._mesh_from_parts([(*wall, 0), (*card, 1)], [0, 1]))
    d.instances.append(scenegen.identity_instance(0))
    d.camera = scenegen.look_at_camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), fov_y=60.0, near=near, far=100.0)
    return d


@pytest.mark.parametrize("near, card_visible", [(1.0, False), (0.5, True)])
def test_clip_range_is_the_vulkan_depth_range(oracle_mod, near, card_visible):
    """Card at view depth 1.5, wall at 10, far 100: 2fn/(f+n) = 1.98 for n = 1
    clips the card; 0.995 for n = 0.5 keeps it."""
    o = oracle_mod.OracleScene(_clip_scene(near))
    v = o.output_buffer(OUTPUT_BUFFER_ALBEDO)
    centre = v[H // 2, W // 2, :3]
    card = np.asarray([0.9, 0.5, 0.1])
    wall = np.asarray([0.2, 0.4, 0.6])
    np.testing.assert_allclose(centre, card if card_visible else wall, atol=1e-6)


def test_far_plane_clips_to_background(oracle_mod):
    d = _clip_scene(0.5)
    d.camera["far_plane"] = 8.0
    v = oracle_mod.OracleScene(d).output_buffer(OUTPUT_BUFFER_ALBEDO)
    assert np.array_equal(v[0, 0], [0, 0, 0, 1])                     # wall beyond far: cleared
    np.testing.assert_allclose(v[H // 2, W // 2, :3], [0.9, 0.5, 0.1], atol=1e-6)


def test_alpha_tested_geometry_is_drawn_solid(oracle_mod):
    """A card whose texture has alpha 0 everywhere: the path's primary rays
    pass through it (any-hit ignores), the view shows its texture colour."""
    d = _clip_scene(0.5)
    tex = np.zeros((2, 2, 4), np.uint8)
    tex[..., :3] = (200, 100, 50)
    d.textures.append((2, 2, TEXTURE_RGBA8_SRGB, tex))
    d.materials[1] = scenegen.material(albedo_texture=0, alpha_test=True)
    o = oracle_mod.OracleScene(d)
    v = o.output_buffer(OUTPUT_BUFFER_ALBEDO)

    def srgb(c):
        c = c / 255.0
        return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4

    np.testing.assert_allclose(v[H // 2, W // 2, :3], [srgb(200), srgb(100), srgb(50)], rtol=1e-5)
    ph = o.primary_hits(width=W, height=H)
    assert ph[H // 2, W // 2]["geometry"] == 0       # the path tracer sees the wall there


def closed_box():
    d = scenegen.cornell_box(width=64, height=64, max_ray_bounces=6)
    front = scenegen._quad((1, 0, 5), (-1, 0, 5), (-1, 2, 5), (1, 2, 5), (0, 0, -1))
    sides = [scenegen._quad((-1, 0, 5), (-1, 0, 1), (-1, 2, 1), (-1, 2, 5), (1, 0, 0)),
             scenegen._quad((1, 0, 1), (1, 0, 5), (1, 2, 5), (1, 2, 1), (-1, 0, 0)),
             scenegen._quad((-1, 0, 1), (1, 0, 1), (1, 0, 5), (-1, 0, 5), (0, 1, 0)),
             scenegen._quad((-1, 2, 5), (1, 2, 5), (1, 2, 1), (-1, 2, 1), (0, -1, 0))]
    d.meshes.append(scenegen._mesh_from_parts([(*q, 0) for q in [front] + sides], [0]))
    d.instances.append(scenegen.identity_instance(len(d.meshes) - 1))
    return d


def by_ray(seg):
    rays = {}
    for s in seg:
        rays.setdefault(tuple(np.asarray(s[6:9], np.float32).tolist()), []).append(s)
    return rays


def test_ray_debug_paths_run_to_the_bounce_cap(oracle_mod):
    d = closed_box()
    o = oracle_mod.OracleScene(d)
    n, frame, bounces = 64, 5, 6
    seg = o.ray_debug((20, 30), n, frame, max_ray_bounces=bounces)
    paths = by_ray(seg)
    assert len(paths) == n
    for i in range(n):
        rng = oracle_mod.RNG(i, 0, frame)
        col = tuple(np.float32(rng.next_float() * np.float32(0.5) + np.float32(0.5)).item() for _ in range(3))
        p = paths[col]
        assert len(p) == bounces - 1           # no Russian roulette in the debug view
        for a, b in zip(p, p[1:]):
            np.testing.assert_allclose(a[3:6], b[0:3], atol=2e-4)   # chained: end = next origin


def test_ray_debug_first_origin_is_the_primary_hit(oracle_mod):
    d = closed_box()
    o = oracle_mod.OracleScene(d)
    w, h, px, py, frame, n = 64, 64, 20, 30, 9, 16
    seg = o.ray_debug((px, py), n, frame, width=w, height=h, max_ray_bounces=4)
    paths = by_ray(seg)
    cam = d.camera
    R = quat_matrix(cam["orientation"])
    th = math.tan(math.radians(cam["fov_y_deg"]) / 2)
    for i in range(n):
        rng = oracle_mod.RNG(i, 0, frame)
        col = tuple(np.float32(rng.next_float() * np.float32(0.5) + np.float32(0.5)).item() for _ in range(3))
        jx, jy = rng.next_float(), rng.next_float()
        nx = (px + 0.5 + jx) / w * 2 - 1
        ny = ((h - py) + 0.5 + jy) / h * 2 - 1          # the shader sees row height - y
        dvec = -R[:, 2] + nx * th * R[:, 0] + ny * th * R[:, 1]
        dvec /= np.linalg.norm(dvec)
        ray = np.zeros(1, RAY_DTYPE)
        ray["origin"] = cam["position"]
        ray["direction"] = dvec
        ray["tmin"], ray["tmax"] = 0.001, 1e4
        hit = o.trace_rays(ray, flags=0, brute=True, fp64=True)[0]
        p0 = np.asarray(cam["position"]) + dvec * hit["t"]
        np.testing.assert_allclose(paths[col][0][0:3], p0, atol=1e-3)


def test_ray_debug_miss_segments_end_at_tmax(oracle_mod):
    d = scenegen.cornell_box(width=64, height=64, max_ray_bounces=8)
    o = oracle_mod.OracleScene(d)
    seg = o.ray_debug((32, 32), 128, 1, max_ray_bounces=8)
    lengths = np.linalg.norm(seg[:, 3:6] - seg[:, 0:3], axis=1)
    escaped = lengths > 100
    assert escaped.any() and (~escaped).any()
    np.testing.assert_allclose(lengths[escaped], 10000.0, rtol=1e-5)
    # an escaped segment is the last of its path
    for p in by_ray(seg).values():
        ls = [np.linalg.norm(s[3:6] - s[0:3]) for s in p]
        assert all(x < 100 for x in ls[:-1])
