"""CPU tests of the oracle: pins each restated function against mathematics
that fixes it (periods, integrals, closed forms, invariants) and against the
committed golden fixtures. No GPU needed."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from helios_amd import scenegen
from helios_amd.abi import HH_TRACE_OPAQUE, HH_TRACE_TERMINATE_ON_FIRST

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------------------
# RNG (random.glsl:11-50, sampling.glsl:6-26)
# ---------------------------------------------------------------------------
def _state_step(oracle_mod, s0, s1):
    st = np.array([s0, s1], np.uint32)
    out = oracle_mod.lib().orc_rng_next(oracle_mod.ptr(st)) if hasattr(oracle_mod, "ptr") else None
    return st


def _step(oracle_mod, s):
    from helios_amd.abi import ptr
    st = np.array(s, np.uint32)
    out = oracle_mod.lib().orc_rng_next(ptr(st))
    return (int(st[0]), int(st[1])), out


def _gf2_matmul(a, b):
    return (a.astype(np.int64) @ b.astype(np.int64)) & 1


def _gf2_pow(m, e):
    r = np.eye(m.shape[0], dtype=np.int64)
    while e:
        if e & 1:
            r = _gf2_matmul(r, m)
        m = _gf2_matmul(m, m)
        e >>= 1
    return r


def test_xoroshiro64_state_is_full_period_linear_map(oracle_mod):
    """The xoroshiro64 state update (rotl 26, shift 9, rotl 13) is a GF(2)-linear
    map of period 2^64 - 1: a wrong rotation or shift constant breaks it."""
    cols = []
    for k in range(64):
        s = [0, 0]
        s[k // 32] = 1 << (k % 32)
        (n0, n1), _ = _step(oracle_mod, s)
        v = (n1 << 32) | n0
        cols.append([(v >> b) & 1 for b in range(64)])
    M = np.array(cols, dtype=np.int64).T  # column k = image of basis vector k
    # state bit b: b < 32 -> s0 bit b, else s1 bit b-32
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = [int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2 ** 32))]
        b = [int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2 ** 32))]
        sa, _ = _step(oracle_mod, a)
        sb, _ = _step(oracle_mod, b)
        sab, _ = _step(oracle_mod, [a[0] ^ b[0], a[1] ^ b[1]])
        assert sab == (sa[0] ^ sb[0], sa[1] ^ sb[1])
    order = 2 ** 64 - 1
    I = np.eye(64, dtype=np.int64)
    assert np.array_equal(_gf2_pow(M, order), I)
    for p in (3, 5, 17, 257, 641, 65537, 6700417):
        assert not np.array_equal(_gf2_pow(M, order // p), I)


def test_xoroshiro64star_output_multiplier(oracle_mod):
    # the * scrambler multiplies s0 by the published constant 0x9E3779BB
    _, out = _step(oracle_mod, [1, 0])
    assert out == 0x9E3779BB
    _, out = _step(oracle_mod, [3, 12345])
    assert out == (3 * 0x9E3779BB) & 0xFFFFFFFF


def test_wang_hash_is_a_permutation_sample(oracle_mod):
    xs = np.random.default_rng(1).choice(2 ** 32, size=200000, replace=False)
    hs = {oracle_mod.rng_hash(int(x)) for x in xs[:20000]}
    assert len(hs) == 20000
    # avalanche: a single flipped input bit flips ~half the output bits
    flips = []
    for x in xs[:2000]:
        h = oracle_mod.rng_hash(int(x))
        for b in (0, 7, 16, 31):
            flips.append(bin(h ^ oracle_mod.rng_hash(int(x) ^ (1 << b))).count("1"))
    assert 14.0 < np.mean(flips) < 18.0


def test_next_float_is_exact_23_bit_grid(oracle_mod):
    r = oracle_mod.RNG(17, 33, 5)
    vals = np.array([r.next_float() for _ in range(5000)], np.float64)
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert np.all(vals * 2 ** 23 == np.floor(vals * 2 ** 23))
    # uniformity (chi-square over 16 bins, 5000 draws)
    h, _ = np.histogram(vals, bins=16, range=(0, 1))
    chi2 = ((h - 5000 / 16) ** 2 / (5000 / 16)).sum()
    assert chi2 < 40.0


def test_rng_golden_vectors(oracle_mod):
    data = json.loads((GOLDEN / "rng_vectors.json").read_text())
    for case in data:
        r = oracle_mod.RNG(case["x"], case["y"], case["frame"])
        assert [int(v) for v in r.s] == case["state"]
        assert [r.next() for _ in range(len(case["draws"]))] == case["draws"]


# ---------------------------------------------------------------------------
# BRDF (brdf.glsl)
# ---------------------------------------------------------------------------
def _sphere_grid(n_theta=400, n_phi=800, hemisphere=False):
    th_max = math.pi / 2 if hemisphere else math.pi
    th = (np.arange(n_theta) + 0.5) * th_max / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * math.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dw = np.sin(T) * (th_max / n_theta) * (2 * math.pi / n_phi)
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)  # y = normal
    return d.reshape(-1, 3), dw.ravel()


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.6, 1.0])
def test_ggx_ndf_normalized(oracle_mod, alpha):
    """∫ D(h) cos(theta_h) dω = 1 where the EPSILON clamp does not bind."""
    d, dw = _sphere_grid(2000, 4, hemisphere=True)
    L = oracle_mod.lib()
    D = np.array([L.orc_D_ggx(float(c), alpha) for c in d[:, 1]])
    assert abs((D * d[:, 1] * dw).sum() - 1.0) < 2e-3


def test_ggx_ndf_clamped_for_smooth_surfaces(oracle_mod):
    # roughness 0.1 -> alpha 0.01: pi * denom^2 < EPSILON at the peak, so D(1) = a2 / EPSILON
    assert oracle_mod.lib().orc_D_ggx(1.0, 0.01) == pytest.approx(1e-4 / 1e-4)


def _surf(albedo=(0.8, 0.5, 0.3), rough=0.5, metal=0.0, n=(0, 1, 0)):
    return np.array([*albedo, rough, metal, *n], np.float64)


def _eval(oracle_mod, s, wo, wi):
    from helios_amd.abi import ptr
    out = np.zeros(4)
    oracle_mod.lib().orc_eval_uber(ptr(s), ptr(np.asarray(wo, np.float64)), ptr(np.asarray(wi, np.float64)), ptr(out))
    return out


@pytest.mark.parametrize("rough,metal", [(0.3, 0.0), (0.7, 1.0), (1.0, 0.5)])
def test_uber_pdf_integrates_to_one(oracle_mod, rough, metal):
    """pdf_uber = ½ cos/π + ½ D·NH/(4·VH): the 1/(4·VH) Jacobian maps the
    integral over Wi to ∫ D(h)(n·h) dω_h over the half vectors h with
    Wo·h > 0, computed independently on a grid of h."""
    wo = np.array([0.3, 0.8, 0.2])
    wo /= np.linalg.norm(wo)
    d, dw = _sphere_grid(300, 600)
    s = _surf(rough=rough, metal=metal)
    pdf = np.array([_eval(oracle_mod, s, wo, wi)[3] for wi in d])
    L = oracle_mod.lib()
    a = max(rough, 0.1) ** 2
    nh = d[:, 1]
    keep = (nh > 0) & (d @ wo > 0)
    Dh = np.array([L.orc_D_ggx(float(c), a) for c in nh[keep]])
    expect = 0.5 * 1.0 + 0.5 * (Dh * nh[keep] * dw[keep]).sum()
    assert abs((pdf * dw).sum() - expect) < 2e-3


def test_uber_brdf_reciprocal(oracle_mod):
    rng = np.random.default_rng(3)
    s = _surf(rough=0.4, metal=0.3)
    for _ in range(50):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[1], b[1] = abs(a[1]), abs(b[1])
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        assert np.allclose(_eval(oracle_mod, s, a, b)[:3], _eval(oracle_mod, s, b, a)[:3], rtol=1e-12)


def test_uber_lambert_limit(oracle_mod):
    """Grazing-free, rough, dielectric: f -> (1 - F)·albedo/π + spec with F ≈ F0 = 0.03."""
    s = _surf(albedo=(1.0, 1.0, 1.0), rough=1.0, metal=0.0)
    wo = np.array([0.0, 1.0, 0.0])
    f = _eval(oracle_mod, s, wo, wo)
    # normal incidence: VdotH = 1 -> F = F0 = 0.03; D(1) = 1/π (alpha 1); G1 = 1 / (1 - k + k) = 1
    spec = (1 / math.pi) * 0.03 * 1.0 / 4.0
    assert f[0] == pytest.approx((1 - 0.03) / math.pi + spec, rel=1e-12)


def test_cosine_lobe_sampling_pdf(oracle_mod):
    """Directions from sample_cosine_lobe have E[cos] = 2/3 and lie in the hemisphere."""
    from helios_amd.abi import ptr
    L = oracle_mod.lib()
    rng = np.random.default_rng(2)
    n = np.array([0.0, 0.0, 1.0])
    cs = []
    out = np.zeros(3)
    for r0, r1 in rng.random((20000, 2)):
        L.orc_sample_cosine_lobe(ptr(n), float(r0), float(r1), ptr(out))
        cs.append(out @ n)
    cs = np.array(cs)
    assert cs.min() > 0
    assert abs(cs.mean() - 2 / 3) < 5e-3


# ---------------------------------------------------------------------------
# Intersection
# ---------------------------------------------------------------------------
def _tri_hit(oracle_mod, o, d, v, tmin=0.0, tmax=1e30):
    from helios_amd.abi import ptr
    out = np.zeros(3, np.float32)
    h = oracle_mod.lib().orc_tri_hit(ptr(np.asarray(o, np.float32)), ptr(np.asarray(d, np.float32)),
                                     ptr(np.asarray(v, np.float32).ravel()), tmin, tmax, ptr(out))
    return bool(h), out


def test_watertight_shared_edges_and_vertices(oracle_mod):
    """Rays through a shared edge / the shared vertex of a triangle fan never slip through."""
    k = 7
    ang = np.linspace(0, 2 * np.pi, k + 1)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(k)], 1) * 1.3
    fan = [np.stack([[0, 0, 0], ring[i], ring[(i + 1) % k]]) for i in range(k)]
    rng = np.random.default_rng(4)
    targets = [np.zeros(3)] + [ring[i] * t for i in range(k) for t in rng.random(30)]
    for tgt in targets:
        o = np.array([*(rng.normal(size=2) * 0.3), 2.0]) + np.array([tgt[0], tgt[1], 0])
        o = o.astype(np.float32)
        d = (np.asarray(tgt, np.float32) - o)
        d /= np.linalg.norm(d)
        hits = [_tri_hit(oracle_mod, o, d, tri)[0] for tri in fan]
        assert any(hits)


def test_degenerate_triangles_never_hit(oracle_mod):
    v = np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2]], np.float32)
    assert not _tri_hit(oracle_mod, [1, 1, -1], [0, 0, 1], v)[0]


def test_axis_aligned_quad_closed_form(oracle_mod):
    v = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0]], np.float32)
    h, (t, u, w) = _tri_hit(oracle_mod, [0.5, -0.25, 2.0], [0, 0, -1], v)
    assert h and t == 2.0
    # point = (1-u-w) v0 + u v1 + w v2 -> (0.5, -0.25)
    p = (1 - u - w) * v[0] + u * v[1] + w * v[2]
    assert np.allclose(p[:2], [0.5, -0.25], atol=1e-7)
    # tmin / tmax are inclusive bounds
    assert _tri_hit(oracle_mod, [0.5, -0.25, 2.0], [0, 0, -1], v, 2.0, 2.0)[0]
    assert not _tri_hit(oracle_mod, [0.5, -0.25, 2.0], [0, 0, -1], v, 0.0, 1.9999)[0]


def test_bvh_traversal_equals_brute_force(oracle_mod):
    d = scenegen.icosphere_field(seed=2, n_spheres=5, level=3, n_leaves=80, width=64, height=36)
    s = oracle_mod.OracleScene(d)
    rays = scenegen.random_rays(4000, (-90, 0.5, -90), (90, 20, 90), seed=9, tmin=1e-4)
    for flags in (0, HH_TRACE_OPAQUE):
        a = s.trace_rays(rays, flags, brute=False)
        b = s.trace_rays(rays, flags, brute=True)
        assert np.array_equal(a, b)
    a = s.trace_rays(rays, HH_TRACE_OPAQUE | HH_TRACE_TERMINATE_ON_FIRST, brute=False)
    b = s.trace_rays(rays, HH_TRACE_OPAQUE, brute=True)
    assert np.array_equal(a["instance"] == 0xFFFFFFFF, b["instance"] == 0xFFFFFFFF)


def test_alpha_test_lets_rays_through_cutouts(oracle_mod):
    d = scenegen.icosphere_field(seed=2, n_spheres=0, level=1, n_leaves=300, width=8, height=8)
    s = oracle_mod.OracleScene(d)
    rays = scenegen.random_rays(20000, (-90, 0.5, -90), (90, 14, 90), seed=1, tmin=1e-4)
    leaf_inst = len(d.instances) - 1
    any_hit = s.trace_rays(rays, 0)
    opaque = s.trace_rays(rays, HH_TRACE_OPAQUE)
    n_any = (any_hit["instance"] == leaf_inst).sum()
    n_opq = (opaque["instance"] == leaf_inst).sum()
    assert n_opq > 50
    assert 0.25 < n_any / n_opq < 0.6   # ~40 % opaque leaf mask


def _two_quads_one_mesh(second_override):
    """One quad mesh (own material: blue) instanced twice; the first instance
    overrides it to green, the second uses `second_override`."""
    from helios_amd.abi import LIGHT_POINT
    d = scenegen.SceneDesc(width=64, height=32, max_ray_bounces=1)
    d.materials += [scenegen.material((0.9, 0.1, 0.1, 1.0)), scenegen.material((0.1, 0.9, 0.1, 1.0)),
                    scenegen.material((0.1, 0.1, 0.9, 1.0))]
    V = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float64)
    d.meshes.append(scenegen._mesh_from_parts([(V, np.tile([0.0, 0.0, 1.0], (4, 1)), None,
                                                np.array([0, 1, 2, 0, 2, 3], np.uint32), 0)], [2]))
    for x, ov in ((-1.2, 1), (1.2, second_override)):
        m, n = scenegen.trs(translate=(x, 0.0, 0.0))
        d.instances.append((m, n, 0, ov))
    d.lights.append(dict(type=LIGHT_POINT, color=(1.0, 1.0, 1.0), intensity=10.0, radius=0.1, position=(0.0, 0.0, 2.0),
                         forward=(0.0, 0.0, -1.0), inner_cone_deg=40.0, outer_cone_deg=50.0))
    d.camera = scenegen.look_at_camera((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), fov_y=35.0)
    return d


def test_unpacked_material_reads_as_material_zero(oracle_mod):
    """scene.cpp:934,1143 (Q19): materials are packed only when a mesh is first
    met, with that node's override; global_material_indices is an
    unordered_map, so a later node of the same mesh whose own material (blue)
    was never packed reads packed record 0 -- the first material packed,
    here the first node's green override -- not its own blue and not the
    scene's material 0 (red)."""
    a = oracle_mod.OracleScene(_two_quads_one_mesh(-1)).render(frames=4)
    b = oracle_mod.OracleScene(_two_quads_one_mesh(1)).render(frames=4)
    assert np.array_equal(a, b)
    for half in (a[:, 40:, :3], a[:, :24, :3]):
        px = half.reshape(-1, 3)
        lit = px[px.sum(axis=1) > 0.05]
        assert len(lit) > 50 and np.all(lit[:, 1] > 10 * lit[:, 0]) and np.all(lit[:, 1] > 10 * lit[:, 2])


# ---------------------------------------------------------------------------
# Camera, sky, tonemap, accumulation, packing
# ---------------------------------------------------------------------------
def test_camera_consts_closed_form(oracle_mod):
    cam = scenegen.look_at_camera((1.0, 2.0, 5.0), (0.0, 1.0, 0.0), fov_y=50.0, focal_length=4.0)
    c = oracle_mod.launch_consts(cam, 320, 200)
    vpi = np.array(c.view_proj_inverse, np.float64).reshape(4, 4).T
    p = vpi @ np.array([0, 0, 0, 1.0])
    p = p[:3] / p[3]
    eye = np.array(cam["position"])
    view_dir = np.array([0.0, 1.0, 0.0]) - eye
    view_dir /= np.linalg.norm(view_dir)
    assert np.allclose((p - eye) / np.linalg.norm(p - eye), view_dir, atol=1e-5)
    # top edge of NDC is fov/2 above the axis
    q = vpi @ np.array([0, 1, 0, 1.0])
    q = q[:3] / q[3] - eye
    ang = math.degrees(math.acos(q @ view_dir / np.linalg.norm(q)))
    assert ang == pytest.approx(25.0, abs=1e-3)
    # focal plane: points at distance focal_length along the view direction
    fp = np.array(c.focal_plane, np.float64)
    x = eye + view_dir * 4.0
    assert abs(fp[:3] @ x + fp[3]) < 1e-4


def test_sky_normalization_and_spline_endpoints(oracle_mod):
    sun = np.array([math.cos(math.radians(30)), math.sin(math.radians(30)), 0.0])
    coef = oracle_mod.sky_coefficients(3.0, 0.1, 1.15, sun)
    A, B, C, D, E, F, G, H, I, Z = coef
    cos_t = sun[1]
    chi = (1 + 1.0) / (1 + H * H - 2 * H) ** 1.5
    S = Z * (1 + A * np.exp(B / (cos_t + 0.01))) * (C + D + F + G * chi + I * math.sqrt(cos_t))
    assert S @ np.array([0.2126, 0.7152, 0.0722]) == pytest.approx(1.15, rel=1e-5)
    # sky at the sun direction equals the normalized sun radiance
    v = oracle_mod.sky_eval(coef, sun, sun[None, :], fp64=True)[0]
    assert v @ np.array([0.2126, 0.7152, 0.0722]) == pytest.approx(1.15, rel=1e-4)
    # mirror symmetry about the sun's vertical plane (z -> -z)
    dirs = np.random.default_rng(0).normal(size=(200, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    m = dirs * np.array([1, 1, -1])
    assert np.allclose(oracle_mod.sky_eval(coef, sun, dirs, fp64=True), oracle_mod.sky_eval(coef, sun, m, fp64=True),
                       rtol=1e-5)


def test_sky_spline_endpoints_are_table_entries(oracle_mod):
    """Sun at the zenith (elevation^(1/3) = 1) makes every spline equal its last
    control point: with albedo 0 and integer turbidity the coefficient is a
    table entry."""
    table = np.fromfile(oracle_mod.HOSEK, "<f8")
    coef = oracle_mod.sky_coefficients(4.0, 0.0, 0.0, (0.0, 1.0, 0.0))
    # datasetRGB1, albedo 0, turbidity 4, control point 5, coefficient A (index 0)
    assert coef[0, 0] == np.float32(table[(4 - 1) * 54 + 5 * 9 + 0])
    # elevation 0 (sun at the horizon): first control point, coefficient C
    coef0 = oracle_mod.sky_coefficients(4.0, 0.0, 0.0, (1.0, 0.0, 0.0))
    assert coef0[2, 1] == np.float32(table[1080 + (4 - 1) * 54 + 0 * 9 + 2])


def test_tonemap_closed_form(oracle_mod):
    img = np.array([[[0, 1, 100, 1], [0.5, 0.25, 2.0, 1]]], np.float32)
    aces = oracle_mod.tonemap(img, 1.0, 0)
    expect = []
    for x in (0.0, 1.0, 100.0):
        y = min(max((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0), 1) ** (1 / 2.2)
        expect.append(int(math.floor(y * 255 + 0.5)))
    assert list(aces[0, 0, :3]) == expect
    rein = oracle_mod.tonemap(img, 2.0, 1)
    x = 1.0
    assert rein[0, 1, 0] == int(math.floor((x / (1 + x)) ** (1 / 2.2) * 255 + 0.5))


def test_progressive_accumulation_algebra(oracle_mod):
    """num_frames = 0 stores, later frames blend by 1/num_frames (Q5)."""
    d = scenegen.cornell_box(width=32, height=32)
    s = oracle_mod.OracleScene(d)
    full = s.render(frames=4)
    step = s.render(frames=1, first_frame=0)
    for f in (1, 2, 3):
        step = s.render(frames=1, first_frame=f, image=step)
    assert np.array_equal(full, step)
    # frame 1 replaces frame 0 (prev + (c - prev) / 1)
    f01 = s.render(frames=2)
    f1 = s.render(frames=1, first_frame=1, image=np.zeros((32, 32, 4), np.float32))
    assert np.allclose(f01, f1, atol=1e-6)


def test_fp32_transcription_agrees_with_fp64(oracle_mod):
    d = scenegen.cornell_box(width=48, height=48)
    s = oracle_mod.OracleScene(d)
    a = s.render(frames=16)
    b = s.render(frames=16, fp64=True)
    assert oracle_mod.relmse(a, b) < 1e-4


def test_packing_quirks(oracle_mod):
    d = scenegen.cornell_box()
    s = oracle_mod.OracleScene(d)
    L = s.lights()
    # one area light: type 4, mesh node 2, material 3, primitive offset 0; count in .x (Q2)
    assert s.num_lights == 1
    assert list(L[0, :5]) == [4.0, 2.0, 3.0, 0.0, 2.0]
    f = scenegen.icosphere_field(seed=1, n_spheres=2, level=1, n_leaves=4, width=8, height=8, n_point_lights=2)
    f.lights.append(dict(type=1, color=(1, 1, 1), intensity=2.0, radius=5.0, position=(0, 10, 0),
                         forward=(0, -1, 0), inner_cone_deg=40.0, outer_cone_deg=50.0))
    s2 = oracle_mod.OracleScene(f)
    types = list(s2.lights()[:, 0])
    assert types == [3.0, 0.0, 2.0, 2.0, 1.0]   # env, directional, points, spot (Q8)
    assert s2.num_lights == 5


def test_material_albedo_gamma_and_texture_dedupe(oracle_mod):
    """albedo rgb packed as pow(v, 2.2) (Q10); a texture already used by an
    earlier material leaves the later material's slot at -1 (scene.cpp:981-994)."""
    from helios_amd.abi import ptr
    d = scenegen.cornell_box(width=8, height=8)
    tex = np.full((4, 4, 4), 200, np.uint8)
    d.textures.append((4, 4, 1, tex))
    m1 = len(d.materials)
    d.materials.append(scenegen.material((1, 1, 1, 1), albedo_texture=0))
    d.materials.append(scenegen.material((1, 1, 1, 1), albedo_texture=0))
    d.meshes[1].materials = [m1]
    d.instances.append(scenegen.identity_instance(1, override=m1 + 1))
    s = oracle_mod.OracleScene(d)
    mats = np.zeros((16, 20), np.int32)
    n = oracle_mod.lib().orc_scene_read_materials(s.h, ptr(mats), 16)
    fm = mats[:n].view(np.float32)
    assert fm[0, 8] == pytest.approx(0.75, rel=1e-6)            # pow(0.75^(1/2.2), 2.2)
    tex_slots = mats[:n, 0]
    assert 0 in tex_slots                                        # first user gets slot 0
    assert (tex_slots == -1).sum() >= 1                          # the later user gets -1
