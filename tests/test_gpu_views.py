"""GPU: output-buffer views, the ray-debug variant and PNG output against the
oracle (which tests/test_oracle_views.py pins on the CPU).

Bars:
  * views: per pixel, the same values as the oracle to 1e-5 (both evaluate
    the same FP32 expressions) except where the two FP32 traversals resolve a
    silhouette or shared-edge tie differently: at most 0.1 % of the pixels;
  * ray debug: the paths are identified by their colour (three RNG draws, bit
    exact); each path's segment count equals the oracle's and its end points
    agree to 1e-3 for at least 95 % of the paths (later bounces inherit FP32
    differences of the transcendental functions, and a grazing bounce can
    then pick another surface); draw arguments {count, 1, 0, 0}; gathers
    append until cleared; past the capacity the count keeps growing;
  * PNG: the file decodes to hh_tonemap's pixels with the rows reversed.
"""
import numpy as np
import pytest

import pngcheck
from helios_amd import scenegen
from helios_amd.abi import OUTPUT_BUFFER_FINAL, OUTPUT_BUFFER_NORMALS

pytestmark = pytest.mark.gpu


def _scenes():
    box = scenegen.cornell_box(width=96, height=80)
    field = scenegen.icosphere_field(seed=4, n_spheres=8, level=3, n_leaves=200, width=128, height=72,
                                     max_ray_bounces=6)
    return [("cornell", box), ("field", field)]


@pytest.mark.parametrize("name, desc", _scenes())
def test_output_buffers_match_oracle(gpu_ctx, oracle_mod, name, desc):
    from helios_amd import runtime
    scene = runtime.Scene(gpu_ctx, desc)
    o = oracle_mod.OracleScene(desc)
    surf = runtime.Surface(gpu_ctx, desc.width, desc.height)
    for which in range(5):
        runtime.render_output_buffer(scene, desc.camera, which, surf)
        got = surf.read()
        ref = o.output_buffer(which)
        bad = ~np.all(np.abs(got - ref) <= 1e-5 + 1e-5 * np.abs(ref), axis=-1)
        assert bad.mean() <= 1e-3, (name, which, int(bad.sum()))
        assert np.all(got[..., 3] == 1.0)
    surf.close()
    scene.close()


def test_final_buffer_is_rejected(gpu_ctx):
    from helios_amd import runtime
    d = scenegen.cornell_box(width=16, height=16)
    scene = runtime.Scene(gpu_ctx, d)
    surf = runtime.Surface(gpu_ctx, 16, 16)
    with pytest.raises(runtime.HeliosError) as e:
        runtime.render_output_buffer(scene, d.camera, OUTPUT_BUFFER_FINAL, surf)
    assert e.value.status == 1
    runtime.render_output_buffer(scene, d.camera, OUTPUT_BUFFER_NORMALS, surf)   # still usable


def _paths(vertices):
    """GPU vertices -> {colour: [(origin, end)]} in append order."""
    out = {}
    for a, b in zip(vertices[0::2], vertices[1::2]):
        col = tuple(np.asarray(a["color"][:3], np.float32).tolist())
        assert tuple(np.asarray(b["color"][:3], np.float32).tolist()) == col
        assert a["position"][3] == 1.0 and b["position"][3] == 1.0
        out.setdefault(col, []).append((a["position"][:3], b["position"][:3]))
    return out


def _oracle_paths(seg):
    out = {}
    for s in seg:
        out.setdefault(tuple(np.asarray(s[6:9], np.float32).tolist()), []).append((s[0:3], s[3:6]))
    return out


def _close(p, q, tol=1e-3):
    if len(p) != len(q):
        return False
    for (a0, a1), (b0, b1) in zip(p, q):
        if not (np.allclose(a0, b0, rtol=tol, atol=tol) and np.allclose(a1, b1, rtol=tol, atol=tol)):
            return False
    return True


@pytest.mark.parametrize("flags", [0, 0x8])
def test_ray_debug_matches_oracle(gpu_ctx, oracle_mod, flags):
    from helios_amd import runtime
    d = scenegen.cornell_box(width=128, height=96, max_ray_bounces=6)
    scene = runtime.Scene(gpu_ctx, d)
    it = runtime.PathIntegrator(gpu_ctx)
    it.set_max_ray_bounces(6)
    it.set_flags(flags)
    surf = runtime.Surface(gpu_ctx, d.width, d.height)
    it.render(scene, d.camera, surf, frames=3)        # num_accumulated_samples = 3 seeds the paths
    frame = it.num_accumulated_samples()
    assert frame == 3
    buf = runtime.DebugRays(gpu_ctx)
    n, pixel = 256, (70, 40)
    it.gather_debug_rays(scene, d.camera, d.width, d.height, pixel, n, buf)
    verts, args = buf.read()
    assert list(args[1:]) == [1, 0, 0] and args[0] % 2 == 0
    ref = oracle_mod.OracleScene(d).ray_debug(pixel, n, frame, max_ray_bounces=6, flags=flags)
    assert len(verts) == int(args[0]) <= buf.capacity
    got, want = _paths(verts), _oracle_paths(ref)
    assert set(got) <= set(want)
    assert len(want) - len(got) <= 0.05 * n          # only paths without any bounce segment are absent
    agree = sum(_close(got.get(c, []), want[c]) for c in want)
    assert agree >= 0.95 * len(want), (agree, len(want))
    # a second gather appends; clear resets the count
    it.gather_debug_rays(scene, d.camera, d.width, d.height, pixel, n, buf)
    _, args2 = buf.read()
    assert args2[0] == 2 * args[0]
    buf.clear()
    v3, args3 = buf.read()
    assert list(args3) == [0, 1, 0, 0] and len(v3) == 0
    buf.close()


def test_ray_debug_overflow_keeps_counting(gpu_ctx):
    from helios_amd import runtime
    d = scenegen.cornell_box(width=64, height=64, max_ray_bounces=5)
    scene = runtime.Scene(gpu_ctx, d)
    it = runtime.PathIntegrator(gpu_ctx)
    big = runtime.DebugRays(gpu_ctx, max_segments=4096)
    small = runtime.DebugRays(gpu_ctx, max_segments=8)
    for b in (big, small):
        it.gather_debug_rays(scene, d.camera, 64, 64, (32, 32), 200, b)
    vb, ab = big.read()
    vs, as_ = small.read()
    assert ab[0] == as_[0] > 16 and len(vs) == 16
    assert set(_paths(vs)) <= set(_paths(vb))


def test_save_png_is_the_flipped_tonemap(gpu_ctx, tmp_path):
    from helios_amd import runtime
    d = scenegen.cornell_box(width=72, height=40)
    scene = runtime.Scene(gpu_ctx, d)
    surf = runtime.Surface(gpu_ctx, d.width, d.height)
    it = runtime.PathIntegrator(gpu_ctx)
    it.render(scene, d.camera, surf, frames=4)
    for op in (0, 1):
        path = tmp_path / f"shot{op}.png"
        surf.save_png(path, exposure=1.5, operator=op)
        img, _ = pngcheck.decode(path.read_bytes())
        assert np.array_equal(img, surf.tone_map(exposure=1.5, operator=op)[::-1])
