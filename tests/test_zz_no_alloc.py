"""hh_trace_rays_noalloc (no device allocation, fixed per-GPU scratch) against
hh_trace_rays on the same rays: every hit record bit-identical, on two calls in
a row where the second is smaller, and sizes beyond the scratch rejected.

The caller-owned device buffers are RGBA32F surfaces of the library used as
plain device memory (created before the calls; hh_surface_write / _read copy
their bytes unchanged), so the whole test goes through the C ABI.
"""
import numpy as np
import pytest

from helios_amd import abi, scenegen

pytestmark = pytest.mark.gpu

MAX_RAYS = 32768   # HH_TRACE_NOALLOC_MAX_RAYS
ROW = 1024         # surface width in 16-byte texels


@pytest.fixture(scope="module")
def field_scene(gpu_ctx):
    from helios_amd import runtime
    d = scenegen.icosphere_field(seed=3, n_spheres=6, level=3, n_leaves=120, width=160, height=90,
                                 max_ray_bounces=8)
    s = runtime.Scene(gpu_ctx, d)
    yield s
    s.close()


@pytest.fixture(scope="module")
def buffers(gpu_ctx):
    """Device buffers for MAX_RAYS rays (32 B each) and hits (24 B each)."""
    from helios_amd import runtime
    rays = runtime.Surface(gpu_ctx, ROW, MAX_RAYS * 32 // 16 // ROW)
    hits = runtime.Surface(gpu_ctx, ROW, -(-MAX_RAYS * 24 // 16 // ROW))
    return rays, hits


def _device_trace(gpu_ctx, scene, buffers, rays, flags):
    rs, hs = buffers
    raw = np.zeros(rs.height * rs.width * 4, np.float32)
    raw[: len(rays) * 8] = np.ascontiguousarray(rays, abi.RAY_DTYPE).view(np.float32)
    rs.write(raw.reshape(rs.height, rs.width, 4))
    # a sentinel in every hit word: the records past n must stay untouched
    hs.write(np.full((hs.height, hs.width, 4), -7, np.int32).view(np.float32))
    scene.trace_rays_device(rs.device_ptr(), len(rays), flags, hs.device_ptr())
    gpu_ctx.sync()
    out = hs.read().reshape(-1).view(np.int32)
    return out[: len(rays) * 6].view(abi.HIT_DTYPE).copy(), out[len(rays) * 6:]


def test_noalloc_matches_trace_rays_two_calls(gpu_ctx, field_scene, buffers):
    from helios_amd.abi import HH_TRACE_OPAQUE
    big = scenegen.random_rays(MAX_RAYS, (-90, 0.5, -90), (90, 25, 90), seed=5, tmin=1e-4)
    small = scenegen.random_rays(5000, (-90, 0.5, -90), (90, 25, 90), seed=9, tmin=1e-4)
    print(f"noalloc: call 1, {len(big)} rays, flags 0", flush=True)
    h1, _ = _device_trace(gpu_ctx, field_scene, buffers, big, 0)
    ref1 = field_scene.trace_rays(big, 0)
    assert h1.tobytes() == ref1.tobytes()
    assert (ref1["instance"] != 0xFFFFFFFF).mean() > 0.2   # the rays hit the scene
    print(f"noalloc: call 2, {len(small)} rays, flags OPAQUE", flush=True)
    h2, tail = _device_trace(gpu_ctx, field_scene, buffers, small, HH_TRACE_OPAQUE)
    ref2 = field_scene.trace_rays(small, HH_TRACE_OPAQUE)
    assert h2.tobytes() == ref2.tobytes()
    assert (tail == -7).all()
    assert field_scene.traversal_faults() == dict(lane_stack=0, packet_stack=0, trips=0)


def test_noalloc_rejects_beyond_scratch(gpu_ctx, field_scene, buffers):
    from helios_amd.runtime import HeliosError
    rs, hs = buffers
    print("noalloc: size check", flush=True)
    with pytest.raises(HeliosError) as e:
        # the size check comes before any launch, so the buffers are never read
        field_scene.trace_rays_device(rs.device_ptr(), MAX_RAYS + 1, 0, hs.device_ptr())
    assert e.value.status == 3   # HH_ERR_LIMIT_EXCEEDED
    field_scene.trace_rays_device(rs.device_ptr(), 0, 0, hs.device_ptr())   # n = 0 is a no-op
    gpu_ctx.sync()
