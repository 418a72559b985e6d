"""CPU: environment maps from Radiance .hdr files (hh_ibl_from_hdr).

The files are written here from the format's description (Ward, "Real
Pixels", Graphics Gems II): RGBE pixels, flat or run-length scanlines. What
pins the loader:
  * a constant map decodes to (m + 0.5) 2^(e - 136) on every texel of every
    face, for both scanline forms and both row orders;
  * a map whose value is a known function of direction, resampled to a cube,
    and looked up by the oracle's own cube sampler through camera rays of an
    empty scene (six cameras, one per axis), returns that function of the
    ray direction -- so the lat-long convention, the face layout and the
    texel directions all agree with the sampler;
  * malformed or unsupported files are refused with HH_ERR_IO.
"""
import math

import numpy as np
import pytest

from helios_amd import runtime, scenegen
from helios_amd.abi import HH_FLAG_NO_RADIANCE_CLAMP


def rgbe(rgb):
    """float RGB (..., 3) -> RGBE bytes (..., 4), Ward's float2rgbe."""
    rgb = np.asarray(rgb, np.float64)
    v = rgb.max(axis=-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    m, e = np.frexp(v)
    ok = v > 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    out[..., :3] = np.where(ok[..., None], np.floor(rgb * scale[..., None]), 0).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def decode(px):
    px = np.asarray(px, np.float64)
    f = np.where(px[..., 3] > 0, np.ldexp(1.0, (px[..., 3] - 136).astype(int)), 0.0)
    return (px[..., :3] + 0.5) * f[..., None]


def rle_channel(b):
    out = bytearray()
    i, n = 0, len(b)
    while i < n:
        j = i
        while j < n and j - i < 127 and b[j] == b[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), b[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (j + 2 < n and b[j] == b[j + 1] == b[j + 2]):
            j += 1
        out += bytes([j - i]) + bytes(b[i:j])
        i = j
    return bytes(out)


def write_hdr(path, rgb, rle=True, bottom_up=False, fmt="32-bit_rle_rgbe"):
    h, w, _ = rgb.shape
    px = rgbe(rgb)
    rows = px[::-1] if bottom_up else px
    out = bytearray(f"#?RADIANCE\n# written by a test\nFORMAT={fmt}\nEXPOSURE=1.0\n\n"
                    f"{'+Y' if bottom_up else '-Y'} {h} +X {w}\n".encode())
    for r in rows:
        if rle:
            out += bytes([2, 2, w >> 8, w & 0xFF])
            for c in range(4):
                out += rle_channel(r[:, c].tobytes())
        else:
            out += r.tobytes()
    path.write_bytes(bytes(out))
    return px


def latlong_dirs(w, h):
    j, i = np.mgrid[0:h, 0:w]
    t = np.pi * (j + 0.5) / h
    p = 2 * np.pi * ((i + 0.5) / w - 0.5)
    return np.stack([np.sin(t) * np.sin(p), np.cos(t), -np.sin(t) * np.cos(p)], -1)


@pytest.mark.parametrize("rle, bottom_up", [(True, False), (False, False), (True, True)])
def test_constant_map(tmp_path, rle, bottom_up):
    rgb = np.broadcast_to(np.array([0.3, 1.7, 12.5]), (16, 40, 3))
    px = write_hdr(tmp_path / "c.hdr", np.array(rgb), rle=rle, bottom_up=bottom_up)
    want = decode(px[0, 0])
    faces = runtime.ibl_from_hdr(tmp_path / "c.hdr", 8)
    assert faces.shape == (6, 8, 8, 4)
    np.testing.assert_allclose(faces[..., :3].reshape(-1, 3), np.broadcast_to(want, (6 * 64, 3)), rtol=2e-7)
    assert np.all(faces[..., 3] == 1.0)


def test_rows_keep_their_order(tmp_path):
    """Top rows of a lat-long image look up (+Y): a map bright in its first
    rows lights the +Y face only, whichever way the file stores its rows."""
    for bottom_up in (False, True):
        rgb = np.full((32, 64, 3), 0.1)
        rgb[:4] = 5.0
        write_hdr(tmp_path / "top.hdr", rgb, bottom_up=bottom_up)
        faces = runtime.ibl_from_hdr(tmp_path / "top.hdr", 8)
        centre = faces[:, 3:5, 3:5, 0].mean(axis=(1, 2))
        assert np.argmax(centre) == 2 and centre[2] > 4.0 and centre[3] < 0.2


def _env_seen_by_cameras(oracle_mod, faces, size):
    d = scenegen.SceneDesc(width=24, height=24, max_ray_bounces=1)
    d.ibl = (size, faces)
    o = oracle_mod.OracleScene(d)
    looks = [(1, 0, 0), (-1, 0, 0), (0, 1, 0.001), (0, -1, 0.001), (0, 0, 1), (0, 0, -1)]
    out = []
    for look in looks:
        cam = scenegen.look_at_camera((0.0, 0.0, 0.0), look, fov_y=80.0, near=0.1)
        img = o.render(camera=cam, frames=1, flags=HH_FLAG_NO_RADIANCE_CLAMP)
        hits = o.primary_hits(camera=cam, frame=0)
        assert np.all(hits["instance"] == 0xFFFFFFFF)
        out.append((cam, img))
    return out


def test_cube_agrees_with_the_sampler(tmp_path, oracle_mod):
    """value(d) = 0.5 + 0.3 d: the oracle's cube lookup along each camera ray
    returns value(ray direction) to the RGBE step (1/256 of the largest
    channel) plus the footprint averaging."""
    w, h, size = 512, 256, 32
    dirs = latlong_dirs(w, h)
    write_hdr(tmp_path / "f.hdr", 0.5 + 0.3 * dirs)
    faces = runtime.ibl_from_hdr(tmp_path / "f.hdr", size)
    for cam, img in _env_seen_by_cameras(oracle_mod, faces, size):
        q = np.array(cam["orientation"], np.float64)
        x, y, z, wq = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * wq), 2 * (x * z + y * wq)],
                      [2 * (x * y + z * wq), 1 - 2 * (x * x + z * z), 2 * (y * z - x * wq)],
                      [2 * (x * z - y * wq), 2 * (y * z + x * wq), 1 - 2 * (x * x + y * y)]])
        th = math.tan(math.radians(cam["fov_y_deg"]) / 2)
        # the traced ray: pixel + 0.5 + the frame's two jitter draws (Q11)
        jit = np.array([[(lambda g: (g.next_float(), g.next_float()))(oracle_mod.RNG(x, y, 0)) for x in range(24)]
                        for y in range(24)])
        ys, xs = np.mgrid[0:24, 0:24]
        nx, ny = (xs + 0.5 + jit[..., 0]) / 24 * 2 - 1, (ys + 0.5 + jit[..., 1]) / 24 * 2 - 1
        dv = -R[:, 2] + (nx * th)[..., None] * R[:, 0] + (ny * th)[..., None] * R[:, 1]
        dv /= np.linalg.norm(dv, axis=-1, keepdims=True)
        want = 0.5 + 0.3 * dv
        err = np.abs(img[..., :3] - want)
        assert err.max() < 0.01, (cam["orientation"], err.max())
        assert err.mean() < 0.003


@pytest.mark.parametrize("text, msg", [
    (b"P6\n1 1\n255\n\x00\x00\x00", "not a Radiance"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x00\x00\x00\x00", "xyze"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+X 1 -Y 1\n\x00\x00\x00\x00", "resolution"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n\x00\x00\x00\x00", "truncated"),
])
def test_bad_files_are_refused(tmp_path, text, msg):
    p = tmp_path / "bad.hdr"
    p.write_bytes(text)
    with pytest.raises(runtime.HeliosError, match=msg) as e:
        runtime.ibl_from_hdr(p, 4)
    assert e.value.status == 7
