"""Drives the AddressSanitizer + UBSan builds of the oracle and of the file
readers (tests/sanitize/Makefile) through the same Python entry points the
CPU tests use. Run by tests/test_sanitizers.py in a child process with
libasan preloaded; any sanitizer report aborts the process with a non-zero
status. Prints one line per stage and "SANITIZED OK" at the end."""
import ctypes as C
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
from helios_amd import abi, runtime, scenegen  # noqa: E402


def stage(name):
    print(f"stage: {name}", flush=True)


def oracle_stages():
    oracle.LIB_PATH = HERE / "_build" / "liboracle_san.so"
    oracle._lib = None
    stage("oracle cornell render + hits")
    d = scenegen.cornell_box(width=24, height=24)
    o = oracle.OracleScene(d)
    o.render(frames=3)
    o.render(frames=2, fp64=True, threads=2)
    o.primary_hits(frame=1)
    stage("oracle showcase (textures, spot, area, sky)")
    d = scenegen.textured_showcase(width=24, height=16)
    d.sky = (3.0, 0.1, 1.15, 16)
    o = oracle.OracleScene(d)
    o.render(frames=2)
    for which in range(5):
        o.output_buffer(which)
    o.ray_debug((5, 5), 4, 0)
    rng = np.random.default_rng(0)
    st = np.array([12345, 678], np.uint32)
    for k in range(o.num_lights):
        o.sample_light(k, (0.1, 0.2, 0.3), (0.0, 1.0, 0.0), st)
    stage("oracle field (alpha leaves, BVH)")
    d = scenegen.icosphere_field(seed=2, n_spheres=3, level=2, n_leaves=30, width=16, height=12, max_ray_bounces=4)
    d.sky = (3.0, 0.1, 1.15, 16)
    o = oracle.OracleScene(d)
    o.render(frames=2)
    rays = scenegen.random_rays(500, (-10, -1, -10), (10, 5, 10), seed=1)
    for flags in (0, abi.HH_TRACE_OPAQUE, abi.HH_TRACE_OPAQUE | abi.HH_TRACE_TERMINATE_ON_FIRST):
        o.trace_rays(rays, flags=flags)
        o.trace_rays(rays, flags=flags, brute=True)
    stage("oracle seamless cube")
    n = 8
    faces = rng.random((6, n, n, 4)).astype(np.float32)
    e = scenegen.SceneDesc(width=4, height=4)
    e.ibl = (n, faces)
    o = oracle.OracleScene(e)
    dirs = rng.normal(size=(4000, 3)).astype(np.float32)
    dirs[:8] = [[1, 1, 1], [-1, 1, 1], [1, -1, 1], [1, 1, -1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1], [-1, -1, -1]]
    o.sample_env(dirs)
    o.sample_env(dirs, fp64=True)
    stage("oracle light power + equirect sky")
    d, cams = scenegen.combined_workload(width=16, height=12, n_meshes=3, n_instances=6, equirect_width=64)
    o = oracle.OracleScene(d)
    o.render(frames=2, flags=abi.HH_FLAG_LIGHT_POWER)


def hostio_stages(tmp):
    L = C.CDLL(str(HERE / "_build" / "libhostio_san.so"))
    for name in ("hh_gltf_load", "hh_ibl_from_hdr", "hh_write_png"):
        getattr(L, name).argtypes = runtime._SIGS[name]
        getattr(L, name).restype = C.c_int
    L.hh_last_error.restype = C.c_char_p
    L.hh_last_error.argtypes = []
    runtime._lib = L
    from test_gltf import Capture, _one_triangle_doc, data_uri, png  # noqa: E402
    from test_hdr_ibl import write_hdr  # noqa: E402

    def expect_error(fn):
        try:
            fn()
        except runtime.HeliosError:
            return
        raise AssertionError("expected the loader to refuse the file")

    stage("glTF exported scenes")
    d = scenegen.cornell_box(width=8, height=8)
    scenegen.write_gltf(d, tmp / "cornell.gltf")
    runtime.gltf_load(tmp / "cornell.gltf", Capture())
    d = scenegen.icosphere_field(seed=1, n_spheres=2, level=1, n_leaves=4, width=8, height=8)
    scenegen.write_gltf(d, tmp / "field.glb")
    runtime.gltf_load(tmp / "field.glb", Capture())
    stage("glTF malformed accessors and views")
    bad = [({"byteOffset": -8}, None), ({"byteOffset": (1 << 52) + 4}, None), ({"byteOffset": 1 << 62}, None),
           (None, {"byteStride": 1 << 40}), (None, {"byteStride": -12}), (None, {"byteStride": 8}),
           ({"count": 1e300}, None), ({"count": -1}, None), ({"count": 4}, None), (None, {"byteOffset": -4}),
           (None, {"byteLength": 1 << 62}), ({"bufferView": 7}, None), ({"componentType": 1}, None)]
    for k, (acc, view) in enumerate(bad):
        p = tmp / f"bad{k}.gltf"
        p.write_text(json.dumps(_one_triangle_doc(acc, view)))
        expect_error(lambda: runtime.gltf_load(p, Capture()))
    doc = _one_triangle_doc()
    doc["nodes"] = [{"mesh": 0, "children": [1e300, -1, 0.5]}]
    (tmp / "kids.gltf").write_text(json.dumps(doc))
    expect_error(lambda: runtime.gltf_load(tmp / "kids.gltf", Capture()))
    for text in ('{"asset": {"version": "2.0"}, "nodes": [ }', '{"a": "\\u12', "[" * 10000, '{"x": 1e99999}'):
        (tmp / "j.gltf").write_text(text)
        try:
            runtime.gltf_load(tmp / "j.gltf", Capture())
        except runtime.HeliosError:
            pass
    stage("PNG textures (valid, truncated, bad CRC)")
    good = png(2, 6, 8, [bytes([1, 2, 3, 4, 5, 6, 7, 8]), bytes([9, 10, 11, 12, 13, 14, 15, 16])])
    for k, blob in enumerate([good, good[:40], good[:-8] + b"\x00" * 8, good[:16] + b"\xff" * 17 + good[33:]]):
        doc = _one_triangle_doc()
        doc["images"] = [{"uri": data_uri(blob, "image/png")}]
        doc["textures"] = [{"source": 0}]
        doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}]
        doc["meshes"][0]["primitives"][0]["material"] = 0
        (tmp / f"png{k}.gltf").write_text(json.dumps(doc))
        try:
            runtime.gltf_load(tmp / f"png{k}.gltf", Capture())
        except runtime.HeliosError:
            pass
    stage("PNG writer")
    img = (np.arange(5 * 7 * 4) % 251).astype(np.uint8).reshape(5, 7, 4)
    runtime.check(L.hh_write_png(str(tmp / "w.png").encode(), img.ctypes.data, 7, 5, 0))
    stage("Radiance .hdr (flat, RLE, truncated)")
    rgb = np.random.default_rng(1).random((8, 20, 3)) * 4
    write_hdr(tmp / "a.hdr", rgb, rle=True)
    runtime.ibl_from_hdr(tmp / "a.hdr", 4)
    write_hdr(tmp / "b.hdr", rgb, rle=False, bottom_up=True)
    runtime.ibl_from_hdr(tmp / "b.hdr", 4)
    data = (tmp / "a.hdr").read_bytes()
    for cut in (len(data) - 1, len(data) // 2, 60):
        (tmp / "c.hdr").write_bytes(data[:cut])
        expect_error(lambda: runtime.ibl_from_hdr(tmp / "c.hdr", 4))
    (tmp / "d.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 40000\n" + b"\x02\x02\xff\xff" * 4)
    expect_error(lambda: runtime.ibl_from_hdr(tmp / "d.hdr", 4))


if __name__ == "__main__":
    oracle_stages()
    with tempfile.TemporaryDirectory() as t:
        hostio_stages(Path(t))
    print("SANITIZED OK", flush=True)
