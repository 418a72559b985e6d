"""CPU: glTF 2.0 ingestion (hh_gltf_load through a Python sink).

The reference's own scene format (AssetCore) is absent, so its loader cannot
be pinned; these tests pin the glTF loader against the glTF 2.0
specification and against the scenes it must reproduce:
  * a hand-written file exercises each mapping rule, with expected values
    computed here from the specification (TRS order, quaternion convention,
    factors, emissive strength, alpha cutoff, channel assignment, flat normals
    for primitives without NORMAL, uint16 / absent indices, light direction
    -Z, cone angles in degrees);
  * scenes exported with scenegen.write_gltf and read back render like the
    scenes they came from (oracle: identical primary hits, radiance equal to
    float rounding);
  * PNG textures of every colour type and several bit depths decode to the
    values the PNG specification defines.
"""
import base64
import json
import math
import struct
import zlib

import numpy as np
import pytest

from helios_amd import abi, runtime, scenegen


class Capture:
    def __init__(self):
        self.textures, self.materials, self.meshes, self.instances, self.lights = [], [], [], [], []

    def add_texture(self, w, h, fmt, texels):
        self.textures.append((w, h, fmt, texels))
        return len(self.textures) - 1

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_mesh(self, v, i, s, m):
        self.meshes.append((v, i, s, m))
        return len(self.meshes) - 1

    def add_instance(self, d):
        self.instances.append(d)
        return len(self.instances) - 1

    def add_light(self, d):
        self.lights.append(d)
        return len(self.lights) - 1


def png(width, ctype, depth, rows, plte=None, trns=None):
    """A PNG from raw scanline bytes (filter 0), assembled by hand."""
    h = len(rows)
    raw = b"".join(b"\x00" + r for r in rows)

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", width, h, depth, ctype, 0, 0, 0))
    if plte is not None:
        out += chunk(b"PLTE", plte)
    if trns is not None:
        out += chunk(b"tRNS", trns)
    return out + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")


def data_uri(b, mime="application/octet-stream"):
    return f"data:{mime};base64," + base64.b64encode(b).decode()


def quat_y(deg):
    a = math.radians(deg) / 2
    return [0.0, math.sin(a), 0.0, math.cos(a)]


def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def srgb(c):
    c = c / 255.0
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


@pytest.fixture
def hand_file(tmp_path):
    # quad (two triangles, uint16 indices) with normals and UVs; a lone triangle without normals or indices
    P0 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    N0 = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    U0 = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    I0 = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    P1 = np.array([[0, 0, 0], [0, 0, -1], [0, 2, 0]], np.float32)      # face normal = (1, 0, 0)... up to sign
    parts = [P0.tobytes(), N0.tobytes(), U0.tobytes(), I0.tobytes() + b"\x00\x00", P1.tobytes()]
    offs = np.cumsum([0] + [len(p) for p in parts]).tolist()
    blob = b"".join(parts)
    image = png(2, 2, 8, [bytes([200, 100, 50, 10, 20, 30]), bytes([0, 255, 128, 255, 0, 64])])
    views = [{"buffer": 0, "byteOffset": offs[k], "byteLength": len(parts[k]) - (2 if k == 3 else 0)}
             for k in range(5)]
    views.append({"buffer": 1, "byteOffset": 0, "byteLength": len(image)})
    acc = [
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 2, "componentType": 5126, "count": 4, "type": "VEC2"},
        {"bufferView": 3, "componentType": 5123, "count": 6, "type": "SCALAR"},
        {"bufferView": 4, "componentType": 5126, "count": 3, "type": "VEC3"},
    ]
    doc = {
        "asset": {"version": "2.0"},
        "buffers": [{"byteLength": len(blob), "uri": data_uri(blob)},
                    {"byteLength": len(image), "uri": "tex%20image.bin"}],
        "bufferViews": views, "accessors": acc,
        "images": [{"bufferView": 5, "mimeType": "image/png"}],
        "textures": [{"source": 0}],
        "materials": [
            {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 0.125, 1.0], "metallicFactor": 0.3,
                                      "roughnessFactor": 0.6},
             "emissiveFactor": [1.0, 0.5, 0.0],
             "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0}}},
            {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 0.5,
                                      "metallicRoughnessTexture": {"index": 0}},
             "alphaMode": "MASK", "alphaCutoff": 0.5},
        ],
        "meshes": [{"primitives": [
            {"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3, "material": 0},
            {"attributes": {"POSITION": 4}, "material": 1}]}],
        "cameras": [{"type": "perspective", "perspective": {"yfov": 0.5, "znear": 0.05}}],
        "nodes": [
            {"translation": [1, 2, 3], "rotation": quat_y(90), "scale": [2, 2, 2], "children": [1]},
            {"translation": [0, 0, -1], "mesh": 0},
            {"translation": [5, 6, 7], "rotation": quat_y(30), "camera": 0},
            {"rotation": quat_y(90), "extensions": {"KHR_lights_punctual": {"light": 0}}},
            {"translation": [0, 4, 0], "extensions": {"KHR_lights_punctual": {"light": 1}}},
            {"translation": [1, 1, 1], "extensions": {"KHR_lights_punctual": {"light": 2}}},
        ],
        "scenes": [{"nodes": [0, 2, 3, 4, 5]}],
        "extensions": {"KHR_lights_punctual": {"lights": [
            {"type": "directional", "color": [1, 0.9, 0.8], "intensity": 3},
            {"type": "point", "intensity": 20},
            {"type": "spot", "spot": {"innerConeAngle": 0.2, "outerConeAngle": 0.4}}]}},
    }
    (tmp_path / "tex image.bin").write_bytes(image)
    path = tmp_path / "hand.gltf"
    path.write_text(json.dumps(doc))
    return path


def test_hand_written_file_maps_every_rule(hand_file):
    cap = Capture()
    info = runtime.gltf_load(hand_file, cap)
    assert (info["num_meshes"], info["num_instances"], info["num_lights"], info["num_cameras"]) == (1, 1, 3, 1)
    assert info["num_triangles"] == 3 and info["warnings"] == 0

    # materials: linear factors survive the packer's pow(v, 2.2)
    m0, m1 = cap.materials
    np.testing.assert_allclose(np.asarray(m0["albedo_value"][:3], np.float64) ** 2.2, [0.5, 0.25, 0.125], rtol=1e-6)
    assert m0["metallic_value"] == pytest.approx(0.3) and m0["roughness_value"] == pytest.approx(0.6)
    np.testing.assert_allclose(m0["emissive_value"][:3], [4.0, 2.0, 0.0])
    assert m0["alpha_test"] == 0 and m0["type"] == 0 and m0["albedo_texture"] == -1
    assert m1["alpha_test"] == 1 and m1["type"] == 0
    assert (m1["roughness_channel"], m1["metallic_channel"]) == (1, 2)
    ids = [m1["albedo_texture"], m1["roughness_texture"], m1["metallic_texture"]]
    assert len(set(ids)) == 3                       # one texture object per use
    # albedo: sRGB decoded, alpha thresholded at 0.5 (opaque RGB image -> 1)
    w, h, fmt, t = cap.textures[m1["albedo_texture"]]
    assert (w, h, fmt) == (2, 2, abi.TEXTURE_RGBA32F)
    np.testing.assert_allclose(t[0, 0], [srgb(200), srgb(100), srgb(50), 1.0], rtol=1e-6)
    np.testing.assert_allclose(t[1, 1], [srgb(255), srgb(0), srgb(64), 1.0], rtol=1e-6)
    # metallic-roughness: linear, metallic factor 0.5 on B, roughness factor 1 on G
    _, _, fmt, t = cap.textures[m1["metallic_texture"]]
    np.testing.assert_allclose(t[0, 1], [10 / 255, 20 / 255, 30 / 255 * 0.5, 1.0], rtol=1e-6)

    # mesh: two submeshes, absolute indices, submesh index in position.w
    v, i, s, mats = cap.meshes[0]
    assert list(mats) == [0, 1]
    assert s[0].tolist() == (0, 6, 4, 0, 0) and s[1].tolist() == (1, 3, 3, 4, 6)
    assert i.tolist() == [0, 1, 2, 0, 2, 3, 4, 5, 6]
    assert v["position"][:4, 3].tolist() == [0, 0, 0, 0] and v["position"][4:, 3].tolist() == [1, 1, 1]
    # flat normal of the triangle without NORMAL: normalize((v1-v0) x (v2-v0))
    e1 = np.array([0, 0, -1.0])
    e2 = np.array([0, 2, 0.0])
    fn = np.cross(e1, e2) / np.linalg.norm(np.cross(e1, e2))
    for k in range(4, 7):
        np.testing.assert_allclose(v["normal"][k, :3], fn, atol=1e-7)
    # tangent frames are orthonormal and right-handed around the normal
    for k in range(7):
        n, t, b = v["normal"][k, :3], v["tangent"][k, :3], v["bitangent"][k, :3]
        assert abs(np.dot(n, t)) < 1e-6 and abs(np.linalg.norm(t) - 1) < 1e-6
        np.testing.assert_allclose(np.abs(b), np.abs(np.cross(n, t)), atol=1e-6)

    # instance: parent T R S times child T (column-major, glTF's order)
    M = np.eye(4)
    M[:3, :3] = rot_y(90) * 2.0
    M[:3, 3] = [1, 2, 3]
    C = np.eye(4)
    C[:3, 3] = [0, 0, -1]
    G = M @ C
    inst = cap.instances[0]
    np.testing.assert_allclose(np.asarray(inst["model"]).reshape(4, 4).T, G, atol=1e-6)
    Nm = G.copy()
    Nm[:3, :3] /= 2.0
    np.testing.assert_allclose(np.asarray(inst["normal"]).reshape(4, 4).T, Nm, atol=1e-6)
    assert inst["mesh"] == 0 and inst["material_override"] == -1

    # camera: node pose, yfov in degrees, default zfar; CameraNode's lens defaults
    cam = info["camera"]
    np.testing.assert_allclose(cam["position"], [5, 6, 7])
    np.testing.assert_allclose(np.abs(np.dot(cam["orientation"], quat_y(30))), 1.0, atol=1e-6)
    assert cam["fov_y_deg"] == pytest.approx(math.degrees(0.5), rel=1e-6)
    assert (cam["near_plane"], cam["far_plane"]) == (pytest.approx(0.05), 1000.0)
    assert (cam["focal_length"], cam["aperture_radius"]) == (8.0, pytest.approx(0.1))

    # lights: forward = node -Z, punctual (radius 0), cones in degrees
    d, p, s_ = cap.lights
    assert d["type"] == abi.LIGHT_DIRECTIONAL and d["radius"] == 0.0
    np.testing.assert_allclose(d["forward"], rot_y(90) @ [0, 0, -1], atol=1e-6)
    np.testing.assert_allclose(d["color"], [1, 0.9, 0.8])
    assert d["intensity"] == 3.0
    assert p["type"] == abi.LIGHT_POINT and p["intensity"] == 20.0
    np.testing.assert_allclose(p["position"], [0, 4, 0])
    assert s_["type"] == abi.LIGHT_SPOT
    assert s_["inner_cone_deg"] == pytest.approx(math.degrees(0.2)) and s_["outer_cone_deg"] == pytest.approx(
        math.degrees(0.4))
    np.testing.assert_allclose(s_["forward"], [0, 0, -1], atol=1e-7)


def _png_file(tmp_path, name, data):
    doc = {"asset": {"version": "2.0"}, "images": [{"uri": name}], "textures": [{"source": 0}],
           "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}],
           "buffers": [{"byteLength": 36, "uri": data_uri(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]],
                                                                   np.float32).tobytes())}],
           "bufferViews": [{"buffer": 0, "byteLength": 36}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "material": 0}]}],
           "nodes": [{"mesh": 0}]}
    (tmp_path / name).write_bytes(data)
    p = tmp_path / (name + ".gltf")
    p.write_text(json.dumps(doc))
    cap = Capture()
    runtime.gltf_load(p, cap)
    w, h, fmt, t = cap.textures[0]
    assert fmt == abi.TEXTURE_RGBA8_SRGB
    return t


def test_png_colour_types(tmp_path):
    # palette, 4-bit, with tRNS alpha for entry 1
    t = _png_file(tmp_path, "pal.png", png(3, 3, 4, [bytes([0x01, 0x20]), bytes([0x21, 0x00])],
                                            plte=bytes([255, 0, 0, 0, 255, 0, 0, 0, 255]), trns=bytes([255, 7])))
    assert t[0].tolist() == [[255, 0, 0, 255], [0, 255, 0, 7], [0, 0, 255, 255]]
    assert t[1].tolist() == [[0, 0, 255, 255], [0, 255, 0, 7], [255, 0, 0, 255]]
    # 16-bit grey with a colour key: high byte kept, key -> alpha 0
    t = _png_file(tmp_path, "g16.png", png(2, 0, 16, [bytes([0x12, 0x34, 0xAB, 0xCD])], trns=bytes([0xAB, 0xCD])))
    assert t[0].tolist() == [[0x12, 0x12, 0x12, 255], [0xAB, 0xAB, 0xAB, 0]]
    # 2-bit grey scales to 0..255
    t = _png_file(tmp_path, "g2.png", png(4, 0, 2, [bytes([0b00011011])]))
    assert t[0, :, 0].tolist() == [0, 85, 170, 255]
    # grey + alpha, 8-bit; RGBA 16-bit
    t = _png_file(tmp_path, "ga.png", png(1, 4, 8, [bytes([9, 200])]))
    assert t[0].tolist() == [[9, 9, 9, 200]]
    t = _png_file(tmp_path, "rgba16.png", png(1, 6, 16, [bytes([1, 2, 3, 4, 5, 6, 7, 8])]))
    assert t[0].tolist() == [[1, 3, 5, 7]]


def _oracle_pair(oracle_mod, d, path):
    scenegen.write_gltf(d, path)
    d2, info = runtime.gltf_scene_desc(path, width=d.width, height=d.height, max_ray_bounces=d.max_ray_bounces)
    d2.camera = dict(d.camera)
    d2.sky = d.sky
    return oracle_mod.OracleScene(d), oracle_mod.OracleScene(d2), info


def test_exported_cornell_box_renders_the_same(oracle_mod, tmp_path):
    d = scenegen.cornell_box(width=48, height=40)
    a, b, info = _oracle_pair(oracle_mod, d, tmp_path / "cb.gltf")
    assert info["num_instances"] == 3 and info["num_triangles"] == d.num_triangles
    ha, hb = a.primary_hits(), b.primary_hits()
    for k in ("instance", "geometry", "primitive"):
        assert np.array_equal(ha[k], hb[k])
    np.testing.assert_allclose(a.render(frames=3), b.render(frames=3), atol=2e-5)


def test_exported_field_glb_renders_the_same(oracle_mod, tmp_path):
    d = scenegen.icosphere_field(seed=2, n_spheres=6, level=2, n_leaves=60, width=40, height=24, max_ray_bounces=4)
    for l in d.lights:
        l["radius"] = 0.0          # glTF lights are punctual
    a, b, info = _oracle_pair(oracle_mod, d, tmp_path / "field.glb")
    assert info["num_lights"] == len(d.lights) and info["num_textures"] >= 1
    ha, hb = a.primary_hits(), b.primary_hits()
    assert np.mean(ha["instance"] == hb["instance"]) == 1.0
    assert np.mean(ha["primitive"] == hb["primitive"]) == 1.0
    np.testing.assert_allclose(a.render(frames=2), b.render(frames=2), atol=5e-5)


def test_exported_textured_materials_render_the_same(oracle_mod, tmp_path):
    """Normal textures, roughness and metallic textures (one
    metallicRoughnessTexture, G and B) and a float emission texture (sRGB
    bytes x emissiveFactor x emissive strength) survive export and load: the
    same hits, every texel the shader reads equal, and the same radiance (the
    lights' radius set to 0: glTF lights are punctual)."""
    d = scenegen.textured_showcase(width=48, height=32)
    for l in d.lights:
        l["radius"] = 0.0
    scenegen.write_gltf(d, tmp_path / "tex.gltf")
    d2, info = runtime.gltf_scene_desc(tmp_path / "tex.gltf", width=d.width, height=d.height,
                                       max_ray_bounces=d.max_ray_bounces)
    m0, m1 = d.materials[0], d2.materials[0]

    def texels(desc, ti):
        w, h, fmt, data = desc.textures[ti]
        return fmt, np.asarray(data).reshape(h, w, 4)

    # what the shader reads: albedo bytes, normal bytes, roughness / metallic channels, emission
    assert np.array_equal(texels(d, m0["albedo_texture"])[1], texels(d2, m1["albedo_texture"])[1])
    assert np.array_equal(texels(d, m0["normal_texture"])[1], texels(d2, m1["normal_texture"])[1])
    for k in ("roughness", "metallic"):
        a_ = texels(d, m0[k + "_texture"])[1][..., m0[k + "_channel"]]
        b_ = texels(d2, m1[k + "_texture"])[1][..., m1[k + "_channel"]]
        assert np.array_equal(a_, b_), k
    e0 = texels(d, d.materials[1]["emissive_texture"])[1][..., :3]
    fmt, e1 = texels(d2, d2.materials[1]["emissive_texture"])
    assert np.array_equal(np.asarray(e0, np.float32), np.asarray(e1[..., :3], np.float32))
    d2.camera = dict(d.camera)
    d2.sky = d.sky
    a, b = oracle_mod.OracleScene(d), oracle_mod.OracleScene(d2)
    ha, hb = a.primary_hits(), b.primary_hits()
    for k in ("instance", "geometry", "primitive"):
        assert np.array_equal(ha[k], hb[k])
    np.testing.assert_allclose(a.render(frames=2), b.render(frames=2), atol=2e-5)


def test_unrepresentable_emission_is_refused(tmp_path):
    """A float emission texture whose texels are not sRGB bytes x one factor
    per channel is refused rather than quantized."""
    d = scenegen.textured_showcase(width=16, height=16)
    w, h, fmt, data = d.textures[d.materials[1]["emissive_texture"]]
    em = np.array(data, np.float32).reshape(h, w, 4)
    em[0, 0, 0] = 1.2345   # between two sRGB steps of the 3.0 peak
    d.textures[d.materials[1]["emissive_texture"]] = (w, h, fmt, em)
    with pytest.raises(NotImplementedError):
        scenegen.write_gltf(d, tmp_path / "bad.gltf")


def test_errors_are_reported(tmp_path):
    with pytest.raises(runtime.HeliosError) as e:
        runtime.gltf_load(tmp_path / "missing.gltf", Capture())
    assert e.value.status == 7
    bad = tmp_path / "bad.gltf"
    bad.write_text('{"asset": {"version": "2.0"}, "nodes": [ }')
    with pytest.raises(runtime.HeliosError, match="JSON"):
        runtime.gltf_load(bad, Capture())
    req = tmp_path / "req.gltf"
    req.write_text(json.dumps({"asset": {"version": "2.0"}, "extensionsRequired": ["KHR_draco_mesh_compression"]}))
    with pytest.raises(runtime.HeliosError, match="KHR_draco_mesh_compression"):
        runtime.gltf_load(req, Capture())
    with pytest.raises(runtime.HeliosError, match="JPEG"):
        _png_file(tmp_path, "photo.jpg", b"\xff\xd8\xff\xe0" + b"\x00" * 16)


def _one_triangle_doc(acc_extra=None, view_extra=None, blob_len=None):
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32).tobytes()
    blob = P + b"\x00" * 16
    view = {"buffer": 0, "byteOffset": 0, "byteLength": len(P)}
    view.update(view_extra or {})
    acc = {"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}
    acc.update(acc_extra or {})
    return {"asset": {"version": "2.0"},
            "buffers": [{"byteLength": blob_len if blob_len is not None else len(blob), "uri": data_uri(blob)}],
            "bufferViews": [view], "accessors": [acc],
            "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}],
            "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}]}


@pytest.mark.parametrize("acc_extra,view_extra,match", [
    ({"byteOffset": -8}, None, "byteOffset"),                          # negative accessor offset
    ({"byteOffset": 2}, None, "byteOffset"),                           # not a multiple of the component size
    ({"byteOffset": 1 << 62}, None, "byteOffset"),                     # beyond 2^53: not an exact integer
    ({"byteOffset": (1 << 52) + 4}, None, "out of range"),             # would wrap off + stride * (count - 1)
    (None, {"byteStride": 1 << 40}, "byteStride"),                     # huge stride
    (None, {"byteStride": -12}, "byteStride"),                         # negative stride
    (None, {"byteStride": 254}, "byteStride"),                         # above 252
    (None, {"byteStride": 6}, "byteStride"),                           # not a multiple of 4
    (None, {"byteStride": 8}, "element size"),                         # below the 12-B VEC3 element
    ({"count": 1e300}, None, "count"),                                 # beyond long's range
    ({"count": 2.5}, None, "count"),                                   # not an integer
    (None, {"byteOffset": -4}, "bufferView out of range"),             # negative view offset
    (None, {"byteLength": 1 << 62}, "bufferView out of range"),        # offset + length wraps
    ({"count": 4}, None, "accessor out of range"),                     # one element past the view
])
def test_malformed_accessors_are_refused(tmp_path, acc_extra, view_extra, match):
    """Offsets, strides and counts outside glTF 2.0's rules (or that would
    overflow the range check) are refused before any byte is read."""
    path = tmp_path / "bad_accessor.gltf"
    path.write_text(json.dumps(_one_triangle_doc(acc_extra, view_extra)))
    with pytest.raises(runtime.HeliosError, match=match):
        runtime.gltf_load(path, Capture())


def test_valid_stride_and_offset_are_read(tmp_path):
    """A strided view (stride 16 > element 12) with an accessor offset reads
    the elements it names."""
    P = np.zeros((4, 4), np.float32)
    P[1:, :3] = [[0, 0, 0], [2, 0, 0], [0, 3, 0]]
    blob = P.tobytes()
    doc = _one_triangle_doc({"byteOffset": 16}, {"byteStride": 16, "byteLength": len(blob)})
    doc["buffers"] = [{"byteLength": len(blob), "uri": data_uri(blob)}]
    path = tmp_path / "strided.gltf"
    path.write_text(json.dumps(doc))
    cap = Capture()
    runtime.gltf_load(path, cap)
    v = cap.meshes[0][0]
    assert np.allclose(v["position"][:, :3], [[0, 0, 0], [2, 0, 0], [0, 3, 0]])
