"""ctypes mirrors of the C structs declared in include/helios_hip.h.

These are the byte layouts of the C ABI (and, where noted there, of the
reference's own GPU records). Shared by the product wrapper and the tests.
"""
import ctypes as C

import numpy as np


class Vertex(C.Structure):
    _fields_ = [("position", C.c_float * 4), ("tex_coord", C.c_float * 4), ("normal", C.c_float * 4),
                ("tangent", C.c_float * 4), ("bitangent", C.c_float * 4)]


class SubMesh(C.Structure):
    _fields_ = [("mat_idx", C.c_uint32), ("index_count", C.c_uint32), ("vertex_count", C.c_uint32),
                ("base_vertex", C.c_uint32), ("base_index", C.c_uint32)]


class MaterialDesc(C.Structure):
    _fields_ = [("albedo_value", C.c_float * 4), ("emissive_value", C.c_float * 4),
                ("metallic_value", C.c_float), ("roughness_value", C.c_float),
                ("albedo_texture", C.c_int32), ("normal_texture", C.c_int32),
                ("roughness_texture", C.c_int32), ("roughness_channel", C.c_int32),
                ("metallic_texture", C.c_int32), ("metallic_channel", C.c_int32),
                ("emissive_texture", C.c_int32), ("type", C.c_uint32), ("alpha_test", C.c_uint32)]


class InstanceDesc(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("normal", C.c_float * 16), ("mesh", C.c_uint32),
                ("material_override", C.c_int32)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("color", C.c_float * 3), ("intensity", C.c_float), ("radius", C.c_float),
                ("position", C.c_float * 3), ("forward", C.c_float * 3), ("inner_cone_deg", C.c_float),
                ("outer_cone_deg", C.c_float)]


class SkyParams(C.Structure):
    _fields_ = [("turbidity", C.c_float), ("albedo", C.c_float), ("normalized_sun_y", C.c_float),
                ("cube_size", C.c_uint32), ("equirect_width", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("orientation", C.c_float * 4), ("fov_y_deg", C.c_float),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("focal_length", C.c_float),
                ("aperture_radius", C.c_float)]


class Settings(C.Structure):
    _fields_ = [("max_ray_bounces", C.c_uint32), ("max_samples", C.c_uint32), ("shadow_ray_bias", C.c_float),
                ("tiled", C.c_uint32), ("rr_start_depth", C.c_uint32), ("flags", C.c_uint32)]


class LaunchConsts(C.Structure):
    _fields_ = [("view_proj_inverse", C.c_float * 16), ("camera_pos", C.c_float * 4), ("up_direction", C.c_float * 4),
                ("right_direction", C.c_float * 4), ("focal_plane", C.c_float * 4),
                ("ray_debug_pixel_coord", C.c_int32 * 4), ("launch_id_size", C.c_uint32 * 4),
                ("accumulation", C.c_float), ("num_lights", C.c_uint32), ("num_frames", C.c_uint32),
                ("debug_vis", C.c_uint32), ("max_ray_bounces", C.c_uint32), ("shadow_ray_bias", C.c_float),
                ("focal_length", C.c_float), ("aperture_radius", C.c_float)]


class Ray(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("tmin", C.c_float), ("direction", C.c_float * 3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("instance", C.c_uint32),
                ("geometry", C.c_uint32), ("primitive", C.c_uint32)]


class View(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("surface", C.c_void_p), ("camera", Camera), ("spp", C.c_uint32),
                ("max_ray_bounces", C.c_uint32), ("shadow_ray_bias", C.c_float), ("first_frame", C.c_uint32),
                ("flags", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("bvh_width", C.c_uint32), ("extend_block", C.c_uint32), ("shade_block", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("persistent_blocks_per_cu", C.c_uint32),
                ("frames_in_flight", C.c_uint32), ("bvh_compressed", C.c_uint32), ("refill_lanes", C.c_uint32),
                ("traversal", C.c_uint32), ("bvh_builder", C.c_uint32), ("continuation_depth", C.c_uint32)]


assert C.sizeof(Vertex) == 80
assert C.sizeof(LaunchConsts) == 192
assert C.sizeof(Hit) == 24
assert C.sizeof(Ray) == 32

# glTF ingestion (hh_scene_sink / hh_gltf_info)
SINK_ADD_TEXTURE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                               C.POINTER(C.c_uint32))
SINK_ADD_MATERIAL = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(MaterialDesc), C.POINTER(C.c_uint32))
SINK_ADD_MESH = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                            C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32))
SINK_ADD_INSTANCE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(InstanceDesc), C.POINTER(C.c_uint32))
SINK_ADD_LIGHT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(LightDesc), C.POINTER(C.c_uint32))


class SceneSink(C.Structure):
    _fields_ = [("user", C.c_void_p), ("add_texture", SINK_ADD_TEXTURE), ("add_material", SINK_ADD_MATERIAL),
                ("add_mesh", SINK_ADD_MESH), ("add_instance", SINK_ADD_INSTANCE), ("add_light", SINK_ADD_LIGHT)]


class GltfInfo(C.Structure):
    _fields_ = [("num_meshes", C.c_uint32), ("num_instances", C.c_uint32), ("num_materials", C.c_uint32),
                ("num_textures", C.c_uint32), ("num_lights", C.c_uint32), ("num_cameras", C.c_uint32),
                ("has_camera", C.c_uint32), ("warnings", C.c_uint32), ("num_triangles", C.c_uint64),
                ("camera", Camera)]


GLTF_WARN_TEXCOORD1 = 0x1
GLTF_WARN_NON_TRIANGLES = 0x2
SUBMESH_DTYPE = np.dtype([("mat_idx", "<u4"), ("index_count", "<u4"), ("vertex_count", "<u4"),
                          ("base_vertex", "<u4"), ("base_index", "<u4")])

# numpy views of the array-of-struct records
VERTEX_DTYPE = np.dtype([("position", "<f4", 4), ("tex_coord", "<f4", 4), ("normal", "<f4", 4),
                         ("tangent", "<f4", 4), ("bitangent", "<f4", 4)])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instance", "<u4"), ("geometry", "<u4"),
                      ("primitive", "<u4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])

HH_TRACE_OPAQUE = 0x1
HH_TRACE_TERMINATE_ON_FIRST = 0x4
HH_FLAG_NO_RADIANCE_CLAMP = 0x1
HH_FLAG_DIRECT_ONLY = 0x2
HH_FLAG_MEGAKERNEL = 0x4
HH_FLAG_LIGHT_POWER = 0x8

# Output-buffer views (Renderer::OutputBuffer, renderer.h:27-35)
OUTPUT_BUFFER_ALBEDO = 0
OUTPUT_BUFFER_NORMALS = 1
OUTPUT_BUFFER_ROUGHNESS = 2
OUTPUT_BUFFER_METALLIC = 3
OUTPUT_BUFFER_EMISSIVE = 4
OUTPUT_BUFFER_FINAL = 5

# DebugRayVertex (common.glsl:54-58); MAX_DEBUG_RAY_DRAW_COUNT (renderer.h:9)
DEBUG_RAY_VERTEX_DTYPE = np.dtype([("position", "<f4", 4), ("color", "<f4", 4)])
HH_MAX_DEBUG_RAY_DRAW_COUNT = 1024

TEXTURE_RGBA8_UNORM = 0
TEXTURE_RGBA8_SRGB = 1
TEXTURE_RGBA32F = 2

LIGHT_DIRECTIONAL = 0
LIGHT_SPOT = 1
LIGHT_POINT = 2

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "LIMIT_EXCEEDED", 4: "DEVICE_ERROR",
          5: "NAN_DETECTED", 6: "NOT_READY", 7: "IO"}


def ptr(a, ctype=C.c_void_p):
    """ctypes pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctype)
