"""Python mirror of the reference's render-path interface over the C ABI.

Class and method names follow the reference:
  PathIntegrator  include/gfx/path_integrator.h:9-62 (render, restart_bake,
                  set_max_ray_bounces, set_max_samples, set_shadow_ray_bias,
                  set_tiled, num_accumulated_samples, num_target_samples, ...)
  Scene           include/resource/scene.h:386-470 (built from a SceneDesc)
  Renderer        tone map / exposure / operator (include/gfx/renderer.h:75-97)
The GPU library (libhelios_hip.so, HIP for gfx950) does all the work; this
module only marshals arguments. It raises HeliosError on any non-OK status and
refuses to run when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import abi, scenegen
from .abi import ptr

LIB_PATH = Path(__file__).resolve().parent / "libhelios_hip.so"

_lib = None


class HeliosError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{abi.STATUS.get(status, status)}: {msg}")
        self.status = status


# name -> argtypes (restype is int status unless noted)
_SIGS = {
    "hh_context_create": [C.c_int, C.POINTER(C.c_void_p)],
    "hh_context_destroy": [C.c_void_p],
    "hh_sync": [C.c_void_p],
    "hh_context_stream": [C.c_void_p, C.POINTER(C.c_void_p)],
    "hh_context_set_tuning": [C.c_void_p, C.POINTER(abi.Tuning)],
    "hh_context_get_tuning": [C.c_void_p, C.POINTER(abi.Tuning)],
    "hh_scene_create": [C.c_void_p, C.POINTER(C.c_void_p)],
    "hh_scene_destroy": [C.c_void_p],
    "hh_scene_add_texture": [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)],
    "hh_scene_add_material": [C.c_void_p, C.POINTER(abi.MaterialDesc), C.POINTER(C.c_uint32)],
    "hh_scene_add_mesh": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                          C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)],
    "hh_scene_add_instance": [C.c_void_p, C.POINTER(abi.InstanceDesc), C.POINTER(C.c_uint32)],
    "hh_scene_set_instance_transform": [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p],
    "hh_scene_add_light": [C.c_void_p, C.POINTER(abi.LightDesc), C.POINTER(C.c_uint32)],
    "hh_scene_set_ibl": [C.c_void_p, C.c_uint32, C.c_void_p],
    "hh_scene_set_sky": [C.c_void_p, C.POINTER(abi.SkyParams)],
    "hh_scene_commit": [C.c_void_p],
    "hh_scene_num_lights": [C.c_void_p, C.POINTER(C.c_uint32)],
    "hh_scene_read_lights": [C.c_void_p, C.c_void_p, C.c_uint32],
    "hh_scene_read_env": [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)],
    "hh_scene_sky_coefficients": [C.c_void_p, C.c_void_p],
    "hh_scene_read_env_equirect": [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)],
    "hh_scene_bvh_stats": [C.c_void_p, C.c_void_p],
    "hh_scene_traversal_faults": [C.c_void_p, C.c_void_p, C.c_int],
    "hh_surface_create": [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)],
    "hh_surface_destroy": [C.c_void_p],
    "hh_surface_clear": [C.c_void_p],
    "hh_surface_read": [C.c_void_p, C.c_void_p],
    "hh_surface_write": [C.c_void_p, C.c_void_p],
    "hh_surface_device_ptr": [C.c_void_p, C.POINTER(C.c_void_p)],
    "hh_surface_enable_features": [C.c_void_p],
    "hh_surface_validate": [C.c_void_p, C.POINTER(C.c_uint64)],
    "hh_surface_read_features": [C.c_void_p, C.c_void_p],
    "hh_tonemap": [C.c_void_p, C.c_float, C.c_uint32, C.c_void_p],
    "hh_integrator_create": [C.c_void_p, C.POINTER(C.c_void_p)],
    "hh_integrator_destroy": [C.c_void_p],
    "hh_integrator_set_settings": [C.c_void_p, C.POINTER(abi.Settings)],
    "hh_integrator_get_settings": [C.c_void_p, C.POINTER(abi.Settings)],
    "hh_integrator_restart_bake": [C.c_void_p],
    "hh_integrator_render": [C.c_void_p, C.c_void_p, C.POINTER(abi.Camera), C.c_void_p],
    "hh_integrator_render_frames": [C.c_void_p, C.c_void_p, C.POINTER(abi.Camera), C.c_void_p, C.c_uint32],
    "hh_integrator_counters": [C.c_void_p, C.c_void_p],
    "hh_integrator_ray_stats": [C.c_void_p, C.c_void_p],
    "hh_compute_launch_consts": [C.POINTER(abi.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.POINTER(abi.Settings), C.POINTER(abi.LaunchConsts)],
    "hh_render_batch": [C.c_void_p, C.c_void_p, C.c_uint32],
    "hh_context_autotune": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(abi.Tuning),
                            C.POINTER(C.c_double)],
    "hh_trace_rays": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p],
    "hh_trace_rays_noalloc": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p],
    "hh_primary_hits": [C.c_void_p, C.POINTER(abi.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p],
    "hh_sky_fit": [C.POINTER(abi.SkyParams), C.c_void_p, C.c_void_p],
    "hh_context_set_profiling": [C.c_void_p, C.c_uint32],
    "hh_context_read_profile": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "hh_context_reset_profile": [C.c_void_p],
    "hh_context_profile_rays": [C.c_void_p, C.c_void_p, C.c_void_p],
    "hh_bandwidth_probe": [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)],
    "hh_scene_load_gltf": [C.c_void_p, C.c_char_p, C.POINTER(abi.GltfInfo)],
    "hh_ibl_from_hdr": [C.c_char_p, C.c_uint32, C.c_void_p],
    "hh_scene_load_ibl_hdr": [C.c_void_p, C.c_char_p, C.c_uint32],
    "hh_gltf_load": [C.c_char_p, C.POINTER(abi.SceneSink), C.POINTER(abi.GltfInfo)],
    "hh_render_output_buffer": [C.c_void_p, C.POINTER(abi.Camera), C.c_uint32, C.c_void_p],
    "hh_surface_save_png": [C.c_void_p, C.c_float, C.c_uint32, C.c_char_p],
    "hh_write_png": [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32],
    "hh_debug_rays_create": [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)],
    "hh_debug_rays_destroy": [C.c_void_p],
    "hh_debug_rays_clear": [C.c_void_p],
    "hh_debug_rays_read": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "hh_debug_rays_device_ptr": [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    "hh_integrator_gather_debug_rays": [C.c_void_p, C.c_void_p, C.POINTER(abi.Camera), C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p],
}

KERNEL_KINDS = ("generate", "extend", "shade", "shadow", "accumulate", "megakernel")


def lib():
    """Load libhelios_hip.so. Raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HeliosError(4, f"{LIB_PATH} is missing; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.hh_last_error.restype = C.c_char_p
        L.hh_last_error.argtypes = []
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise HeliosError(status, lib().hh_last_error().decode(errors="replace"))


def _struct(cls, d):
    s = cls()
    for k, v in d.items():
        f = getattr(s, k)
        if isinstance(f, C.Array):
            for i, x in enumerate(v):
                f[i] = x
        else:
            setattr(s, k, v)
    return s


def camera_struct(cam):
    return cam if isinstance(cam, abi.Camera) else _struct(abi.Camera, cam)


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().hh_context_create(device, C.byref(self.h)))

    def sync(self):
        check(lib().hh_sync(self.h))

    def set_tuning(self, bvh_width=4, extend_block=256, queue_capacity=0, persistent_blocks_per_cu=8,
                   frames_in_flight=4, bvh_compressed=False, refill_lanes=24, traversal=0, bvh_builder=3,
                   continuation_depth=0):
        t = abi.Tuning(bvh_width, extend_block, extend_block, queue_capacity, persistent_blocks_per_cu,
                       frames_in_flight, int(bool(bvh_compressed)), refill_lanes, traversal, bvh_builder,
                       continuation_depth)
        check(lib().hh_context_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = abi.Tuning()
        check(lib().hh_context_get_tuning(self.h, C.byref(t)))
        return dict(bvh_width=t.bvh_width, extend_block=t.extend_block, queue_capacity=t.queue_capacity,
                    persistent_blocks_per_cu=t.persistent_blocks_per_cu, frames_in_flight=t.frames_in_flight,
                    bvh_compressed=bool(t.bvh_compressed), refill_lanes=t.refill_lanes, traversal=t.traversal,
                    bvh_builder=t.bvh_builder, continuation_depth=t.continuation_depth)

    def stream(self):
        s = C.c_void_p()
        check(lib().hh_context_stream(self.h, C.byref(s)))
        return s.value

    def set_profiling(self, timing=False, counting=False):
        check(lib().hh_context_set_profiling(self.h, (1 if timing else 0) | (2 if counting else 0)))

    def reset_profile(self):
        check(lib().hh_context_reset_profile(self.h))

    def bandwidth_probe(self, nbytes=1 << 30, iters=4):
        """Read+write GB/s of a plain device copy of nbytes (HBM calibration)."""
        out = C.c_double()
        check(lib().hh_bandwidth_probe(self.h, nbytes, iters, C.byref(out)))
        return out.value

    def read_profile(self):
        """Per kernel kind: total ms, launches (HIP events on the context
        stream) and the extension / shadow rays those launches traced, plus
        traversal fetch counts of the counting variant."""
        ms = np.zeros(6, np.float64)
        n = np.zeros(6, np.uint64)
        c = np.zeros(6, np.uint64)
        check(lib().hh_context_read_profile(self.h, ptr(ms), ptr(n), ptr(c)))
        er = np.zeros(6, np.uint64)
        sr = np.zeros(6, np.uint64)
        check(lib().hh_context_profile_rays(self.h, ptr(er), ptr(sr)))
        out = {k: dict(ms=float(ms[i]), launches=int(n[i]), extension_rays=int(er[i]), shadow_rays=int(sr[i]))
               for i, k in enumerate(KERNEL_KINDS)}
        out["counts"] = dict(ext_nodes=int(c[0]), ext_tris=int(c[1]), ext_insts=int(c[2]), shadow_nodes=int(c[3]),
                             shadow_tris=int(c[4]), shadow_insts=int(c[5]))
        return out

    def close(self):
        if self.h:
            lib().hh_context_destroy(self.h)
            self.h = None


class Scene:
    """helios::Scene built from a scenegen.SceneDesc and committed to the GPU."""

    def __init__(self, ctx, desc=None, commit=True):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().hh_scene_create(ctx.h, C.byref(self.h)))
        self.desc = desc
        if desc is not None:
            self.load(desc)
            if commit:
                self.commit()

    def load(self, d):
        L = lib()
        out = C.c_uint32()
        for (w, h, fmt, data) in d.textures:
            data = np.ascontiguousarray(data)
            check(L.hh_scene_add_texture(self.h, w, h, fmt, ptr(data), C.byref(out)))
        for m in d.materials:
            check(L.hh_scene_add_material(self.h, C.byref(_struct(abi.MaterialDesc, m)), C.byref(out)))
        for me in d.meshes:
            sm = np.ascontiguousarray(np.array(me.submeshes, np.uint32).reshape(-1, 5))
            mats = np.ascontiguousarray(np.array(me.materials, np.uint32))
            v = np.ascontiguousarray(me.vertices)
            i = np.ascontiguousarray(me.indices, dtype=np.uint32)
            check(L.hh_scene_add_mesh(self.h, ptr(v), len(v), ptr(i), len(i), ptr(sm), len(sm), ptr(mats), len(mats),
                                      C.byref(out)))
        for (model, normal, mesh, ov) in d.instances:
            check(L.hh_scene_add_instance(self.h, C.byref(_struct(abi.InstanceDesc, dict(
                model=list(model), normal=list(normal), mesh=mesh, material_override=ov))), C.byref(out)))
        for l in d.lights:
            check(L.hh_scene_add_light(self.h, C.byref(_struct(abi.LightDesc, l)), C.byref(out)))
        if d.ibl is not None:
            size, faces = d.ibl
            faces = np.ascontiguousarray(faces, np.float32)
            check(L.hh_scene_set_ibl(self.h, size, ptr(faces)))
        check(L.hh_scene_set_sky(self.h, C.byref(abi.SkyParams(*d.sky))))

    def commit(self):
        check(lib().hh_scene_commit(self.h))

    def load_ibl_hdr(self, path, size=512):
        """IBL cube (size x size faces) from a Radiance .hdr lat-long image."""
        check(lib().hh_scene_load_ibl_hdr(self.h, str(path).encode(), size))

    def load_gltf(self, path):
        """Add the contents of a glTF 2.0 file (.gltf / .glb); returns gltf_info
        as a dict (counts, warnings, 'camera' or None). Commit afterwards."""
        info = abi.GltfInfo()
        check(lib().hh_scene_load_gltf(self.h, str(path).encode(), C.byref(info)))
        return _gltf_info(info)

    def num_lights(self):
        n = C.c_uint32()
        check(lib().hh_scene_num_lights(self.h, C.byref(n)))
        return n.value

    def lights(self):
        n = self.num_lights()
        out = np.zeros((max(n, 1), 16), np.float32)
        check(lib().hh_scene_read_lights(self.h, ptr(out), n))
        return out[:n]

    def env(self):
        n = C.c_uint32()
        check(lib().hh_scene_read_env(self.h, None, C.byref(n)))
        out = np.zeros((6, n.value, n.value, 4), np.float32)
        check(lib().hh_scene_read_env(self.h, ptr(out), C.byref(n)))
        return out

    def env_equirect(self):
        """Equirectangular sky as float16 (H, W, 4), or None for the cube."""
        w = C.c_uint32()
        check(lib().hh_scene_read_env_equirect(self.h, None, C.byref(w)))
        if not w.value:
            return None
        out = np.zeros((w.value // 2, w.value, 4), np.uint16)
        check(lib().hh_scene_read_env_equirect(self.h, ptr(out), C.byref(w)))
        return out.view(np.float16)

    def sky_coefficients(self):
        out = np.zeros(30, np.float32)
        check(lib().hh_scene_sky_coefficients(self.h, ptr(out)))
        return out.reshape(10, 3)

    def bvh_stats(self):
        out = np.zeros(4, np.uint64)
        check(lib().hh_scene_bvh_stats(self.h, ptr(out)))
        return dict(blas_nodes=int(out[0]), triangles=int(out[1]), tlas_nodes=int(out[2]), meshes=int(out[3]))

    def traversal_faults(self, reset=False):
        """Rays cut short by a traversal bound since commit (or the last reset):
        dropped lane-stack pushes, dropped packet-stack pushes, visit-bound stops."""
        out = np.zeros(3, np.uint64)
        check(lib().hh_scene_traversal_faults(self.h, ptr(out), int(reset)))
        return dict(lane_stack=int(out[0]), packet_stack=int(out[1]), trips=int(out[2]))

    def trace_rays(self, rays, flags=0):
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        out = np.zeros(len(rays), abi.HIT_DTYPE)
        check(lib().hh_trace_rays(self.h, ptr(rays), len(rays), flags, ptr(out)))
        return out

    def trace_rays_device(self, d_rays, n, flags, d_hits):
        """hh_trace_rays_noalloc on device buffers (addresses as integers): n
        hh_ray records in, n hh_hit records out, asynchronous on the context's
        stream; nothing is allocated."""
        check(lib().hh_trace_rays_noalloc(self.h, C.c_void_p(d_rays), n, flags, C.c_void_p(d_hits)))

    def primary_hits(self, camera, width, height, frame=0):
        out = np.zeros((height, width), abi.HIT_DTYPE)
        check(lib().hh_primary_hits(self.h, C.byref(camera_struct(camera)), width, height, frame, ptr(out)))
        return out

    def close(self):
        if self.h:
            lib().hh_scene_destroy(self.h)
            self.h = None


class Surface:
    """The RGBA32F output image (row 0 = NDC y = -1)."""

    def __init__(self, ctx, width, height):
        self.ctx = ctx
        self.width, self.height = width, height
        self.h = C.c_void_p()
        check(lib().hh_surface_create(ctx.h, width, height, C.byref(self.h)))

    def clear(self):
        check(lib().hh_surface_clear(self.h))

    def read(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(lib().hh_surface_read(self.h, ptr(out)))
        return out

    def write(self, img):
        img = np.ascontiguousarray(img, np.float32)
        check(lib().hh_surface_write(self.h, ptr(img)))

    def validate(self):
        """Number of non-finite components (raises HH_ERR_NAN_DETECTED when any)."""
        n = C.c_uint64()
        check(lib().hh_surface_validate(self.h, C.byref(n)))
        return n.value

    def enable_features(self):
        check(lib().hh_surface_enable_features(self.h))

    def read_features(self):
        """(albedo rgb + 1/(1+t), normal xyz + M2) planes, each (H, W, 4)."""
        out = np.zeros((2, self.height, self.width, 4), np.float32)
        check(lib().hh_surface_read_features(self.h, ptr(out)))
        return out[0], out[1]

    def device_ptr(self):
        p = C.c_void_p()
        check(lib().hh_surface_device_ptr(self.h, C.byref(p)))
        return p.value

    def tone_map(self, exposure=1.0, operator=0):
        """tone_map.frag: ACES (0) / Reinhard (1), exposure, gamma 1/2.2 -> RGBA8."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().hh_tonemap(self.h, exposure, operator, ptr(out)))
        return out

    def save_png(self, path, exposure=1.0, operator=0):
        """Renderer::save_image_to_disk: tone-mapped RGBA8 PNG, top row first."""
        check(lib().hh_surface_save_png(self.h, exposure, operator, str(path).encode()))

    def close(self):
        if self.h:
            lib().hh_surface_destroy(self.h)
            self.h = None


class PathIntegrator:
    """helios::PathIntegrator over the GPU wavefront integrator."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().hh_integrator_create(ctx.h, C.byref(self.h)))

    def _settings(self):
        s = abi.Settings()
        check(lib().hh_integrator_get_settings(self.h, C.byref(s)))
        return s

    def _set(self, **kw):
        s = self._settings()
        for k, v in kw.items():
            setattr(s, k, v)
        check(lib().hh_integrator_set_settings(self.h, C.byref(s)))

    # path_integrator.h:18-37
    def max_ray_bounces(self):
        return self._settings().max_ray_bounces

    def max_samples(self):
        return self._settings().max_samples

    def shadow_ray_bias(self):
        return self._settings().shadow_ray_bias

    def is_tiled(self):
        return bool(self._settings().tiled)

    def set_max_ray_bounces(self, n):
        self._set(max_ray_bounces=n)

    def set_max_samples(self, n):
        self._set(max_samples=n)

    def set_shadow_ray_bias(self, b):
        self._set(shadow_ray_bias=b)

    def set_tiled(self, tiled):
        self._set(tiled=int(bool(tiled)))

    def set_flags(self, flags):
        self._set(flags=flags)

    def set_rr_start_depth(self, depth):
        """First bounce at which Russian roulette may end a path (0 = the reference)."""
        self._set(rr_start_depth=depth)

    def restart_bake(self):
        check(lib().hh_integrator_restart_bake(self.h))

    def _counters(self):
        out = np.zeros(4, np.uint32)
        check(lib().hh_integrator_counters(self.h, ptr(out)))
        return out

    def num_accumulated_samples(self):
        return int(self._counters()[0])

    def num_target_samples(self):
        return int(self._counters()[1])

    def tile_idx(self):
        return int(self._counters()[2])

    def render(self, scene, camera, surface, frames=1):
        check(lib().hh_integrator_render_frames(self.h, scene.h, C.byref(camera_struct(camera)), surface.h, frames))

    def gather_debug_rays(self, scene, camera, width, height, pixel, num_debug_rays, buffer):
        """PathIntegrator::gather_debug_rays: append the segments of num_debug_rays
        paths through window pixel (x, y) (rows top-down) to a DebugRays buffer."""
        check(lib().hh_integrator_gather_debug_rays(self.h, scene.h, C.byref(camera_struct(camera)), width, height,
                                                    int(pixel[0]), int(pixel[1]), num_debug_rays, buffer.h))

    def ray_stats(self):
        out = np.zeros(3, np.uint64)
        check(lib().hh_integrator_ray_stats(self.h, ptr(out)))
        return dict(extension_rays=int(out[0]), shadow_rays=int(out[1]), frames=int(out[2]))

    def close(self):
        if self.h:
            lib().hh_integrator_destroy(self.h)
            self.h = None


def ibl_from_hdr(path, size):
    """The six (size, size, 4) RGBA32F cube faces hh_scene_set_ibl takes, from
    a Radiance .hdr lat-long image (host only)."""
    out = np.zeros((6, size, size, 4), np.float32)
    check(lib().hh_ibl_from_hdr(str(path).encode(), size, ptr(out)))
    return out


def _gltf_info(info):
    d = {k: int(getattr(info, k)) for k, _ in abi.GltfInfo._fields_ if k != "camera"}
    cam = info.camera
    d["camera"] = dict(position=tuple(cam.position), orientation=tuple(cam.orientation), fov_y_deg=cam.fov_y_deg,
                       near_plane=cam.near_plane, far_plane=cam.far_plane, focal_length=cam.focal_length,
                       aperture_radius=cam.aperture_radius) if info.has_camera else None
    return d


def gltf_load(path, sink):
    """hh_gltf_load with a Python sink: an object with add_texture(w, h, fmt,
    texels (h, w, 4)), add_material(dict), add_mesh(vertices, indices,
    submeshes, materials), add_instance(dict), add_light(dict), each returning
    the new id. Returns gltf_info as a dict."""
    errors = []

    def guard(fn):
        def cb(*args):
            try:
                return fn(*args)
            except Exception as e:  # reported after the C call returns
                errors.append(e)
                return 1
        return cb

    def fields(struct):
        return {k: (list(getattr(struct, k)) if isinstance(getattr(struct, k), C.Array) else getattr(struct, k))
                for k, _ in struct._fields_}

    def array(addr, ctype, n):
        return np.ctypeslib.as_array((ctype * n).from_address(addr)).copy()

    def add_texture(user, w, h, fmt, texels, out):
        ct = C.c_float if fmt == abi.TEXTURE_RGBA32F else C.c_uint8
        out[0] = sink.add_texture(w, h, fmt, array(texels, ct, w * h * 4).reshape(h, w, 4))
        return 0

    def add_material(user, d, out):
        out[0] = sink.add_material(fields(d.contents))
        return 0

    def add_mesh(user, v, nv, ix, ni, sm, ns, mats, nm, out):
        verts = array(v, C.c_uint8, nv * abi.VERTEX_DTYPE.itemsize).view(abi.VERTEX_DTYPE)
        subs = array(sm, C.c_uint8, ns * abi.SUBMESH_DTYPE.itemsize).view(abi.SUBMESH_DTYPE)
        out[0] = sink.add_mesh(verts, array(ix, C.c_uint32, ni), subs, array(mats, C.c_uint32, nm))
        return 0

    def add_instance(user, d, out):
        out[0] = sink.add_instance(fields(d.contents))
        return 0

    def add_light(user, d, out):
        out[0] = sink.add_light(fields(d.contents))
        return 0

    cbs = (abi.SINK_ADD_TEXTURE(guard(add_texture)), abi.SINK_ADD_MATERIAL(guard(add_material)),
           abi.SINK_ADD_MESH(guard(add_mesh)), abi.SINK_ADD_INSTANCE(guard(add_instance)),
           abi.SINK_ADD_LIGHT(guard(add_light)))
    s = abi.SceneSink(None, *cbs)
    info = abi.GltfInfo()
    st = lib().hh_gltf_load(str(path).encode(), C.byref(s), C.byref(info))
    if errors:
        raise errors[0]
    check(st)
    return _gltf_info(info)


class _DescSink:
    def __init__(self, desc):
        self.d = desc

    def add_texture(self, w, h, fmt, texels):
        self.d.textures.append((w, h, fmt, texels))
        return len(self.d.textures) - 1

    def add_material(self, m):
        self.d.materials.append(m)
        return len(self.d.materials) - 1

    def add_mesh(self, verts, indices, subs, mats):
        sm = [tuple(int(x) for x in r) for r in subs[["mat_idx", "index_count", "vertex_count", "base_vertex",
                                                      "base_index"]].tolist()]
        self.d.meshes.append(scenegen.MeshDesc(verts, indices, sm, [int(m) for m in mats]))
        return len(self.d.meshes) - 1

    def add_instance(self, inst):
        self.d.instances.append((np.asarray(inst["model"], np.float32), np.asarray(inst["normal"], np.float32),
                                 inst["mesh"], inst["material_override"]))
        return len(self.d.instances) - 1

    def add_light(self, light):
        self.d.lights.append(light)
        return len(self.d.lights) - 1


def gltf_scene_desc(path, width=256, height=256, max_ray_bounces=5, spp=1):
    """A scenegen.SceneDesc holding what the glTF loader produces for `path`
    (camera = the file's first camera when it has one); returns (desc, info)."""
    d = scenegen.SceneDesc(width=width, height=height, max_ray_bounces=max_ray_bounces, spp=spp)
    info = gltf_load(path, _DescSink(d))
    if info["camera"] is not None:
        d.camera = dict(info["camera"])
    return d, info


def render_output_buffer(scene, camera, which, surface):
    """Output-buffer view (abi.OUTPUT_BUFFER_*) of `scene` into `surface`."""
    check(lib().hh_render_output_buffer(scene.h, C.byref(camera_struct(camera)), which, surface.h))


def write_png(path, rgba8, flip_rows=False):
    """Write an (H, W, 4) uint8 array as a PNG (row 0 first unless flip_rows)."""
    a = np.ascontiguousarray(rgba8, np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) uint8 array")
    check(lib().hh_write_png(str(path).encode(), ptr(a), a.shape[1], a.shape[0], int(bool(flip_rows))))


class DebugRays:
    """The ray-debug vertex buffer and its draw arguments (renderer.cpp:1473-1479)."""

    def __init__(self, ctx, max_segments=abi.HH_MAX_DEBUG_RAY_DRAW_COUNT):
        self.ctx = ctx
        self.capacity = 2 * max_segments
        self.h = C.c_void_p()
        check(lib().hh_debug_rays_create(ctx.h, max_segments, C.byref(self.h)))

    def clear(self):
        check(lib().hh_debug_rays_clear(self.h))

    def read(self):
        """(vertices (n,) DEBUG_RAY_VERTEX_DTYPE, draw_args[4]); n = min(count, capacity)."""
        args = np.zeros(4, np.uint32)
        out = np.zeros(self.capacity, abi.DEBUG_RAY_VERTEX_DTYPE)
        check(lib().hh_debug_rays_read(self.h, ptr(out), self.capacity, ptr(args)))
        return out[:min(int(args[0]), self.capacity)], args

    def device_ptrs(self):
        v, a = C.c_void_p(), C.c_void_p()
        check(lib().hh_debug_rays_device_ptr(self.h, C.byref(v), C.byref(a)))
        return v.value, a.value

    def close(self):
        if self.h:
            lib().hh_debug_rays_destroy(self.h)
            self.h = None


def render_batch(ctx, views):
    """views: list of dict(scene, surface, camera, spp, max_ray_bounces, shadow_ray_bias, first_frame, flags)."""
    arr = _view_array(views)
    check(lib().hh_render_batch(ctx.h, arr, len(views)))


def _view_array(views):
    arr = (abi.View * len(views))()
    for i, v in enumerate(views):
        arr[i].scene = v["scene"].h.value
        arr[i].surface = v["surface"].h.value
        arr[i].camera = camera_struct(v["camera"])
        arr[i].spp = v.get("spp", 1)
        arr[i].max_ray_bounces = v.get("max_ray_bounces", 8)
        arr[i].shadow_ray_bias = v.get("shadow_ray_bias", 0.0)
        arr[i].first_frame = v.get("first_frame", 0)
        arr[i].flags = v.get("flags", 0)
    return arr


def autotune(ctx, views, frames=8, bvh_width=False):
    """hh_context_autotune: keeps the fastest tuning on ctx; returns (tuning, ms per frame)."""
    arr = _view_array(views)
    t = abi.Tuning()
    ms = C.c_double()
    check(lib().hh_context_autotune(ctx.h, arr, len(views), frames, 1 if bvh_width else 0, C.byref(t), C.byref(ms)))
    return ctx.tuning(), ms.value


def launch_consts(camera, width, height, num_lights=1, frame=0, max_ray_bounces=5):
    c = abi.LaunchConsts()
    st = abi.Settings(max_ray_bounces, 5000, 0.0, 0, 0, 0)
    check(lib().hh_compute_launch_consts(C.byref(camera_struct(camera)), width, height, num_lights, frame,
                                         C.byref(st), C.byref(c)))
    return c


def sky_fit(turbidity, albedo, normalized_sun_y, sun_dir):
    out = np.zeros(30, np.float32)
    sd = np.asarray(sun_dir, np.float32)
    check(lib().hh_sky_fit(C.byref(abi.SkyParams(turbidity, albedo, normalized_sun_y, 512)), ptr(sd), ptr(out)))
    return out.reshape(10, 3)
