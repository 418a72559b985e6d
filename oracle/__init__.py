"""Python handle on the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a scalar C++ restatement of the reference shaders
(oracle.cpp). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product library never does.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from helios_amd import abi
from helios_amd.abi import ptr

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
HOSEK = HERE.parent / "helios_amd" / "data" / "hosek_rgb.bin"

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.orc_scene_create.restype = vp
        L.orc_scene_destroy.argtypes = [vp]
        for name in ("orc_scene_add_texture", "orc_scene_add_material", "orc_scene_add_mesh",
                     "orc_scene_add_instance", "orc_scene_add_light", "orc_scene_set_ibl", "orc_scene_set_sky",
                     "orc_scene_commit", "orc_render", "orc_primary_hits", "orc_trace_rays", "orc_load_hosek"):
            getattr(L, name).restype = C.c_int
        L.orc_scene_add_texture.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.orc_scene_add_material.argtypes = [vp, C.POINTER(abi.MaterialDesc)]
        L.orc_scene_add_mesh.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32]
        L.orc_scene_add_instance.argtypes = [vp, C.POINTER(abi.InstanceDesc)]
        L.orc_scene_add_light.argtypes = [vp, C.POINTER(abi.LightDesc)]
        L.orc_scene_set_ibl.argtypes = [vp, C.c_uint32, vp]
        L.orc_scene_set_sky.argtypes = [vp, C.POINTER(abi.SkyParams)]
        L.orc_scene_commit.argtypes = [vp]
        L.orc_scene_num_lights.argtypes = [vp]
        L.orc_scene_num_lights.restype = C.c_uint32
        L.orc_scene_read_lights.argtypes = [vp, vp, C.c_uint32]
        L.orc_scene_read_lights.restype = C.c_uint32
        L.orc_scene_read_materials.argtypes = [vp, vp, C.c_uint32]
        L.orc_scene_read_materials.restype = C.c_uint32
        L.orc_scene_env_size.argtypes = [vp]
        L.orc_scene_env_size.restype = C.c_uint32
        L.orc_scene_env_equirect.argtypes = [vp, vp]
        L.orc_scene_env_equirect.restype = C.c_uint32
        L.orc_scene_read_env.argtypes = [vp, vp]
        L.orc_scene_sky_coefficients.argtypes = [vp, vp]
        L.orc_sky_coefficients.argtypes = [C.POINTER(abi.SkyParams), vp, vp]
        L.orc_sky_eval.argtypes = [vp, vp, vp, C.c_uint32, C.c_int, vp]
        L.orc_launch_consts.argtypes = [C.POINTER(abi.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(abi.Settings), C.POINTER(abi.LaunchConsts)]
        L.orc_render.argtypes = [vp, C.POINTER(abi.Camera), C.POINTER(abi.Settings), C.c_uint32, C.c_uint32,
                                 C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                 C.c_int, vp, vp]
        L.orc_primary_hits.argtypes = [vp, C.POINTER(abi.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, vp]
        L.orc_trace_rays.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, vp]
        L.orc_output_buffer.argtypes = [vp, C.POINTER(abi.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, vp]
        L.orc_output_buffer.restype = C.c_int
        L.orc_ray_debug.argtypes = [vp, C.POINTER(abi.Camera), C.POINTER(abi.Settings), C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, C.c_uint32]
        L.orc_ray_debug.restype = C.c_uint32
        L.orc_rng_hash.argtypes = [C.c_uint32]
        L.orc_rng_hash.restype = C.c_uint32
        L.orc_rng_init.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.orc_rng_next.argtypes = [vp]
        L.orc_rng_next.restype = C.c_uint32
        L.orc_next_float.argtypes = [vp]
        L.orc_next_float.restype = C.c_float
        L.orc_next_uint.argtypes = [vp, C.c_uint32]
        L.orc_next_uint.restype = C.c_uint32
        L.orc_eval_uber.argtypes = [vp, vp, vp, vp]
        L.orc_sample_uber.argtypes = [vp, vp, vp, vp]
        L.orc_D_ggx.argtypes = [C.c_double, C.c_double]
        L.orc_D_ggx.restype = C.c_double
        L.orc_sample_cosine_lobe.argtypes = [vp, C.c_double, C.c_double, vp]
        L.orc_sample_ggx.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp]
        L.orc_tonemap.argtypes = [vp, C.c_uint32, C.c_float, C.c_uint32, vp]
        L.orc_sample_env.argtypes = [vp, vp, C.c_uint32, C.c_int, vp]
        L.orc_sample_light_probe.argtypes = [vp, C.c_uint32, vp, vp, C.c_uint32, vp, vp]
        L.orc_sample_light_probe.restype = C.c_int
        L.orc_tri_hit.argtypes = [vp, vp, vp, C.c_float, C.c_float, vp]
        L.orc_tri_hit.restype = C.c_int
        table = np.fromfile(HOSEK, dtype="<f8")
        if L.orc_load_hosek(ptr(table), len(table)) != 0:
            raise RuntimeError("bad Hosek table")
        _lib = L
    return _lib


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
    return _struct(abi.Camera, cam)


def settings_struct(max_ray_bounces=5, max_samples=5000, shadow_ray_bias=0.0, tiled=0, rr_start_depth=0, flags=0):
    return abi.Settings(max_ray_bounces, max_samples, shadow_ray_bias, tiled, rr_start_depth, flags)


class OracleScene:
    """CPU scene mirroring helios::Scene + its GPU packing."""

    def __init__(self, desc):
        L = lib()
        self.desc = desc
        self.h = C.c_void_p(L.orc_scene_create())
        for (w, h, fmt, data) in desc.textures:
            data = np.ascontiguousarray(data)
            L.orc_scene_add_texture(self.h, w, h, fmt, ptr(data))
        for m in desc.materials:
            L.orc_scene_add_material(self.h, C.byref(_struct(abi.MaterialDesc, m)))
        for me in desc.meshes:
            sm = np.array(me.submeshes, np.uint32).reshape(-1, 5)
            mats = np.array(me.materials, np.uint32)
            v = np.ascontiguousarray(me.vertices)
            i = np.ascontiguousarray(me.indices, dtype=np.uint32)
            L.orc_scene_add_mesh(self.h, ptr(v), len(v), ptr(i), len(i), ptr(sm), len(sm), ptr(mats), len(mats))
        for (model, normal, mesh, ov) in desc.instances:
            L.orc_scene_add_instance(self.h, C.byref(_struct(abi.InstanceDesc, dict(
                model=list(model), normal=list(normal), mesh=mesh, material_override=ov))))
        for l in desc.lights:
            L.orc_scene_add_light(self.h, C.byref(_struct(abi.LightDesc, l)))
        if desc.ibl is not None:
            size, faces = desc.ibl
            faces = np.ascontiguousarray(faces, np.float32)
            L.orc_scene_set_ibl(self.h, size, ptr(faces))
        L.orc_scene_set_sky(self.h, C.byref(abi.SkyParams(*desc.sky)))
        rc = L.orc_scene_commit(self.h)
        if rc != 0:
            raise RuntimeError(f"oracle commit failed: {rc}")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_destroy(self.h)
            self.h = None

    @property
    def num_lights(self):
        return lib().orc_scene_num_lights(self.h)

    def lights(self):
        out = np.zeros((max(1, self.num_lights), 16), np.float32)
        n = lib().orc_scene_read_lights(self.h, ptr(out), len(out))
        return out[:n]

    def env(self):
        n = lib().orc_scene_env_size(self.h)
        out = np.zeros((6, n, n, 4), np.float32)
        lib().orc_scene_read_env(self.h, ptr(out))
        return out

    def env_equirect(self):
        """Equirectangular sky as float16 (H, W, 4), or None for the cube."""
        w = lib().orc_scene_env_equirect(self.h, None)
        if not w:
            return None
        out = np.zeros((w // 2, w, 4), np.uint16)
        lib().orc_scene_env_equirect(self.h, ptr(out))
        return out.view(np.float16)

    def sky_coefficients(self):
        out = np.zeros(30, np.float32)
        lib().orc_scene_sky_coefficients(self.h, ptr(out))
        return out.reshape(10, 3)

    def render(self, camera=None, width=None, height=None, frames=1, first_frame=0, max_ray_bounces=None,
               crop=None, fp64=False, threads=None, flags=0, shadow_ray_bias=0.0, image=None, rr_start_depth=0):
        """Progressive samples first_frame..first_frame+frames-1 into an RGBA32F image."""
        d = self.desc
        width = width or d.width
        height = height or d.height
        cam = camera_struct(camera or d.camera)
        st = settings_struct(max_ray_bounces or d.max_ray_bounces, flags=flags, shadow_ray_bias=shadow_ray_bias,
                             rr_start_depth=rr_start_depth)
        x0, y0, x1, y1 = crop or (0, 0, width, height)
        img = image if image is not None else np.zeros((height, width, 4), np.float32)
        stats = np.zeros(2, np.uint64)
        threads = threads or os.cpu_count() or 1
        rc = lib().orc_render(self.h, C.byref(cam), C.byref(st), width, height, x0, y0, x1, y1, first_frame, frames,
                              int(fp64), threads, ptr(img), ptr(stats))
        if rc:
            raise RuntimeError("oracle render failed")
        self.last_stats = stats
        return img

    def primary_hits(self, camera=None, width=None, height=None, frame=0, brute=True):
        d = self.desc
        width = width or d.width
        height = height or d.height
        out = np.zeros((height, width), abi.HIT_DTYPE)
        lib().orc_primary_hits(self.h, C.byref(camera_struct(camera or d.camera)), width, height, frame, int(brute),
                               ptr(out))
        return out

    def output_buffer(self, which, camera=None, width=None, height=None, brute=False):
        """Output-buffer view `which` (0 albedo, 1 normals, 2 roughness, 3 metallic,
        4 emissive) as an (height, width, 4) float32 image."""
        d = self.desc
        width = width or d.width
        height = height or d.height
        out = np.zeros((height, width, 4), np.float32)
        if lib().orc_output_buffer(self.h, C.byref(camera_struct(camera or d.camera)), width, height, which,
                                   int(brute), ptr(out)):
            raise RuntimeError("oracle output_buffer failed")
        return out

    def ray_debug(self, pixel, num_rays, frame, camera=None, width=None, height=None, max_ray_bounces=None,
                  flags=0):
        """Ray-debug segments of num_rays paths through window pixel (x, y):
        (n, 9) float32 = origin xyz, end xyz, colour rgb, ray by ray."""
        d = self.desc
        width = width or d.width
        height = height or d.height
        st = settings_struct(max_ray_bounces or d.max_ray_bounces, flags=flags)
        cap = num_rays * max(st.max_ray_bounces, 1)
        out = np.zeros((cap, 9), np.float32)
        n = lib().orc_ray_debug(self.h, C.byref(camera_struct(camera or d.camera)), C.byref(st), width, height,
                                pixel[0], pixel[1], num_rays, frame, ptr(out), cap)
        return out[:min(n, cap)]

    def sample_env(self, dirs, fp64=False):
        """Environment radiance along each direction (n, 3)."""
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((len(dirs), 3), np.float32)
        lib().orc_sample_env(self.h, ptr(dirs), len(dirs), int(fp64), ptr(out))
        return out

    def sample_light(self, light, pos, normal, rng_state, depth=0):
        """One FP64 sample_light call on packed light `light`: returns
        (Wi, Li, pdf, shadow tmax, draws) and advances rng_state (uint32[2])."""
        p = np.ascontiguousarray(pos, np.float64)
        n = np.ascontiguousarray(normal, np.float64)
        out = np.zeros(9, np.float64)
        if lib().orc_sample_light_probe(self.h, light, ptr(p), ptr(n), depth, ptr(rng_state), ptr(out)):
            raise ValueError("bad light index or scene not committed")
        return out[0:3].copy(), out[3:6].copy(), float(out[6]), float(out[7]), int(out[8])

    def trace_rays(self, rays, flags=0, brute=False, fp64=False):
        rays = np.ascontiguousarray(rays)
        out = np.zeros(len(rays), abi.HIT_DTYPE)
        lib().orc_trace_rays(self.h, ptr(rays), len(rays), flags, int(brute), int(fp64), ptr(out))
        return out


def launch_consts(camera, width, height, num_lights=1, frame=0, max_ray_bounces=5):
    c = abi.LaunchConsts()
    lib().orc_launch_consts(C.byref(camera_struct(camera)), width, height, num_lights, frame,
                            C.byref(settings_struct(max_ray_bounces)), C.byref(c))
    return c


def tonemap(rgba, exposure=1.0, op=0):
    rgba = np.ascontiguousarray(rgba, np.float32)
    n = rgba.size // 4
    out = np.zeros(rgba.shape[:-1] + (4,), np.uint8)
    lib().orc_tonemap(ptr(rgba), n, exposure, op, ptr(out))
    return out


def sky_coefficients(turbidity, albedo, normalized_sun_y, sun_dir):
    out = np.zeros(30, np.float32)
    sd = np.asarray(sun_dir, np.float32)
    lib().orc_sky_coefficients(C.byref(abi.SkyParams(turbidity, albedo, normalized_sun_y, 512)), ptr(sd), ptr(out))
    return out.reshape(10, 3)


def sky_eval(coef, sun_dir, dirs, fp64=False):
    coef = np.ascontiguousarray(coef, np.float32).ravel()
    sd = np.asarray(sun_dir, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    out = np.zeros((len(dirs), 3), np.float32)
    lib().orc_sky_eval(ptr(coef), ptr(sd), ptr(dirs), len(dirs), int(fp64), ptr(out))
    return out


class RNG:
    def __init__(self, x, y, frame):
        self.s = np.zeros(2, np.uint32)
        lib().orc_rng_init(x, y, frame, ptr(self.s))

    def next(self):
        return lib().orc_rng_next(ptr(self.s))

    def next_float(self):
        return lib().orc_next_float(ptr(self.s))

    def next_uint(self, n):
        return lib().orc_next_uint(ptr(self.s), n)


def rng_hash(x):
    return lib().orc_rng_hash(x)


def relmse(x, ref, eps=0.01):
    """mean over pixels and channels of (x - r)^2 / (r^2 + eps)."""
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))
