"""CPU tests of the C ABI: the library loads, exports every symbol the header
declares, validates arguments, and its host-only math (launch constants, sky
fit) equals the CPU restatement. No GPU calls."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from helios_amd import abi, runtime, scenegen

HEADER = Path(__file__).resolve().parents[1] / "include" / "helios_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:hh_status|const char\*)\s+(hh_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = runtime.lib()
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_struct_sizes_match_reference_layouts():
    assert C.sizeof(abi.Vertex) == 80            # mesh.h:10-17
    assert C.sizeof(abi.LaunchConsts) == 192     # path_integrator.cpp:11-28
    assert C.sizeof(abi.Hit) == 24
    assert C.sizeof(abi.Ray) == 32


def test_launch_consts_match_oracle(oracle_mod):
    for cam, w, h in [(scenegen.cornell_box().camera, 256, 256),
                      (scenegen.icosphere_field(seed=1, n_spheres=1, level=0, n_leaves=0).camera, 1920, 1080),
                      (scenegen.look_at_camera((3, -2, 7), (0.5, 0.5, -1), fov_y=75, focal_length=2.5,
                                               aperture=0.2), 640, 480)]:
        a = runtime.launch_consts(cam, w, h)
        b = oracle_mod.launch_consts(cam, w, h)
        for f in ("view_proj_inverse", "camera_pos", "up_direction", "right_direction", "focal_plane"):
            assert list(getattr(a, f)) == list(getattr(b, f)), f
        assert a.aperture_radius == b.aperture_radius


@pytest.mark.parametrize("turb,alb,elev", [(3.0, 0.1, 30.0), (4.0, 0.1, 5.0), (7.5, 0.6, 70.0), (1.0, 0.0, 89.0)])
def test_sky_fit_matches_oracle(oracle_mod, turb, alb, elev):
    sun = -np.array(scenegen.sun_forward(elev, 20.0))
    a = runtime.sky_fit(turb, alb, 1.15, sun)
    b = oracle_mod.sky_coefficients(turb, alb, 1.15, sun)
    assert np.array_equal(a, b)


def test_argument_validation_without_device():
    L = runtime.lib()
    # null pointers are rejected with INVALID_ARG and a message, not a crash
    assert L.hh_context_create(0, None) == 1
    assert b"null" in L.hh_last_error()
    assert L.hh_sky_fit(None, None, None) == 1
    cam = runtime.camera_struct(scenegen.cornell_box().camera)
    st = abi.Settings(5, 5000, 0.0, 0, 0, 0)
    out = abi.LaunchConsts()
    assert L.hh_compute_launch_consts(C.byref(cam), 0, 10, 1, 0, C.byref(st), C.byref(out)) == 1


def test_denoiser_parameter_count_without_device():
    import ctypes as C
    from helios_amd.runtime import lib
    from oracle import unet
    n = C.c_uint64()
    lib().hh_denoiser_num_params.argtypes = [C.POINTER(C.c_uint64)]
    assert lib().hh_denoiser_num_params(C.byref(n)) == 0
    assert n.value == unet.num_params() == 480675
