"""Build helios_amd/libhelios_hip.so (HIP for gfx950) with hipcc.

Every translation unit is compiled with:
  -ffp-contract=off                        exact-parity arithmetic (ray
                                           generation, watertight test,
                                           alpha test) rounds like the CPU
                                           restatement; hot paths call fmaf
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed-FP32 instructions
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libhelios_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HH_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.hip", "bvh_build.hip", "integrator.hip", "sky_host.hip", "denoise.hip"]
HOST_SOURCES = ["image_io.cpp", "scene_io.cpp"]    # host-only C++ (g++), linked with zlib
CXX = os.environ.get("CXX", "g++")


def flags():
    hosek = PKG / "data" / "hosek_rgb.bin"
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-result", f'-DHH_HOSEK_BIN="{hosek}"',
            f"-I{PKG.parent / 'include'}"]


def compile_one(src):
    obj = BUILD / (Path(src).stem + ".o")
    if src in HOST_SOURCES:
        cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{PKG.parent / 'include'}", "-c", str(CSRC / src),
               "-o", str(obj)]
    else:
        cmd = [HIPCC, *flags(), "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True):
    BUILD.mkdir(exist_ok=True)
    srcs = SOURCES + HOST_SOURCES
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
