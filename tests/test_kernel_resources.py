"""Register and scratch budgets of the hot kernels, read from the built code.

The traversal, shading and continuation kernels are held to a number of waves
per SIMD by their VGPR budget (k_extend and k_extend_packet 5, k_shadow 6,
k_shade and k_megakernel 4). A kernel that outgrows its budget spills to
scratch, which puts private-memory traffic into the inner loops (a single
spilled address cost measurable time here). This reads the AMDGPU kernel
metadata of the gfx950 code object inside helios_amd/_build/integrator.o
(the object the library is linked from) and checks that the default BVH4
kernels of every block size use no scratch and fit their budgets.
CPU-only: it needs the built object and the ROCm LLVM tools, and skips
without them.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
OBJ = ROOT / "helios_amd" / "_build" / "integrator.o"
LLVM = Path("/opt/rocm/llvm/bin")
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

# mangled-name prefix -> VGPR budget (512 VGPRs per SIMD lane / waves, in
# allocation granules of 8)
BUDGETS = {}
for bs in (64, 128, 256):
    BUDGETS[f"_ZN2hh8k_extendILi4ELi{bs}ELi0ELb0E"] = 96
    BUDGETS[f"_ZN2hh8k_shadowILi4ELi{bs}ELi0ELb0E"] = 80
    BUDGETS[f"_ZN2hh15k_extend_packetILi4ELi{bs}EE"] = 96
    BUDGETS[f"_ZN2hh7k_shadeILi{bs}EE"] = 128
    BUDGETS[f"_ZN2hh12k_megakernelILi4ELi{bs}EE"] = 128


def _kernel_metadata(tmp_path):
    tools = [shutil.which("objcopy"), LLVM / "clang-offload-bundler", LLVM / "llvm-readelf"]
    if not OBJ.exists() or tools[0] is None or not all(Path(t).exists() for t in tools[1:]):
        pytest.skip("needs helios_amd/_build/integrator.o and the ROCm LLVM tools")
    fat, dev = tmp_path / "fatbin.bin", tmp_path / "device.o"
    subprocess.run([tools[0], f"--dump-section=.hip_fatbin={fat}", str(OBJ), str(tmp_path / "host.o")], check=True,
                   capture_output=True)
    subprocess.run([str(tools[1]), "--unbundle", "--type=o", f"--input={fat}", f"--targets={TARGET}",
                    f"--output={dev}"], check=True, capture_output=True)
    notes = subprocess.run([str(tools[2]), "--notes", str(dev)], check=True, capture_output=True, text=True).stdout
    meta, cur = {}, None
    for line in notes.splitlines():
        m = re.match(r"\s*\.name:\s+(_Z\S+)", line)
        if m:
            cur = meta.setdefault(m.group(1), {})
            continue
        m = re.match(r"\s*\.(private_segment_fixed_size|vgpr_count):\s+(\d+)", line)
        if m and cur is not None:
            cur[m.group(1)] = int(m.group(2))
    return meta


def test_hot_kernels_fit_their_budgets_without_scratch(tmp_path):
    meta = _kernel_metadata(tmp_path)
    checked = 0
    for prefix, budget in BUDGETS.items():
        names = [n for n in meta if n.startswith(prefix)]
        assert len(names) == 1, (prefix, names)
        k = meta[names[0]]
        assert k.get("private_segment_fixed_size") == 0, (names[0], k)
        assert k.get("vgpr_count", 1 << 10) <= budget, (names[0], k, budget)
        checked += 1
    assert checked == len(BUDGETS)
