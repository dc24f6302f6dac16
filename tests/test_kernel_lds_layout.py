"""Static LDS of the built frame kernels, read from the library's own gfx950
code object (ff-srl_amd/_build/libffsrl.so: its .hip_fatbin section,
unbundled, AMDGPU metadata notes).  The dynamic LDS of a kernel starts right
after its static segment, and the partitioned and shared frames keep float4
positions there: a static segment that is not a multiple of 16 bytes leaves
every 16-byte LDS access misaligned, which ran the partitioned frame 47%
slower with identical results.  The partitioned kernels' static LDS must
also stay within the 1 KB their launch reserves."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ff-srl_amd", "_build", "libffsrl.so")
LLVM = "/opt/rocm/llvm/bin"


def _kernel_lds(tmp):
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf"):
        if not os.path.exists(os.path.join(LLVM, tool)):
            pytest.skip(f"{tool} not found")
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    fat, co = os.path.join(tmp, "fatbin.bin"), os.path.join(tmp, "gfx950.co")
    shutil.copy(LIB, os.path.join(tmp, "lib.so"))
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section", f".hip_fatbin={fat}",
                    os.path.join(tmp, "lib.so"), os.path.join(tmp, "out.so")], check=True, capture_output=True)
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True, capture_output=True)
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], check=True, capture_output=True,
                           text=True).stdout
    out, pending = {}, None
    for line in notes.splitlines():
        m = re.match(r"\s*\.group_segment_fixed_size:\s+(\d+)", line)
        if m:
            pending = int(m.group(1))
            continue
        m = re.match(r"\s*\.symbol:\s+(\S+)\.kd", line)
        if m and pending is not None:
            out[m.group(1)] = pending
            pending = None
    return out


def test_float4_lds_kernels_start_dynamic_lds_aligned(tmp_path):
    lds = _kernel_lds(str(tmp_path))
    parts = {k: v for k, v in lds.items() if "k_frame_parts" in k}
    shared = {k: v for k, v in lds.items() if "k_frame_shared" in k}
    print("\n".join(f"{v:6d} {k[:90]}" for k, v in sorted(lds.items()) if "k_frame" in k))
    assert len(parts) == 4 and len(shared) >= 12
    for k, v in {**parts, **shared}.items():
        assert v % 16 == 0, (k, v)
    for k, v in parts.items():
        assert v <= 1024, (k, v)
