"""Properties of the built device code (no GPU needed).

The HIP library is compiled with the packed-FP32 target feature switched off
(``-Xclang -target-feature -Xclang -packed-fp32-ops``, microsim_amd/build.py). hipcc's host pass warns that it
does not know that feature, which is harmless: the device pass honours it (without the flag, ordinary float2 code
already compiles to packed FP32 FMAs). This test disassembles the gfx950 code object of libgp_hip.so and checks
that no packed FP32 arithmetic instruction is in it, so the property is enforced, not assumed.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from microsim_amd import _abi

LLVM = "/opt/rocm/llvm/bin"
PACKED_FP32 = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def _tool(name):
    p = os.path.join(LLVM, name)
    return p if os.path.exists(p) else shutil.which(name)


def test_device_code_has_no_packed_fp32_arithmetic():
    so = os.path.join(_abi.LIBDIR, "libgp_hip.so")
    if not os.path.exists(so):
        pytest.skip("libgp_hip.so is not built")
    objcopy, bundler, objdump = _tool("llvm-objcopy"), _tool("clang-offload-bundler"), _tool("llvm-objdump")
    if not (objcopy and bundler and objdump):
        pytest.skip("LLVM binutils not found")
    d = tempfile.mkdtemp(prefix="gp_dis_")
    fat, co = os.path.join(d, "fatbin.bin"), os.path.join(d, "co.o")
    subprocess.run([objcopy, "--dump-section", ".hip_fatbin=" + fat, so, os.path.join(d, "stripped.so")],
                   check=True)
    subprocess.run([bundler, "--unbundle", "--type=o", "--input=" + fat,
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + co], check=True)
    dis = subprocess.run([objdump, "-d", "--mcpu=gfx950", co], check=True, capture_output=True, text=True).stdout
    shutil.rmtree(d, ignore_errors=True)
    assert "v_" in dis, "no device instructions found"
    bad = sorted(set(m.group(0) for m in PACKED_FP32.finditer(dis)))
    assert not bad, "packed FP32 arithmetic in the device code: %s" % bad
