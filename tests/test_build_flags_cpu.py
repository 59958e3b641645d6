"""The device code as built: compiled to gfx950 assembly with the build's own flags (microsim_amd/build.py), the
kernels contain no packed FP32 arithmetic (v_pk_fma_f32 / v_pk_mul_f32 / v_pk_add_f32: the build switches the
packed-fp32-ops target feature off; without that flag the same unit compiles to several thousand such instructions)
and no scalar-memory stores."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from microsim_amd import build

CSRC = os.path.join(os.path.dirname(build.__file__), "csrc")


def _assembly(src, defs):
    out = os.path.join(tempfile.mkdtemp(prefix="gp_asm_"), "k.s")
    flags = [f for f in build.HIP_FLAGS if f not in ("-shared", "-fPIC")]
    cmd = [build.hipcc()] + flags + defs + ["-I" + build.INC, "-I" + CSRC, "--cuda-device-only", "-S", "-o", out, src]
    subprocess.run(cmd, check=True, capture_output=True, timeout=900)
    text = open(out).read()
    shutil.rmtree(os.path.dirname(out), ignore_errors=True)
    return text


@pytest.mark.parametrize("unit,defs", [("gp_ops.hip", ["-DGP_N=3", "-DGP_NC=2"]), ("gp_hip.hip", [])])
def test_device_code_has_no_packed_fp32_and_no_scalar_stores(unit, defs):
    try:
        build.hipcc()
    except RuntimeError:
        pytest.skip("hipcc is not installed")
    asm = _assembly(os.path.join(CSRC, unit), defs)
    assert re.search(r"^\s*_ZN3gpk|^\s*k_", asm, re.M), "no kernels in the assembly"
    packed = re.findall(r"^\s*(v_pk_(?:fma|mul|add)_f32)\b", asm, re.M)
    assert not packed, "%d packed FP32 instructions" % len(packed)
    sstores = re.findall(r"^\s*(s_(?:buffer_)?store_\w+|s_(?:buffer_)?atomic_\w+|s_dcache_wb\w*)\b", asm, re.M)
    assert not sstores, "scalar stores: %s" % sorted(set(sstores))
