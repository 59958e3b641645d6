"""In-tree build of the native pieces (no cmake): explicit gcc / hipcc command lines.

    python -m microsim_amd.build            # host library, HIP library, host driver
    python -m microsim_amd.build --host     # CPU parts only

Outputs go to microsim_amd/lib/: libgp_host.so, libgp_hip.so (gfx950 code object), microsim_gp.
"""
from __future__ import annotations

import concurrent.futures
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
INC = os.path.join(REPO, "include")
ARCH = os.environ.get("GP_OFFLOAD_ARCH", "gfx950")

HOST_CFLAGS = ["-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function"]
# FP64 parity needs IEEE division/sqrt and no contraction; packed FP32 instructions are switched off.
# -fno-hip-fp32-correctly-rounded-divide-sqrt affects single precision only (the FP32 production mode).
HIP_FLAGS = ["-O3", "-fPIC", "-shared", "-std=c++17", "--offload-arch=%s" % ARCH, "-ffp-contract=off",
             "-fno-fast-math", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-munsafe-fp-atomics"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_host():
    os.makedirs(LIB, exist_ok=True)
    _run(["gcc"] + HOST_CFLAGS + ["-shared", "-I" + INC, "-o", os.path.join(LIB, "libgp_host.so"),
                                  os.path.join(HERE, "host", "gp_host.c"), os.path.join(HERE, "host", "gp_h5.c"),
                                  os.path.join(HERE, "host", "gp_ranks.c"), "-lm"])


def combos():
    """The (phases, mu fields) pairs the kernels are instantiated for (GP_COMBOS in csrc/gp_internal.h)."""
    txt = open(os.path.join(HERE, "csrc", "gp_internal.h")).read()
    m = re.search(r"#define GP_COMBOS\(X\)(.*)", txt)
    return [(int(a), int(b)) for a, b in re.findall(r"X\((\d+),\s*(\d+)\)", m.group(1))]


def build_hip(jobs: int = 0):
    """Compile gp_hip.hip (C ABI) and gp_ops.hip once per (phases, mu fields) pair in parallel, then link."""
    os.makedirs(LIB, exist_ok=True)
    csrc = os.path.join(HERE, "csrc")
    objdir = tempfile.mkdtemp(prefix="gp_hip_obj_")
    units = [(os.path.join(csrc, "gp_hip.hip"), [], "gp_hip.o")]
    units += [(os.path.join(csrc, "gp_ops.hip"), ["-DGP_N=%d" % n, "-DGP_NC=%d" % nc], "gp_ops_%d_%d.o" % (n, nc))
              for n, nc in combos()]
    cflags = [f for f in HIP_FLAGS if f != "-shared"]

    def compile_unit(u):
        src, defs, obj = u
        _run([hipcc()] + cflags + defs + ["-I" + INC, "-I" + csrc, "-c", "-o", os.path.join(objdir, obj), src])

    jobs = jobs or min(len(units), max(1, int(os.environ.get("MAX_JOBS") or os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(compile_unit, u) for u in units]:
            f.result()
    out = os.path.join(LIB, "libgp_hip.so")
    tmp = out + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=%s" % ARCH, "-o", tmp] +
         [os.path.join(objdir, u[2]) for u in units])
    os.replace(tmp, out)
    shutil.rmtree(objdir, ignore_errors=True)


def build_driver():
    src = os.path.join(HERE, "host", "microsim_gp.c")
    if not os.path.exists(src):
        return
    _run(["gcc"] + HOST_CFLAGS + ["-I" + INC, "-o", os.path.join(LIB, "microsim_gp"), src,
                                  "-L" + LIB, "-lgp_host", "-lgp_hip", "-Wl,-rpath,$ORIGIN", "-lm"])


def build_reconstruct():
    """The MPI layout's post-processing tools on the host library: reconstruct (joins the per-rank snapshots) and
    write_xdmf (the XDMF descriptions of the HDF5 snapshots)."""
    for tool in ("reconstruct", "write_xdmf"):
        _run(["gcc"] + HOST_CFLAGS + ["-I" + INC, "-o", os.path.join(LIB, tool),
                                      os.path.join(HERE, "host", tool + ".c"), "-L" + LIB, "-lgp_host",
                                      "-Wl,-rpath,$ORIGIN", "-lm"])


def build_all(host_only: bool = False):
    build_host()
    build_reconstruct()
    if not host_only:
        build_hip()
        build_driver()


if __name__ == "__main__":
    build_all(host_only="--host" in sys.argv)
