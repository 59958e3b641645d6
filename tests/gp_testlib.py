"""Shared test helpers: the oracle library (checker), golden fixtures, case loading."""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess
import tempfile

import numpy as np

from microsim_amd import _abi, configs, host

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle_gp.so")

_orc = None


def ensure_host_built():
    if not os.path.exists(os.path.join(_abi.LIBDIR, "libgp_host.so")):
        from microsim_amd import build
        build.build_host()


def oracle():
    """The FP64 CPU restatement (built with make -C oracle when missing)."""
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "CC=gcc", "liboracle_gp.so"], check=True)
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_step.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, _abi._dp, C.c_int64,
                                    C.POINTER(_abi.GPStats), C.c_int]
        lib.oracle_smooth.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, C.c_int64, C.c_int]
        lib.oracle_apply_bc.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp]
        lib.oracle_apply_bc.restype = None
        lib.oracle_stats_init.argtypes = [C.POINTER(_abi.GPStats)]
        lib.oracle_stats_init.restype = None
        lib.oracle_invariants.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, _abi._dp, _abi._dp]
        lib.oracle_invariants.restype = None
        _orc = lib
    return _orc


_model = None


def model_lib():
    """Host build (g++) of the kernels' per-cell arithmetic, tests/cpu_model/model_step.cpp: model_step (the step
    in the kernels' decomposition) and model_jvp (its forward-mode directional derivative)."""
    global _model
    if _model is None:
        out = os.path.join(tempfile.mkdtemp(prefix="gp_model_"), "libmodel_step.so")
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(REPO, "microsim_amd", "csrc"),
                        "-o", out, os.path.join(REPO, "tests", "cpu_model", "model_step.cpp")], check=True)
        lib = C.CDLL(out)
        lib.model_step.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, _abi._dp, C.c_int64, C.c_int,
                                   C.c_int]
        lib.model_jvp.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, _abi._dp, _abi._dp, _abi._dp,
                                  _abi._dp, _abi._dp, C.c_int64, C.c_int]
        lib.model_vjp.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, _abi._dp, _abi._dp, _abi._dp,
                                  _abi._dp, _abi._dp, C.c_int64, C.c_int]
        lib.model_step_c32.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, C.c_int64, C.c_int]
        lib.cell_check.argtypes = [C.POINTER(_abi.GPParams), _abi._dp, _abi._dp, _abi._dp, C.c_int64, C.c_int,
                                   _abi._dp, _abi._dp, _abi._dp]
        lib.c32_roundtrip.argtypes = [_abi._dp, C.c_int64, C.c_void_p, C.c_void_p, _abi._dp]
        lib.c32_roundtrip.restype = None
        lib.df_roundtrip.argtypes = [_abi._dp, C.c_int64, C.c_void_p, C.c_void_p, _abi._dp]
        lib.df_roundtrip.restype = None
        _model = lib
    return _model


def model_vjp(params, phi, mu, wphi, wmu, nsteps):
    """(gphi0, gmu0, gtheta): the gradient of <wphi, phi_T> + <wmu, mu_T> with respect to the initial state and
    theta, by the reverse-mode per-cell functions on the host (gp_reverse.h), FP64."""
    from microsim_amd import adjoint
    nth = len(adjoint.theta_from_params(params))
    g0, m0, th = np.zeros_like(phi), np.zeros_like(mu), np.zeros(nth)
    rc = model_lib().model_vjp(C.byref(params), _abi.dptr(np.ascontiguousarray(phi)),
                               _abi.dptr(np.ascontiguousarray(mu)),
                               _abi.dptr(np.ascontiguousarray(wphi, dtype=np.float64)),
                               _abi.dptr(np.ascontiguousarray(wmu, dtype=np.float64)), _abi.dptr(g0),
                               _abi.dptr(m0), _abi.dptr(th), nsteps, 0)
    assert rc == 0
    return g0, m0, th


def cell_check(params, phi, mu, dphi, c, stage, w):
    """(vjp, jvp) of one cell's stage (0: phi_delta, 1: currents and update) at flat index c of a halo-inclusive
    state: the reverse-mode transposed Jacobian applied to w, and the same products column by column in dual
    numbers (tests/cpu_model/model_step.cpp cell_check)."""
    vjp, jvp = np.zeros(512), np.zeros(512)
    n = model_lib().cell_check(C.byref(params), _abi.dptr(np.ascontiguousarray(phi)), _abi.dptr(np.ascontiguousarray(mu)),
                               _abi.dptr(np.ascontiguousarray(dphi)), int(c), int(stage),
                               _abi.dptr(np.ascontiguousarray(w, dtype=np.float64)), _abi.dptr(vjp), _abi.dptr(jvp))
    assert n > 0
    return vjp[:n], jvp[:n]


def pow2_spacing(params) -> bool:
    """dx and dy are powers of two (the kernels then multiply by the reciprocal spacings, bit for bit the same)."""
    import math
    return math.frexp(params.dx)[0] == 0.5 and math.frexp(params.dy)[0] == 0.5


def model_step_c32(params, phi, mu, nsteps):
    """(phi_T, mu_T) after nsteps steps of the compensated FP32 mode in a host build of the device arithmetic
    (tests/cpu_model/model_step.cpp run_c32), decoded to FP64."""
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    mu = np.ascontiguousarray(mu, dtype=np.float64).copy()
    rc = model_lib().model_step_c32(C.byref(params), _abi.dptr(phi), _abi.dptr(mu), nsteps, int(pow2_spacing(params)))
    assert rc == 0, rc
    return phi, mu


def model_jvp(params, phi, mu, vphi, vmu, vtheta, nsteps):
    """(phi_T, mu_T, tangent of phi_T, tangent of mu_T) along (vphi, vmu, vtheta), host build, FP64."""
    phi, mu = phi.copy(), mu.copy()
    tphi, tmu = np.zeros_like(phi), np.zeros_like(mu)
    vphi = np.ascontiguousarray(vphi, dtype=np.float64)
    vmu = np.ascontiguousarray(vmu, dtype=np.float64)
    vtheta = np.ascontiguousarray(vtheta, dtype=np.float64)
    rc = model_lib().model_jvp(C.byref(params), _abi.dptr(phi), _abi.dptr(mu), _abi.dptr(vphi), _abi.dptr(vmu),
                               _abi.dptr(vtheta), _abi.dptr(tphi), _abi.dptr(tmu), nsteps, 0)
    assert rc == 0
    return phi, mu, tphi, tmu


def oracle_step(params, phi, mu, nsteps, dphi=None, stats=None, threads=0):
    rc = oracle().oracle_step(C.byref(params), _abi.dptr(phi), _abi.dptr(mu),
                              _abi.dptr(dphi) if dphi is not None else None, nsteps,
                              C.byref(stats) if stats is not None else None, threads)
    assert rc == 0


def oracle_smooth(params, phi, mu, nsteps, threads=0):
    assert oracle().oracle_smooth(C.byref(params), _abi.dptr(phi), _abi.dptr(mu), nsteps, threads) == 0


def _frame_deck(d):
    """Whether a golden deck is non-isothermal (ISOTHERMAL = 0) or runs the moving frame (Shift = 1)."""
    keys = {}
    for line in open(os.path.join(d, "Input.in")):
        if "=" in line:
            k, v = line.split("=", 1)
            keys[k.strip()] = v.strip().rstrip(";").strip()
    return keys.get("ISOTHERMAL", "1") == "0" or keys.get("Shift", "0") == "1"


def golden_cases(frame: bool = False):
    """Golden decks: the isothermal fixed-frame ones, or with frame=True also the non-isothermal / moving-frame
    ones (which the FP64 device path, the host stage and the driver cover; the oracle, the FP32 modes and the
    adjoint do not take them)."""
    dirs = sorted(os.path.dirname(p) for p in glob.glob(os.path.join(GOLDEN, "*", "ref.npz")))
    return [os.path.basename(d) for d in dirs if frame or not _frame_deck(d)]


def frame_cases():
    """The non-isothermal / moving-frame golden decks."""
    return [n for n in golden_cases(frame=True) if n not in golden_cases()]


def load_golden(name):
    """(Case with the initial state, ref.npz contents)."""
    ensure_host_built()
    d = os.path.join(GOLDEN, name)
    z = np.load(os.path.join(d, "ref.npz"))
    restart = None
    if os.path.exists(os.path.join(d, "restart0.npz")):
        r = np.load(os.path.join(d, "restart0.npz"))
        inp = host.read_input(os.path.join(d, "Input.in"))
        tmp = tempfile.mkdtemp(prefix="gp_rs_")
        restart = os.path.join(tmp, "out_0.vtk")
        configs.write_vtk(restart, r["phi"], r["mu"], ["p%d" % i for i in range(inp.p.nphases)],
                          ["c%d" % i for i in range(inp.p.ncomp)], binary=not inp.ascii)
    case = host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"), restart)
    ref = {"steps": list(z["steps"]), "phi": z["phi"], "mu": z["mu"]}
    if "T" in z.files:
        ref["T"] = z["T"]
    return case, ref


def max_rel_err(a, b, floor=1e-300):
    """max |a-b| / max(|b|, tiny) over elements, treating exact equality as 0 (covers +-0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), floor)
    r = np.where(d == 0, 0.0, d / s)
    return float(r.max()) if r.size else 0.0
