"""Host library and C ABI checks that need no GPU.

- Every function declared in include/*.h is exported by the library that implements it, both libraries load,
  and the ctypes mirrors have the C sizes.
- The host path before the time loop (Input.in parser, fill_domain, boundary conditions, VTK writer) reproduces
  the reference's t = 0 snapshot DATA/out_0.vtk byte for byte, and the echo files <stem>.out / <stem>.bd the
  reference writes next to Input.in (Grand_potential_Finite_difference_2D_serial/microsim_gp.c:42-84,
  solverloop/file_writer.h:71-255).
"""
import ctypes as C
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from gp_testlib import GOLDEN, REPO, ensure_host_built, golden_cases, load_golden
from microsim_amd import _abi, host

HEADERS = {"gp_hip.h": "libgp_hip.so", "gp_host.h": "libgp_host.so"}


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \t\*]*?\b(gp_\w+)\s*\(", text, flags=re.M)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(_abi.LIBDIR, lib)], check=True,
                         capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_every_declared_symbol(header):
    lib = HEADERS[header]
    if not os.path.exists(os.path.join(_abi.LIBDIR, lib)):
        pytest.fail("%s is not built: run __graft_entry__.build()" % lib)
    names = _declared(header)
    assert len(names) >= 10, names
    missing = [n for n in names if n not in _exported(lib)]
    assert not missing, "%s does not export %s" % (lib, missing)


def test_libraries_load_without_gpu():
    ensure_host_built()
    _abi.host_lib()  # also checks the ctypes layouts against gp_sizeof_*
    lib = _abi.hip_lib()
    assert lib.gp_version().decode().startswith("gp_hip")
    names = [lib.gp_kernel_name(i).decode() for i in range(_abi.GP_NKERNELS)]
    assert len(set(names)) == _abi.GP_NKERNELS and lib.gp_kernel_name(_abi.GP_NKERNELS) == b"?"


def test_create_rejects_bad_arguments_before_touching_the_device():
    """Argument checks run before any HIP call, so they hold on a machine without a GPU as well."""
    lib = _abi.hip_lib()
    h = C.c_void_p()
    case, _ = load_golden("tern_c1_mini")
    p = _abi.GPParams()
    C.memmove(C.byref(p), C.byref(case.params), C.sizeof(p))
    p.nphases = 7
    assert lib.gp_create(C.byref(p), _abi.GP_FP64, C.byref(h)) == _abi.GP_E_ARG
    assert b"nphases" in lib.gp_last_error(None)
    C.memmove(C.byref(p), C.byref(case.params), C.sizeof(p))
    assert lib.gp_create(C.byref(p), 7, C.byref(h)) == _abi.GP_E_ARG
    assert lib.gp_create(None, _abi.GP_FP64, C.byref(h)) == _abi.GP_E_ARG


def _copy_case(name):
    d = tempfile.mkdtemp(prefix="gp_host_")
    for f in ("Input.in", "Filling.in"):
        shutil.copy(os.path.join(GOLDEN, name, f), d)
    return d


@pytest.mark.parametrize("name", golden_cases(frame=True))
def test_initial_snapshot_matches_reference_bytes(name):
    case, _ = load_golden(name)
    d = tempfile.mkdtemp(prefix="gp_vtk_")
    try:
        out = os.path.join(d, "out_0.vtk")
        host.write_vtk(case.inp, out, case.phi, case.mu, host.temperature_rows(case.inp.p))
        got = open(out, "rb").read()
        want = open(os.path.join(GOLDEN, name, "out_0.vtk"), "rb").read()
        assert got == want, "%s: %d vs %d bytes" % (name, len(got), len(want))
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name", golden_cases(frame=True))
def test_echo_files_match_reference(name):
    d = _copy_case(name)
    try:
        inp = host.read_input(os.path.join(d, "Input.in"))
        assert _abi.host_lib().gp_write_echo(C.byref(inp), os.path.join(d, "Input.in").encode()) == 0
        # The reference never sets some diagonal entries (e.g. dab[a][a] of the last phase) and prints whatever
        # the allocation held, typically a subnormal such as 4.640697e-310; the host prints 0 there.
        garbage = re.compile(r"-?\d\.\d{6}e-3[0-2]\d")
        for ext in ("out", "bd"):
            got = open(os.path.join(d, "Input." + ext)).read()
            want = garbage.sub("0.000000e+00", open(os.path.join(GOLDEN, name, "Input." + ext)).read())
            assert got == want, "%s: Input.%s differs" % (name, ext)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_vtk_round_trip_binary_and_ascii():
    case, _ = load_golden("tern_aniso_per")
    d = tempfile.mkdtemp(prefix="gp_rt_")
    try:
        rng = np.random.default_rng(5)
        phi = rng.random(case.phi.shape)
        mu = rng.standard_normal(case.mu.shape)
        for ascii_flag in (0, 1):
            case.inp.ascii = ascii_flag
            path = os.path.join(d, "o_%d.vtk" % ascii_flag)
            host.write_vtk(case.inp, path, phi, mu)
            p2, m2 = host.empty_state(case.inp)
            assert _abi.host_lib().gp_read_vtk(C.byref(case.inp), path.encode(), _abi.dptr(p2), _abi.dptr(m2)) == 0
            sl = (slice(None), slice(3, -3), slice(3, -3))
            if ascii_flag:
                np.testing.assert_allclose(p2[sl], phi[sl], rtol=1e-6)
                np.testing.assert_allclose(m2[sl], mu[sl], rtol=1e-6)
            else:
                assert np.array_equal(p2[sl], phi[sl]) and np.array_equal(m2[sl], mu[sl])
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("text,needle", [
    ("ISOTHERMAL = 0;", "ISOTHERMAL"),
    ("Function_F = 2;", "Function_F"),
    ("Noise_phasefield = 1;", "Noise"),
])
def test_unsupported_inputs_are_rejected(text, needle):
    d = _copy_case("tern_c1_mini")
    try:
        with open(os.path.join(d, "Input.in"), "a") as f:
            f.write(text + "\n")
        with pytest.raises(host.InputError, match=needle):
            host.read_input(os.path.join(d, "Input.in"))
    finally:
        shutil.rmtree(d, ignore_errors=True)
