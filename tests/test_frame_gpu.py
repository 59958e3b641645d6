"""Non-isothermal runs and the moving frame on the GPU, against the reference's own runs of the frame decks.

ISOTHERMAL = 0 with Tempgrady: the temperature depends on y only and is reset from the gradient after every step
(Grand_potential_Finite_difference_2D_serial/microsim_gp.c:107-112); Shift = 1: every 100 steps the frame moves
down by the rows the solid front has passed Shiftj (microsim_gp.c:114-125, functions/shift.h). The FP64 device
path must reproduce phi, mu and the Temperature blocks bit for bit at every stored step, and the frame offset of
every DATA/shift.dat line. The drop-in driver is checked on the same decks, file by file.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from gp_testlib import GOLDEN, frame_cases, load_golden
from microsim_amd import _abi, configs

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(_abi.LIBDIR, "microsim_gp")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _shift_positions(name):
    out = {}
    for line in open(os.path.join(GOLDEN, name, "shift.dat")):
        t, pos = line.split()
        out[int(t)] = int(pos)
    return out


@pytest.mark.parametrize("name", frame_cases())
def test_frame_deck_bit_exact(name):
    from microsim_amd.solver import GPSolver
    _say("[gpu] frame deck", name)
    case, ref = load_golden(name)
    pos = _shift_positions(name)
    my = int(case.params.mesh_y)
    with GPSolver(case.params) as s:
        s.upload(case.phi, case.mu)
        t = 0
        for i, step in enumerate(ref["steps"]):
            if step > t:
                s.step(step - t)
                t = step
            phi, mu = s.download()
            off, T = s.frame()
            p, m = phi[:, 3:-3, 3:-3], mu[:, 3:-3, 3:-3]
            nd = int((p != ref["phi"][i]).sum() + (m != ref["mu"][i]).sum())
            _say("   step %d: differing values %d, frame offset %d" % (step, nd, off))
            assert nd == 0, "step %d: max |dphi| %g |dmu| %g" % (step, np.abs(p - ref["phi"][i]).max(),
                                                                 np.abs(m - ref["mu"][i]).max())
            assert np.array_equal(np.broadcast_to(T[3:my + 3], ref["T"][i].shape), ref["T"][i]), "T at %d" % step
            if step > 0:
                assert off == pos[step]


def test_frame_rejected_where_unsupported():
    """FP32C and the adjoint do not take the frame decks: creation / recording fails with a message."""
    from microsim_amd.solver import GPError, GPSolver
    _say("[gpu] frame rejections")
    case, _ = load_golden("bin_tempgrad_shift")
    with pytest.raises(GPError):
        GPSolver(case.params, _abi.GP_FP32C)
    with GPSolver(case.params) as s:
        s.upload(case.phi, case.mu)
        with pytest.raises(GPError):
            s.adjoint_forward(4, 2)


def _set_keys(path, **keys):
    text = open(path).read()
    for k, v in keys.items():
        text, n = re.subn(r"\b%s\s*=\s*[^;]*;" % k, "%s = %s;" % (k, v), text)
        assert n == 1, k
    open(path, "w").write(text)


def _run_driver(d):
    r = subprocess.run([DRIVER, "Input.in", "Filling.in", "out"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("name", frame_cases())
def test_driver_frame_deck(name):
    """The driver's snapshots (phi, mu, Temperature) at every stored step and its DATA/shift.dat, byte for byte
    against the reference's run (SAVET = 1 in both)."""
    _say("[gpu] driver frame deck", name)
    if not os.path.exists(DRIVER):
        pytest.fail("microsim_gp is not built: run __graft_entry__.build()")
    case, ref = load_golden(name)
    mx, my = int(case.params.mesh_x), int(case.params.mesh_y)
    d = tempfile.mkdtemp(prefix="gp_frame_")
    try:
        for f in ("Input.in", "Filling.in"):
            shutil.copy(os.path.join(GOLDEN, name, f), d)
        _run_driver(d)
        for i, s in enumerate(ref["steps"]):
            f = configs.read_vtk_fields(os.path.join(d, "DATA", "out_%d.vtk" % s), mx, my)
            assert list(f)[-1] == "Temperature"
            assert np.array_equal(f["Temperature"], ref["T"][i]), "T at step %d" % s
            phi, mu = configs.read_vtk(os.path.join(d, "DATA", "out_%d.vtk" % s), case.N, case.NC + 1, mx, my)
            assert np.array_equal(phi, ref["phi"][i]) and np.array_equal(mu, ref["mu"][i]), "step %d" % s
        got = open(os.path.join(d, "DATA", "shift.dat")).read()
        assert got == open(os.path.join(GOLDEN, name, "shift.dat")).read()
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_driver_frame_restart():
    """A moving-frame run restarted at step 200 from its snapshot and DATA/shift.dat (microsim_gp.c:56-74: the
    frame position comes back from shift.dat, the gradient offset moves by floor(velocity STARTTIME dt) minus that
    position) continues the uninterrupted run: same frame offsets, the same Temperature block up to the rounding of
    the two time sums, and the same state to 1e-12."""
    _say("[gpu] driver frame restart")
    if not os.path.exists(DRIVER):
        pytest.fail("microsim_gp is not built: run __graft_entry__.build()")
    name = "bin_tempgrad_shift"
    case, _ = load_golden(name)
    mx, my = int(case.params.mesh_x), int(case.params.mesh_y)
    d1, d2 = tempfile.mkdtemp(prefix="gp_fr1_"), tempfile.mkdtemp(prefix="gp_fr2_")
    try:
        for d in (d1, d2):
            for f in ("Input.in", "Filling.in"):
                shutil.copy(os.path.join(GOLDEN, name, f), d)
        _set_keys(os.path.join(d1, "Input.in"), NTIMESTEPS=400, SAVET=100, TRACK_PROGRESS=100)
        _run_driver(d1)
        os.makedirs(os.path.join(d2, "DATA"))
        shutil.copy(os.path.join(d1, "DATA", "out_200.vtk"), os.path.join(d2, "DATA"))
        shutil.copy(os.path.join(d1, "DATA", "shift.dat"), os.path.join(d2, "DATA"))
        _set_keys(os.path.join(d2, "Input.in"), NTIMESTEPS=200, SAVET=100, TRACK_PROGRESS=100, RESTART=1,
                  STARTTIME=200)
        _run_driver(d2)
        lines1 = open(os.path.join(d1, "DATA", "shift.dat")).read().split("\n")
        lines2 = open(os.path.join(d2, "DATA", "shift.dat")).read().split("\n")
        assert lines2[4:6] == lines1[2:4], (lines1, lines2)
        for s in (300, 400):
            a = configs.read_vtk_fields(os.path.join(d1, "DATA", "out_%d.vtk" % s), mx, my)
            b = configs.read_vtk_fields(os.path.join(d2, "DATA", "out_%d.vtk" % s), mx, my)
            assert list(a) == list(b)
            for k in a:
                np.testing.assert_allclose(b[k], a[k], rtol=1e-12, atol=1e-14, err_msg="%s at %d" % (k, s))
    finally:
        shutil.rmtree(d1, ignore_errors=True)
        shutil.rmtree(d2, ignore_errors=True)
