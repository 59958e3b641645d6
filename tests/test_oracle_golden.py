"""Pin the FP64 oracle against the serial reference's own output.

The golden vectors were written by the reference binary built from its sources (oracle/Makefile "ref",
tests/golden/make_golden.py). The oracle must reproduce them exactly (full-precision binary VTK cases) at every
stored step. Where the reference tree is present, the same comparison also runs live against a fresh reference
run of a perturbed case, so the fixtures are not the only anchor.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from gp_testlib import REPO, golden_cases, load_golden, oracle_smooth, oracle_step
from microsim_amd import configs

CASES = golden_cases()


def _run_oracle_to(case, steps):
    phi, mu = case.phi.copy(), case.mu.copy()
    out = {}
    t = 0
    for s in steps:
        if s > t:
            if t == 0 and case.inp.nsmooth > 1:
                oracle_smooth(case.params, phi, mu, case.inp.nsmooth - 1)
            oracle_step(case.params, phi, mu, s - t)
            t = s
        out[s] = (phi[:, 3:-3, 3:-3].copy(), mu[:, 3:-3, 3:-3].copy())
    return out


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    case, ref = load_golden(name)
    got = _run_oracle_to(case, ref["steps"])
    ascii_out = bool(case.inp.ascii)
    for i, s in enumerate(ref["steps"]):
        p, m = got[s]
        if ascii_out:
            # the ASCII writer keeps 7 significant digits (%le)
            np.testing.assert_allclose(p, ref["phi"][i], rtol=0, atol=1e-6 * max(1, np.abs(ref["phi"][i]).max()))
            np.testing.assert_allclose(m, ref["mu"][i], rtol=1e-6, atol=1e-12)
        else:
            assert np.array_equal(p, ref["phi"][i]), "%s step %d: phi differs (max %g)" % (
                name, s, np.abs(p - ref["phi"][i]).max())
            assert np.array_equal(m, ref["mu"][i]), "%s step %d: mu differs (max %g)" % (
                name, s, np.abs(m - ref["mu"][i]).max())


def test_golden_covers_every_branch():
    names = set(CASES)
    for required in ("bin_aniso_neu", "bin_tanh_restart", "tern_aniso_per", "tern_iso_mixed", "bin_well_rot",
                     "tern_lu_undercool", "tern_dilute", "bin_smooth", "quad_k3", "tern_c1_mini"):
        assert required in names


REF_BIN = os.path.join(REPO, "oracle", "_ref", "microsim_gp")


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference binary not built (make -C oracle ref)")
def test_oracle_matches_live_reference_perturbed():
    """A deck that is not among the fixtures (different seed layout, dx != dy, 3 phases, periodic y only)."""
    from microsim_amd import host
    rng = np.random.default_rng(99)
    inp = configs.GPInput(mesh_x=30, mesh_y=26, nphases=3, ncomp=3, dx=1.1, dy=0.9, dt=0.02, ntimesteps=25,
                          savet=25, gamma=[1.0, 0.8, 1.2], tau_ab=[0.3, 0.25, 0.35], dab=[0.02, 0.03, 0.01],
                          gamma_abc=[0.5], components=["A", "B", "C"], phases=["s1", "s2", "liq"],
                          bc_phi=(1, 1, 3, 3), rotations={(0, 2): (0.0, 0.0, 15.0)})
    configs.ternary_material(inp, rng)
    fills = [configs.Fill("cylinder", 0, (10, 12, 5.0)), configs.Fill("cylinder", 1, (20, 14, 6.0))]
    work = tempfile.mkdtemp(prefix="gp_live_")
    try:
        configs.write_case(work, inp, fills)
        r = subprocess.run([REF_BIN, "Input.in", "Filling.in", "o"], cwd=work, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        gp, gm = configs.read_vtk(os.path.join(work, "DATA", "o_25.vtk"), 3, 3, 30, 26)
        case = host.load_case(os.path.join(work, "Input.in"), os.path.join(work, "Filling.in"))
        phi, mu = case.phi.copy(), case.mu.copy()
        oracle_step(case.params, phi, mu, 25)
        assert np.array_equal(phi[:, 3:-3, 3:-3], gp)
        assert np.array_equal(mu[:, 3:-3, 3:-3], gm)
    finally:
        shutil.rmtree(work, ignore_errors=True)
