"""FP32 production mode against the FP64 oracle.

The FP32 mode stores phi and mu in FP32 and computes the step in FP32 (statistics and set-up in FP64). Its error
against the FP64 oracle is bounded below by the sensitivity of the trajectory itself: the same FP64 run started
from the initial state rounded to FP32 (the perturbation FP32 storage makes before the first step) already
drifts, and on the growing-seed workload the drift grows with the step count. The bounds below are for 1000
steps of that workload at 256^2 (8 seeds): the FP32 mode measured phi 1.7e-5, mu 3.1e-6 in a host build of the
same arithmetic; the bound leaves a factor of about 5.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from gp_testlib import load_golden, oracle_step
from microsim_amd import _abi, configs, host

pytestmark = pytest.mark.gpu


def _run(case, precision, n):
    from microsim_amd.solver import GPSolver
    with GPSolver(case.params, precision) as s:
        s.upload(case.phi, case.mu)
        st = s.step(n, stats=True)
        assert st["nonfinite"] == 0
        return s.download()


def test_fp32_seed_growth_within_bounds():
    print("[gpu] fp32 256^2 8 seeds 1000 steps", flush=True)
    sys.stdout.flush()
    inp, fills, _ = configs.config_ternary(256, 8, seed=1234, min_spacing=60.0)
    d = tempfile.mkdtemp(prefix="gp_fp32_")
    configs.write_case(d, inp, fills)
    case = host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))
    n = 1000
    phi, mu = _run(case, _abi.GP_FP32, n)
    ophi, omu = case.phi.copy(), case.mu.copy()
    oracle_step(case.params, ophi, omu, n, threads=int(os.environ.get("OMP_NUM_THREADS", "8")))
    sl = (slice(None), slice(3, -3), slice(3, -3))
    ep, em = np.abs(phi[sl] - ophi[sl]).max(), np.abs(mu[sl] - omu[sl]).max()
    c32 = host.composition(case.params, np.ascontiguousarray(phi[sl]), np.ascontiguousarray(mu[sl]))
    c64 = host.composition(case.params, np.ascontiguousarray(ophi[sl]), np.ascontiguousarray(omu[sl]))
    ec = np.abs(c32 - c64).max()
    print("   max abs error phi %.3g mu %.3g c %.3g" % (ep, em, ec), flush=True)
    assert ep <= 1e-4 and em <= 2e-5 and ec <= 2e-5
    # the seeds grew: the workload is not degenerate
    assert ophi[:2][sl].sum() > case.phi[:2][sl].sum() * 1.2


@pytest.mark.parametrize("name", ["tern_c1_mini", "bin_aniso_neu", "quad_k3"])
def test_fp32_golden_short(name):
    """A few steps of the reference decks: FP32 stays within single-precision rounding of the oracle."""
    print("[gpu] fp32 golden", name, flush=True)
    case, _ = load_golden(name)
    phi, mu = _run(case, _abi.GP_FP32, 10)
    ophi, omu = case.phi.copy(), case.mu.copy()
    oracle_step(case.params, ophi, omu, 10)
    sl = (slice(None), slice(3, -3), slice(3, -3))
    ep = np.abs(phi[sl] - ophi[sl]).max()
    em = np.abs(mu[sl] - omu[sl]).max() / max(1.0, np.abs(omu[sl]).max())
    print("   max abs error phi %.3g, mu (relative to max |mu|) %.3g" % (ep, em), flush=True)
    assert ep <= 1e-5 and em <= 1e-5
