"""FP64 parity at the size of configuration 1: the ternary three-phase 1024^2 workload with 64 seeds (periodic sides,
random SPD Hessians, 4-fold anisotropy, obstacle potential), 200 steps, against the threaded FP64 oracle.

Gate: max relative error <= 1e-12 per field over the whole halo-inclusive state, and the reference loop's
cumulative max/min statistics equal to the oracle's (they are extrema of bit-identical values).
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from gp_testlib import max_rel_err, oracle, oracle_step
from microsim_amd import _abi, configs, host

pytestmark = pytest.mark.gpu

GATE = 1e-12


def test_fp64_matches_oracle_at_config1_size():
    from microsim_amd.solver import GPSolver
    print("[gpu] config-1 size: ternary 1024^2, 64 seeds, 200 FP64 steps", flush=True)
    sys.stdout.flush()
    inp, fills, _ = configs.config_ternary(1024, 64, seed=1234, min_spacing=60.0)
    d = tempfile.mkdtemp(prefix="gp_c1_")
    configs.write_case(d, inp, fills)
    case = host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))
    n = 200
    with GPSolver(case.params, _abi.GP_FP64) as s:
        s.upload(case.phi, case.mu)
        got = s.step(n, stats=True)
        phi, mu = s.download()
    st = _abi.GPStats()
    oracle().oracle_stats_init(st)
    ophi, omu = case.phi.copy(), case.mu.copy()
    oracle_step(case.params, ophi, omu, n, stats=st, threads=int(os.environ.get("OMP_NUM_THREADS", "8")))
    ep, em = max_rel_err(phi, ophi), max_rel_err(mu, omu)
    nd = int((phi != ophi).sum() + (mu != omu).sum())
    print("   max rel err phi %.3g mu %.3g, differing values %d" % (ep, em, nd), flush=True)
    assert ep <= GATE and em <= GATE
    N, NC = case.N, case.NC
    assert got["phi_max"] == list(st.phi_max[:N]) and got["phi_min"] == list(st.phi_min[:N])
    assert got["mu_max"] == list(st.mu_max[:NC]) and got["mu_min"] == list(st.mu_min[:NC])
    assert got["nonfinite"] == 0 and got["step"] == n
    # the workload is not degenerate: the solids grew and interfaces moved
    sl = (slice(None), slice(3, -3), slice(3, -3))
    assert ophi[:2][sl].sum() > case.phi[:2][sl].sum() * 1.02
