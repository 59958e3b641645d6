"""Compensated FP32 mode (GP_FP32C) on the GPU.

1. Bit for bit against a host build of the same arithmetic (tests/cpu_model/model_step.cpp, run_c32): the
   8-bit extensions, the FP32 bulk closed form and the FP64 interface arithmetic are deterministic IEEE operations,
   so the device's decoded state equals the host's exactly, on every golden deck (both spacing forms, every solve
   mode, both potentials, periodic / Neumann / mixed sides) and on a growing-seed ternary case.
2. Error budget against the FP64 oracle on the ternary growing-seed workload (256^2, 8 seeds, 1000 steps): max
   |phi| error at most 1e-5 and within 3x of the floor, the FP64 run from the initial state rounded to FP32 (the
   perturbation plain FP32 storage makes before the first step). mu and the composition c likewise.
3. The paths that run FP64 kernels on the decoded state: delta-phi downloads, smoothing steps, solute/energy.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from gp_testlib import golden_cases, load_golden, model_step_c32, oracle, oracle_step
from microsim_amd import _abi, configs, host

pytestmark = pytest.mark.gpu
SL = (slice(None), slice(3, -3), slice(3, -3))


def _gpu(case, n, smooth=0, dphi=False):
    from microsim_amd.solver import GPSolver
    with GPSolver(case.params, _abi.GP_FP32C) as s:
        s.upload(case.phi, case.mu)
        if smooth:
            s.smooth(smooth)
        st = s.step(n, stats=True) if n else s.stats()
        assert st["nonfinite"] == 0
        return s.download(with_dphi=dphi) + (st,)


def _seed_case(mesh, seeds, spacing):
    inp, fills, _ = configs.config_ternary(mesh, seeds, seed=1234, min_spacing=spacing)
    d = tempfile.mkdtemp(prefix="gp_fp32c_")
    configs.write_case(d, inp, fills)
    return host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))


@pytest.mark.parametrize("name", golden_cases())
def test_fp32c_golden_bit_exact_to_host_build(name):
    print("[gpu] fp32c bit-exact", name, flush=True)
    case, _ = load_golden(name)
    n = 20
    phi, mu, _ = _gpu(case, n)
    hphi, hmu = model_step_c32(case.params, case.phi, case.mu, n)
    assert np.array_equal(phi, hphi), np.abs(phi - hphi).max()
    assert np.array_equal(mu, hmu), np.abs(mu - hmu).max()


def test_fp32c_seed_case_bit_exact_to_host_build():
    print("[gpu] fp32c bit-exact ternary seeds 128^2 x 150 steps", flush=True)
    case = _seed_case(128, 6, 40.0)
    phi, mu, _ = _gpu(case, 150)
    hphi, hmu = model_step_c32(case.params, case.phi, case.mu, 150)
    assert np.array_equal(phi, hphi), np.abs(phi - hphi).max()
    assert np.array_equal(mu, hmu), np.abs(mu - hmu).max()
    # the interfaces moved: the seeds grow
    assert phi[:2][SL].sum() > case.phi[:2][SL].sum()


def test_fp32c_error_budget_seed_growth():
    print("[gpu] fp32c 256^2 8 seeds 1000 steps against the oracle", flush=True)
    sys.stdout.flush()
    case = _seed_case(256, 8, 60.0)
    n = 1000
    phi, mu, _ = _gpu(case, n)
    th = int(os.environ.get("OMP_NUM_THREADS", "8"))
    ophi, omu = case.phi.copy(), case.mu.copy()
    oracle_step(case.params, ophi, omu, n, threads=th)
    # the floor: the FP64 run from the initial state rounded to FP32
    fphi = case.phi.astype(np.float32).astype(np.float64)
    fmu = case.mu.astype(np.float32).astype(np.float64)
    oracle_step(case.params, fphi, fmu, n, threads=th)

    def comp(p, m):
        return host.composition(case.params, np.ascontiguousarray(p[SL]), np.ascontiguousarray(m[SL]))
    c, oc, fc = comp(phi, mu), comp(ophi, omu), comp(fphi, fmu)
    e = {"phi": np.abs(phi[SL] - ophi[SL]).max(), "mu": np.abs(mu[SL] - omu[SL]).max(), "c": np.abs(c - oc).max()}
    f = {"phi": np.abs(fphi[SL] - ophi[SL]).max(), "mu": np.abs(fmu[SL] - omu[SL]).max(), "c": np.abs(fc - oc).max()}
    print("   max abs error  phi %.3g  mu %.3g  c %.3g" % (e["phi"], e["mu"], e["c"]), flush=True)
    print("   floor          phi %.3g  mu %.3g  c %.3g" % (f["phi"], f["mu"], f["c"]), flush=True)
    assert e["phi"] <= 1e-5
    for k in e:
        assert e[k] <= 3.0 * f[k], k
    assert ophi[:2][SL].sum() > case.phi[:2][SL].sum() * 1.2


def test_fp32c_delta_phi_download_and_invariants():
    """delta-phi of the last step and the solute/energy: FP64 arithmetic on the decoded state, equal to the oracle's
    on the downloaded (decoded) states: delta-phi to 1e-12 of its largest value (the mode's anisotropy terms take
    one reciprocal of |q|^2 where the oracle takes the reference's five quotients), the sums to 1e-12 (their
    order differs)."""
    import ctypes as C
    from microsim_amd.solver import GPSolver
    case, _ = load_golden("tern_c1_mini")
    with GPSolver(case.params, _abi.GP_FP32C) as s:
        s.upload(case.phi, case.mu)
        s.step(9)
        p9, m9 = s.download()
        s.step(1)
        p10, m10, d10 = s.download(with_dphi=True)
        st = s.stats()
    op, om, od = p9.copy(), m9.copy(), np.zeros_like(p9)
    oracle_step(case.params, op, om, 1, dphi=od)
    sl = (slice(None), slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(d10[sl], od[sl], rtol=0, atol=1e-12 * np.abs(od[sl]).max())
    assert np.abs(od[sl]).max() > 1e-4
    sol, en = (C.c_double * 4)(), C.c_double()
    oracle().oracle_invariants(C.byref(case.params), _abi.dptr(np.ascontiguousarray(p10)),
                               _abi.dptr(np.ascontiguousarray(m10)), sol, C.byref(en))
    assert np.allclose(st["solute"], list(sol)[:case.NC], rtol=1e-12, atol=0)
    assert abs(st["energy"] - en.value) <= 1e-12 * abs(en.value)


def test_fp32c_smoothing_steps():
    """The smoothing steps run the FP64 two-pass kernels on the decoded state: after them the state equals the
    oracle's smoothing of the same decoded state, up to the encoding (about 2^-31 relative)."""
    from microsim_amd.solver import GPSolver
    case, _ = load_golden("bin_smooth")
    with GPSolver(case.params, _abi.GP_FP32C) as s:
        s.upload(case.phi, case.mu)
        p0, m0 = s.download()
        s.smooth(5)
        p5, m5 = s.download()
    op, om = p0.copy(), m0.copy()
    from gp_testlib import oracle_smooth
    oracle_smooth(case.params, op, om, 5)
    assert np.abs(p5[SL] - op[SL]).max() <= 1e-8
    assert np.abs(m5[SL] - om[SL]).max() <= 1e-8 * max(1.0, np.abs(om[SL]).max())


def test_fp32c_adjoint_matches_fp64_adjoint():
    """The adjoint of a compensated trajectory (FP64 Jacobians at the decoded states) against the FP64 adjoint: the
    trajectories differ by the encoding only, so the gradients agree to about 1e-6 of their largest entry."""
    from microsim_amd.solver import GPSolver
    case, _ = load_golden("tern_c1_mini")
    rng = np.random.default_rng(3)
    wphi, wmu = rng.standard_normal(case.phi.shape), rng.standard_normal(case.mu.shape)
    out = {}
    for prec in (_abi.GP_FP64, _abi.GP_FP32C):
        with GPSolver(case.params, prec) as s:
            s.upload(case.phi, case.mu)
            s.adjoint_forward(20, 5)
            out[prec] = s.adjoint_backward(wphi, wmu)
    for a, b in zip(out[_abi.GP_FP64], out[_abi.GP_FP32C]):
        assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
