"""Compensated FP32 mode (GP_FP32C) on the host: the storage forms and the host build of its step.

The device kernels of the mode are checked bit for bit against this host build (tests/test_fp32c_gpu.py), so the
checks here pin the mode's arithmetic itself:
  - the storage forms (gp_model.h): mu as an FP32 part plus an 8-bit extension in units of ulp32/128, phi as two
    FP32 parts; exact values (0, 1, any FP32 number) are kept exactly, the rest to the stated precision;
  - the step's error against the FP64 oracle on the golden decks whose trajectory depends smoothly on the state:
    below the floor of plain FP32 storage (the FP64 run from the initial state rounded to FP32) for phi, within
    a small factor of it for mu.
Three decks (quad_k3, tern_iso_mixed, tern_lu_undercool) have trajectories that change by 1e-4..1e-2 under a
relative perturbation of 1e-15 of the initial mu (a phase appears or vanishes where the simplex projection meets
a value within rounding of 0); no perturbed precision can follow them, and the test shows that property instead
of gating an error there.
"""
import ctypes as C

import numpy as np
import pytest

import gp_testlib
from gp_testlib import golden_cases, load_golden, model_step_c32, oracle_step
from microsim_amd import _abi

SL = (slice(None), slice(3, -3), slice(3, -3))
CHAOTIC = ("quad_k3", "tern_iso_mixed", "tern_lu_undercool")


def _roundtrip(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    hi = np.empty(v.size, np.float32)
    q = np.empty(v.size, np.int8)
    gp_testlib.model_lib().c32_roundtrip(_abi.dptr(v), v.size, hi.ctypes.data, q.ctypes.data, _abi.dptr(out))
    return hi, q, out


def test_mu_extension_precision_and_exact_values():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.uniform(-3, 3, 20000), rng.uniform(-1e-3, 1e-3, 5000),
                        np.ldexp(rng.uniform(0.5, 1, 2000), rng.integers(-120, 120, 2000))])
    hi, q, out = _roundtrip(v)
    assert np.all(hi == v.astype(np.float32))              # the FP32 part is the rounded value
    assert np.all(np.abs(q.astype(int)) <= 64)            # |v - hi| <= ulp/2 = 64 units
    ulp = np.spacing(np.abs(hi)).astype(np.float64)       # ulp32 of the FP32 part (above it)
    assert np.all(np.abs(out - v) <= ulp / 256 * (1 + 1e-12))
    # values an FP32 number holds exactly come back exactly, 0 and 1 included
    exact = np.concatenate([[0.0, 1.0, -1.0, 0.5], rng.uniform(-2, 2, 1000).astype(np.float32).astype(np.float64)])
    _, q2, out2 = _roundtrip(exact)
    assert np.all(q2 == 0) and np.array_equal(out2, exact)


def test_phi_pair_keeps_one_minus_phi():
    """phi near 1: the two-part form keeps 1 - phi to FP32 relative precision (the step forms 1 - phi), and FP32
    numbers (0 and 1 among them) exactly."""
    eps = np.logspace(-15, -1, 400)
    v = np.ascontiguousarray(1.0 - eps)
    hi, lo = np.empty(v.size, np.float32), np.empty(v.size, np.float32)
    out = np.empty_like(v)
    gp_testlib.model_lib().df_roundtrip(_abi.dptr(v), v.size, hi.ctypes.data, lo.ctypes.data, _abi.dptr(out))
    one_minus = 1.0 - v   # exact (Sterbenz): the reference value of 1 - phi
    assert np.all(np.abs((1.0 - out) - one_minus) <= 2.0 ** -23 * one_minus)
    ex = np.ascontiguousarray([0.0, 1.0, 0.25, np.float64(np.float32(0.3))])
    h2, l2 = np.empty(4, np.float32), np.empty(4, np.float32)
    o2 = np.empty_like(ex)
    gp_testlib.model_lib().df_roundtrip(_abi.dptr(ex), 4, h2.ctypes.data, l2.ctypes.data, _abi.dptr(o2))
    assert np.all(l2 == 0.0) and np.array_equal(o2, ex)


@pytest.mark.parametrize("name", [n for n in golden_cases() if n not in CHAOTIC])
def test_c32_step_error_against_oracle(name):
    case, ref = load_golden(name)
    n = int(ref["steps"][-1])
    cphi, cmu = model_step_c32(case.params, case.phi, case.mu, n)
    ophi, omu = case.phi.copy(), case.mu.copy()
    oracle_step(case.params, ophi, omu, n)
    fphi = case.phi.astype(np.float32).astype(np.float64)
    fmu = case.mu.astype(np.float32).astype(np.float64)
    oracle_step(case.params, fphi, fmu, n)
    scale = max(1.0, np.abs(omu[SL]).max())
    ep, fp = np.abs(cphi[SL] - ophi[SL]).max(), np.abs(fphi[SL] - ophi[SL]).max()
    em, fm = np.abs(cmu[SL] - omu[SL]).max() / scale, np.abs(fmu[SL] - omu[SL]).max() / scale
    print("%s: phi %.3g (fp32 floor %.3g)  mu %.3g (floor %.3g)" % (name, ep, fp, em, fm))
    assert ep <= max(fp, 1e-12) and ep <= 1e-8
    assert em <= max(4.0 * fm, 1e-9) and em <= 1e-7


@pytest.mark.parametrize("name", CHAOTIC)
def test_chaotic_decks_follow_no_perturbed_run(name):
    """The FP64 oracle itself, from the initial state with mu perturbed by 1e-15 relative (a few ulp; the filled
    phi is exactly 0 or 1 everywhere), ends 1e-4 or more away from the unperturbed run on these decks in most of
    three tries (measured: quad_k3 7e-3, tern_iso_mixed 2.5e-4, tern_lu_undercool 1.3e-2)."""
    case, ref = load_golden(name)
    n = int(ref["steps"][-1])
    ophi, omu = case.phi.copy(), case.mu.copy()
    oracle_step(case.params, ophi, omu, n)
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(3):
        p, m = case.phi.copy(), case.mu.copy()
        m *= 1.0 + 1e-15 * rng.standard_normal(m.shape)
        oracle_step(case.params, p, m, n)
        worst = max(worst, np.abs(p[SL] - ophi[SL]).max())
    assert worst >= 1e-4
