"""Derivative references for the device adjoint, on the host.

model_jvp (tests/cpu_model/model_step.cpp) steps the kernels' per-cell functions in dual numbers; torch_step.py is
an independent whole-grid torch restatement of the step, differentiated by torch.autograd. Here:
  * model_jvp's values are the FP64 oracle's, bit for bit (the dual arithmetic leaves the value path unchanged);
  * torch_step's values agree with the oracle to rounding on every golden deck;
  * the two derivatives agree: <grad L, v> from torch.autograd equals <w, J v> from model_jvp to rounding, for a
    tangent v over every interior cell of phi and mu and every parameter, on every golden deck;
  * model_jvp's tangent matches central finite differences of the oracle over one step on every golden deck,
    and over twenty steps on the decks whose trajectories are smooth at the finite-difference scale.
Over many steps a trajectory that starts from a sharp interface is not smooth at that scale: cells of the
spreading tail cross the anisotropy's |q|^2 > 1e-15 cut-off and the simplex projection's clamp, so central
differences of a deck such as bin_well_rot wander by 1e-4 relative as the step size changes. Those decks are
checked through the dot-product test of the device adjoint against this tangent (tests/test_adjoint_gpu.py),
which holds whatever the branches.
"""
import numpy as np
import pytest

import torch_step
import gp_testlib
from gp_testlib import golden_cases, load_golden, model_jvp, oracle_step
from microsim_amd import adjoint


def directions(case, seed, phi_cells="interface"):
    """A tangent (vphi, vmu, vtheta): mu on every interior cell, phi on the interface cells (every phase strictly
    inside (0, 1)) or on every interior cell, theta relative to its values."""
    rng = np.random.default_rng(seed)
    inner = np.zeros(case.phi.shape[1:], dtype=bool)
    inner[3:-3, 3:-3] = True
    cells = inner
    if phi_cells == "interface":
        cells = inner & np.all((case.phi > 1e-3) & (case.phi < 1 - 1e-3), axis=0)
    vphi = rng.standard_normal(case.phi.shape) * cells
    vmu = rng.standard_normal(case.mu.shape) * inner
    th = adjoint.theta_from_params(case.params)
    vth = rng.standard_normal(th.shape) * np.abs(th)
    return vphi, vmu, vth


def _oracle_run(case, params, phi, mu, n):
    phi, mu = phi.copy(), mu.copy()
    oracle_step(params, phi, mu, n)
    return phi, mu


def _fd(case, vphi, vmu, vth, n, h):
    th = adjoint.theta_from_params(case.params)
    out = []
    for s in (h, -h):
        p = adjoint.params_with_theta(case.params, th + s * vth)
        out.append(_oracle_run(case, p, case.phi + s * vphi, case.mu + s * vmu, n))
    return [(a - b) / (2 * h) for a, b in zip(out[0], out[1])]


@pytest.mark.parametrize("name", golden_cases())
def test_jvp_values_are_the_oracle(name):
    case, _ = load_golden(name)
    vphi, vmu, vth = directions(case, 1, phi_cells="all")
    phi, mu, _, _ = model_jvp(case.params, case.phi, case.mu, vphi, vmu, vth, 12)
    rphi, rmu = _oracle_run(case, case.params, case.phi, case.mu, 12)
    assert np.array_equal(phi, rphi) and np.array_equal(mu, rmu)


def _check_tangent(case, n, seed, h, rtol):
    vphi, vmu, vth = directions(case, seed)
    _, _, tphi, tmu = model_jvp(case.params, case.phi, case.mu, vphi, vmu, vth, n)
    fphi, fmu = _fd(case, vphi, vmu, vth, n, h)
    for t, f in ((tphi, fphi), (tmu, fmu)):
        scale = np.abs(f).max()
        assert scale > 0
        assert np.abs(t - f).max() <= rtol * scale, (np.abs(t - f).max(), scale)


@pytest.mark.parametrize("name", golden_cases())
def test_jvp_one_step_matches_central_differences(name):
    case, _ = load_golden(name)
    _check_tangent(case, 1, 2, 1e-7, 1e-6)


@pytest.mark.parametrize("name", [c for c in golden_cases() if c != "bin_well_rot"])
def test_jvp_trajectory_matches_central_differences(name):
    case, _ = load_golden(name)
    _check_tangent(case, 20, 3, 1e-6, 1e-6)


def torch_trajectory(case, n, theta=None, phi0=None, mu0=None):
    """n torch steps from (phi0, mu0) (the case's state by default), boundary copies applied first as an upload
    does; returns (phi, mu)."""
    import torch
    m = torch_step.TorchModel(case.params, theta)
    phi = phi0 if phi0 is not None else torch.tensor(case.phi)
    mu = mu0 if mu0 is not None else torch.tensor(case.mu)
    phi = torch.stack(torch_step.apply_bc(m, list(phi), 0))
    mu = torch.stack(torch_step.apply_bc(m, list(mu), 1))
    for _ in range(n):
        phi, mu = torch_step.step(m, phi, mu)
    return phi, mu


@pytest.mark.parametrize("name", golden_cases())
def test_torch_step_matches_oracle(name):
    case, _ = load_golden(name)
    phi, mu = torch_trajectory(case, 5)
    rphi, rmu = _oracle_run(case, case.params, case.phi, case.mu, 5)
    assert np.abs(phi.numpy() - rphi).max() <= 1e-14 * max(np.abs(rphi).max(), 1.0)
    assert np.abs(mu.numpy() - rmu).max() <= 1e-14 * max(np.abs(rmu).max(), 1.0)


@pytest.mark.parametrize("name", golden_cases())
def test_torch_autograd_matches_forward_mode(name):
    import torch
    case, _ = load_golden(name)
    rng = np.random.default_rng(5)
    wphi, wmu = rng.standard_normal(case.phi.shape), rng.standard_normal(case.mu.shape)
    vphi, vmu, vth = directions(case, 6, phi_cells="all")
    th = torch.tensor(adjoint.theta_from_params(case.params), requires_grad=True)
    phi0 = torch.tensor(case.phi, requires_grad=True)
    mu0 = torch.tensor(case.mu, requires_grad=True)
    phi, mu = torch_trajectory(case, 5, th, phi0, mu0)
    ((phi * torch.from_numpy(wphi)).sum() + (mu * torch.from_numpy(wmu)).sum()).backward()
    lhs = np.concatenate([(phi0.grad.numpy() * vphi).ravel(), (mu0.grad.numpy() * vmu).ravel(), th.grad.numpy() * vth])
    _, _, tphi, tmu = model_jvp(case.params, case.phi, case.mu, vphi, vmu, vth, 5)
    rhs = np.concatenate([(wphi * tphi).ravel(), (wmu * tmu).ravel()])
    assert abs(lhs.sum() - rhs.sum()) <= 1e-13 * np.abs(rhs).sum()


@pytest.mark.parametrize("name", golden_cases())
def test_reverse_mode_is_the_transposed_tangent(name):
    """The reverse-mode per-cell functions (gp_reverse.h, which the device adjoint runs), composed over every cell of
    five steps on the host (model_vjp), against the forward-mode tangent (model_jvp): the dot-product identity
    <J^T w, v> = <w, J v> for a random tangent over every interior cell and every parameter, checked for the phi,
    mu and theta blocks separately (a wrong term in any block cannot cancel against another)."""
    case, _ = load_golden(name)
    n = 5
    vphi, vmu, vth = directions(case, 1, phi_cells="all")
    rng = np.random.default_rng(7)
    wphi, wmu = rng.standard_normal(case.phi.shape), rng.standard_normal(case.mu.shape)
    g0, m0, gth = gp_testlib.model_vjp(case.params, case.phi, case.mu, wphi, wmu, n)
    zero_p, zero_m, zero_t = np.zeros_like(vphi), np.zeros_like(vmu), np.zeros_like(vth)
    for vp, vm, vt, got in ((vphi, zero_m, zero_t, (g0 * vphi).sum()), (zero_p, vmu, zero_t, (m0 * vmu).sum()),
                            (zero_p, zero_m, vth, (gth * vth).sum())):
        _, _, tp, tm = model_jvp(case.params, case.phi, case.mu, vp, vm, vt, n)
        want = (wphi * tp).sum() + (wmu * tm).sum()
        assert got == pytest.approx(want, rel=1e-11, abs=1e-13 * (np.abs(wphi * tp).sum() + np.abs(wmu * tm).sum()))
    # the halo entries of the initial-state gradient are zero (they are copies of the interior)
    assert not g0[:, :3].any() and not g0[:, -3:].any() and not g0[:, :, :3].any() and not m0[:, :, -3:].any()


def test_reverse_projection_when_no_phase_lies_inside():
    """A cell where the projection clamps a vanishing phase while no active phase lay strictly inside (0, 1) before
    the clamping (simplex_projection.h: count = 0, so delta / count is an infinity that the phase pushed above 1 then
    overwrites with -phi). The forward value is finite; the transposed projection must be too and must equal the
    dual-number tangent. The 3x3 neighbourhood is from a ternary 2048^2 seed run at step 3702, where a phase-1 grain
    meets the liquid with phase 0 absent."""
    import os
    import tempfile
    from microsim_amd import configs, host
    inp, fills, _ = configs.config_ternary(58, 1, seed=1234, min_spacing=30.0)
    d = tempfile.mkdtemp(prefix="gp_proj_")
    configs.write_case(d, inp, fills)
    case = host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))
    phi, mu = case.phi.copy(), case.mu.copy()
    phi[0, 32:35, 32:35] = 0.0
    phi[1, 32:35, 32:35] = [[0.9719499370683231, 1.0, 1.0], [0.9593802331723985, 0.9999994421461349, 1.0],
                            [0.9391134117412449, 0.9906544827095958, 1.0]]
    phi[2, 32:35, 32:35] = [[0.028050062931676667, 0.0, 0.0], [0.04061976682760295, 5.578538646156249e-07, 0.0],
                            [0.06088658825875734, 0.009345517290404339, 0.0]]
    mu[:, 33, 33] = [0.3959179828929696, 0.5021716343457124]
    ry = phi.shape[2]
    dphi = np.zeros_like(phi)
    for o in range(case.N):
        w = np.zeros(case.N)
        w[o] = 1.0
        vjp, jvp = gp_testlib.cell_check(case.params, phi, mu, dphi, 33 * ry + 33, 0, w)
        assert np.isfinite(jvp).all()
        assert np.isfinite(vjp).all(), "output %d: non-finite transposed projection" % o
        np.testing.assert_allclose(vjp, jvp, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(jvp).max()))
