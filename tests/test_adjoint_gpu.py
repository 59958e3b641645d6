"""Adjoint of the FP64 solver loop against forward-mode derivatives and central finite differences.

The loss is a fixed random linear functional of the final state, L = <w_phi, phi_T> + <w_mu, mu_T> over every
halo-inclusive entry (so the transposed boundary conditions at the final state are exercised too). The checks:
  * dot-product test on every golden deck: <adjoint gradient, v> against <w, J v>, where J v is the host
    forward-mode tangent of the same trajectory (tests/cpu_model/model_step.cpp, itself checked against finite
    differences of the oracle in tests/test_adjoint_cpu.py), for a random v over every interior cell of phi and
    mu and every parameter; and dL/dtheta component by component against the tangent along each parameter.
    Both sides take the branches the values take, so these hold to rounding on any deck;
  * dL/dtheta for every parameter, against central differences of the GPU forward with a relative step of 1e-6,
    and the state gradient along a random direction (all interior mu cells and the interface cells of phi)
    against the central difference of L along it, on the decks whose trajectories are smooth at that scale
    (bin_well_rot is not: cells of its spreading tail cross the anisotropy cut-off and the projection's clamp);
  * independence of the checkpoint interval (the reverse sweep recomputes segments; the gradients must agree to
    rounding, which the atomic accumulation makes order-dependent);
  * the forward of the differentiable trajectory equals plain stepping bit for bit;
  * the full gradients (initial phi, mu and theta) equal those of torch.autograd through torch_step.py, an
    independent FP64 torch restatement of the step, to rounding;
  * FP32 handles: the reverse sweep over the FP32 trajectory equals the chain of torch.autograd vector-Jacobian
    products of the FP64 torch step taken at the FP32 states (what the FP32 adjoint is defined to compute), to
    rounding; its gradients stay close to the FP64 adjoint's.
The decks are the golden ones (binary, ternary and four-phase; isotropic and anisotropic, rotated; obstacle and
well potentials; periodic, Neumann and mixed sides; dilute and LU solves; undercooling).
"""
import sys

import numpy as np
import pytest

from gp_testlib import golden_cases, load_golden, model_jvp
from microsim_amd import _abi

CASES = ["bin_tanh_restart", "bin_aniso_neu", "tern_aniso_per", "tern_c1_mini", "quad_k3"]
NSTEPS = 20
EVERY = 6
RTOL = 1e-4
FP32_CALIB_RTOL = 5e-4   # FP32 gradients at 256^2 over 2000 steps against FP64 differences: measured 8.9e-5, 2.3e-4


def _say(msg):
    print("[gpu] adjoint " + msg, flush=True)
    sys.stdout.flush()


def _weights(case, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(case.phi.shape), rng.standard_normal(case.mu.shape)


def _loss(params, phi0, mu0, wphi, wmu, nsteps=NSTEPS):
    from microsim_amd.solver import GPSolver
    with GPSolver(params, _abi.GP_FP64) as s:
        s.upload(phi0, mu0)
        s.step(nsteps)
        phi, mu = s.download()
    return float(np.sum(wphi * phi) + np.sum(wmu * mu))


def _adjoint(params, phi0, mu0, wphi, wmu, nsteps=NSTEPS, every=EVERY, precision=_abi.GP_FP64):
    from microsim_amd.solver import GPSolver
    with GPSolver(params, precision) as s:
        s.upload(phi0, mu0)
        s.adjoint_forward(nsteps, every)
        phiT, muT = s.download()
        g0, m0, th = s.adjoint_backward(wphi, wmu)
    return g0, m0, th, phiT, muT


@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases())
def test_adjoint_dot_product_with_forward_mode(name):
    from microsim_amd import adjoint
    _say("dot product " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 21)
    g0, m0, gth, phiT, muT = _adjoint(case.params, case.phi, case.mu, wphi, wmu)
    rng = np.random.default_rng(22)
    inner = np.zeros(case.phi.shape[1:], dtype=bool)
    inner[3:-3, 3:-3] = True
    vphi = rng.standard_normal(case.phi.shape) * inner
    vmu = rng.standard_normal(case.mu.shape) * inner
    th0 = adjoint.theta_from_params(case.params)
    vth = rng.standard_normal(th0.shape) * np.abs(th0)
    hphi, hmu, tphi, tmu = model_jvp(case.params, case.phi, case.mu, vphi, vmu, vth, NSTEPS)
    assert np.array_equal(hphi, phiT) and np.array_equal(hmu, muT)
    terms_adj = np.concatenate([(g0 * vphi).ravel(), (m0 * vmu).ravel(), gth * vth])
    terms_fwd = np.concatenate([(wphi * tphi).ravel(), (wmu * tmu).ravel()])
    adj, fwd = terms_adj.sum(), terms_fwd.sum()
    scale = max(np.abs(terms_adj).sum(), np.abs(terms_fwd).sum())
    print("  <g, v> % .15e  <w, Jv> % .15e  (scale %.3e)" % (adj, fwd, scale))
    assert abs(adj - fwd) <= 1e-10 * scale
    # every parameter on its own: the tangent along theta_i alone is dL/dtheta_i
    names = adjoint.theta_names(case.N, case.NC + 1)
    zphi, zmu = np.zeros_like(vphi), np.zeros_like(vmu)
    for i in range(th0.size):
        e = np.zeros_like(th0)
        e[i] = 1.0
        _, _, tphi, tmu = model_jvp(case.params, case.phi, case.mu, zphi, zmu, e, NSTEPS)
        ref = np.sum(wphi * tphi) + np.sum(wmu * tmu)
        s = np.abs(wphi * tphi).sum() + np.abs(wmu * tmu).sum()
        assert abs(gth[i] - ref) <= 1e-10 * max(s, 1e-300), (names[i], gth[i], ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_theta_gradient_matches_central_differences(name):
    from microsim_amd import adjoint
    _say("theta fd " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 11)
    _, _, gth, _, _ = _adjoint(case.params, case.phi, case.mu, wphi, wmu)
    th0 = adjoint.theta_from_params(case.params)
    names = adjoint.theta_names(case.N, case.NC + 1)
    assert gth.shape == th0.shape
    fd = np.empty_like(th0)
    for i in range(th0.size):
        h = 1e-6 * abs(th0[i])
        assert h > 0, names[i]
        tp, tm = th0.copy(), th0.copy()
        tp[i] += h
        tm[i] -= h
        lp = _loss(adjoint.params_with_theta(case.params, tp), case.phi, case.mu, wphi, wmu)
        lm = _loss(adjoint.params_with_theta(case.params, tm), case.phi, case.mu, wphi, wmu)
        fd[i] = (lp - lm) / ((th0[i] + h) - (th0[i] - h))
    scale = np.abs(fd) * np.abs(th0)          # sensitivity of L to a relative change of each parameter
    err = np.abs(gth - fd) * np.abs(th0)
    for i in range(th0.size):
        print("  %-14s adj % .10e  fd % .10e" % (names[i], gth[i], fd[i]))
    bad = err > RTOL * np.maximum(scale, 1e-3 * scale.max())
    assert not bad.any(), [(names[i], gth[i], fd[i]) for i in np.nonzero(bad)[0]]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_state_gradient_matches_directional_difference(name):
    _say("state fd " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 12)
    g0, m0, _, _, _ = _adjoint(case.params, case.phi, case.mu, wphi, wmu)
    rng = np.random.default_rng(13)
    inner = np.zeros(case.phi.shape[1:], dtype=bool)
    inner[3:-3, 3:-3] = True
    # phi directions only where every phase is strictly inside (0, 1): the per-cell arithmetic branches on
    # exactly pure values, where a perturbation changes the branch
    iface = inner & np.all((case.phi > 1e-3) & (case.phi < 1 - 1e-3), axis=0)
    vphi = rng.standard_normal(case.phi.shape) * iface
    vmu = rng.standard_normal(case.mu.shape) * inner
    adj = float(np.sum(g0 * vphi) + np.sum(m0 * vmu))
    h = 1e-6
    lp = _loss(case.params, case.phi + h * vphi, case.mu + h * vmu, wphi, wmu)
    lm = _loss(case.params, case.phi - h * vphi, case.mu - h * vmu, wphi, wmu)
    fd = (lp - lm) / (2 * h)
    print("  cells phi %d mu %d: adj % .10e fd % .10e" % (iface.sum(), inner.sum(), adj, fd))
    assert abs(adj - fd) <= RTOL * max(abs(fd), 1e-8)
    # the halo entries of the initial state are overwritten from the interior by the upload: zero gradient
    halo = ~inner
    assert np.all(g0[:, halo] == 0) and np.all(m0[:, halo] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tern_aniso_per", "bin_aniso_neu"])
def test_checkpoint_interval_does_not_change_gradients(name):
    _say("checkpoint intervals " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 14)
    ref = _adjoint(case.params, case.phi, case.mu, wphi, wmu, every=NSTEPS)
    for every in (1, 0, 7):
        got = _adjoint(case.params, case.phi, case.mu, wphi, wmu, every=every)
        for a, b in zip(got[:3], ref[:3]):
            tol = 1e-12 * np.abs(b).max()
            assert np.abs(a - b).max() <= tol, every
        assert np.array_equal(got[3], ref[3]) and np.array_equal(got[4], ref[4])


@pytest.mark.gpu
def test_trajectory_function_forward_and_backward():
    import torch
    from microsim_amd import adjoint
    from microsim_amd.solver import GPSolver
    _say("torch trajectory tern_c1_mini")
    case, _ = load_golden("tern_c1_mini")
    th = torch.tensor(adjoint.theta_from_params(case.params), requires_grad=True)
    phi0 = torch.tensor(case.phi, requires_grad=True)
    mu0 = torch.tensor(case.mu, requires_grad=True)
    phi, mu = adjoint.trajectory(case.params, th, phi0, mu0, NSTEPS, every=5)
    with GPSolver(case.params, _abi.GP_FP64) as s:
        s.upload(case.phi, case.mu)
        s.step(NSTEPS)
        phi_ref, mu_ref = s.download()
    assert np.array_equal(phi.detach().numpy(), phi_ref) and np.array_equal(mu.detach().numpy(), mu_ref)
    wphi, wmu = _weights(case, 15)
    loss = (phi * torch.from_numpy(wphi)).sum() + (mu * torch.from_numpy(wmu)).sum()
    loss.backward()
    g0, m0, gth, _, _ = _adjoint(case.params, case.phi, case.mu, wphi, wmu, every=5)
    # a second reverse sweep accumulates in another atomic order: equal to rounding
    for a, b in ((th.grad.numpy(), gth), (phi0.grad.numpy(), g0), (mu0.grad.numpy(), m0)):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_theta_round_trip_cpu():
    from microsim_amd import adjoint
    for name in CASES:
        case, _ = load_golden(name)
        th = adjoint.theta_from_params(case.params)
        assert len(th) == len(adjoint.theta_names(case.N, case.NC + 1))
        p2 = adjoint.params_with_theta(case.params, th)
        assert bytes(p2) == bytes(case.params), name
        p3 = adjoint.params_with_theta(case.params, th * 1.5)
        assert np.array_equal(adjoint.theta_from_params(p3), th * 1.5)
    with pytest.raises(ValueError):
        adjoint.params_with_theta(case.params, th[:-1])


def _c0_case(mesh, steps, radius=60.0):
    """The binary anisotropic seed material (configs.config_c0) at mesh^2: tanh seed, liquid mu. The workload's
    own radius-10 seed lies below the critical radius (interface energy 1 over a driving force of 0.024: about 42
    cells) and melts within 500 steps, after which every gradient of a loss on the final phi is zero; a seed of
    radius 60 keeps a moving interface through the run."""
    import os
    import tempfile
    from microsim_amd import configs, host
    inp, _ = configs.config_c0(mesh, steps, steps)
    d = tempfile.mkdtemp(prefix="gp_c0_")
    configs.write_case(d, inp, [])
    phi_i, mu_i = configs.tanh_seed_state(mesh, mesh, radius, 2.0, [2.0 * 1.0 * 0.34])
    case = host.state_case(os.path.join(d, "Input.in"), phi_i, mu_i)
    return case.params, case.phi, case.mu


@pytest.mark.gpu
def test_calibration_gradient_binary_256():
    """Binary anisotropic seed (radius 60) at 256^2 over 2000 steps, loss = mean squared difference between the final phi and
    a target made by a forward run with dab 0.03, tau 0.25 and liquid diffusivity 1.2; gradients with respect to
    dab, tau_ab, Gamma, the liquid diffusivity, tau and the initial phi (checkpoints every 100 steps), against
    central differences with a relative step of 1e-6."""
    import torch
    from microsim_amd import adjoint
    from microsim_amd.solver import GPSolver
    _say("calibration gradient binary 256^2 2000 steps")
    steps = 2000
    params, phi0, mu0 = _c0_case(256, steps)
    names = adjoint.theta_names(2, 2)
    th0 = adjoint.theta_from_params(params)
    tt = th0.copy()
    tt[names.index("dab[0][1]")] = 0.03
    tt[names.index("tau_ab[0][1]")] = 0.25
    tt[names.index("tau")] = 0.25
    tt[names.index("D[1][0][0]")] = 1.2

    def final_phi(th, phi):
        with GPSolver(adjoint.params_with_theta(params, th), _abi.GP_FP64) as s:
            s.upload(phi, mu0)
            s.step(steps)
            return s.download()[0]

    target = final_phi(tt, phi0)[:, 3:-3, 3:-3]

    def loss_np(th, phi):
        return float(np.mean((final_phi(th, phi)[:, 3:-3, 3:-3] - target) ** 2))

    th = torch.tensor(th0, requires_grad=True)
    p0 = torch.tensor(phi0, requires_grad=True)
    phiT, _ = adjoint.trajectory(params, th, p0, torch.tensor(mu0), steps, every=100)
    loss = torch.mean((phiT[:, 3:-3, 3:-3] - torch.from_numpy(target)) ** 2)
    loss.backward()
    lval = float(loss.detach())
    assert lval > 0 and abs(lval - loss_np(th0, phi0)) <= 1e-12 * lval
    gth = th.grad.numpy()
    fd = np.empty_like(th0)
    for i in range(th0.size):
        h = 1e-6 * abs(th0[i])
        tp, tm = th0.copy(), th0.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (loss_np(tp, phi0) - loss_np(tm, phi0)) / (tp[i] - tm[i])
        print("  %-14s adj % .10e  fd % .10e" % (names[i], gth[i], fd[i]))
    scale = np.abs(fd * th0)
    err = np.abs(gth - fd) * np.abs(th0)
    assert np.all(err <= 1e-4 * np.maximum(scale, 1e-3 * scale.max())), list(zip(names, gth, fd))
    # the initial phi: directional difference over the interface cells. The loss is strongly curved along this
    # direction (central differences of the oracle's trajectory move from -2.0e-4 at h = 1e-4 to -3.0104e-4 at
    # 1e-6, -3.01129e-4 at 1e-7 and -3.011304e-4 at 1e-8), so the step here is 1e-8.
    rng = np.random.default_rng(31)
    iface = np.zeros(phi0.shape[1:], dtype=bool)
    iface[3:-3, 3:-3] = True
    iface &= np.all((phi0 > 1e-3) & (phi0 < 1 - 1e-3), axis=0)
    v = rng.standard_normal(phi0.shape) * iface
    adj = float(np.sum(p0.grad.numpy() * v))
    hh = 1e-8
    fdv = (loss_np(th0, phi0 + hh * v) - loss_np(th0, phi0 - hh * v)) / (2 * hh)
    print("  initial phi along v (%d cells): adj % .10e fd % .10e" % (iface.sum(), adj, fdv))
    assert abs(adj - fdv) <= 1e-4 * abs(fdv)
    # FP32: the same loss over the FP32 trajectory (FP32 checkpoints and recompute, FP64 Jacobians and gradient
    # accumulation), against the FP64 central differences above
    th32 = torch.tensor(th0, requires_grad=True)
    p32 = torch.tensor(phi0, requires_grad=True)
    phiT, _ = adjoint.trajectory(params, th32, p32, torch.tensor(mu0), steps, every=100, precision=_abi.GP_FP32)
    loss32 = torch.mean((phiT[:, 3:-3, 3:-3] - torch.from_numpy(target)) ** 2)
    loss32.backward()
    g32 = th32.grad.numpy()
    for i in range(th0.size):
        print("  fp32 %-14s adj % .10e  fd % .10e  rel %.2e" % (names[i], g32[i], fd[i],
                                                                abs(g32[i] - fd[i]) / abs(fd[i])))
    adj32 = float(np.sum(p32.grad.numpy() * v))
    print("  fp32 loss %.10e (fp64 %.10e); initial phi along v: adj % .10e rel %.2e"
          % (float(loss32.detach()), lval, adj32, abs(adj32 - fdv) / abs(fdv)))
    err32 = np.abs(g32 - fd) * np.abs(th0)
    assert np.all(err32 <= FP32_CALIB_RTOL * np.maximum(scale, 1e-3 * scale.max())), list(zip(names, g32, fd))
    assert abs(adj32 - fdv) <= FP32_CALIB_RTOL * abs(fdv)


@pytest.mark.gpu
@pytest.mark.parametrize("name,rtol", [("bin_tanh_restart", 1e-10), ("tern_c1_mini", 1e-10), ("quad_k3", 1e-10),
                                       ("bin_well_rot", 1e-6)])
def test_gradients_match_torch_autograd(name, rtol):
    """Element-wise agreement, relative to the largest entry. bin_well_rot's initial-phi gradients reach 2e7 (the
    1/|q| factors of the anisotropy at the cells of its sharp initial interface), and the rounding differences of
    two evaluation orders grow with them: measured 2.8e-7 there."""
    import torch
    from microsim_amd import adjoint
    from test_adjoint_cpu import torch_trajectory
    _say("torch autograd reference " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 41)
    g0, m0, gth, _, _ = _adjoint(case.params, case.phi, case.mu, wphi, wmu)
    th = torch.tensor(adjoint.theta_from_params(case.params), requires_grad=True)
    phi0 = torch.tensor(case.phi, requires_grad=True)
    mu0 = torch.tensor(case.mu, requires_grad=True)
    phi, mu = torch_trajectory(case, NSTEPS, th, phi0, mu0)
    ((phi * torch.from_numpy(wphi)).sum() + (mu * torch.from_numpy(wmu)).sum()).backward()
    for got, ref in ((g0, phi0.grad.numpy()), (m0, mu0.grad.numpy()), (gth, th.grad.numpy())):
        assert np.abs(got - ref).max() <= rtol * np.abs(ref).max()


def _fp32_states(params, phi0, mu0, nsteps):
    """The FP32 trajectory's states s_0 .. s_nsteps, widened to float64 (s_0 is the uploaded state after its
    rounding to FP32 and the boundary copies)."""
    from microsim_amd.solver import GPSolver
    out = []
    with GPSolver(params, _abi.GP_FP32) as s:
        s.upload(phi0, mu0)
        out.append(s.download())
        for _ in range(nsteps):
            s.step(1)
            out.append(s.download())
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["bin_tanh_restart", "bin_aniso_neu", "tern_aniso_per", "tern_c1_mini", "quad_k3",
                                  "bin_well_rot"])
def test_fp32_adjoint_is_the_fp64_transpose_at_fp32_states(name):
    """The FP32 reverse sweep against torch.autograd: one FP64 torch step per FP32 state, its vector-Jacobian
    product taken with the cotangent of the next state, chained from the final state back to the upload (the
    boundary copies of the upload transposed last). The FP32 rounding of the upload is treated as the identity.
    Also: recording for the adjoint leaves the FP32 trajectory bit for bit as plain stepping makes it."""
    import torch
    import torch_step
    from microsim_amd import adjoint
    _say("fp32 transpose " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 51)
    g0, m0, gth, phiT, muT = _adjoint(case.params, case.phi, case.mu, wphi, wmu, precision=_abi.GP_FP32)
    states = _fp32_states(case.params, case.phi, case.mu, NSTEPS)
    assert np.array_equal(states[-1][0], phiT) and np.array_equal(states[-1][1], muT)
    th = torch.tensor(adjoint.theta_from_params(case.params), requires_grad=True)
    gp, gm = torch.from_numpy(wphi), torch.from_numpy(wmu)
    for t in range(NSTEPS - 1, -1, -1):
        m = torch_step.TorchModel(case.params, th)
        x = torch.tensor(states[t][0], requires_grad=True)
        y = torch.tensor(states[t][1], requires_grad=True)
        phi, mu = torch_step.step(m, x, y)
        ((phi * gp).sum() + (mu * gm).sum()).backward()
        gp, gm = x.grad, y.grad
    m = torch_step.TorchModel(case.params)
    x = torch.tensor(case.phi, requires_grad=True)
    y = torch.tensor(case.mu, requires_grad=True)
    bphi = torch.stack(torch_step.apply_bc(m, list(x), 0))
    bmu = torch.stack(torch_step.apply_bc(m, list(y), 1))
    ((bphi * gp).sum() + (bmu * gm).sum()).backward()
    rtol = 1e-6 if name == "bin_well_rot" else 1e-10
    for got, ref in ((g0, x.grad.numpy()), (m0, y.grad.numpy()), (gth, th.grad.numpy())):
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("  max |fp32 adjoint - torch| / max |torch| = %.3e" % err)
        assert err <= rtol


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_fp32_adjoint_close_to_fp64(name):
    """The FP32 gradients against the FP64 adjoint's, relative to the largest entry of each (measured at most
    2.1e-6, on tern_aniso_per's initial phi): the FP32 trajectory stays within 2.1e-7 of the FP64 one over these
    steps, and the linearisations follow it. The initial-phi gradient is compared at the cells whose phases the
    FP32 rounding leaves on the same branch: every phase exactly 0 or 1, or farther than 1e-6 from both. A tanh
    tail cell such as (6.7e-10, 1 - 6.7e-10) becomes (6.7e-10, 1) in FP32, where the step's derivative is another
    one (bin_tanh_restart: -6.0 against 2.6e-9 in FP64); there the FP32 gradient is its own trajectory's exact
    one, as the transpose test above shows."""
    _say("fp32 vs fp64 gradients " + name)
    case, _ = load_golden(name)
    wphi, wmu = _weights(case, 61)
    r64 = _adjoint(case.params, case.phi, case.mu, wphi, wmu)
    r32 = _adjoint(case.params, case.phi, case.mu, wphi, wmu, precision=_abi.GP_FP32)
    p = case.phi
    same = np.all((p == 0) | (p == 1) | ((p > 1e-6) & (p < 1 - 1e-6)), axis=0)
    print("  initial phi compared on %d of %d cells" % (same.sum(), same.size))
    for label, a, b in zip(("phi0", "mu0", "theta"), r32[:3], r64[:3]):
        if label == "phi0":
            a, b = a[:, same], b[:, same]
        err = np.abs(a - b).max() / np.abs(b).max()
        print("  %-6s max rel diff %.3e" % (label, err))
        assert err <= 1e-5
