"""Differentiable trajectories of the solver loop (torch.autograd over libgp_hip's hand-written adjoint).

``trajectory(params, theta, phi0, mu0, nsteps)`` runs ``nsteps`` steps of the grand-potential loop
(``Grand_potential_Finite_difference_2D_serial/microsim_gp.c:102-106`` over ``solverloop/solverloop.h:21-71``:
the phase-field update, its boundary conditions, the chemical-potential update, its boundary conditions) on the
GPU in FP64 (or in the FP32 production mode) and returns the final ``(phi, mu)`` as float64 tensors. Gradients
flow to ``theta`` (the model parameters listed by ``theta_names``), to ``phi0`` and to ``mu0``. The backward pass
is the library's reverse sweep: checkpoints every ``every`` steps, recomputation of each segment, and the
transposed stencil of every step with its local Jacobians evaluated exactly. In FP32 the trajectory, its checkpoints and its recompute are FP32; each transposed step
evaluates the FP64 model's Jacobians at the FP32 states, and adjoints and parameter gradients accumulate in FP64.
There is no CPU path; a missing library or device raises.

theta holds the parameters as the solver loop reads them: the pair mobility coefficients ``dab``, the pair
relaxation times ``tau_ab`` and the interface energies ``Gamma`` (one value per pair a < b, used symmetrically),
the liquid's diagonal diffusivities, and ``tau``. Quantities the input stage derives from them before the loop
(for example tau_ab from the diffusivities) are not re-derived: theta is the loop's own parameter set.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np
import torch

from . import _abi
from .solver import GPSolver

P, NCM = _abi.P, _abi.NCM


def theta_names(nphases: int, ncomp: int) -> List[str]:
    N, NC = nphases, ncomp - 1
    names = []
    for a in range(N):
        for b in range(a + 1, N):
            names += ["dab[%d][%d]" % (a, b), "tau_ab[%d][%d]" % (a, b), "Gamma[%d][%d]" % (a, b)]
    names += ["D[%d][%d][%d]" % (N - 1, k, k) for k in range(NC)]
    names.append("tau")
    return names


def theta_from_params(params: _abi.GPParams) -> np.ndarray:
    """theta of a parameter set, in the order theta_names gives."""
    N, NC = int(params.nphases), int(params.ncomp) - 1
    th = []
    for a in range(N):
        for b in range(a + 1, N):
            th += [params.dab[a * P + b], params.tau_ab[a * P + b], params.gamma[a * P + b]]
    L = N - 1
    th += [params.diffusivity[(L * NCM + k) * NCM + k] for k in range(NC)]
    th.append(params.tau)
    return np.array(th, dtype=np.float64)


def params_with_theta(params: _abi.GPParams, theta) -> _abi.GPParams:
    """A copy of params with theta written in (pair values into both [a][b] and [b][a])."""
    N, NC = int(params.nphases), int(params.ncomp) - 1
    theta = np.asarray(theta, dtype=np.float64).ravel()
    if theta.size != len(theta_names(N, NC + 1)):
        raise ValueError("theta has %d entries, the model has %d" % (theta.size, len(theta_names(N, NC + 1))))
    p = _abi.GPParams.from_buffer_copy(params)
    i = 0
    for a in range(N):
        for b in range(a + 1, N):
            for arr in (p.dab, p.tau_ab, p.gamma):
                arr[a * P + b] = arr[b * P + a] = float(theta[i])
                i += 1
    L = N - 1
    for k in range(NC):
        p.diffusivity[(L * NCM + k) * NCM + k] = float(theta[i])
        i += 1
    p.tau = float(theta[i])
    return p


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, phi0, mu0, params, nsteps, every, precision):
        p = params_with_theta(params, theta.detach().cpu().numpy())
        s = GPSolver(p, precision)
        try:
            s.upload(phi0.detach().cpu().numpy(), mu0.detach().cpu().numpy())
            s.adjoint_forward(int(nsteps), int(every))
            phi, mu = s.download()
        except BaseException:
            s.close()
            raise
        ctx.solver = s
        ctx.devices = (theta.device, phi0.device, mu0.device)
        return torch.from_numpy(phi).to(phi0.device), torch.from_numpy(mu).to(mu0.device)

    @staticmethod
    def backward(ctx, gphi, gmu):
        s = ctx.solver
        if s is None:
            raise RuntimeError("trajectory: backward ran already; the recorded trajectory and its device buffers "
                               "are released after the first backward (pass retain_graph=False)")
        # the handle (checkpoints, segment states, adjoint buffers) is released as soon as the gradients are
        # out, so that a caller holding the outputs (e.g. the previous iteration's loss) keeps no device memory
        ctx.solver = None
        try:
            shp_phi = (s.N, s.rx, s.ry)
            shp_mu = (s.NC, s.rx, s.ry)
            gphi = np.zeros(shp_phi) if gphi is None else gphi.detach().to(torch.float64).cpu().numpy()
            gmu = np.zeros(shp_mu) if gmu is None else gmu.detach().to(torch.float64).cpu().numpy()
            g0, m0, th = s.adjoint_backward(gphi, gmu)
        finally:
            s.close()
        dt, dp, dm = ctx.devices
        return (torch.from_numpy(th).to(dt), torch.from_numpy(g0).to(dp), torch.from_numpy(m0).to(dm),
                None, None, None, None)


def trajectory(params: _abi.GPParams, theta: torch.Tensor, phi0: torch.Tensor, mu0: torch.Tensor, nsteps: int,
               every: int = 0, precision: int = _abi.GP_FP64) -> Tuple[torch.Tensor, torch.Tensor]:
    """Final (phi, mu) after nsteps steps from (phi0, mu0) with the parameters params overridden by theta.

    phi0 [N, rows_x, rows_y] and mu0 [NC, rows_x, rows_y] are halo-inclusive, as the host stage lays them out;
    their halo entries are overwritten from the interior by the boundary conditions and get zero gradient.
    every: checkpoint interval of the reverse sweep (0: about sqrt(nsteps)); device memory for the adjoint is
    (nsteps / every + every + 3) copies of the state (the checkpoints in the handle's precision).
    precision: _abi.GP_FP64 (default), _abi.GP_FP32 or _abi.GP_FP32C."""
    return _Trajectory.apply(theta, phi0, mu0, params, int(nsteps), int(every), int(precision))
