"""Python handle over libgp_hip's C ABI (include/gp_hip.h), through ctypes.

This is a thin host-side mirror of the serial driver's use of its solver loop
(``Grand_potential_Finite_difference_2D_serial/microsim_gp.c:97-106``): upload the state once, step, read
back at output times. Every call goes to the HIP library; there is no CPU fallback, and a missing library or
device raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _abi


class GPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("gp_hip error %d: %s" % (code, msg))
        self.code = code


def stats_dict(st: _abi.GPStats, N: int, NC: int) -> dict:
    return {
        "phi_max": list(st.phi_max[:N]), "phi_min": list(st.phi_min[:N]),
        "mu_max": list(st.mu_max[:NC]), "mu_min": list(st.mu_min[:NC]),
        "rel_change_phi": list(st.rel_change_phi[:N]), "rel_change_mu": list(st.rel_change_mu[:NC]),
        "solute": list(st.solute[:NC]), "energy": st.energy, "step": st.step, "nonfinite": st.nonfinite,
    }


class GPSolver:
    """One libgp_hip handle (one HIP stream) for a fixed parameter set."""

    def __init__(self, params: _abi.GPParams, precision: int = _abi.GP_FP64):
        self.lib = _abi.hip_lib()
        self.params = params
        self.precision = precision
        self.N = int(params.nphases)
        self.NC = int(params.ncomp) - 1
        self.rx = int(params.mesh_x) + 6
        self.ry = int(params.mesh_y) + 6
        h = C.c_void_p()
        rc = self.lib.gp_create(C.byref(params), precision, C.byref(h))
        if rc != 0:
            raise GPError(rc, self.lib.gp_last_error(None).decode())
        self.h = h

    def _check(self, rc: int):
        if rc != 0:
            raise GPError(rc, self.lib.gp_last_error(self.h).decode())

    def upload(self, phi: np.ndarray, mu: np.ndarray) -> None:
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        assert phi.shape == (self.N, self.rx, self.ry) and mu.shape == (self.NC, self.rx, self.ry)
        self._check(self.lib.gp_upload(self.h, _abi.dptr(phi), _abi.dptr(mu)))

    def step(self, nsteps: int, stats: bool = False) -> Optional[dict]:
        if stats:
            st = _abi.GPStats()
            self._check(self.lib.gp_step(self.h, int(nsteps), C.byref(st)))
            return stats_dict(st, self.N, self.NC)
        self._check(self.lib.gp_step(self.h, int(nsteps), None))
        return None

    def smooth(self, nsteps: int) -> None:
        self._check(self.lib.gp_smooth(self.h, int(nsteps)))

    def stats(self) -> dict:
        st = _abi.GPStats()
        self._check(self.lib.gp_get_stats(self.h, C.byref(st)))
        return stats_dict(st, self.N, self.NC)

    def synchronize(self) -> None:
        self._check(self.lib.gp_synchronize(self.h))

    def download(self, with_dphi: bool = False) -> Tuple[np.ndarray, ...]:
        phi = np.empty((self.N, self.rx, self.ry))
        mu = np.empty((self.NC, self.rx, self.ry))
        dphi = np.empty((self.N, self.rx, self.ry)) if with_dphi else None
        self._check(self.lib.gp_download(self.h, _abi.dptr(phi), _abi.dptr(mu),
                                         _abi.dptr(dphi) if with_dphi else None))
        return (phi, mu, dphi) if with_dphi else (phi, mu)

    def frame(self) -> Tuple[int, np.ndarray]:
        """(rows the moving frame has shifted since upload, temperature of every allocated y row)."""
        off = C.c_int64()
        T = np.empty(self.ry)
        self._check(self.lib.gp_get_frame(self.h, C.byref(off), _abi.dptr(T)))
        return int(off.value), T

    def profile_enable(self, on=True) -> None:
        """True / _abi.GP_PROFILE_LAUNCHES: events around every kernel launch; _abi.GP_PROFILE_CALLS: around every
        whole step / smooth call only (kind "call"); False: off."""
        mode = on if not isinstance(on, bool) else (_abi.GP_PROFILE_LAUNCHES if on else _abi.GP_PROFILE_OFF)
        self._check(self.lib.gp_profile_enable(self.h, int(mode)))

    def profile_read(self) -> dict:
        """{kernel name: (mean ms per launch, launches)} since the previous read (HIP events on the handle's
        own stream)."""
        avg = (C.c_double * _abi.GP_NKERNELS)()
        cnt = (C.c_int64 * _abi.GP_NKERNELS)()
        self._check(self.lib.gp_profile_read(self.h, avg, cnt))
        return {self.lib.gp_kernel_name(k).decode(): (avg[k], cnt[k]) for k in range(_abi.GP_NKERNELS) if cnt[k]}

    # ---------------------------------------------------------------- adjoint
    def theta(self) -> np.ndarray:
        """Differentiable parameters (gp_hip.h GP_THETA layout): per phase pair a < b dab, tau_ab, Gamma; the
        liquid's diagonal diffusivities; tau."""
        n = self.lib.gp_theta_count(self.h)
        th = np.empty(n)
        self._check(self.lib.gp_theta_values(self.h, _abi.dptr(th)))
        return th

    def adjoint_forward(self, nsteps: int, every: int = 0) -> None:
        """nsteps steps from the current state, keeping a checkpoint every `every` steps (0: about sqrt)."""
        self._check(self.lib.gp_adjoint_forward(self.h, int(nsteps), int(every)))

    def adjoint_backward(self, gphi: np.ndarray, gmu: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Gradients of a scalar loss with respect to the state uploaded before adjoint_forward and to theta,
        given its gradients with respect to the final state (halo-inclusive arrays). The handle holds the initial
        state afterwards."""
        gphi = np.ascontiguousarray(gphi, dtype=np.float64)
        gmu = np.ascontiguousarray(gmu, dtype=np.float64)
        assert gphi.shape == (self.N, self.rx, self.ry) and gmu.shape == (self.NC, self.rx, self.ry)
        g0 = np.empty_like(gphi)
        m0 = np.empty_like(gmu)
        th = np.empty(self.lib.gp_theta_count(self.h))
        self._check(self.lib.gp_adjoint_backward(self.h, _abi.dptr(gphi), _abi.dptr(gmu), _abi.dptr(g0),
                                                 _abi.dptr(m0), _abi.dptr(th)))
        return g0, m0, th

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
