"""The AoS adapters gp_upload_fields / gp_download_fields through the C ABI.

The reference host keeps its state as an array of `struct fields { double *phia, *compi, *deltaphi, temperature; }`
(GPs/functions/global_vars.h:327-332), one heap block per array and cell. A maintainer who drops libgp_hip into the
reference driver passes that array as is. These tests build such an array (one separate ctypes buffer per cell and
field, the reference's layout), run steps through the adapters and compare with the SoA entry points and the
oracle.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from gp_testlib import load_golden, max_rel_err, oracle_step
from microsim_amd import _abi

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class AoS:
    """gridinfo as the reference allocates it: per cell, malloc'd phia[N], compi[NC], deltaphi[N]."""

    def __init__(self, n, N, NC, with_dphi=True):
        self.n, self.N, self.NC = n, N, NC
        self.cells = (_abi.GPCell * n)()
        self.keep = []
        for i in range(n):
            pa, ca = (C.c_double * N)(), (C.c_double * NC)()
            da = (C.c_double * N)() if with_dphi else None
            self.keep.append((pa, ca, da))
            self.cells[i].phia = C.cast(pa, C.POINTER(C.c_double))
            self.cells[i].compi = C.cast(ca, C.POINTER(C.c_double))
            self.cells[i].deltaphi = C.cast(da, C.POINTER(C.c_double)) if da is not None else None
            self.cells[i].temperature = 0.0

    def set(self, phi, mu):
        """phi [N, rx, ry], mu [NC, rx, ry] (SoA, y fastest) into the cells (index x*ry + y)"""
        pf, mf = phi.reshape(self.N, -1), mu.reshape(self.NC, -1)
        for i, (pa, ca, _) in enumerate(self.keep):
            for a in range(self.N):
                pa[a] = pf[a, i]
            for k in range(self.NC):
                ca[k] = mf[k, i]

    def get(self, shape_phi, shape_mu):
        phi, mu = np.empty(shape_phi), np.empty(shape_mu)
        dphi = np.empty(shape_phi)
        pf, mf, df = phi.reshape(self.N, -1), mu.reshape(self.NC, -1), dphi.reshape(self.N, -1)
        for i, (pa, ca, da) in enumerate(self.keep):
            pf[:, i] = pa[:self.N]
            mf[:, i] = ca[:self.NC]
            if da is not None:
                df[:, i] = da[:self.N]
        return phi, mu, dphi


@pytest.mark.parametrize("name", ["bin_aniso_neu", "tern_c1_mini"])
def test_aos_adapters_round_trip_and_step(name):
    from microsim_amd.solver import GPSolver
    _say("[gpu] aos adapters", name)
    case, _ = load_golden(name)
    N, NC = case.N, case.NC
    rx, ry = case.phi.shape[1:]
    n = rx * ry
    aos = AoS(n, N, NC)
    aos.set(case.phi, case.mu)
    steps = 12
    with GPSolver(case.params) as s:
        lib = s.lib
        # upload through the adapter, step, download through the SoA call and the adapter
        rc = lib.gp_upload_fields(s.h, aos.cells, n)
        assert rc == 0, lib.gp_last_error(s.h)
        s.step(steps)
        phi_s, mu_s, dphi_s = s.download(with_dphi=True)
        out = AoS(n, N, NC)
        rc = lib.gp_download_fields(s.h, out.cells, n)
        assert rc == 0, lib.gp_last_error(s.h)
        phi_a, mu_a, dphi_a = out.get(case.phi.shape, case.mu.shape)
        # a wrong cell count is rejected with an error, not read out of bounds
        assert lib.gp_upload_fields(s.h, aos.cells, n - 1) == _abi.GP_E_ARG
        assert lib.gp_download_fields(s.h, out.cells, n + 1) == _abi.GP_E_ARG
        # deltaphi may be absent (NULL) in the caller's cells: only phi and mu are written then
        nod = AoS(n, N, NC, with_dphi=False)
        assert lib.gp_download_fields(s.h, nod.cells, n) == 0
        phi_n, mu_n, _ = nod.get(case.phi.shape, case.mu.shape)
    assert np.array_equal(phi_a, phi_s) and np.array_equal(mu_a, mu_s) and np.array_equal(dphi_a, dphi_s)
    assert np.array_equal(phi_n, phi_s) and np.array_equal(mu_n, mu_s)
    ophi, omu = case.phi.copy(), case.mu.copy()
    odphi = np.zeros_like(ophi)
    oracle_step(case.params, ophi, omu, steps, dphi=odphi)
    ep, em = max_rel_err(phi_a, ophi), max_rel_err(mu_a, omu)
    sl = (slice(None), slice(2, -2), slice(2, -2))
    ed = max_rel_err(dphi_a[sl], odphi[sl])
    _say("   adapter state vs oracle: max rel err phi %.3g mu %.3g dphi %.3g" % (ep, em, ed))
    assert ep <= 1e-12 and em <= 1e-12 and ed <= 1e-12
