"""The kernels' per-cell arithmetic (microsim_amd/csrc/gp_model.h) compiled for the host with g++ and stepped in
the kernels' decomposition, against the FP64 oracle. Catches arithmetic mistakes in the kernel source on the
CPU, before any GPU run; the GPU tests then only have to show the device reproduces its own host build."""
import ctypes as C

import numpy as np
import pytest

from gp_testlib import golden_cases, load_golden, model_lib, oracle_smooth, oracle_step
from microsim_amd import _abi

@pytest.mark.parametrize("pow2", [0, 1])
@pytest.mark.parametrize("name", golden_cases())
def test_kernel_math_matches_oracle(name, pow2):
    case, _ = load_golden(name)
    p = case.params
    a_phi, a_mu = case.phi.copy(), case.mu.copy()
    b_phi, b_mu = case.phi.copy(), case.mu.copy()
    lib = model_lib()
    if case.inp.nsmooth > 1:
        rc = lib.model_step(C.byref(p), _abi.dptr(a_phi), _abi.dptr(a_mu), None, case.inp.nsmooth - 1, 1, pow2)
        if rc == -3:
            pytest.skip("spacings are not powers of two")
        assert rc == 0
        oracle_smooth(p, b_phi, b_mu, case.inp.nsmooth - 1)
    a_d = np.zeros_like(a_phi)
    rc = lib.model_step(C.byref(p), _abi.dptr(a_phi), _abi.dptr(a_mu), _abi.dptr(a_d), 12, 0, pow2)
    if rc == -3:
        pytest.skip("spacings are not powers of two")
    assert rc == 0
    b_d = np.zeros_like(b_phi)
    oracle_step(p, b_phi, b_mu, 12, dphi=b_d)
    assert np.array_equal(a_phi, b_phi) and np.array_equal(a_mu, b_mu)
    assert np.array_equal(a_d[:, 2:-2, 2:-2], b_d[:, 2:-2, 2:-2])
