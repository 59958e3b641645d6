"""gp_invariants_noalloc (the entry point that allocates no device memory) against gp_get_stats.

Both evaluate the same kernels (k_invariants + k_fold_sum) over the same block grid, gp_get_stats on the handle's own
plane layout with the handle's partial-sum buffer, gp_invariants_noalloc on a caller's device arrays in the
reference's plain layout with the library's static scratch: the per-cell values and the summation order are the
same, so the results are equal bit for bit. Two calls in a row on one GPU, the second on a smaller grid, check
that the scratch the calls share carries nothing over from the first call. The grid bound is checked on the host
side (the call is rejected before any launch).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import gp_testlib  # noqa: E402
from microsim_amd import _abi  # noqa: E402
from microsim_amd.solver import GPSolver  # noqa: E402

pytestmark = pytest.mark.gpu


def _torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _noalloc(s, phi, mu, torch, stream=None):
    """gp_invariants_noalloc on device copies of the halo-inclusive host arrays; returns (solute[NC], energy)."""
    dt = torch.float64 if s.precision == _abi.GP_FP64 else torch.float32
    dphi = torch.from_numpy(np.ascontiguousarray(phi)).to(device="cuda", dtype=dt).contiguous()
    dmu = torch.from_numpy(np.ascontiguousarray(mu)).to(device="cuda", dtype=dt).contiguous()
    out = torch.full((s.NC + 1,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = s.lib.gp_invariants_noalloc(s.h, C.c_void_p(dphi.data_ptr()), C.c_void_p(dmu.data_ptr()),
                                     C.c_void_p(out.data_ptr()), stream)
    s._check(rc)
    s.synchronize()
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o[:s.NC], float(o[s.NC])


def _case(mesh, seeds, seed):
    from microsim_amd import configs, host
    import tempfile
    inp, fills, _ = configs.config_ternary(mesh, seeds, seed=seed, min_spacing=16.0)
    d = tempfile.mkdtemp(prefix="gp_noalloc_")
    configs.write_case(d, inp, fills)
    return host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))


@pytest.mark.parametrize("precision", [_abi.GP_FP64, _abi.GP_FP32], ids=["fp64", "fp32"])
def test_noalloc_invariants_equal_get_stats_on_two_calls(precision):
    torch = _torch_cuda()
    results = []
    # the first call on a larger grid, the second on a smaller one: the shared scratch keeps the first call's block
    # sums beyond the second call's range, which must not leak into its result
    for mesh, seeds, seed in ((160, 12, 7), (72, 4, 1234)):
        case = _case(mesh, seeds, seed)
        with GPSolver(case.params, precision) as s:
            s.upload(case.phi, case.mu)
            s.step(25)
            phi, mu = s.download()
            ref = s.stats()
            print("noalloc: %s mesh %d" % ("fp64" if precision == _abi.GP_FP64 else "fp32", mesh), flush=True)
            sol, en = _noalloc(s, phi, mu, torch)
            # on a stream of the caller's as well
            st = torch.cuda.Stream()
            sol2, en2 = _noalloc(s, phi, mu, torch, C.c_void_p(st.cuda_stream))
        assert np.all(np.isfinite(sol)) and np.isfinite(en)
        assert np.array_equal(sol, np.array(ref["solute"])), (sol, ref["solute"])
        assert en == ref["energy"], (en, ref["energy"])
        assert np.array_equal(sol2, sol) and en2 == en
        results.append((mesh, sol, en))
    # the two grids give different values (the test would not see a stale result otherwise)
    assert results[0][2] != results[1][2]


def test_noalloc_invariants_match_oracle_fp64():
    """The FP64 result against the oracle's invariants of the same state (FP64 restatement, 1e-12 relative)."""
    torch = _torch_cuda()
    case = _case(96, 6, 3)
    with GPSolver(case.params, _abi.GP_FP64) as s:
        s.upload(case.phi, case.mu)
        s.step(10)
        phi, mu = s.download()
        sol, en = _noalloc(s, phi, mu, torch)
    osol_c, oen_c = (C.c_double * 4)(), C.c_double()
    gp_testlib.oracle().oracle_invariants(C.byref(case.params), _abi.dptr(np.ascontiguousarray(phi)),
                                          _abi.dptr(np.ascontiguousarray(mu)), osol_c, C.byref(oen_c))
    osol, oen = np.array(osol_c[:case.NC]), oen_c.value
    assert np.allclose(sol, osol, rtol=1e-12, atol=0), (sol, osol)
    assert abs(en - oen) <= 1e-12 * abs(oen), (en, oen)


def test_noalloc_rejects_grids_beyond_the_scratch():
    """A grid whose block sums exceed the scratch is rejected before any launch (null state pointers are never
    dereferenced: the size check comes first)."""
    torch = _torch_cuda()
    from microsim_amd import configs, host
    import tempfile
    mesh = 12288   # binary (NC = 1): 2 * 192 * 3072 block sums > 2^20
    inp, _ = configs.config_c0(mesh, 1, 1)
    d = tempfile.mkdtemp(prefix="gp_noalloc_big_")
    configs.write_case(d, inp, [])
    params = host.read_input(os.path.join(d, "Input.in")).p
    nb = ((mesh + 63) // 64) * ((mesh + 3) // 4)
    assert int(params.ncomp) * nb > _abi.GP_NOALLOC_SCRATCH_DOUBLES   # (NC + 1) block sums per block
    with GPSolver(params, _abi.GP_FP32) as s:
        dummy = torch.zeros(1, device="cuda")
        rc = s.lib.gp_invariants_noalloc(s.h, C.c_void_p(dummy.data_ptr()), C.c_void_p(dummy.data_ptr()),
                                         C.c_void_p(dummy.data_ptr()), None)
        assert rc == _abi.GP_E_ARG
        assert b"GP_NOALLOC_SCRATCH_DOUBLES" in s.lib.gp_last_error(s.h)
