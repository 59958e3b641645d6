"""The drop-in driver microsim_gp (microsim_amd/host/microsim_gp.c over libgp_hip) against the reference's own run
of the same deck: same snapshots DATA/<out>_<t>.vtk (FP64, bit-exact) and the same progress block on stdout
(Grand_potential_Finite_difference_2D_serial/microsim_gp.c:41-151).

The progress block's Max/Min are cumulative over every step in both programs and must print identically. The
reference's Relative_Change sums also visit the halo cells its sweep updates (with stale neighbours there), which
the driver leaves out by design: on these decks that moves them by up to 5 %, so they are only compared to 10 %
here; the interior sums themselves are pinned against the oracle in test_gpu_parity.py.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from gp_testlib import GOLDEN, load_golden
from microsim_amd import _abi, configs

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(_abi.LIBDIR, "microsim_gp")
GATE = 1e-12


def _progress(text):
    """[(time, [(name, max, min, relchange), ...])] of the progress blocks in a stdout text."""
    blocks = []
    for m in re.finditer(r"Time=(\S+)\n((?:\s*\S+, Max = \S+, Min = \S+, Relative_Change=\S+\n)+)", text):
        rows = re.findall(r"\s*(\S+), Max = (\S+), Min = (\S+), Relative_Change=(\S+)", m.group(2))
        blocks.append((float(m.group(1)), [(n, float(a), float(b), float(c)) for n, a, b, c in rows]))
    return blocks


@pytest.mark.parametrize("name", ["tern_c1_mini", "bin_tanh_restart", "tern_aniso_per"])
def test_driver_reproduces_reference_run(name):
    print("[gpu] driver", name, flush=True)
    sys.stdout.flush()
    if not os.path.exists(DRIVER):
        pytest.fail("microsim_gp is not built: run __graft_entry__.build()")
    case, ref = load_golden(name)
    d = tempfile.mkdtemp(prefix="gp_drv_")
    try:
        for f in ("Input.in", "Filling.in"):
            shutil.copy(os.path.join(GOLDEN, name, f), d)
        text = open(os.path.join(d, "Input.in")).read()
        nt = int(re.search(r"NTIMESTEPS\s*=\s*(\d+)", text).group(1))
        saves = [s for s in ref["steps"] if s > 0]
        savet = int(np.gcd.reduce(saves))
        text = re.sub(r"SAVET\s*=\s*\d+;", "SAVET = %d;" % savet, text)
        open(os.path.join(d, "Input.in"), "w").write(text)
        if case.inp.restart:
            os.makedirs(os.path.join(d, "DATA"))
            r = np.load(os.path.join(GOLDEN, name, "restart0.npz"))
            configs.write_vtk(os.path.join(d, "DATA", "out_0.vtk"), r["phi"], r["mu"],
                              ["p%d" % i for i in range(case.N)], ["c%d" % i for i in range(case.NC + 1)],
                              binary=not case.inp.ascii)
        r = subprocess.run([DRIVER, "Input.in", "Filling.in", "out"], cwd=d, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        for i, s in enumerate(ref["steps"]):
            if s % savet:
                continue
            phi, mu = configs.read_vtk(os.path.join(d, "DATA", "out_%d.vtk" % s), case.N, case.NC + 1,
                                       int(case.params.mesh_x), int(case.params.mesh_y))
            assert np.array_equal(phi, ref["phi"][i]) and np.array_equal(mu, ref["mu"][i]), "step %d" % s
        lines = open(os.path.join(d, "DATA", "shift.dat")).read().split("\n")
        assert lines[0] == "%d 0" % savet and len([x for x in lines if x]) == nt // savet
        got = _progress(r.stdout)
        want = _progress(open(os.path.join(GOLDEN, name, "stdout.txt")).read())
        assert len(got) == len(want) > 0
        for (tg, rg), (tw, rw) in zip(got, want):
            assert tg == tw and [x[0] for x in rg] == [x[0] for x in rw]
            for (n, a, b, c), (_, a2, b2, c2) in zip(rg, rw):
                assert (a, b) == (a2, b2), "%s: max/min %g %g vs reference %g %g" % (n, a, b, a2, b2)
                assert c == pytest.approx(c2, rel=0.1), "%s: relative change %g vs %g" % (n, c, c2)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _run_driver(d, env=None):
    r = subprocess.run([DRIVER, "Input.in", "Filling.in", "out"], cwd=d, capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    return r


def _set_keys(path, **keys):
    text = open(path).read()
    for k, v in keys.items():
        text, n = re.subn(r"\b%s\s*=\s*[^;]*;" % k, "%s = %s;" % (k, v), text)
        assert n == 1, k
    open(path, "w").write(text)


@pytest.mark.parametrize("name,precision", [("tern_aniso_per", "fp64"), ("quad_k3", "fp64"), ("bin_aniso_neu", "fp64"),
                                            ("tern_aniso_per", "fp32")])
def test_driver_restart_equals_uninterrupted_run(name, precision):
    """Run 2n steps in one go; then run n steps, restart from DATA/out_<n>.vtk (RESTART = 1, STARTTIME = n,
    the reader of Grand_potential_Finite_difference_2D_serial/microsim_gp.c:48-56) and run n more: the snapshot at
    2n is the same file byte for byte. Binary VTK holds the FP64 state exactly, so the restarted run starts from
    the state the uninterrupted one holds at n; in FP32 mode the state is FP32 and widens to FP64 exactly in the
    snapshot, and narrows back exactly on upload."""
    print("[gpu] driver restart", name, precision, flush=True)
    sys.stdout.flush()
    if not os.path.exists(DRIVER):
        pytest.fail("microsim_gp is not built: run __graft_entry__.build()")
    env = dict(os.environ, GP_PRECISION=precision)
    n = 20
    d1, d2 = tempfile.mkdtemp(prefix="gp_rs1_"), tempfile.mkdtemp(prefix="gp_rs2_")
    try:
        for d in (d1, d2):
            for f in ("Input.in", "Filling.in"):
                shutil.copy(os.path.join(GOLDEN, name, f), d)
        _set_keys(os.path.join(d1, "Input.in"), NTIMESTEPS=2 * n, SAVET=n, TRACK_PROGRESS=n, RESTART=0,
                  STARTTIME=0)
        _run_driver(d1, env)
        whole = open(os.path.join(d1, "DATA", "out_%d.vtk" % (2 * n)), "rb").read()
        os.makedirs(os.path.join(d2, "DATA"))
        shutil.copy(os.path.join(d1, "DATA", "out_%d.vtk" % n), os.path.join(d2, "DATA"))
        _set_keys(os.path.join(d2, "Input.in"), NTIMESTEPS=n, SAVET=n, TRACK_PROGRESS=n, RESTART=1, STARTTIME=n)
        _run_driver(d2, env)
        assert not os.path.exists(os.path.join(d2, "DATA", "out_0.vtk"))
        resumed = open(os.path.join(d2, "DATA", "out_%d.vtk" % (2 * n)), "rb").read()
        assert resumed == whole
        assert open(os.path.join(d2, "DATA", "shift.dat")).read() == "%d 0\n" % (2 * n)
    finally:
        shutil.rmtree(d1, ignore_errors=True)
        shutil.rmtree(d2, ignore_errors=True)


def test_driver_hdf5_snapshots_and_restart():
    """WRITEHDF5 = 1 (the MPI driver's key): the snapshots are DATA/out_<t>.h5 with their DATA/out_<t>.xmf, hold the
    reference run's state bit for bit, and a restart from the HDF5 snapshot at n continues to the same file at 2n."""
    from test_h5_cpu import H5
    print("[gpu] driver hdf5", flush=True)
    sys.stdout.flush()
    if not os.path.exists(DRIVER):
        pytest.fail("microsim_gp is not built: run __graft_entry__.build()")
    name = "tern_aniso_per"
    case, ref = load_golden(name)
    d1, d2 = tempfile.mkdtemp(prefix="gp_h5d1_"), tempfile.mkdtemp(prefix="gp_h5d2_")
    try:
        for d in (d1, d2):
            for f in ("Input.in", "Filling.in"):
                shutil.copy(os.path.join(GOLDEN, name, f), d)
            with open(os.path.join(d, "Input.in"), "a") as fh:
                fh.write("WRITEHDF5 = 1;\n")
        saves = [s for s in ref["steps"] if s > 0]
        savet = int(np.gcd.reduce(saves))
        n = max(saves)
        _set_keys(os.path.join(d1, "Input.in"), NTIMESTEPS=n, SAVET=savet, TRACK_PROGRESS=n)
        _run_driver(d1)
        names = [case.inp.phases[a].value.decode() for a in range(case.N)]
        comps = [case.inp.components[k].value.decode() for k in range(case.NC)]
        for i, s in enumerate(ref["steps"]):
            f = H5(os.path.join(d1, "DATA", "out_%d.h5" % s))
            data = {nm: f.dataset(oh)[1] for nm, oh in f.links()}
            for a, nm in enumerate(names):
                assert np.array_equal(data[nm][3:-3, 3:-3], ref["phi"][i][a]), "step %d %s" % (s, nm)
            for k, c in enumerate(comps):
                assert np.array_equal(data["Mu_" + c][3:-3, 3:-3], ref["mu"][i][k]), "step %d Mu_%s" % (s, c)
            assert os.path.exists(os.path.join(d1, "DATA", "out_%d.xmf" % s))
        # restart from the HDF5 snapshot at the first save
        m = saves[0]
        os.makedirs(os.path.join(d2, "DATA"))
        shutil.copy(os.path.join(d1, "DATA", "out_%d.h5" % m), os.path.join(d2, "DATA"))
        _set_keys(os.path.join(d2, "Input.in"), NTIMESTEPS=n - m, SAVET=savet, TRACK_PROGRESS=n, RESTART=1,
                  STARTTIME=m)
        _run_driver(d2)
        a = open(os.path.join(d1, "DATA", "out_%d.h5" % n), "rb").read()
        b = open(os.path.join(d2, "DATA", "out_%d.h5" % n), "rb").read()
        assert a == b
    finally:
        shutil.rmtree(d1, ignore_errors=True)
        shutil.rmtree(d2, ignore_errors=True)


def test_driver_mpi_form_per_rank_snapshots_restart_and_reconstruct():
    """The MPI form `microsim_gp Input.in Filling.in out 2 3` writes the MPI driver's per-rank files
    DATA/Processor_<r>/out_<t>.vtk (Grand_potential_Finite_difference_2D_MPI/microsim_gp.c:113-118, 248-285,
    368-376): they hold the reference run's state at every saved step, the files at t = 0 hold the state
    after the NSMOOTH - 1 smoothing steps (the MPI main writes t = 0 again after smoothing), a restart from the
    per-rank files at m continues to the same files at n, and the reconstruct tool joins them into DATA/out_<t>.vtk
    with the same values. State comparisons use the FP64 gate of the parity tests (relative 1e-12)."""
    from gp_testlib import max_rel_err, oracle_smooth
    from test_ranks_cpu import RECONSTRUCT, _parse_joined
    from microsim_amd import host
    print("[gpu] driver mpi form", flush=True)
    sys.stdout.flush()
    if not os.path.exists(DRIVER) or not os.path.exists(RECONSTRUCT):
        pytest.fail("microsim_gp / reconstruct are not built: run __graft_entry__.build()")
    name = "bin_smooth"
    case, ref = load_golden(name)
    inp = case.inp
    saves = [s for s in ref["steps"] if s > 0]
    n, m = max(saves), saves[0]
    d1, d2 = tempfile.mkdtemp(prefix="gp_mpi1_"), tempfile.mkdtemp(prefix="gp_mpi2_")
    try:
        for d in (d1, d2):
            for f in ("Input.in", "Filling.in"):
                shutil.copy(os.path.join(GOLDEN, name, f), d)
        _set_keys(os.path.join(d1, "Input.in"), NTIMESTEPS=n, SAVET=1, TRACK_PROGRESS=n)
        r = subprocess.run([DRIVER, "Input.in", "Filling.in", "out", "2", "3"], cwd=d1, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        data = os.path.join(d1, "DATA")
        assert not os.path.exists(os.path.join(data, "out_0.vtk"))
        for i, s in enumerate(ref["steps"]):
            if s == 0:
                continue
            phi, mu = host.read_rank_vtk(inp, data, "out", s, 6)
            assert max_rel_err(phi[:, 3:-3, 3:-3], ref["phi"][i]) <= GATE, "step %d" % s
            assert max_rel_err(mu[:, 3:-3, 3:-3], ref["mu"][i]) <= GATE, "step %d" % s
        phi0, mu0 = case.phi.copy(), case.mu.copy()
        oracle_smooth(case.params, phi0, mu0, int(inp.nsmooth) - 1)
        phi, mu = host.read_rank_vtk(inp, data, "out", 0, 6)
        assert max_rel_err(phi[:, 3:-3, 3:-3], phi0[:, 3:-3, 3:-3]) <= GATE
        assert max_rel_err(mu[:, 3:-3, 3:-3], mu0[:, 3:-3, 3:-3]) <= GATE
        assert max_rel_err(phi[:, 3:-3, 3:-3], case.phi[:, 3:-3, 3:-3]) > 1e-6, "t = 0 holds the smoothed state"
        # restart from the per-rank files at m
        for rk in range(6):
            os.makedirs(os.path.join(d2, "DATA", "Processor_%d" % rk))
            shutil.copy(os.path.join(data, "Processor_%d" % rk, "out_%d.vtk" % m),
                        os.path.join(d2, "DATA", "Processor_%d" % rk))
        _set_keys(os.path.join(d2, "Input.in"), NTIMESTEPS=n - m, SAVET=1, TRACK_PROGRESS=n, RESTART=1,
                  STARTTIME=m, NSMOOTH=1)
        r = subprocess.run([DRIVER, "Input.in", "Filling.in", "out", "2", "3"], cwd=d2, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        for rk in range(6):
            a = open(os.path.join(data, "Processor_%d" % rk, "out_%d.vtk" % n), "rb").read()
            b = open(os.path.join(d2, "DATA", "Processor_%d" % rk, "out_%d.vtk" % n), "rb").read()
            assert a == b, "rank %d" % rk
        # reconstruct every saved step
        r = subprocess.run([RECONSTRUCT, "Input.in", "out", "6", "1", str(n)], cwd=d1, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        names = [inp.phases[a].value.decode() for a in range(case.N)]
        comps = [inp.components[k].value.decode() for k in range(case.NC)]
        for i, s in enumerate(ref["steps"]):
            if s == 0:
                continue
            got = _parse_joined(os.path.join(data, "out_%d.vtk" % s), inp, bool(inp.ascii))
            phi, mu = host.read_rank_vtk(inp, data, "out", s, 6)
            for a, nm in enumerate(names):
                assert np.array_equal(got[nm], phi[a, 3:-3, 3:-3]), "step %d %s" % (s, nm)
                assert max_rel_err(got[nm], ref["phi"][i][a]) <= GATE, "step %d %s" % (s, nm)
            for k, c in enumerate(comps):
                assert np.array_equal(got["Mu_" + c], mu[k, 3:-3, 3:-3]), "step %d Mu_%s" % (s, c)
                assert max_rel_err(got["Mu_" + c], ref["mu"][i][k]) <= GATE, "step %d Mu_%s" % (s, c)
    finally:
        shutil.rmtree(d1, ignore_errors=True)
        shutil.rmtree(d2, ignore_errors=True)
