"""Python face of the host library: Input.in / Filling.in -> (gp_input, initial state) as numpy arrays.

Mirrors what the serial driver does before its time loop (``Grand_potential_Finite_difference_2D_serial/
microsim_gp.c:42-84``): read the parameters, build the grid, fill the domain (or read the restart snapshot)
and apply the boundary conditions.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _abi


class InputError(RuntimeError):
    pass


@dataclass
class Case:
    inp: _abi.GPInput
    phi: np.ndarray   # [N, rows_x, rows_y]
    mu: np.ndarray    # [NC, rows_x, rows_y]

    @property
    def params(self) -> _abi.GPParams:
        return self.inp.p

    @property
    def N(self) -> int:
        return self.inp.p.nphases

    @property
    def NC(self) -> int:
        return self.inp.p.ncomp - 1

    @property
    def rows(self):
        return int(self.inp.p.mesh_x) + 6, int(self.inp.p.mesh_y) + 6

    def interior(self, arr: np.ndarray) -> np.ndarray:
        return arr[..., 3:-3, 3:-3]


def read_input(input_path: str, quiet: bool = True, validate: bool = True) -> _abi.GPInput:
    lib = _abi.host_lib()
    inp = _abi.GPInput()
    err = C.create_string_buffer(512)
    rc = lib.gp_read_input(input_path.encode(), C.byref(inp), int(quiet), err, 512)
    if rc != 0:
        raise InputError(err.value.decode())
    if validate:
        rc = lib.gp_validate_input(C.byref(inp), err, 512)
        if rc != 0:
            raise InputError(err.value.decode())
    return inp


def empty_state(inp: _abi.GPInput):
    rx, ry = int(inp.p.mesh_x) + 6, int(inp.p.mesh_y) + 6
    N, NC = inp.p.nphases, inp.p.ncomp - 1
    return np.zeros((N, rx, ry)), np.zeros((max(NC, 0), rx, ry))


def load_case(input_path: str, filling_path: str, restart_vtk: str = None, quiet: bool = True) -> Case:
    """Parameters plus initial state: fill_domain, or the restart snapshot when given; halos applied."""
    lib = _abi.host_lib()
    inp = read_input(input_path, quiet=quiet)
    phi, mu = empty_state(inp)
    err = C.create_string_buffer(512)
    if restart_vtk is None:
        rc = lib.gp_fill_domain(C.byref(inp), filling_path.encode(), _abi.dptr(phi), _abi.dptr(mu), None,
                                int(quiet), err, 512)
        if rc != 0:
            raise InputError(err.value.decode())
    else:
        rc = lib.gp_read_vtk(C.byref(inp), restart_vtk.encode(), _abi.dptr(phi), _abi.dptr(mu))
        if rc != 0:
            raise InputError("cannot read %s" % restart_vtk)
    lib.gp_apply_bc_host(C.byref(inp.p), _abi.dptr(phi), _abi.dptr(mu))
    return Case(inp, phi, mu)


def state_case(input_path: str, phi_interior: np.ndarray, mu_interior: np.ndarray) -> Case:
    """Parameters from Input.in with a given interior state (what a restart snapshot supplies), halos applied."""
    inp = read_input(input_path)
    phi, mu = empty_state(inp)
    phi[:, 3:-3, 3:-3] = phi_interior
    mu[:, 3:-3, 3:-3] = mu_interior
    _abi.host_lib().gp_apply_bc_host(C.byref(inp.p), _abi.dptr(phi), _abi.dptr(mu))
    return Case(inp, phi, mu)


def temperature_rows(params: _abi.GPParams, t: int = 0, shift_offset: int = 0) -> np.ndarray:
    """Temperature of every allocated row as the serial driver sets it before the loop (t = 0) or after step t,
    with the moving frame shifted by shift_offset rows (T everywhere for an isothermal run)."""
    T = np.empty(int(params.mesh_y) + 6)
    _abi.host_lib().gp_temperature_rows(C.byref(params), int(t), int(shift_offset), _abi.dptr(T))
    return T


def write_vtk(case_inp: _abi.GPInput, path: str, phi: np.ndarray, mu: np.ndarray, T_rows: np.ndarray = None) -> None:
    """The snapshot file; T_rows (per allocated row) is required when ISOTHERMAL = 0."""
    lib = _abi.host_lib()
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    if case_inp.p.tempgrady and T_rows is None:
        raise ValueError("a non-isothermal snapshot needs the row temperatures (T_rows)")
    T = None if T_rows is None else np.ascontiguousarray(T_rows, dtype=np.float64)
    rc = lib.gp_write_vtk(C.byref(case_inp), path.encode(), _abi.dptr(phi), _abi.dptr(mu),
                          None if T is None else _abi.dptr(T))
    if rc != 0:
        raise IOError("cannot write %s" % path)


def composition(params: _abi.GPParams, phi: np.ndarray, mu: np.ndarray) -> np.ndarray:
    """c_k = sum_b c_b^k(mu) h_b(phi) for every cell (the Composition_<k> output) as [NC, ...] float64."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    n = int(np.prod(phi.shape[1:]))
    c = np.empty((mu.shape[0],) + phi.shape[1:])
    _abi.host_lib().gp_composition_grid(C.byref(params), _abi.dptr(phi), _abi.dptr(mu), _abi.dptr(c), n)
    return c


def rank_block(params: _abi.GPParams, workers_x: int, workers_y: int, rank: int):
    """((rows_X, rows_Y), (offset_X, offset_Y)) of one rank in the MPI driver's decomposition of the halo-inclusive
    grid."""
    rows, off = (C.c_int64 * 2)(), (C.c_int64 * 2)()
    if _abi.host_lib().gp_rank_block(C.byref(params), workers_x, workers_y, rank, rows, off) != 0:
        raise ValueError("a %d x %d decomposition leaves a worker without rows" % (workers_x, workers_y))
    return (rows[0], rows[1]), (off[0], off[1])


def write_rank_vtk(case_inp: _abi.GPInput, data_dir: str, out: str, t: int, workers_x: int, workers_y: int,
                   phi: np.ndarray, mu: np.ndarray, T_rows: np.ndarray = None) -> None:
    """The MPI driver's per-rank snapshot files <data_dir>/Processor_<r>/<out>_<t>.vtk."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    if case_inp.p.tempgrady and T_rows is None:
        raise ValueError("a non-isothermal snapshot needs the row temperatures (T_rows)")
    T = None if T_rows is None else np.ascontiguousarray(T_rows, dtype=np.float64)
    rank_block(case_inp.p, workers_x, workers_y, 0)
    rc = _abi.host_lib().gp_write_rank_vtk(C.byref(case_inp), data_dir.encode(), out.encode(), t, workers_x,
                                           workers_y, _abi.dptr(phi), _abi.dptr(mu),
                                           None if T is None else _abi.dptr(T))
    if rc != 0:
        raise IOError("cannot write the per-rank files of %s_%d in %s (%d)" % (out, t, data_dir, rc))


def read_rank_vtk(case_inp: _abi.GPInput, data_dir: str, out: str, t: int, nranks: int):
    """(phi, mu), halo-inclusive, from nranks per-rank snapshot files."""
    phi, mu = empty_state(case_inp)
    rc = _abi.host_lib().gp_read_rank_vtk(C.byref(case_inp), data_dir.encode(), out.encode(), t, nranks,
                                          _abi.dptr(phi), _abi.dptr(mu))
    if rc != 0:
        raise IOError("cannot read the per-rank files of %s_%d in %s (%d)" % (out, t, data_dir, rc))
    return phi, mu


def reconstruct(case_inp: _abi.GPInput, data_dir: str, out: str, nranks: int, t: int) -> str:
    """Joins the per-rank files of time t into <data_dir>/<out>_<t>.vtk (the reconstruct tool); returns the path."""
    rc = _abi.host_lib().gp_reconstruct(C.byref(case_inp), data_dir.encode(), out.encode(), nranks, t)
    if rc != 0:
        raise IOError("cannot reconstruct %s_%d from %d ranks in %s (%d)" % (out, t, nranks, data_dir, rc))
    return os.path.join(data_dir, "%s_%d.vtk" % (out, t))
