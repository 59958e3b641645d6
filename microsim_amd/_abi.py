"""ctypes mirror of the C ABI in include/gp_hip.h and include/gp_host.h.

The structures are declared field for field in the C order; ``check_layout`` compares their sizes with the
``gp_sizeof_*`` values the compiled host library reports, so a drift between the header and this file fails
loudly instead of corrupting memory.
"""
from __future__ import annotations

import ctypes as C
import os

GP_MAXP = 6
GP_MAXNC = 4
GP_NAMELEN = 51

GP_OK = 0
GP_E_ARG = -1
GP_E_HIP = -2
GP_E_OOM = -3
GP_E_NAN = -4
GP_E_IO = -5
GP_E_NODEV = -6

GP_FP64 = 0
GP_FP32 = 1
GP_FP32C = 2

GP_SOLVE_LU = 0
GP_SOLVE_BINARY = 1
GP_SOLVE_TERNARY = 2
GP_SOLVE_DILUTE = 3

GP_NKERNELS = 12
GP_PROFILE_OFF = 0
GP_PROFILE_LAUNCHES = 1
GP_PROFILE_CALLS = 2

GP_NOALLOC_SCRATCH_DOUBLES = 1 << 20

P, NCM = GP_MAXP, GP_MAXNC


class GPParams(C.Structure):
    _fields_ = [
        ("nphases", C.c_int32), ("ncomp", C.c_int32),
        ("mesh_x", C.c_int64), ("mesh_y", C.c_int64),
        ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double),
        ("epsilon", C.c_double), ("tau", C.c_double), ("T", C.c_double), ("Teq", C.c_double),
        ("fn_w", C.c_int32), ("fn_aniso", C.c_int32), ("fold", C.c_int32), ("solve", C.c_int32),
        ("jat_skip_row", C.c_int32), ("reserved0", C.c_int32),
        ("halo_src_x", (C.c_int32 * 6) * 2), ("halo_src_y", (C.c_int32 * 6) * 2),
        ("tau_ab", C.c_double * (P * P)), ("gamma", C.c_double * (P * P)),
        ("gamma_abc", C.c_double * (P * P * P)), ("dab", C.c_double * (P * P)),
        ("rot", C.c_double * (P * P * 4)), ("rot_inv", C.c_double * (P * P * 4)),
        ("diffusivity", C.c_double * (P * NCM * NCM)), ("A", C.c_double * (P * NCM * NCM)),
        ("cmu", C.c_double * (P * NCM * NCM)), ("B", C.c_double * (P * NCM)), ("C", C.c_double * P),
        ("Beq", C.c_double * (P * NCM)), ("dBbdT", C.c_double * (P * NCM)),
        ("tempgrady", C.c_int32), ("shift", C.c_int32), ("shiftj", C.c_int64),
        ("tg_base_temp", C.c_double), ("tg_deltaT", C.c_double), ("tg_distance", C.c_double),
        ("tg_velocity", C.c_double), ("tg_offset", C.c_double),
        ("halo_src_T_y", C.c_int32 * 6), ("reserved1", C.c_int32 * 2),
        ("ceq_sol", C.c_double * (P * NCM)), ("ceq_liq", C.c_double * (P * NCM)), ("delta_c", C.c_double * (P * NCM)),
        ("delta_T_sol", C.c_double * P), ("delta_T_liq", C.c_double * P),
        ("dcbdT_phase", C.c_double * (P * NCM)), ("mu_fill", C.c_double * NCM),
    ]


class GPStats(C.Structure):
    _fields_ = [
        ("phi_max", C.c_double * P), ("phi_min", C.c_double * P),
        ("mu_max", C.c_double * NCM), ("mu_min", C.c_double * NCM),
        ("rel_change_phi", C.c_double * P), ("rel_change_mu", C.c_double * NCM),
        ("solute", C.c_double * NCM), ("energy", C.c_double),
        ("step", C.c_int64), ("nonfinite", C.c_int64),
    ]


class GPInput(C.Structure):
    _fields_ = [
        ("p", GPParams),
        ("dimension", C.c_int32), ("mesh_z", C.c_int64), ("dz", C.c_double),
        ("ntimesteps", C.c_int64), ("savet", C.c_int64), ("nsmooth", C.c_int64), ("starttime", C.c_int64),
        ("restart", C.c_int64), ("track_progress", C.c_int64),
        ("ascii", C.c_int32), ("writecomposition", C.c_int32), ("isothermal", C.c_int32),
        ("noise_phasefield", C.c_int32), ("amp_noise_phase", C.c_double),
        ("shift", C.c_int32), ("shiftj", C.c_int64),
        ("fn_f", C.c_int32), ("binary", C.c_int32), ("ternary", C.c_int32), ("dilute", C.c_int32),
        ("R", C.c_double), ("V", C.c_double), ("Tfill", C.c_double), ("temperature_scale", C.c_double),
        ("tilt_angle", C.c_double), ("rtheta", C.c_double),
        ("phases", (C.c_char * GP_NAMELEN) * P), ("components", (C.c_char * GP_NAMELEN) * (NCM + 1)),
        ("max_length", C.c_int32),
        ("ceq", ((C.c_double * NCM) * P) * P), ("cfill", ((C.c_double * NCM) * P) * P),
        ("slopes", ((C.c_double * NCM) * P) * P),
        ("rot3", (((C.c_double * 3) * 3) * P) * P), ("rot3_inv", (((C.c_double * 3) * 3) * P) * P),
        ("bc_type", (C.c_int32 * 3) * 6), ("bc_points", ((C.c_int32 * 3) * 3) * 6),
        ("bc_proxy", ((C.c_int32 * 3) * 3) * 6), ("bc_proxy_set", (C.c_int32 * 3) * 6),
        ("seen", C.c_uint32), ("seen_A", C.c_uint8 * P), ("seen_D", C.c_uint8 * P),
        ("seen_rot", (C.c_uint8 * P) * P), ("eigen_dummy", C.c_double),
        ("writehdf5", C.c_int32), ("reserved_h5", C.c_int32),
    ]


class GPCell(C.Structure):
    _fields_ = [("phia", C.POINTER(C.c_double)), ("compi", C.POINTER(C.c_double)),
                ("deltaphi", C.POINTER(C.c_double)), ("temperature", C.c_double)]


HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

_dp = C.POINTER(C.c_double)


def _declare_host(lib):
    lib.gp_read_input.argtypes = [C.c_char_p, C.POINTER(GPInput), C.c_int, C.c_char_p, C.c_size_t]
    lib.gp_validate_input.argtypes = [C.POINTER(GPInput), C.c_char_p, C.c_size_t]
    lib.gp_property_matrices.argtypes = [C.POINTER(GPInput)]
    lib.gp_property_matrices.restype = None
    lib.gp_rows_x.argtypes = [C.POINTER(GPParams)]
    lib.gp_rows_x.restype = C.c_int64
    lib.gp_rows_y.argtypes = [C.POINTER(GPParams)]
    lib.gp_rows_y.restype = C.c_int64
    lib.gp_fill_domain.argtypes = [C.POINTER(GPInput), C.c_char_p, _dp, _dp, _dp, C.c_int, C.c_char_p, C.c_size_t]
    lib.gp_apply_bc_host.argtypes = [C.POINTER(GPParams), _dp, _dp]
    lib.gp_apply_bc_host.restype = None
    lib.gp_write_vtk.argtypes = [C.POINTER(GPInput), C.c_char_p, _dp, _dp, _dp]
    lib.gp_write_h5.argtypes = [C.POINTER(GPInput), C.c_char_p, _dp, _dp, _dp]
    lib.gp_read_h5.argtypes = [C.POINTER(GPInput), C.c_char_p, _dp, _dp]
    lib.gp_write_xdmf.argtypes = [C.POINTER(GPInput), C.c_char_p, C.c_char_p]
    lib.gp_rank_block.argtypes = [C.POINTER(GPParams), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]
    lib.gp_write_rank_vtk.argtypes = [C.POINTER(GPInput), C.c_char_p, C.c_char_p, C.c_int64, C.c_int, C.c_int, _dp,
                                      _dp, _dp]
    lib.gp_read_rank_vtk.argtypes = [C.POINTER(GPInput), C.c_char_p, C.c_char_p, C.c_int64, C.c_int, _dp, _dp]
    lib.gp_reconstruct.argtypes = [C.POINTER(GPInput), C.c_char_p, C.c_char_p, C.c_int, C.c_int64]
    lib.gp_composition_at.argtypes = [C.POINTER(GPParams), _dp, _dp, C.c_int, C.c_double]
    lib.gp_composition_at.restype = C.c_double
    lib.gp_temperature_rows.argtypes = [C.POINTER(GPParams), C.c_int64, C.c_int64, _dp]
    lib.gp_temperature_rows.restype = None
    lib.gp_read_vtk.argtypes = [C.POINTER(GPInput), C.c_char_p, _dp, _dp]
    lib.gp_write_echo.argtypes = [C.POINTER(GPInput), C.c_char_p]
    lib.gp_composition.argtypes = [C.POINTER(GPParams), _dp, _dp, C.c_int]
    lib.gp_composition.restype = C.c_double
    lib.gp_composition_grid.argtypes = [C.POINTER(GPParams), _dp, _dp, _dp, C.c_int64]
    lib.gp_composition_grid.restype = None
    for f in ("gp_sizeof_input", "gp_sizeof_params", "gp_sizeof_stats"):
        getattr(lib, f).restype = C.c_size_t
        getattr(lib, f).argtypes = []


def _declare_hip(lib):
    h = C.c_void_p
    lib.gp_create.argtypes = [C.POINTER(GPParams), C.c_int, C.POINTER(h)]
    lib.gp_upload.argtypes = [h, _dp, _dp]
    lib.gp_upload_fields.argtypes = [h, C.POINTER(GPCell), C.c_int64]
    lib.gp_download.argtypes = [h, _dp, _dp, _dp]
    lib.gp_download_fields.argtypes = [h, C.POINTER(GPCell), C.c_int64]
    lib.gp_smooth.argtypes = [h, C.c_int64]
    lib.gp_step.argtypes = [h, C.c_int64, C.POINTER(GPStats)]
    lib.gp_get_stats.argtypes = [h, C.POINTER(GPStats)]
    lib.gp_synchronize.argtypes = [h]
    lib.gp_last_error.argtypes = [h]
    lib.gp_last_error.restype = C.c_char_p
    lib.gp_destroy.argtypes = [h]
    lib.gp_destroy.restype = None
    lib.gp_version.argtypes = []
    lib.gp_version.restype = C.c_char_p
    lib.gp_profile_enable.argtypes = [h, C.c_int]
    lib.gp_profile_read.argtypes = [h, _dp, C.POINTER(C.c_int64)]
    lib.gp_kernel_name.argtypes = [C.c_int]
    lib.gp_kernel_name.restype = C.c_char_p
    lib.gp_theta_count.argtypes = [h]
    lib.gp_theta_values.argtypes = [h, _dp]
    lib.gp_adjoint_forward.argtypes = [h, C.c_int64, C.c_int64]
    lib.gp_adjoint_backward.argtypes = [h, _dp, _dp, _dp, _dp, _dp]
    lib.gp_invariants_noalloc.argtypes = [h, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gp_get_frame.argtypes = [h, C.POINTER(C.c_int64), _dp]


_host = None
_hip = None


def host_lib():
    """libgp_host.so (CPU). Raises if it has not been built (python -c 'import __graft_entry__ as g; g.build()')."""
    global _host
    if _host is None:
        path = os.path.join(LIBDIR, "libgp_host.so")
        if not os.path.exists(path):
            raise RuntimeError("libgp_host.so is not built: run __graft_entry__.build()")
        lib = C.CDLL(path)
        _declare_host(lib)
        check_layout(lib)
        _host = lib
    return _host


def hip_lib():
    """libgp_hip.so (HIP, gfx950). Raises if it has not been built; there is no CPU fallback."""
    global _hip
    if _hip is None:
        path = os.path.join(LIBDIR, "libgp_hip.so")
        if not os.path.exists(path):
            raise RuntimeError("libgp_hip.so is not built: run __graft_entry__.build()")
        lib = C.CDLL(path)
        _declare_hip(lib)
        _hip = lib
    return _hip


def check_layout(lib) -> None:
    for name, cls in (("gp_sizeof_input", GPInput), ("gp_sizeof_params", GPParams), ("gp_sizeof_stats", GPStats)):
        n = getattr(lib, name)()
        if n != C.sizeof(cls):
            raise RuntimeError("ABI mismatch: %s = %d, ctypes %s = %d" % (name, n, cls.__name__, C.sizeof(cls)))


def dptr(arr):
    """double* of a C-contiguous float64 numpy array."""
    import numpy as np
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(_dp)
