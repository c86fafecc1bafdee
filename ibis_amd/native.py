"""ctypes binding of libibis_gpu.so (include/ibis_gpu.h).

The library is built in-tree by ``ibis_amd.build``.  Importing this module
never falls back to anything: if the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

MAX_TAGS, MAX_ACT, NPAR = 16, 4, 8
NKERNELS = 8
# timing slots of ibis_residual_kernel_times.  On the fused second-order
# path: boundary = k_bc (+ k_pack) + k_ghost_fs, cell_stencil = k_cell_fs
# (LSQ gradients, limiter, face values), face_flux = k_face_*_fs (flux and
# surface integral).  On the general path: k_bc, k_cell_grad_var,
# k_face_flux, k_post_flux, the viscous kernels and k_integral.
KERNEL_NAMES = ("boundary", "cell_stencil", "face_flux", "post_flux", "boundary_viscous",
                "cell_gradient_viscous", "viscous_flux", "surface_integral")

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "ibis_gpu.h")


class IbisError(RuntimeError):
    pass


class GridIOStruct(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ntags", C.c_int32), ("nv", C.c_int64),
                ("xyz", C.c_void_p), ("nc", C.c_int64), ("ctype", C.c_void_p),
                ("cvtx", C.c_void_p), ("tag_ghost", C.c_void_p), ("nb", C.c_int64),
                ("btag", C.c_void_p), ("btype", C.c_void_p), ("bvtx", C.c_void_p)]


class MeshDescStruct(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ntags", C.c_int32), ("n_cells", C.c_int64),
                ("n_ghost", C.c_int64), ("n_faces", C.c_int64), ("face_left", C.c_void_p),
                ("face_right", C.c_void_p), ("face_geom", C.c_void_p),
                ("cell_face_off", C.c_void_p), ("cell_faces", C.c_void_p),
                ("cell_outsigns", C.c_void_p), ("cell_neighbours", C.c_void_p),
                ("centroid", C.c_void_p), ("volume", C.c_void_p), ("lsq", C.c_void_p),
                ("tag_off", C.c_void_p), ("bfaces", C.c_void_p)]


class FVDescStruct(C.Structure):
    _fields_ = [("flux", C.c_int32), ("order", C.c_int32), ("thermo", C.c_int32),
                ("limiter", C.c_int32), ("viscous", C.c_int32), ("ntags", C.c_int32),
                ("ldfss_delta", C.c_double), ("lim_eps", C.c_double), ("venkat_K", C.c_double),
                ("signal_factor", C.c_double), ("R", C.c_double), ("Cv", C.c_double),
                ("Cp", C.c_double), ("gamma", C.c_double), ("mu0", C.c_double),
                ("T0", C.c_double), ("Ts", C.c_double), ("Pr", C.c_double),
                ("pre_kind", C.c_int32 * (MAX_TAGS * MAX_ACT)),
                ("pre_par", C.c_double * (MAX_TAGS * MAX_ACT * NPAR)),
                ("visc_kind", C.c_int32 * (MAX_TAGS * MAX_ACT)),
                ("visc_par", C.c_double * (MAX_TAGS * MAX_ACT * NPAR)),
                ("post_kind", C.c_int32 * (MAX_TAGS * MAX_ACT)),
                ("post_par", C.c_double * (MAX_TAGS * MAX_ACT * NPAR))]


MAX_CFL_POINTS = 16
NNORMS = 6


class LinearResultStruct(C.Structure):
    """ibis_linear_result (LinearSolveResult)."""
    _fields_ = [("success", C.c_int32), ("n_iters", C.c_int32), ("tol", C.c_double),
                ("residual", C.c_double)]


class SteadyOpStruct(C.Structure):
    _fields_ = [("r", C.c_void_p), ("cons", C.c_void_p), ("dt_star", C.c_double),
                ("allow_reconstruction", C.c_int32)]


class DenseOpStruct(C.Structure):
    _fields_ = [("n", C.c_int64), ("a", C.c_void_p)]


class JfnkDescStruct(C.Structure):
    _fields_ = [("linear_solver", C.c_int32), ("max_iters", C.c_int32), ("tol", C.c_double),
                ("max_precond_iters", C.c_int32), ("precond_tol", C.c_double),
                ("cfl_kind", C.c_int32), ("n_cfl_points", C.c_int32), ("cfl_value", C.c_double),
                ("cfl_times", C.c_double * MAX_CFL_POINTS),
                ("cfl_values", C.c_double * MAX_CFL_POINTS),
                ("growth_threshold", C.c_double), ("power", C.c_double),
                ("start_cfl", C.c_double), ("max_cfl", C.c_double)]


_P = C.c_void_p
_SIGS = {
    "ibis_grid_build": (C.c_int, [C.POINTER(GridIOStruct), C.c_int, C.POINTER(C.c_void_p)]),
    "ibis_grid_destroy": (None, [_P]),
    "ibis_grid_sizes": (C.c_int, [_P, _P]),
    "ibis_grid_permutation": (C.c_int, [_P, _P, _P]),
    "ibis_grid_export": (C.c_int, [_P] * 13),
    "ibis_mesh_create": (C.c_int, [C.POINTER(MeshDescStruct), C.POINTER(C.c_void_p)]),
    "ibis_mesh_from_grid": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ibis_mesh_destroy": (None, [_P]),
    "ibis_residual_create": (C.c_int, [_P, C.POINTER(FVDescStruct), C.POINTER(C.c_void_p)]),
    "ibis_residual_destroy": (None, [_P]),
    "ibis_compute_dudt": (C.c_int, [_P, _P, _P, C.c_int, _P]),
    "ibis_compute_dudt_jvp": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P]),
    "ibis_cons_to_prim_jvp": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "ibis_estimate_dt": (C.c_int, [_P, _P, C.POINTER(C.c_double), _P]),
    "ibis_count_bad_cells": (C.c_int, [_P, _P, C.POINTER(C.c_uint64), _P]),
    "ibis_cons_to_prim": (C.c_int, [_P, _P, _P, _P]),
    "ibis_prim_to_cons": (C.c_int, [_P, _P, _P, _P]),
    "ibis_apply_time_derivative": (C.c_int, [_P, _P, _P, _P, C.c_double, _P]),
    "ibis_rk_step": (C.c_int, [_P, C.c_int, _P, _P, C.c_double, _P, _P,
                               C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double,
                               C.POINTER(C.c_double), _P]),
    "ibis_rk_step_noalloc": (C.c_int, [_P, C.c_int, _P, _P, C.c_double, _P, _P,
                               C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double,
                               C.POINTER(C.c_double), _P]),
    "ibis_residual_scratch": (C.c_int, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_int)]),
    "ibis_compute_convective_gradient": (C.c_int, [_P, _P, _P]),
    "ibis_compute_viscous_gradient": (C.c_int, [_P, _P, _P]),
    "ibis_cell_gradients": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ibis_entry_indices_fit": (C.c_int, [C.c_int64, C.c_int64, C.c_int64]),
    "ibis_residual_profile": (C.c_int, [_P, C.c_int]),
    "ibis_residual_set_option": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "ibis_residual_kernel_times": (C.c_int, [_P, _P, _P, C.c_int]),
    "ibis_residual_sizes": (C.c_int, [_P, _P]),
    "ibis_residual_set_precision": (C.c_int, [_P, C.c_int]),
    "ibis_steady_matvec": (C.c_int, [_P, _P, _P, _P, C.c_double, C.c_int, _P]),
    "ibis_steady_rhs": (C.c_int, [_P, _P, _P, _P, C.c_int, _P]),
    "ibis_steady_op_matvec": (C.c_int, [_P, _P, _P, _P]),
    "ibis_dense_matvec": (C.c_int, [_P, _P, _P, _P]),
    "ibis_gmres_create": (C.c_int, [C.c_int64, C.c_int32, C.c_double, C.POINTER(C.c_void_p)]),
    "ibis_gmres_destroy": (None, [_P]),
    "ibis_gmres_solve": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "ibis_fgmres_solve": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ibis_jfnk_create": (C.c_int, [_P, _P, C.POINTER(C.c_void_p)]),
    "ibis_jfnk_destroy": (None, [_P]),
    "ibis_jfnk_initialise": (C.c_int, [_P, _P, _P, _P, _P]),
    "ibis_jfnk_step": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P, _P, _P]),
    "ibis_cfl_eval": (C.c_int, [_P, C.c_double, _P, _P, _P]),
    "ibis_compute_dudt_vjp": (C.c_int, [_P, _P, _P, C.c_uint64, _P, _P, C.c_int, _P]),
    "ibis_cons_to_prim_vjp": (C.c_int, [_P, _P, _P, _P, _P]),
    "ibis_wall_force": (C.c_int, [_P, _P, C.c_uint64, _P, _P]),
    "ibis_batch_create": (C.c_int, [_P, C.c_int, _P, _P, _P, C.POINTER(C.c_void_p)]),
    "ibis_batch_destroy": (None, [_P]),
    "ibis_batch_residual": (C.c_void_p, [_P]),
    "ibis_batch_rk_init": (C.c_int, [_P, C.c_int, _P, _P, C.c_double, C.c_double, C.c_double,
                                     _P, _P]),
    "ibis_batch_rk_steps": (C.c_int, [_P, C.c_int, C.c_int]),
    "ibis_batch_times": (C.c_int, [_P, _P, _P]),
    "ibis_batch_set_option": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "ibis_last_error": (C.c_char_p, []),
}

_lib = None


def lib_path() -> str:
    return _build.LIB


def lib():
    """The loaded libibis_gpu.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise IbisError(f"{path} is missing: run ibis_amd.build.build() (or "
                            "__graft_entry__.build()) first")
        # torch ships its own HIP runtime.  Import it first so that this
        # library's libamdhip64.so.7 dependency binds to the runtime already
        # in the process; loading a second HIP runtime leaves torch without
        # a device.
        import torch  # noqa: F401
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().ibis_last_error().decode(errors="replace")
        raise IbisError(f"{what} failed ({rc}): {msg}")


def ptr(a) -> C.c_void_p:
    """Address of a numpy array or a torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(a.data_ptr())


def header_symbols():
    """Function names declared in include/ibis_gpu.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ibis_[a-z0-9_]+)\s*\((?!\*)", txt)))
