"""Steady-state solver surface: the pseudo-transient linearisation, GMRES /
FGMRES and the JFNK step, all running on the MI355X through libibis_gpu.

Mirrors the reference classes
  SteadyStateLinearisation  src/solvers/solvers/steady_state.cpp:13-174
  Gmres, FGmres             src/linear_algebra/linear_algebra/gmres.cpp:87-318
  Jfnk                      src/solvers/solvers/jfnk.cpp:7-93, jfnk.h:39-53
  CflSchedule               src/solvers/solvers/cfl.cpp:9-83
and read the same json keys.  Vectors are float64 CUDA tensors of
n_cells * (dim+2) entries in the reference's order (cell * n_cons + q).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .native import (MAX_CFL_POINTS, NNORMS, DenseOpStruct, IbisError, JfnkDescStruct,
                     LinearResultStruct, SteadyOpStruct, check, lib, ptr)
from .finite_volume import FiniteVolume, _stream_handle

NORM_NAMES = ("mass", "momentum_x", "momentum_y", "momentum_z", "energy", "global")


@dataclass
class LinearSolveResult:
    """LinearSolveResult (linear_algebra/linear_system.h)."""
    success: bool
    n_iters: int
    tol: float
    residual: float

    @staticmethod
    def _from(s: LinearResultStruct) -> "LinearSolveResult":
        return LinearSolveResult(bool(s.success), int(s.n_iters), float(s.tol), float(s.residual))


def _fn(name: str) -> C.c_void_p:
    """Address of an exported ibis_matvec_fn."""
    return C.cast(getattr(lib(), name), C.c_void_p)


class DenseOperator:
    """w = A v for a dense n x n CUDA matrix (the reference's GMRES test
    system, gmres.cpp:320-500)."""

    def __init__(self, a):
        self.a = a.contiguous()
        self._s = DenseOpStruct(self.a.shape[0], self.a.data_ptr())
        self.fn = _fn("ibis_dense_matvec")
        self.ctx = C.cast(C.pointer(self._s), C.c_void_p)

    @property
    def num_vars(self):
        return self.a.shape[0]


class SteadyStateLinearisation:
    """(I/dt* - dR/dU) as an operator on the MI355X (steady_state.cpp:52-128)."""

    def __init__(self, fv: FiniteVolume, cons, allow_reconstruction: bool = True):
        self.fv = fv
        self.cons = cons
        self.allow_reconstruction = allow_reconstruction
        self.n_vars = fv.grid.N * fv.ncons
        self._s = SteadyOpStruct(fv._h.value, cons.data_ptr(), 1.0,
                                 1 if allow_reconstruction else 0)
        self.fn = _fn("ibis_steady_op_matvec")
        self.ctx = C.cast(C.pointer(self._s), C.c_void_p)

    def preconditioner(self) -> "SteadyStateLinearisation":
        """The same system without reconstruction (steady_state.cpp:47-50)."""
        return SteadyStateLinearisation(self.fv, self.cons, False)

    def set_pseudo_time_step(self, dt_star: float):
        self._s.dt_star = float(dt_star)

    def num_vars(self) -> int:
        return self.n_vars

    def matrix_vector_product(self, vec, result, stream=None):
        check(lib().ibis_steady_matvec(self.fv._h, ptr(self.cons), ptr(vec), ptr(result),
                                       self._s.dt_star, self._s.allow_reconstruction,
                                       _stream_handle(stream)), "ibis_steady_matvec")

    def eval_rhs(self, prim, dudt, rhs, stream=None):
        check(lib().ibis_steady_rhs(self.fv._h, ptr(prim), ptr(dudt), ptr(rhs),
                                    self._s.allow_reconstruction, _stream_handle(stream)),
              "ibis_steady_rhs")


class Gmres:
    """Gmres(system, max_iters, tol) (gmres.cpp:87-182)."""

    def __init__(self, num_vars: int, max_iters: int, tol: float):
        h = C.c_void_p()
        check(lib().ibis_gmres_create(int(num_vars), int(max_iters), float(tol), C.byref(h)),
              "ibis_gmres_create")
        self._h = h

    def solve(self, system, rhs, x, stream=None) -> LinearSolveResult:
        res = LinearResultStruct()
        check(lib().ibis_gmres_solve(self._h, system.fn, system.ctx, ptr(rhs), ptr(x),
                                     C.byref(res), _stream_handle(stream)), "ibis_gmres_solve")
        return LinearSolveResult._from(res)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ibis_gmres_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FGmres(Gmres):
    """FGmres(system, max_iters, tol, preconditioner, max_precondition_iters,
    precondition_tol) (gmres.cpp:185-303)."""

    def __init__(self, num_vars: int, max_iters: int, tol: float, max_precondition_iters: int,
                 precondition_tol: float):
        super().__init__(num_vars, max_iters, tol)
        self.inner = Gmres(num_vars, max_precondition_iters, precondition_tol)

    def solve(self, system, rhs, x, preconditioner=None, stream=None) -> LinearSolveResult:
        if preconditioner is None:
            raise IbisError("FGmres.solve needs a preconditioner system")
        res = LinearResultStruct()
        check(lib().ibis_fgmres_solve(self._h, system.fn, system.ctx, ptr(rhs), ptr(x),
                                      self.inner._h, preconditioner.fn, preconditioner.ctx,
                                      C.byref(res), _stream_handle(stream)),
              "ibis_fgmres_solve")
        return LinearSolveResult._from(res)


def jfnk_desc(solver_config: dict) -> JfnkDescStruct:
    """Encode the reference's solver json (jfnk.cpp:7-22, gmres.cpp:114-116,
    229-232, 305-318, cfl.cpp:71-83)."""
    d = JfnkDescStruct()
    ls = solver_config["linear_solver"]
    kind = ls["type"]
    if kind == "gmres":
        d.linear_solver = 0
        d.max_iters = int(ls["max_iters"])
        d.tol = float(ls["tol"])
    elif kind == "fgmres":
        d.linear_solver = 1
        d.max_iters = int(ls["max_iters"])
        d.tol = float(ls["tolerance"])
        d.max_precond_iters = int(ls["max_preconditioner_iters"])
        d.precond_tol = float(ls["preconditioner_tolerance"])
    else:
        raise IbisError(f"Unknown linear solver {kind}")
    cfl = solver_config["cfl"]
    ctype = cfl["type"]
    if ctype == "constant":
        d.cfl_kind = 0
        d.cfl_value = float(cfl["value"])
    elif ctype == "linear_interpolate":
        d.cfl_kind = 1
        times, cfls = list(cfl["times"]), list(cfl["cfls"])
        if not 1 <= len(times) <= MAX_CFL_POINTS or len(times) != len(cfls):
            raise IbisError("linear_interpolate CFL: times and cfls must match (<= 16 points)")
        d.n_cfl_points = len(times)
        for i, (t, c) in enumerate(zip(times, cfls)):
            d.cfl_times[i] = float(t)
            d.cfl_values[i] = float(c)
    elif ctype == "residual_based":
        d.cfl_kind = 2
        d.growth_threshold = float(cfl["growth_threshold"])
        d.power = float(cfl["power"])
        d.start_cfl = float(cfl["start_cfl"])
        d.max_cfl = float(cfl["max_cfl"])
    else:
        raise IbisError(f"Unknown CFL schedule {ctype}")
    return d


class CflSchedule:
    """CflSchedule::eval with the residual-based state (cfl.cpp:9-68)."""

    def __init__(self, cfl_config: dict):
        self._d = jfnk_desc({"linear_solver": {"type": "gmres", "max_iters": 1, "tol": 0.0},
                             "cfl": cfl_config})
        self._prev_cfl = C.c_double(self._d.start_cfl)
        self._prev_res = C.c_double(1.0)

    def eval(self, t: float) -> float:
        out = C.c_double()
        check(lib().ibis_cfl_eval(C.byref(self._d), float(t), C.byref(self._prev_cfl),
                                  C.byref(self._prev_res), C.byref(out)), "ibis_cfl_eval")
        return out.value


class Jfnk:
    """Jfnk(system, cfl, residuals, config) (jfnk.cpp:7-93), on the device.

    The state is the caller's (cons, prim) pair of CUDA tensors; ``step``
    updates both in place."""

    def __init__(self, fv: FiniteVolume, solver_config: dict):
        self.fv = fv
        self.max_steps = int(solver_config.get("max_steps", 0))
        self.tolerance = float(solver_config.get("tolerance", 0.0))
        self._desc = jfnk_desc(solver_config)
        h = C.c_void_p()
        check(lib().ibis_jfnk_create(fv._h, C.byref(self._desc), C.byref(h)), "ibis_jfnk_create")
        self._h = h
        self._norms = np.zeros(NNORMS)
        self._initial = np.ones(NNORMS)
        self.last_dt_star = None

    def initialise(self, cons, prim, stream=None) -> dict:
        out = np.zeros(NNORMS)
        check(lib().ibis_jfnk_initialise(self._h, ptr(cons), ptr(prim), ptr(out),
                                         _stream_handle(stream)), "ibis_jfnk_initialise")
        self._norms = out.copy()
        self._initial = out.copy()
        return dict(zip(NORM_NAMES, out))

    def step(self, cons, prim, step: int, stream=None) -> LinearSolveResult:
        res = LinearResultStruct()
        out = np.zeros(NNORMS)
        dts = C.c_double()
        check(lib().ibis_jfnk_step(self._h, ptr(cons), ptr(prim), int(step), C.byref(res),
                                   ptr(out), C.byref(dts), _stream_handle(stream)),
              "ibis_jfnk_step")
        self._norms = out
        self.last_dt_star = dts.value
        return LinearSolveResult._from(res)

    def residual_norms(self) -> dict:
        return dict(zip(NORM_NAMES, self._norms))

    def relative_residual_norms(self) -> dict:
        return dict(zip(NORM_NAMES, self._norms / self._initial))

    def target_residual(self) -> float:
        return self.tolerance

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ibis_jfnk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
