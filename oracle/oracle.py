"""ctypes wrapper of the CPU oracle (oracle/ibis_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Loaded by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product (ibis_amd)
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libibis_oracle.so")

MAX_TAGS, MAX_ACT, NPAR = 16, 4, 8
NPRIM = 7

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class GridIOStruct(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ntags", C.c_int32), ("nv", C.c_int64),
                ("xyz", C.c_void_p), ("nc", C.c_int64), ("ctype", C.c_void_p),
                ("cvtx", C.c_void_p), ("tag_ghost", C.c_void_p), ("nb", C.c_int64),
                ("btag", C.c_void_p), ("btype", C.c_void_p), ("bvtx", C.c_void_p)]


class FVConfigStruct(C.Structure):
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


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_grid_build.restype = C.c_void_p
        L.or_grid_build.argtypes = [C.POINTER(GridIOStruct)]
        L.or_grid_free.argtypes = [C.c_void_p]
        L.or_grid_sizes.argtypes = [C.c_void_p, _i64p]
        L.or_grid_export.argtypes = [C.c_void_p] + [C.c_void_p] * 12
        L.or_grid_gradient.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_fv_create.restype = C.c_void_p
        L.or_fv_create.argtypes = [C.c_void_p, C.POINTER(FVConfigStruct)]
        L.or_fv_free.argtypes = [C.c_void_p]
        L.or_compute_dudt.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int]
        L.or_compute_dudt_jvp.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p, C.c_int]
        L.or_estimate_dt.restype = C.c_double
        L.or_estimate_dt.argtypes = [C.c_void_p, _f64p]
        L.or_count_bad_cells.restype = C.c_int64
        L.or_count_bad_cells.argtypes = [C.c_void_p, _f64p]
        L.or_cons_to_prim.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_cons_to_prim_jvp.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
        L.or_prim_to_cons.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_fv_intermediates.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.or_compute_convective_gradient.argtypes = [C.c_void_p, _f64p]
        L.or_compute_viscous_gradient.argtypes = [C.c_void_p, _f64p]
        L.or_fv_face_gradients.argtypes = [C.c_void_p, _f64p]
        L.or_flux_face.argtypes = [C.POINTER(FVConfigStruct), C.c_int, _f64p, _f64p, _f64p]
        L.or_dual_eval.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.or_rk_step.restype = C.c_double
        L.or_rk_step.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p, C.c_double, _f64p, _f64p,
                                 C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double]
        L.or_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def set_threads(n: int):
    lib().or_set_threads(int(n))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleGrid:
    """GridBlock restatement built from a GridIO and the boundary tags of a
    config (tags in the reference's sorted order)."""

    def __init__(self, grid_io, tags, ghost):
        self.io = grid_io
        self.tags = list(tags)
        btag, btype, bvtx = grid_io.flat_markers(self.tags)
        self._keep = [np.ascontiguousarray(grid_io.xyz, dtype=np.float64),
                      np.ascontiguousarray(grid_io.ctype, dtype=np.int32),
                      np.ascontiguousarray(grid_io.cvtx, dtype=np.int64),
                      np.ascontiguousarray(ghost, dtype=np.int32), btag, btype, bvtx]
        s = GridIOStruct(grid_io.dim, len(self.tags), grid_io.xyz.shape[0],
                         _ptr(self._keep[0]), grid_io.num_cells, _ptr(self._keep[1]),
                         _ptr(self._keep[2]), _ptr(self._keep[3]), btag.shape[0], _ptr(btag),
                         _ptr(btype), _ptr(bvtx))
        self.h = lib().or_grid_build(C.byref(s))
        if not self.h:
            raise RuntimeError("oracle grid build failed")
        sz = np.zeros(8, np.int64)
        lib().or_grid_sizes(self.h, sz)
        (self.dim, self.N, self.G, self.F, self.nnz, self.n_lsq, self.ntags,
         self.nbf) = [int(v) for v in sz]
        self.Nt = self.N + self.G
        self._export()

    def _export(self):
        F, N, Nt = self.F, self.N, self.Nt
        self.face_geom = np.zeros((13, F))
        self.face_lr = np.zeros((2, F), np.int64)
        self.cell_off = np.zeros(N + 1, np.int64)
        self.cell_faces = np.zeros(self.nnz, np.int64)
        self.outsigns = np.zeros(self.nnz, np.int32)
        self.neighbours = np.zeros(self.nnz, np.int64)
        self.centroid = np.zeros((3, Nt))
        self.volume = np.zeros(Nt)
        self.lsq = np.zeros((self.n_lsq, N))
        self.tag_off = np.zeros(self.ntags + 1, np.int64)
        self.bfaces = np.zeros(self.nbf, np.int64)
        self.ghosts = np.zeros(self.nbf, np.int64)
        lib().or_grid_export(self.h, *[_ptr(a) for a in (
            self.face_geom, self.face_lr, self.cell_off, self.cell_faces, self.outsigns,
            self.neighbours, self.centroid, self.volume, self.lsq, self.tag_off, self.bfaces,
            self.ghosts)])

    def norm(self):
        return self.face_geom[0:3]

    def tan1(self):
        return self.face_geom[3:6]

    def tan2(self):
        return self.face_geom[6:9]

    def face_centre(self):
        return self.face_geom[9:12]

    def area(self):
        return self.face_geom[12]

    def ghost_cells(self, tag):
        t = self.tags.index(tag)
        return self.ghosts[self.tag_off[t]:self.tag_off[t + 1]]

    def boundary_faces(self, tag):
        t = self.tags.index(tag)
        return self.bfaces[self.tag_off[t]:self.tag_off[t + 1]]

    def gradient(self, values):
        out = np.zeros(3 * self.N)
        lib().or_grid_gradient(self.h, np.ascontiguousarray(values, dtype=np.float64), out)
        return out.reshape(3, self.N)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_grid_free(self.h)
            self.h = None


def flux_face(cfg, dim, left, right):
    """The configured flux calculator on one face.  left/right: rho, p, T,
    e, u_n, u_t1, u_t2 in the face frame.  Returns (mass, n-, t1-, t2-momentum,
    energy)."""
    s = fv_struct(cfg)
    out = np.zeros(5)
    lib().or_flux_face(C.byref(s), int(dim), np.ascontiguousarray(left, dtype=np.float64),
                       np.ascontiguousarray(right, dtype=np.float64), out)
    return out


DUAL_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "abs": 5, "pow": 6, "min": 7,
            "max": 8, "floor": 9, "copysign": 10, "add_real": 11, "sub_real": 12,
            "real_sub": 13, "mul_real": 14, "real_mul": 15, "div_real": 16, "real_div": 17,
            "iadd": 18, "isub": 19}


def dual_eval(op, a, b=(0.0, 0.0)):
    """The oracle's Dual<double> operation `op` on (real, dual) pairs a, b
    (for the *_real forms only b's real part is used)."""
    out = np.zeros(2)
    lib().or_dual_eval(DUAL_OPS[op], np.array(a, dtype=np.float64),
                       np.array(b, dtype=np.float64), out)
    return out[0], out[1]


def fv_struct(cfg) -> FVConfigStruct:
    from ibis_amd.config import FLUX, LIMITER, THERMO  # input encoding only
    enc = cfg.encode()
    s = FVConfigStruct()
    s.flux = FLUX[cfg.flux]
    s.order = cfg.reconstruction_order
    s.thermo = THERMO[cfg.thermo_interpolator]
    s.limiter = LIMITER[cfg.limiter]
    s.viscous = 1 if cfg.viscous else 0
    s.ntags = len(enc["tags"])
    s.ldfss_delta = cfg.ldfss_delta
    s.lim_eps = cfg.epsilon
    s.venkat_K = cfg.venkat_K
    s.signal_factor = cfg.signal_factor
    g, t = cfg.gas, cfg.transport
    s.R, s.Cv, s.Cp, s.gamma = g.R, g.Cv, g.Cp, g.gamma
    s.mu0, s.T0, s.Ts, s.Pr = t.mu0, t.T0, t.Ts, t.Pr
    for name in ("pre", "visc", "post"):
        k, p = enc[name]
        getattr(s, name + "_kind")[:] = [int(v) for v in k]
        getattr(s, name + "_par")[:] = [float(v) for v in p]
    return s


class OracleFV:
    """FiniteVolume<double> restatement (plus the Dual<double> JVP)."""

    def __init__(self, grid: OracleGrid, cfg):
        self.grid = grid
        self.cfg = cfg
        self._s = fv_struct(cfg)
        self.h = lib().or_fv_create(grid.h, C.byref(self._s))
        self.ncons = grid.dim + 2

    def new_prim(self):
        return np.zeros((NPRIM, self.grid.Nt))

    def new_cons(self):
        return np.zeros((self.ncons, self.grid.Nt))

    def compute_dudt(self, prim, dudt=None, allow_reconstruction=True):
        if dudt is None:
            dudt = self.new_cons()
        lib().or_compute_dudt(self.h, prim, dudt, 1 if allow_reconstruction else 0)
        return dudt

    def compute_dudt_jvp(self, prim, prim_dot, allow_reconstruction=True):
        dudt, dudt_dot = self.new_cons(), self.new_cons()
        lib().or_compute_dudt_jvp(self.h, prim, prim_dot, dudt, dudt_dot,
                                  1 if allow_reconstruction else 0)
        return dudt, dudt_dot

    def estimate_dt(self, prim):
        return lib().or_estimate_dt(self.h, prim)

    def count_bad_cells(self, prim):
        return lib().or_count_bad_cells(self.h, prim)

    def cons_to_prim(self, cons, prim=None):
        if prim is None:
            prim = self.new_prim()
        lib().or_cons_to_prim(self.h, cons, prim)
        return prim

    def cons_to_prim_jvp(self, cons, cons_dot):
        prim, prim_dot = self.new_prim(), self.new_prim()
        lib().or_cons_to_prim_jvp(self.h, cons, cons_dot, prim, prim_dot)
        return prim, prim_dot

    def prim_to_cons(self, prim, cons=None):
        if cons is None:
            cons = self.new_cons()
        lib().or_prim_to_cons(self.h, prim, cons)
        return cons

    def compute_convective_gradient(self, prim):
        """Boundary actions on prim (ghost rows written) and the convective
        gradients into the gradient slots (FV/finite_volume.cpp:283-288)."""
        lib().or_compute_convective_gradient(self.h, prim)

    def compute_viscous_gradient(self, prim):
        """FV/finite_volume.cpp:273-280."""
        lib().or_compute_viscous_gradient(self.h, prim)

    def intermediates(self):
        g = self.grid
        out = {"grads": np.zeros((7, 3, g.N)), "lim": np.zeros((7, g.N)),
               "flux": np.zeros((self.ncons, g.F)), "left": np.zeros((7, g.F)),
               "right": np.zeros((7, g.F))}
        lib().or_fv_intermediates(self.h, *[_ptr(out[k]) for k in
                                            ("grads", "lim", "flux", "left", "right")])
        return out

    def face_gradients(self):
        """Face gradients of T, vx, vy, vz (4 x 3 x F) of the last viscous
        residual (copy_gradients_to_face / hasselbacher_average)."""
        out = np.zeros((4, 3, self.grid.F))
        lib().or_fv_face_gradients(self.h, out)
        return out

    def rk_step(self, a, b, cfl, cons, prim, dt_prev, t, max_time, plot_remaining=-1.0):
        d = C.c_double(dt_prev)
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        dt = lib().or_rk_step(self.h, len(b), a, b, cfl, cons, prim, C.byref(d), t, max_time,
                              plot_remaining)
        return dt

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_fv_free(self.h)
            self.h = None
