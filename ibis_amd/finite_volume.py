"""Host-side operator surface of the MI355X residual.

Mirrors the reference's FiniteVolume<T> interface
(src/finite_volume/finite_volume/finite_volume.h:28-142):

    grid = GridBlock(grid_io, config)                 # GridBlock<T>(GridIO, json)
    fv = FiniteVolume(grid, config)                   # FiniteVolume<T>(grid, json)
    fv.compute_dudt(flow_state, grid, dudt, gas_model, trans_prop,
                    allow_reconstruction=True)        # finite_volume.h:44-47
    fv.estimate_dt(flow_state, grid, gas_model, trans_prop)   # :66-67
    fv.count_bad_cells(flow_state, num_cells)         # :98

All arithmetic runs in the gfx950 kernels of libibis_gpu.so; this module
only allocates device buffers (torch) and calls the C ABI.  There is no CPU
path: without a GPU and the built library the calls raise.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import config as cfgmod
from .gridio import GridIO
from .native import (FVDescStruct, GridIOStruct, IbisError, MeshDescStruct, check, lib, ptr)

NPRIM = 7
ROWS = ("rho", "p", "T", "e", "vx", "vy", "vz")


def _torch():
    import torch
    return torch


def _stream_handle(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _device_view(addr: int, shape):
    """A float64 CUDA tensor over library-owned device memory (no copy)."""
    torch = _torch()
    n = int(np.prod(shape))

    class _Ptr:
        __cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (addr, False),
                                    "version": 3, "strides": None}
    return torch.as_tensor(_Ptr(), device="cuda").view(*shape)


class GridBlock:
    """Host grid built by libibis_gpu (reference numbering, then optionally
    renumbered for locality) plus its device copy."""

    def __init__(self, grid_io: GridIO, config: cfgmod.FVConfig, reorder: bool = True,
                 device: bool = True):
        enc = config.encode()
        self.tags = enc["tags"]
        btag, btype, bvtx = grid_io.flat_markers(self.tags)
        keep = [np.ascontiguousarray(grid_io.xyz, dtype=np.float64),
                np.ascontiguousarray(grid_io.ctype, dtype=np.int32),
                np.ascontiguousarray(grid_io.cvtx, dtype=np.int64),
                np.ascontiguousarray(enc["ghost"], dtype=np.int32), btag, btype, bvtx]
        s = GridIOStruct(grid_io.dim, len(self.tags), grid_io.xyz.shape[0], ptr(keep[0]),
                         grid_io.num_cells, ptr(keep[1]), ptr(keep[2]), ptr(keep[3]),
                         btag.shape[0], ptr(btag), ptr(btype), ptr(bvtx))
        h = C.c_void_p()
        check(lib().ibis_grid_build(C.byref(s), 1 if reorder else 0, C.byref(h)),
              "ibis_grid_build")
        self._grid = h
        sz = np.zeros(8, np.int64)
        check(lib().ibis_grid_sizes(self._grid, ptr(sz)), "ibis_grid_sizes")
        (self.dim, self.N, self.G, self.F, self.nnz, self.n_lsq, self.ntags,
         self.nbf) = [int(v) for v in sz]
        self.Nt = self.N + self.G
        self.cell_new_of_old = np.zeros(self.Nt, np.int64)
        self.face_new_of_old = np.zeros(self.F, np.int64)
        check(lib().ibis_grid_permutation(self._grid, ptr(self.cell_new_of_old),
                                          ptr(self.face_new_of_old)), "ibis_grid_permutation")
        self._mesh = None
        if device:
            m = C.c_void_p()
            check(lib().ibis_mesh_from_grid(self._grid, C.byref(m)), "ibis_mesh_from_grid")
            self._mesh = m

    # sizes, named as the reference's GridBlock accessors (grid.h:280-337)
    def num_cells(self):
        return self.N

    def num_ghost_cells(self):
        return self.G

    def num_total_cells(self):
        return self.Nt

    def num_interfaces(self):
        return self.F

    def export(self) -> dict:
        F, N, Nt = self.F, self.N, self.Nt
        out = {"face_geom": np.zeros((13, F)), "face_lr": np.zeros((2, F), np.int64),
               "cell_off": np.zeros(N + 1, np.int64), "cell_faces": np.zeros(self.nnz, np.int64),
               "outsigns": np.zeros(self.nnz, np.int32),
               "neighbours": np.zeros(self.nnz, np.int64), "centroid": np.zeros((3, Nt)),
               "volume": np.zeros(Nt), "lsq": np.zeros((self.n_lsq, N)),
               "tag_off": np.zeros(self.ntags + 1, np.int64),
               "bfaces": np.zeros(self.nbf, np.int64), "ghosts": np.zeros(self.nbf, np.int64)}
        check(lib().ibis_grid_export(self._grid, *[ptr(out[k]) for k in (
            "face_geom", "face_lr", "cell_off", "cell_faces", "outsigns", "neighbours",
            "centroid", "volume", "lsq", "tag_off", "bfaces", "ghosts")]), "ibis_grid_export")
        return out

    # state permutation between the reference numbering and this grid's
    def to_grid_order(self, a_ref: np.ndarray) -> np.ndarray:
        out = np.empty_like(a_ref)
        out[..., self.cell_new_of_old] = a_ref
        return out

    def to_reference_order(self, a: np.ndarray) -> np.ndarray:
        return a[..., self.cell_new_of_old]

    def close(self):
        if getattr(self, "_mesh", None):
            lib().ibis_mesh_destroy(self._mesh)
            self._mesh = None
        if getattr(self, "_grid", None):
            lib().ibis_grid_destroy(self._grid)
            self._grid = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fv_desc(config: cfgmod.FVConfig) -> FVDescStruct:
    enc = config.encode()
    s = FVDescStruct()
    s.flux = cfgmod.FLUX[config.flux]
    s.order = config.reconstruction_order
    s.thermo = cfgmod.THERMO[config.thermo_interpolator]
    s.limiter = cfgmod.LIMITER[config.limiter]
    s.viscous = 1 if config.viscous else 0
    s.ntags = len(enc["tags"])
    s.ldfss_delta = config.ldfss_delta
    s.lim_eps = config.epsilon
    s.venkat_K = config.venkat_K
    s.signal_factor = config.signal_factor
    g, t = config.gas, config.transport
    s.R, s.Cv, s.Cp, s.gamma = g.R, g.Cv, g.Cp, g.gamma
    s.mu0, s.T0, s.Ts, s.Pr = t.mu0, t.T0, t.Ts, t.Pr
    for name in ("pre", "visc", "post"):
        k, p = enc[name]
        getattr(s, name + "_kind")[:] = [int(v) for v in k]
        getattr(s, name + "_par")[:] = [float(v) for v in p]
    return s


class FlowStates:
    """Primitive state on the device: a (7, N_total) float64 torch tensor with
    rows rho, p, T, e, vx, vy, vz (GasStates + velocity)."""

    def __init__(self, n_total: int, device="cuda", data=None):
        torch = _torch()
        self.data = data if data is not None else torch.zeros((NPRIM, n_total),
                                                               dtype=torch.float64,
                                                               device=device)

    def __getattr__(self, name):
        if name in ROWS:
            return self.data[ROWS.index(name)]
        raise AttributeError(name)


class ConservedQuantities:
    """Conserved variables on the device: (dim+2, N_total) float64 tensor,
    rows mass, momentum, energy (conserved_quantities.cpp:16-23)."""

    def __init__(self, n_total: int, dim: int, device="cuda", data=None):
        torch = _torch()
        self.dim = dim
        self.data = data if data is not None else torch.zeros((dim + 2, n_total),
                                                               dtype=torch.float64,
                                                               device=device)


class FiniteVolume:
    """The residual operator (FiniteVolume<T>) on an MI355X."""

    def __init__(self, grid: GridBlock, config: cfgmod.FVConfig):
        if grid._mesh is None:
            raise IbisError("FiniteVolume needs a GridBlock with a device mesh")
        self.grid = grid
        self.config = config
        self._desc = fv_desc(config)
        h = C.c_void_p()
        check(lib().ibis_residual_create(grid._mesh, C.byref(self._desc), C.byref(h)),
              "ibis_residual_create")
        self._h = h
        self.ncons = grid.dim + 2

    @staticmethod
    def _check_state(t, rows, n_total, what):
        torch = _torch()
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                and tuple(t.shape) == (rows, n_total)):
            raise IbisError(f"{what} must be a contiguous float64 CUDA tensor of shape "
                            f"({rows}, {n_total}); got {tuple(t.shape)} {t.dtype} {t.device}")

    def compute_dudt(self, flow_state, grid: Optional[GridBlock] = None, dudt=None,
                     gas_model=None, trans_prop=None, allow_reconstruction: bool = True,
                     stream=None):
        """dU/dt of ``flow_state`` into ``dudt``; ghost rows of ``flow_state``
        are overwritten by the boundary actions.  Asynchronous on ``stream``.
        gas_model and trans_prop are fixed at construction (the config)."""
        prim = flow_state.data if isinstance(flow_state, FlowStates) else flow_state
        if dudt is None:
            dudt = ConservedQuantities(self.grid.Nt, self.grid.dim, device=prim.device)
        d = dudt.data if isinstance(dudt, ConservedQuantities) else dudt
        self._check_state(prim, NPRIM, self.grid.Nt, "flow_state")
        self._check_state(d, self.ncons, self.grid.Nt, "dudt")
        check(lib().ibis_compute_dudt(self._h, ptr(prim), ptr(d),
                                      1 if allow_reconstruction else 0, _stream_handle(stream)),
              "ibis_compute_dudt")
        return dudt

    def compute_dudt_jvp(self, prim, prim_dot, dudt, dudt_dot, allow_reconstruction: bool = True,
                         stream=None):
        """Value and directional derivative of dU/dt along ``prim_dot``
        (FiniteVolume<Ibis::dual>::compute_dudt, steady_state.cpp:52-128).
        Ghost rows of ``prim`` and ``prim_dot`` are written."""
        for t, rows, what in ((prim, NPRIM, "prim"), (prim_dot, NPRIM, "prim_dot"),
                              (dudt, self.ncons, "dudt"), (dudt_dot, self.ncons, "dudt_dot")):
            self._check_state(t, rows, self.grid.Nt, what)
        check(lib().ibis_compute_dudt_jvp(self._h, ptr(prim), ptr(prim_dot), ptr(dudt),
                                          ptr(dudt_dot), 1 if allow_reconstruction else 0,
                                          _stream_handle(stream)), "ibis_compute_dudt_jvp")

    def cons_to_prim_jvp(self, cons, cons_dot, prim, prim_dot, stream=None):
        """conserved_to_primatives on (value, tangent) pairs."""
        for t, rows, what in ((cons, self.ncons, "cons"), (cons_dot, self.ncons, "cons_dot"),
                              (prim, NPRIM, "prim"), (prim_dot, NPRIM, "prim_dot")):
            self._check_state(t, rows, self.grid.Nt, what)
        check(lib().ibis_cons_to_prim_jvp(self._h, ptr(cons), ptr(cons_dot), ptr(prim),
                                          ptr(prim_dot), _stream_handle(stream)),
              "ibis_cons_to_prim_jvp")

    def tag_mask(self, names) -> int:
        """Bit mask of boundary tags (by name) for the force functional."""
        tags = list(self.grid.tags)
        mask = 0
        for n in ([names] if isinstance(names, str) else names):
            if n not in tags:
                raise IbisError(f"unknown boundary tag {n!r}; tags are {tags}")
            mask |= 1 << tags.index(n)
        return mask

    def compute_dudt_vjp(self, prim, prim_bar, dudt_bar=None, force_tags: int = 0,
                         force_dir=(1.0, 0.0, 0.0), limiter: str = "active", stream=None):
        """prim_bar = (dR/dprim)^T dudt_bar + (dJ/dprim)^T, J the force on the
        boundaries in force_tags along force_dir (ibis_compute_dudt_vjp).
        Ghost rows of prim are written as by compute_dudt."""
        modes = {"frozen": 0, "active": 1}
        if limiter not in modes:
            raise IbisError(f"limiter must be 'frozen' or 'active', not {limiter!r}")
        self._check_state(prim, NPRIM, self.grid.Nt, "prim")
        self._check_state(prim_bar, NPRIM, self.grid.Nt, "prim_bar")
        if dudt_bar is not None:
            self._check_state(dudt_bar, self.ncons, self.grid.Nt, "dudt_bar")
        fd = np.ascontiguousarray(np.asarray(force_dir, dtype=np.float64).reshape(3))
        check(lib().ibis_compute_dudt_vjp(self._h, ptr(prim), ptr(dudt_bar), int(force_tags),
                                          ptr(fd), ptr(prim_bar), modes[limiter],
                                          _stream_handle(stream)), "ibis_compute_dudt_vjp")

    def cons_to_prim_vjp(self, cons, prim_bar, cons_bar, stream=None):
        check(lib().ibis_cons_to_prim_vjp(self._h, ptr(cons), ptr(prim_bar), ptr(cons_bar),
                                          _stream_handle(stream)), "ibis_cons_to_prim_vjp")

    def wall_force(self, prim, force_tags: int, stream=None) -> np.ndarray:
        """Pressure + viscous force on the boundaries in force_tags (x, y, z)."""
        self._check_state(prim, NPRIM, self.grid.Nt, "prim")
        out = np.zeros(3)
        check(lib().ibis_wall_force(self._h, ptr(prim), int(force_tags), ptr(out),
                                    _stream_handle(stream)), "ibis_wall_force")
        return out

    def estimate_dt(self, flow_state, grid=None, gas_model=None, trans_prop=None, stream=None):
        prim = flow_state.data if isinstance(flow_state, FlowStates) else flow_state
        self._check_state(prim, NPRIM, self.grid.Nt, "flow_state")
        out = C.c_double()
        check(lib().ibis_estimate_dt(self._h, ptr(prim), C.byref(out), _stream_handle(stream)),
              "ibis_estimate_dt")
        return out.value

    def count_bad_cells(self, flow_state, num_cells=None, stream=None):
        prim = flow_state.data if isinstance(flow_state, FlowStates) else flow_state
        self._check_state(prim, NPRIM, self.grid.Nt, "flow_state")
        out = C.c_uint64()
        check(lib().ibis_count_bad_cells(self._h, ptr(prim), C.byref(out),
                                         _stream_handle(stream)), "ibis_count_bad_cells")
        return int(out.value)

    def cons_to_prim(self, cons, prim, stream=None):
        check(lib().ibis_cons_to_prim(self._h, ptr(cons), ptr(prim), _stream_handle(stream)),
              "ibis_cons_to_prim")

    def prim_to_cons(self, prim, cons, stream=None):
        check(lib().ibis_prim_to_cons(self._h, ptr(prim), ptr(cons), _stream_handle(stream)),
              "ibis_prim_to_cons")

    def compute_convective_gradient(self, flow_state, grid=None, gas_model=None,
                                    trans_prop=None, stream=None):
        """FiniteVolume<T>::compute_convective_gradient (finite_volume.cpp:283-288):
        boundary actions on ``flow_state``, then the gradients of the
        reconstruction variables into the gradient storage."""
        prim = flow_state.data if isinstance(flow_state, FlowStates) else flow_state
        self._check_state(prim, NPRIM, self.grid.Nt, "flow_state")
        check(lib().ibis_compute_convective_gradient(self._h, ptr(prim), _stream_handle(stream)),
              "ibis_compute_convective_gradient")

    def compute_viscous_gradient(self, flow_state, grid=None, gas_model=None, trans_prop=None,
                                 stream=None):
        """FiniteVolume<T>::compute_viscous_gradient (finite_volume.cpp:273-280):
        boundary actions, then the gradients of T, vx, vy, vz."""
        prim = flow_state.data if isinstance(flow_state, FlowStates) else flow_state
        self._check_state(prim, NPRIM, self.grid.Nt, "flow_state")
        check(lib().ibis_compute_viscous_gradient(self._h, ptr(prim), _stream_handle(stream)),
              "ibis_compute_viscous_gradient")

    def cell_gradients(self):
        """FiniteVolume<T>::cell_gradients() (finite_volume.h:102): the
        gradient storage as a (7, 3, N) float64 CUDA tensor view, slots p,
        rho, T, u, vx, vy, vz (no copy; valid while this object lives)."""
        torch = _torch()
        g = C.c_void_p()
        check(lib().ibis_cell_gradients(self._h, C.byref(g)), "ibis_cell_gradients")
        return _device_view(g.value, (7, 3, self.grid.N))

    def scratch(self):
        """Device pointers (ints) of gradients, limiters and face flux, and
        whether the flux is pre-multiplied by face area."""
        g, lm, fx = C.c_void_p(), C.c_void_p(), C.c_void_p()
        pa = C.c_int()
        check(lib().ibis_residual_scratch(self._h, C.byref(g), C.byref(lm), C.byref(fx),
                                          C.byref(pa)), "ibis_residual_scratch")
        return g.value, lm.value, fx.value, bool(pa.value)

    def set_precision(self, mode: str):
        """"fp64" (default); "fp32_flux": float face values, thermo closure
        and flux calculator with FP64 accumulation; "fp32_recon": also the
        gradients, limiter and reconstruction in float."""
        modes = {"fp64": 0, "fp32_flux": 1, "fp32_recon": 2}
        if mode not in modes:
            raise IbisError(f"unknown precision mode {mode!r}; expected one of {sorted(modes)}")
        check(lib().ibis_residual_set_precision(self._h, modes[mode]), "ibis_residual_set_precision")

    def set_option(self, key: str, value: int):
        """Implementation choice of the kernels (same results for every value)."""
        check(lib().ibis_residual_set_option(self._h, key.encode(), int(value)),
              "ibis_residual_set_option")

    def profile(self, enable: bool = True):
        """Record HIP events around every kernel launch of compute_dudt."""
        check(lib().ibis_residual_profile(self._h, 1 if enable else 0), "ibis_residual_profile")

    def kernel_times(self, reset: bool = True) -> dict:
        """{kernel name: (total ms, launches)} of the profiled launches."""
        from .native import KERNEL_NAMES, NKERNELS
        ms = np.zeros(NKERNELS)
        n = np.zeros(NKERNELS, np.int64)
        check(lib().ibis_residual_kernel_times(self._h, ptr(ms), ptr(n), 1 if reset else 0),
              "ibis_residual_kernel_times")
        return {KERNEL_NAMES[q]: (float(ms[q]), int(n[q])) for q in range(NKERNELS) if n[q] > 0}

    def close(self):
        if getattr(self, "_h", None):
            lib().ibis_residual_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RungeKutta:
    """Explicit RK driver on the device (SO/runge_kutta.cpp:108-203): one
    compute_dudt per stage, one dt reduction per step."""

    def __init__(self, fv: FiniteVolume, method: str = "ssp-rk3", cfl: float = 0.5,
                 max_time: float = 1e30, plot_frequency: float = -1.0, dt_init: float = -1.0):
        self.fv = fv
        a, b = cfgmod.butcher(method)
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        self.cfl = cfl
        self.max_time = max_time
        self.plot_frequency = plot_frequency
        self.t = 0.0
        self.time_since_last_plot = 0.0
        self.dt = dt_init if dt_init > 0 else np.finfo(np.float64).max

    def take_step(self, cons, prim, stream=None) -> float:
        dt_prev = C.c_double(self.dt)
        taken = C.c_double()
        plot_rem = -1.0
        if self.plot_frequency > 0.0 and self.time_since_last_plot < self.plot_frequency:
            plot_rem = self.plot_frequency - self.time_since_last_plot
        c = cons.data if isinstance(cons, ConservedQuantities) else cons
        p = prim.data if isinstance(prim, FlowStates) else prim
        check(lib().ibis_rk_step(self.fv._h, len(self.b), ptr(self.a), ptr(self.b), self.cfl,
                                 ptr(c), ptr(p), C.byref(dt_prev), self.t, self.max_time,
                                 plot_rem, C.byref(taken), _stream_handle(stream)),
              "ibis_rk_step")
        self.dt = dt_prev.value
        self.t += taken.value
        self.time_since_last_plot += taken.value
        return taken.value
