"""The pybind11 module ``python_api`` (ibis_amd/host/python_api.cpp) and a
torch-facing FiniteVolume over it.

``python_api`` has the reference's binding surface (GasState, PyIdealGas,
Vector3 and the flux calculator classes of
src/python/python/python_api.cpp:14-70) and the MI355X residual: GridBlock
(grid file + grid config) and FiniteVolume (job config), whose methods take
device pointers.  ``PyFiniteVolume`` below wraps them for torch tensors with
the methods the autograd Functions of ibis_amd.autograd call, so that
``autograd.residual`` and ``autograd.wall_force`` run through the pybind
layer as well as through ctypes.
"""
from __future__ import annotations

import importlib.util
import json
import os
from types import SimpleNamespace

import numpy as np

from . import build as _build
from . import config as cfgmod

_mod = None


def python_api():
    """The loaded python_api module.  Raises if it has not been built."""
    global _mod
    if _mod is None:
        # torch first, so that the library binds to the HIP runtime torch loaded
        import torch  # noqa: F401
        path = _build.python_api_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: run ibis_amd.build.build() first")
        spec = importlib.util.spec_from_file_location("python_api", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _mod = mod
    return _mod


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


class PyFiniteVolume:
    """FiniteVolume of python_api on (rows, n_total) float64 CUDA tensors in
    the grid's numbering, built from an SU2 grid file and a job config
    (an FVConfig or the reference's config.json as a dict)."""

    def __init__(self, grid_file: str, config, reorder: bool = True):
        cfg = config.to_json() if isinstance(config, cfgmod.FVConfig) else dict(config)
        api = python_api()
        self._grid = api.GridBlock(grid_file, json.dumps(cfg["grid"]), reorder)
        self._fv = api.FiniteVolume(self._grid, json.dumps(cfg))
        dim = self._grid.dim()
        self.grid = SimpleNamespace(N=self._grid.num_cells(), Nt=self._grid.num_total_cells(),
                                    dim=dim, tags=list(self._grid.boundary_tags()))
        self.ncons = dim + 2
        self.cell_new_of_old = np.asarray(self._grid.cell_new_of_old(), dtype=np.int64)

    def to_grid_order(self, a_ref: np.ndarray) -> np.ndarray:
        out = np.empty_like(a_ref)
        out[..., self.cell_new_of_old] = a_ref
        return out

    def to_reference_order(self, a: np.ndarray) -> np.ndarray:
        return a[..., self.cell_new_of_old]

    def compute_dudt(self, prim, grid=None, dudt=None, allow_reconstruction=True, stream=None):
        import torch
        if dudt is None:
            dudt = torch.zeros((self.ncons, self.grid.Nt), dtype=torch.float64, device=prim.device)
        self._fv.compute_dudt(prim.data_ptr(), dudt.data_ptr(), allow_reconstruction,
                              _stream(stream))
        return dudt

    def estimate_dt(self, prim, stream=None) -> float:
        return self._fv.estimate_dt(prim.data_ptr(), _stream(stream))

    def count_bad_cells(self, prim, stream=None) -> int:
        return self._fv.count_bad_cells(prim.data_ptr(), _stream(stream))

    def tag_mask(self, names) -> int:
        return self._fv.tag_mask([names] if isinstance(names, str) else list(names))

    def compute_dudt_vjp(self, prim, prim_bar, dudt_bar=None, force_tags: int = 0,
                         force_dir=(1.0, 0.0, 0.0), limiter: str = "active", stream=None):
        modes = {"frozen": 0, "active": 1}
        if limiter not in modes:
            raise ValueError(f"limiter must be 'frozen' or 'active', not {limiter!r}")
        self._fv.compute_dudt_vjp(prim.data_ptr(), 0 if dudt_bar is None else dudt_bar.data_ptr(),
                                  int(force_tags), [float(v) for v in force_dir],
                                  prim_bar.data_ptr(), modes[limiter], _stream(stream))

    def wall_force(self, prim, force_tags: int, stream=None) -> np.ndarray:
        return np.array(self._fv.wall_force(prim.data_ptr(), int(force_tags), _stream(stream)))
