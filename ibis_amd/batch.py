"""Many independent wedge/ramp instances advanced together on one MI355X.

Each instance's grid is built as usual (reference numbering, then locality
order); the grids are concatenated into one mesh with no face between
instances, so one launch of each residual kernel covers every instance.
The explicit RK step (RungeKutta::take_step, runge_kutta.cpp:108-203) keeps a
dt per instance on the device and is captured into a HIP graph once; see
ibis_amd/csrc/batch.hip.  Every instance's numbers are the ones it gets
when stepped alone.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import config as cfgmod
from .finite_volume import GridBlock, fv_desc
from .native import FVDescStruct, IbisError, MeshDescStruct, check, lib, ptr


def _concat_grids(blocks: Sequence[GridBlock]):
    """Mesh arrays of the concatenated grids plus each instance's valid-cell
    and ghost ranges."""
    ex = [b.export() for b in blocks]
    Ns = np.array([b.N for b in blocks], np.int64)
    Gs = np.array([b.G for b in blocks], np.int64)
    Fs = np.array([b.F for b in blocks], np.int64)
    cell_off = np.concatenate([[0], np.cumsum(Ns)])
    ghost_off = np.concatenate([[0], np.cumsum(Gs)])
    face_off = np.concatenate([[0], np.cumsum(Fs)])
    N, G = int(cell_off[-1]), int(ghost_off[-1])

    def cmap(i, c):
        c = np.asarray(c, np.int64)
        return np.where(c < Ns[i], c + cell_off[i], c - Ns[i] + N + ghost_off[i])

    face_lr = np.concatenate([cmap(i, e["face_lr"]) for i, e in enumerate(ex)], axis=1)
    face_geom = np.concatenate([e["face_geom"] for e in ex], axis=1)
    nnz = np.array([e["cell_faces"].size for e in ex], np.int64)
    nnz_off = np.concatenate([[0], np.cumsum(nnz)])
    cell_off_faces = np.concatenate(
        [e["cell_off"][:-1] + nnz_off[i] for i, e in enumerate(ex)] + [[nnz_off[-1]]])
    cell_faces = np.concatenate([e["cell_faces"] + face_off[i] for i, e in enumerate(ex)])
    outsigns = np.concatenate([e["outsigns"] for e in ex]).astype(np.int32)
    nbrs = np.concatenate([cmap(i, e["neighbours"]) for i, e in enumerate(ex)])
    cen = np.concatenate([e["centroid"][:, :Ns[i]] for i, e in enumerate(ex)] +
                         [e["centroid"][:, Ns[i]:] for i, e in enumerate(ex)], axis=1)
    vol = np.concatenate([e["volume"][:Ns[i]] for i, e in enumerate(ex)] +
                         [e["volume"][Ns[i]:] for i, e in enumerate(ex)])
    lsq = np.concatenate([e["lsq"] for e in ex], axis=1)
    tag_off, bfaces = [0], []
    for i, e in enumerate(ex):
        for t in range(len(e["tag_off"]) - 1):
            fs = e["bfaces"][e["tag_off"][t]:e["tag_off"][t + 1]] + face_off[i]
            bfaces.append(fs)
            tag_off.append(tag_off[-1] + fs.size)
    arrays = {
        "face_left": np.ascontiguousarray(face_lr[0]), "face_right": np.ascontiguousarray(face_lr[1]),
        "face_geom": np.ascontiguousarray(face_geom), "cell_face_off": cell_off_faces.astype(np.int64),
        "cell_faces": cell_faces.astype(np.int64), "cell_outsigns": outsigns,
        "cell_neighbours": nbrs.astype(np.int64), "centroid": np.ascontiguousarray(cen),
        "volume": np.ascontiguousarray(vol), "lsq": np.ascontiguousarray(lsq),
        "tag_off": np.array(tag_off, np.int64), "bfaces": np.concatenate(bfaces).astype(np.int64),
    }
    return arrays, N, G, int(face_off[-1]), cell_off, ghost_off


class Batch:
    """A batch of independent instances: (GridIO, FVConfig) pairs that share
    the numerics (flux, reconstruction, limiter, gas) and differ in grid and
    boundary conditions."""

    def __init__(self, instances: Sequence, reorder: bool = True, device_blocks: bool = False):
        self.configs = [c for _, c in instances]
        self.blocks: List[GridBlock] = [GridBlock(g, c, reorder=reorder, device=device_blocks)
                                        for g, c in instances]
        self.dim = self.blocks[0].dim
        if any(b.dim != self.dim for b in self.blocks):
            raise IbisError("batch instances must have the same dimension")
        arr, self.N, self.G, self.F, self.cell_off, self.ghost_off = _concat_grids(self.blocks)
        self.Nt = self.N + self.G
        self.n = len(self.blocks)
        self._arr = arr
        d = MeshDescStruct()
        d.dim, d.ntags = self.dim, len(arr["tag_off"]) - 1
        d.n_cells, d.n_ghost, d.n_faces = self.N, self.G, self.F
        for k, v in arr.items():
            setattr(d, k, ptr(v))
        m = C.c_void_p()
        check(lib().ibis_mesh_create(C.byref(d), C.byref(m)), "ibis_mesh_create (batch)")
        self._mesh = m
        descs = (FVDescStruct * self.n)(*[fv_desc(c) for c in self.configs])
        self._descs = descs
        h = C.c_void_p()
        check(lib().ibis_batch_create(self._mesh, self.n, C.cast(descs, C.c_void_p),
                                      ptr(self.cell_off),
                                      ptr(self.ghost_off), C.byref(h)), "ibis_batch_create")
        self._h = h
        self.ncons = self.dim + 2

    # state of instance i (grid order, N_i + G_i columns) <-> batch columns
    def columns(self, i: int) -> np.ndarray:
        b = self.blocks[i]
        return np.concatenate([np.arange(self.cell_off[i], self.cell_off[i + 1]),
                               self.N + np.arange(self.ghost_off[i], self.ghost_off[i + 1])])

    def gather(self, per_instance: Sequence[np.ndarray]) -> np.ndarray:
        rows = per_instance[0].shape[0]
        out = np.zeros((rows, self.Nt))
        for i, a in enumerate(per_instance):
            out[:, self.columns(i)] = a
        return out

    def scatter(self, a: np.ndarray, i: int) -> np.ndarray:
        return a[:, self.columns(i)]

    def residual_handle(self):
        return C.c_void_p(lib().ibis_batch_residual(self._h))

    def rk_init(self, cons, prim, cfl: float = 0.5, max_time: float = 1e30,
                method: str = "ssp-rk3", dt_init: float = np.finfo(np.float64).max):
        a, b = cfgmod.butcher(method)
        self._a = np.ascontiguousarray(a, np.float64)
        self._b = np.ascontiguousarray(b, np.float64)
        self._state = (cons, prim)
        check(lib().ibis_batch_rk_init(self._h, len(self._b), ptr(self._a), ptr(self._b),
                                       float(cfl), float(max_time), float(dt_init), ptr(cons),
                                       ptr(prim)), "ibis_batch_rk_init")

    def rk_steps(self, nsteps: int, use_graph: bool = True):
        check(lib().ibis_batch_rk_steps(self._h, int(nsteps), 1 if use_graph else 0),
              "ibis_batch_rk_steps")

    def set_option(self, key: str, value: int):
        """A residual option, or the batch option "fuse_stages"."""
        if key == "fuse_stages":
            check(lib().ibis_batch_set_option(self._h, key.encode(), int(value)),
                  "ibis_batch_set_option")
            return
        check(lib().ibis_residual_set_option(self.residual_handle(), key.encode(), int(value)),
              "ibis_residual_set_option")

    def times(self):
        t, dt = np.zeros(self.n), np.zeros(self.n)
        check(lib().ibis_batch_times(self._h, ptr(t), ptr(dt)), "ibis_batch_times")
        return t, dt

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ibis_batch_destroy(self._h)
            self._h = None
        if getattr(self, "_mesh", None) is not None and self._mesh.value:
            lib().ibis_mesh_destroy(self._mesh)
            self._mesh = None
        for b in getattr(self, "blocks", []):
            b.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
