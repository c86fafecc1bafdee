"""Flow snapshots in the reference's native formats (src/io/io/native.cpp,
src/io/io/io.cpp:32-158), so that states written here load in ibis and the
other way round.

Layout of one snapshot (time index i, padded to four digits):
    <dir>/<iiii>/meta_data.json   {"time": t}
    <dir>/<iiii>/T, p, vx, vy [, vz]   one value per valid cell, reference
                                        cell order
    <dir>/flows                   one line per snapshot written: "<iiii>"
native_binary stores raw FP64 in the host's byte order (x86-64 and the
MI355X hosts: little endian, write_binary in src/io/io/binary_util.h:8-16);
native_text stores one value per line in fixed notation with 16 decimals
(native.cpp:7-90).  Reading fills rho and e from p and T
(update_thermo_from_pT, src/gas/gas/gas_model.h).
"""
from __future__ import annotations

import json
import os
from typing import Tuple

import numpy as np

from . import config as cfgmod

FORMATS = ("native_binary", "native_text")
ROWS = {"rho": 0, "p": 1, "T": 2, "e": 3, "vx": 4, "vy": 5, "vz": 6}


def pad_time_index(idx: int, length: int = 4) -> str:
    """pad_time_index (io.cpp:32-37)."""
    s = str(idx)
    return "0" * (length - min(length, len(s))) + s


def _variables(dim: int):
    return ["T", "p", "vx", "vy"] + (["vz"] if dim == 3 else [])


class FVIO:
    """FVIO<T> (io.cpp:72-158) for the native formats: write() numbers the
    snapshots 0, 1, 2, ...; read(i) loads snapshot i."""

    def __init__(self, flow_format: str = "native_binary", input_dir: str = "io/flow",
                 output_dir: str = "io/flow", time_index: int = 0):
        if flow_format not in FORMATS:
            raise ValueError(f"Unknown flow format {flow_format}")
        self.flow_format = flow_format
        self.input_dir = input_dir
        self.output_dir = output_dir
        self.time_index = time_index

    def write(self, prim, grid, time: float) -> str:
        """prim: 7 x n_total state (numpy or tensor) in grid's cell order."""
        if hasattr(prim, "detach"):
            prim = prim.detach().cpu().numpy()
        ref = grid.to_reference_order(np.asarray(prim, dtype=np.float64))
        n = grid.N
        tdir = pad_time_index(self.time_index)
        d = os.path.join(self.output_dir, tdir)
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "meta_data.json"), "w") as f:
            json.dump({"time": float(time)}, f, indent=4)
        with open(os.path.join(self.output_dir, "flows"), "a") as f:
            f.write(tdir + "\n")
        for name in _variables(grid.dim):
            values = np.ascontiguousarray(ref[ROWS[name], :n], dtype="<f8")
            path = os.path.join(d, name)
            if self.flow_format == "native_binary":
                values.tofile(path)
            else:
                with open(path, "w") as f:
                    f.writelines(f"{v:.16f}\n" for v in values)
        self.time_index += 1
        return d

    def read(self, grid, time_idx: int = 0, gas: cfgmod.IdealGas = None) -> Tuple[np.ndarray, dict]:
        """Snapshot time_idx as a 7 x n_total state in grid's cell order
        (ghost rows zero) and its meta data."""
        gas = gas or cfgmod.IdealGas()
        d = os.path.join(self.input_dir, pad_time_index(time_idx))
        with open(os.path.join(d, "meta_data.json")) as f:
            meta = json.load(f)
        n = grid.N
        ref = np.zeros((7, grid.Nt))
        for name in _variables(grid.dim):
            path = os.path.join(d, name)
            if self.flow_format == "native_binary":
                values = np.fromfile(path, dtype="<f8", count=n)
            else:
                with open(path) as f:
                    values = np.array([float(line) for line in f][:n])
            if values.size != n:
                raise ValueError(f"{path}: expected {n} values, found {values.size}")
            ref[ROWS[name], :n] = values
        # update_thermo_from_pT
        p, T = ref[1, :n], ref[2, :n]
        ref[0, :n] = p / (gas.R * T)
        ref[3, :n] = gas.Cv * T
        return grid.to_grid_order(ref), meta
