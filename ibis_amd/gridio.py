"""Grid description and SU2 text I/O.

``GridIO`` is the in-memory form of an SU2 block, the same information the
reference's ``GridIO`` holds (src/grid/grid/grid_io.h:55): vertex positions,
cells as SU2 element codes with their vertex ids, and boundary markers as
lists of face elements.  Arrays are fixed-width so they can be handed to the
native grid builder without per-element Python objects.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Tuple

import numpy as np

# SU2 element codes (src/grid/grid/grid_io.cpp:28-45)
LINE, TRI, QUAD, TETRA, HEX, WEDGE, PYRAMID = 3, 5, 9, 10, 12, 13, 14
N_VERTICES = {LINE: 2, TRI: 3, QUAD: 4, TETRA: 4, HEX: 8, WEDGE: 6, PYRAMID: 5}


@dataclasses.dataclass
class GridIO:
    dim: int
    xyz: np.ndarray            # (nv, 3) float64; z = 0 in 2D
    ctype: np.ndarray          # (nc,) int32 SU2 element code
    cvtx: np.ndarray           # (nc, 8) int64, padded with -1
    markers: Dict[str, Tuple[np.ndarray, np.ndarray]]  # tag -> (btype (nb,), bvtx (nb, 4))

    @property
    def num_cells(self) -> int:
        return int(self.ctype.shape[0])

    def flat_markers(self, tags):
        """Boundary elements of ``tags`` (in that order) as flat arrays."""
        btag, btype, bvtx = [], [], []
        for t, name in enumerate(tags):
            if name not in self.markers:
                continue
            bt, bv = self.markers[name]
            btag.append(np.full(bt.shape[0], t, dtype=np.int32))
            btype.append(bt.astype(np.int32))
            bvtx.append(bv.astype(np.int64))
        if not btag:
            return (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 4), np.int64))
        return (np.ascontiguousarray(np.concatenate(btag)),
                np.ascontiguousarray(np.concatenate(btype)),
                np.ascontiguousarray(np.concatenate(bvtx)))

    def permute_cells(self, perm: np.ndarray) -> "GridIO":
        """New GridIO whose cell k is this grid's cell perm[k]."""
        return GridIO(self.dim, self.xyz, self.ctype[perm].copy(), self.cvtx[perm].copy(),
                      self.markers)


def _tokens(line: str):
    return line.replace("=", " ").split()


def read_su2(path: str) -> GridIO:
    """Parse an SU2 file (src/grid/grid/grid_io.cpp:218-268).

    Comment lines start with '%'; blank lines are skipped; markers with a
    repeated tag are appended to the first (grid_io.cpp:254-264).
    """
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    lines = [ln for ln in lines if ln and not ln.startswith("%")]
    dim, xyz, ctype, cvtx = 2, None, None, None
    markers: Dict[str, list] = {}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("NDIME"):
            dim = int(_tokens(ln)[1])
            i += 1
        elif ln.startswith("NPOIN"):
            n = int(_tokens(ln)[1])
            xyz = np.zeros((n, 3))
            for k in range(n):
                tok = lines[i + 1 + k].split()
                xyz[k, 0] = float(tok[0])
                xyz[k, 1] = float(tok[1])
                if dim == 3:
                    xyz[k, 2] = float(tok[2])
            i += n + 1
        elif ln.startswith("NELEM"):
            n = int(_tokens(ln)[1])
            ctype = np.zeros(n, np.int32)
            cvtx = np.full((n, 8), -1, np.int64)
            for k in range(n):
                tok = [int(t) for t in lines[i + 1 + k].split()]
                ctype[k] = tok[0]
                nv = N_VERTICES[tok[0]]
                cvtx[k, :nv] = tok[1:1 + nv]
            i += n + 1
        elif ln.startswith("NMARK"):
            nm = int(_tokens(ln)[1])
            i += 1
            for _ in range(nm):
                tag = _tokens(lines[i])[1]
                ne = int(_tokens(lines[i + 1])[1])
                bt = np.zeros(ne, np.int32)
                bv = np.full((ne, 4), -1, np.int64)
                for k in range(ne):
                    tok = [int(t) for t in lines[i + 2 + k].split()]
                    bt[k] = tok[0]
                    nv = N_VERTICES[tok[0]]
                    bv[k, :nv] = tok[1:1 + nv]
                markers.setdefault(tag, []).append((bt, bv))
                i += ne + 2
        else:
            i += 1
    merged = {t: (np.concatenate([a for a, _ in v]), np.concatenate([b for _, b in v]))
              for t, v in markers.items()}
    return GridIO(dim, xyz, ctype, cvtx, merged)


def write_su2(grid: GridIO, path: str) -> None:
    """Write ``grid`` as SU2 text (same layout the reference writes,
    src/grid/grid/grid_io.cpp:270-296)."""
    with open(path, "w") as fh:
        fh.write(f"NDIME= {grid.dim}\n")
        fh.write(f"NELEM= {grid.num_cells}\n")
        for k in range(grid.num_cells):
            nv = N_VERTICES[int(grid.ctype[k])]
            ids = " ".join(str(int(v)) for v in grid.cvtx[k, :nv])
            fh.write(f"{int(grid.ctype[k])} {ids} {k}\n")
        fh.write(f"NPOIN= {grid.xyz.shape[0]}\n")
        for k in range(grid.xyz.shape[0]):
            # repr of a Python float: the shortest string that reads back
            # to the same double
            c = [float(v) for v in grid.xyz[k]]
            if grid.dim == 3:
                fh.write(f"{c[0]!r} {c[1]!r} {c[2]!r} {k}\n")
            else:
                fh.write(f"{c[0]!r} {c[1]!r} {k}\n")
        fh.write(f"NMARK= {len(grid.markers)}\n")
        for tag, (bt, bv) in grid.markers.items():
            fh.write(f"MARKER_TAG= {tag}\n")
            fh.write(f"MARKER_ELEMS= {bt.shape[0]}\n")
            for k in range(bt.shape[0]):
                nv = N_VERTICES[int(bt[k])]
                ids = " ".join(str(int(v)) for v in bv[k, :nv])
                fh.write(f"{int(bt[k])} {ids}\n")
