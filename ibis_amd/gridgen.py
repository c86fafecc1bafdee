"""Seeded synthetic grids and flow states for the residual workloads.

All randomness comes from one generator, splitmix64 mapped to a double in
[0, 1) as ``(x >> 11) * 2**-53``, so the same seed gives the same grid in any
language.  Grids are structured maps written as unstructured SU2 blocks.

Workloads:
  * ``wedge``  -- 2D 15-degree wedge (geometry of examples/wedge/2D/transient/
    grid.py:13-21 in the reference), quads, tags inflow/outflow/wall.
  * ``ramp``   -- 2D compression ramp of the same family.
  * ``flat_plate`` -- 2D plate of length 1 m and height 0.25 m (examples/
    flat_plate/2D/full_plate/transient/grid.py), geometric wall clustering
    with ratio 1.05 capped at the uniform far-field spacing.
  * ``hex_box`` -- unit cube of n^3 hexes, interior nodes jittered by up to
    +-10% of the spacing, cells stored in a seeded random order.
"""
from __future__ import annotations

import math

import numpy as np

from .gridio import HEX, LINE, PYRAMID, QUAD, TRI, GridIO

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive splitmix64 outputs (uint64) for ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def seeded_permutation(seed: int, n: int) -> np.ndarray:
    """Permutation of range(n): argsort of n splitmix64 draws (stable)."""
    return np.argsort(splitmix64(seed, n), kind="stable").astype(np.int64)


def _quad_block(X: np.ndarray, Y: np.ndarray):
    """Quad cells of an (nx+1, ny+1) vertex map, row-major (j outer)."""
    nxp, nyp = X.shape
    nx, ny = nxp - 1, nyp - 1
    vid = lambda i, j: j * nxp + i  # noqa: E731
    xyz = np.zeros((nxp * nyp, 3))
    xyz[:, 0] = X.T.reshape(-1)
    xyz[:, 1] = Y.T.reshape(-1)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    I, J = I.reshape(-1), J.reshape(-1)
    cvtx = np.full((nx * ny, 8), -1, np.int64)
    cvtx[:, 0] = vid(I, J)
    cvtx[:, 1] = vid(I + 1, J)
    cvtx[:, 2] = vid(I + 1, J + 1)
    cvtx[:, 3] = vid(I, J + 1)
    ctype = np.full(nx * ny, QUAD, np.int32)
    return xyz, ctype, cvtx, vid


def _lines(a: np.ndarray, b: np.ndarray):
    bv = np.full((a.shape[0], 4), -1, np.int64)
    bv[:, 0] = a
    bv[:, 1] = b
    return np.full(a.shape[0], LINE, np.int32), bv


def _cat(*parts):
    return (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]))


def wedge(nx: int = 192, ny: int = 96, angle_deg: float = 15.0, inflow_length: float = 0.05,
          length: float = 0.9, height: float = 0.6) -> GridIO:
    """Structured quad map of the reference's 2D wedge domain.

    Bottom wall: flat for x < inflow_length, then a ramp at ``angle_deg``;
    top boundary: the straight line from (0, height/4) to (length, height).
    """
    n1 = max(1, int(round(nx * inflow_length / length)))
    x = np.concatenate([np.linspace(0.0, inflow_length, n1 + 1)[:-1],
                        np.linspace(inflow_length, length, nx - n1 + 1)])
    ang = math.radians(angle_deg)
    yb = np.where(x <= inflow_length, 0.0,
                  (x - inflow_length) / (length - inflow_length) *
                  ((length - inflow_length) * math.sin(ang)))
    yt = height / 4 + (height - height / 4) * x / length
    s = np.linspace(0.0, 1.0, ny + 1)
    X = np.repeat(x[:, None], ny + 1, axis=1)
    Y = yb[:, None] + (yt - yb)[:, None] * s[None, :]
    xyz, ctype, cvtx, vid = _quad_block(X, Y)
    i = np.arange(nx)
    j = np.arange(ny)
    markers = {
        "inflow": _lines(vid(0, j + 1), vid(0, j)),
        "outflow": _cat(_lines(vid(nx, j), vid(nx, j + 1)),
                        _lines(vid(i + 1, ny), vid(i, ny))),
        "wall": _lines(vid(i, 0), vid(i + 1, 0)),
    }
    return GridIO(2, xyz, ctype, cvtx, markers)


def ramp(nx: int, ny: int, angle_deg: float, length: float = 1.0, height: float = 0.5,
         ramp_start: float = 0.4) -> GridIO:
    """Compression ramp: flat wall to ``ramp_start``, then a ramp at
    ``angle_deg``; flat top at ``height``.  Tags inflow/outflow/wall."""
    n1 = max(1, int(round(nx * ramp_start / length)))
    x = np.concatenate([np.linspace(0.0, ramp_start, n1 + 1)[:-1],
                        np.linspace(ramp_start, length, nx - n1 + 1)])
    yb = np.where(x <= ramp_start, 0.0, (x - ramp_start) * math.tan(math.radians(angle_deg)))
    yt = np.full_like(x, height)
    s = np.linspace(0.0, 1.0, ny + 1)
    X = np.repeat(x[:, None], ny + 1, axis=1)
    Y = yb[:, None] + (yt - yb)[:, None] * s[None, :]
    xyz, ctype, cvtx, vid = _quad_block(X, Y)
    i = np.arange(nx)
    j = np.arange(ny)
    markers = {
        "inflow": _lines(vid(0, j + 1), vid(0, j)),
        "outflow": _cat(_lines(vid(nx, j), vid(nx, j + 1)),
                        _lines(vid(i + 1, ny), vid(i, ny))),
        "wall": _lines(vid(i, 0), vid(i + 1, 0)),
    }
    return GridIO(2, xyz, ctype, cvtx, markers)


def plate_spacing(ny: int, height: float, ratio: float = 1.05, n_stretch: int = 64):
    """Wall-normal node positions: dy_k = dy0 * ratio**k for k < n_stretch,
    then constant dy0 * ratio**n_stretch, scaled to fill ``height``."""
    k = np.arange(ny)
    w = np.where(k < n_stretch, ratio ** k, ratio ** n_stretch)
    dy = w * (height / w.sum())
    y = np.concatenate([[0.0], np.cumsum(dy)])
    y[-1] = height
    return y


def flat_plate(nx: int = 1024, ny: int = 512, length: float = 1.0, height: float = 0.25,
               ratio: float = 1.05, n_stretch: int = 64) -> GridIO:
    x = np.linspace(0.0, length, nx + 1)
    y = plate_spacing(ny, height, ratio, n_stretch)
    X = np.repeat(x[:, None], ny + 1, axis=1)
    Y = np.repeat(y[None, :], nx + 1, axis=0)
    xyz, ctype, cvtx, vid = _quad_block(X, Y)
    i = np.arange(nx)
    j = np.arange(ny)
    markers = {
        "inflow": _lines(vid(0, j + 1), vid(0, j)),
        "outflow": _cat(_lines(vid(nx, j), vid(nx, j + 1)),
                        _lines(vid(i + 1, ny), vid(i, ny))),
        "wall": _lines(vid(i, 0), vid(i + 1, 0)),
    }
    return GridIO(2, xyz, ctype, cvtx, markers)


def hex_box(n: int = 160, jitter: float = 0.1, jitter_seed: int = 3, perm_seed: int = 4,
            permute: bool = True) -> GridIO:
    """Unit cube of n^3 hexes.  Interior nodes move by up to ``jitter`` of the
    spacing on each axis; nodes on a boundary plane keep that coordinate.
    Cells are stored in a seeded random order (``perm_seed``)."""
    nv1 = n + 1
    h = 1.0 / n
    g = np.arange(nv1) * h
    g[-1] = 1.0
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")  # vertex id = i + nv1*(j + nv1*k)
    xyz = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1)
    nvt = xyz.shape[0]
    u = uniform01(jitter_seed, 3 * nvt).reshape(nvt, 3)
    d = (2.0 * u - 1.0) * (jitter * h)
    for ax in range(3):
        c = xyz[:, ax]
        on_plane = (c == 0.0) | (c == 1.0)
        xyz[:, ax] = np.where(on_plane, c, c + d[:, ax])
    vid = lambda i, j, k: i + nv1 * (j + nv1 * k)  # noqa: E731
    K, J, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    I, J, K = I.reshape(-1), J.reshape(-1), K.reshape(-1)
    cvtx = np.empty((n ** 3, 8), np.int64)
    cvtx[:, 0] = vid(I, J, K)
    cvtx[:, 1] = vid(I + 1, J, K)
    cvtx[:, 2] = vid(I + 1, J + 1, K)
    cvtx[:, 3] = vid(I, J + 1, K)
    cvtx[:, 4] = vid(I, J, K + 1)
    cvtx[:, 5] = vid(I + 1, J, K + 1)
    cvtx[:, 6] = vid(I + 1, J + 1, K + 1)
    cvtx[:, 7] = vid(I, J + 1, K + 1)
    ctype = np.full(n ** 3, HEX, np.int32)
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)

    def quads(v0, v1, v2, v3):
        bv = np.stack([v0, v1, v2, v3], axis=1).astype(np.int64)
        return np.full(bv.shape[0], QUAD, np.int32), bv

    markers = {
        "west": quads(vid(0, a, b), vid(0, a + 1, b), vid(0, a + 1, b + 1), vid(0, a, b + 1)),
        "east": quads(vid(n, a, b), vid(n, a + 1, b), vid(n, a + 1, b + 1), vid(n, a, b + 1)),
        "south": quads(vid(a, 0, b), vid(a + 1, 0, b), vid(a + 1, 0, b + 1), vid(a, 0, b + 1)),
        "north": quads(vid(a, n, b), vid(a + 1, n, b), vid(a + 1, n, b + 1), vid(a, n, b + 1)),
        "bottom": quads(vid(a, b, 0), vid(a + 1, b, 0), vid(a + 1, b + 1, 0), vid(a, b + 1, 0)),
        "top": quads(vid(a, b, n), vid(a + 1, b, n), vid(a + 1, b + 1, n), vid(a, b + 1, n)),
    }
    grid = GridIO(3, xyz, ctype, cvtx, markers)
    if permute:
        grid = grid.permute_cells(seeded_permutation(perm_seed, n ** 3))
    return grid


def split_quads(grid: GridIO, every: int = 2) -> GridIO:
    """2D: every ``every``-th quad (storage order) split into the triangles
    (0, 1, 2) and (0, 2, 3) along its diagonal.  The mesh stays conforming
    and the boundary edges do not change: a mixed triangle / quad mesh whose
    cells have 3 or 4 faces."""
    ctype, cvtx = [], []
    for c in range(grid.num_cells):
        v = grid.cvtx[c]
        if grid.ctype[c] == QUAD and c % every == 0:
            for tri in ((v[0], v[1], v[2]), (v[0], v[2], v[3])):
                ctype.append(TRI)
                cvtx.append(list(tri) + [-1] * 5)
        else:
            ctype.append(int(grid.ctype[c]))
            cvtx.append(list(v))
    return GridIO(grid.dim, grid.xyz.copy(), np.array(ctype, np.int32),
                  np.array(cvtx, np.int64), dict(grid.markers))


# each hex face (VTK local vertex order, as hex_box writes the cells) listed
# inward, i.e. counter-clockwise seen from the hex centre: the base of the
# pyramid (base, centre) over that face
_HEX_FACES_INWARD = ((1, 2, 3, 0), (7, 6, 5, 4), (4, 5, 1, 0), (5, 6, 2, 1), (6, 7, 3, 2),
                     (7, 4, 0, 3))


def split_hexes(grid: GridIO, every: int = 3) -> GridIO:
    """3D: every ``every``-th hex (storage order) split into six pyramids
    with their apex at the mean of its vertices (a new vertex).  Each
    pyramid's base is one of the hex's faces, so the mesh stays conforming
    with the neighbouring hexes and the boundary faces do not change: a mixed
    hex / pyramid mesh whose cells have 5 or 6 faces, quad and triangle."""
    xyz = [grid.xyz]
    nv = grid.xyz.shape[0]
    ctype, cvtx = [], []
    for c in range(grid.num_cells):
        v = grid.cvtx[c]
        if grid.ctype[c] == HEX and c % every == 0:
            xyz.append(grid.xyz[v[:8]].mean(axis=0)[None, :])
            apex = nv
            nv += 1
            for f in _HEX_FACES_INWARD:
                ctype.append(PYRAMID)
                cvtx.append([v[k] for k in f] + [apex, -1, -1, -1])
        else:
            ctype.append(int(grid.ctype[c]))
            cvtx.append(list(v))
    return GridIO(grid.dim, np.concatenate(xyz), np.array(ctype, np.int32),
                  np.array(cvtx, np.int64), dict(grid.markers))


def cell_centres(grid: GridIO) -> np.ndarray:
    """Vertex-average cell centres (the reference's centroid definition)."""
    nv = np.array([{3: 2, 5: 3, 9: 4, 10: 4, 12: 8, 13: 6, 14: 5}[int(t)]
                   for t in np.unique(grid.ctype)])
    if nv.size != 1:
        out = np.zeros((grid.num_cells, 3))
        for k in range(grid.num_cells):
            ids = grid.cvtx[k][grid.cvtx[k] >= 0]
            out[k] = grid.xyz[ids].mean(axis=0)
        return out
    ids = grid.cvtx[:, :int(nv[0])]
    return grid.xyz[ids].sum(axis=1) / float(nv[0])


def smooth_field(centres: np.ndarray, seed: int, dim: int):
    """Smooth seeded primitive field: rho, p, T, e, vx, vy, vz (ideal air).

    rho and p are O(1) modulations of (1.2 kg/m^3, 1e5 Pa); the velocity
    is a sum of three sinusoids per component with seeded phases, scaled so
    that the local Mach number spans roughly [0.3, 3]."""
    R, Cv, gamma = 287.0, 717.5, 1.4
    u = uniform01(seed, 64)
    x, y, z = centres[:, 0], centres[:, 1], centres[:, 2]

    def wave(k0):
        s = np.zeros_like(x)
        for m in range(3):
            kx, ky, kz = 1.0 + m, 2.0 - 0.5 * m, 0.5 + m
            ph = 2 * np.pi * u[k0 + m]
            s += np.sin(2 * np.pi * (kx * x + ky * y + (kz * z if dim == 3 else 0.0)) + ph) / 3.0
        return s

    rho = 1.2 * (1.0 + 0.2 * wave(0))
    p = 1.0e5 * (1.0 + 0.2 * wave(3))
    T = p / (rho * R)
    a = np.sqrt(gamma * R * T)
    mach = 1.65 + 1.35 * wave(6)                 # in [0.3, 3.0]
    dirx, diry = wave(9), wave(12)
    dirz = wave(15) if dim == 3 else np.zeros_like(x)
    norm = np.sqrt(dirx ** 2 + diry ** 2 + dirz ** 2) + 1e-3
    speed = mach * a
    vx, vy, vz = speed * dirx / norm, speed * diry / norm, speed * dirz / norm
    e = Cv * T
    return np.stack([rho, p, T, e, vx, vy, vz])
