"""The residual workloads as (GridIO, FVConfig, primitive state) triples.

  wedge_case   -- 2D inviscid 15-degree wedge, Mach 3 air (p=101325 Pa,
                  T=300 K), AUSMDV, 2nd order, Barth-Jespersen, rho_u.
  plate_case   -- 2D viscous flat plate, Mach 2, Re_L=1e5 at T=300 K,
                  Sutherland + Pr=0.72, no-slip isothermal wall.
  hex_case     -- 3D jittered, cell-permuted hex box, Hanel, 2nd order,
                  Venkatakrishnan, all faces supersonic outflow.
The primitive state is laid out as 7 x N_total rows (rho, p, T, e, vx, vy,
vz); ghost rows are zero and are filled by the boundary actions.
"""
from __future__ import annotations

import math

import numpy as np

from . import config as C
from . import gridgen


def freestream(mach: float, p: float = 101325.0, T: float = 300.0, gas=None):
    gas = gas or C.IdealGas()
    a = gas.speed_of_sound(T)
    return C.FlowState.from_pT(p, T, vx=mach * a, gas=gas)


def uniform_prim(fs: C.FlowState, n_valid: int, n_total: int) -> np.ndarray:
    prim = np.zeros((7, n_total))
    prim[:, :n_valid] = np.array([fs.rho, fs.p, fs.T, fs.energy, fs.vx, fs.vy, fs.vz])[:, None]
    return prim


def perturbed_prim(grid_io, fs: C.FlowState, n_total: int, seed: int, amp: float = 0.05,
                   gas=None) -> np.ndarray:
    """Free stream times a smooth seeded modulation of rho, T and velocity."""
    gas = gas or C.IdealGas()
    cc = gridgen.cell_centres(grid_io)
    u = gridgen.uniform01(seed, 16)
    x, y, z = cc[:, 0], cc[:, 1], cc[:, 2]
    L = max(np.ptp(x), np.ptp(y), 1e-30)

    def wave(k):
        return np.sin(2 * np.pi * ((1 + u[k]) * x / L + (1 + u[k + 1]) * y / L +
                                   u[k + 2] * z / L) + 2 * np.pi * u[k + 3])

    n = grid_io.num_cells
    rho = fs.rho * (1.0 + amp * wave(0))
    T = fs.T * (1.0 + amp * wave(4))
    speed = math.hypot(fs.vx, fs.vy) + 1e-30
    vx = fs.vx * (1.0 + amp * wave(8)) + amp * speed * wave(12) * 0.5
    vy = fs.vy + amp * speed * wave(2)
    prim = np.zeros((7, n_total))
    prim[0, :n] = rho
    prim[2, :n] = T
    prim[1, :n] = rho * gas.R * T
    prim[3, :n] = gas.Cv * T
    prim[4, :n] = vx
    prim[5, :n] = vy
    prim[6, :n] = fs.vz
    return prim


def wedge_config(fs: C.FlowState, flux: str = "ausmdv", limiter: str = "barth_jespersen",
                 order: int = 2, thermo: str = "rho_u") -> C.FVConfig:
    return C.FVConfig(flux=flux, reconstruction_order=order, thermo_interpolator=thermo,
                      limiter=limiter, epsilon=1e-25,
                      boundaries={"inflow": C.supersonic_inflow(fs),
                                  "outflow": C.supersonic_outflow(),
                                  "wall": C.slip_wall()})


def wedge_case(nx: int = 192, ny: int = 96, mach: float = 3.0, angle: float = 15.0, **kw):
    fs = freestream(mach)
    return gridgen.wedge(nx, ny, angle_deg=angle), wedge_config(fs, **kw), fs


def ramp_case(nx: int, ny: int, mach: float, angle: float, **kw):
    fs = freestream(mach)
    return gridgen.ramp(nx, ny, angle_deg=angle), wedge_config(fs, **kw), fs


def plate_freestream(mach: float = 2.0, Re: float = 1e5, T: float = 300.0, L: float = 1.0):
    gas = C.IdealGas()
    tp = C.TransportProperties()
    u = mach * gas.speed_of_sound(T)
    mu = tp.viscosity(T)
    rho = Re * mu / (u * L)
    p = rho * gas.R * T
    return C.FlowState(T, p, rho, gas.Cv * T, u, 0.0, 0.0)


def plate_config(fs: C.FlowState, flux: str = "ausmdv") -> C.FVConfig:
    return C.FVConfig(flux=flux, reconstruction_order=2, thermo_interpolator="rho_u",
                      limiter="barth_jespersen", epsilon=1e-25, viscous=True, signal_factor=4.0,
                      boundaries={"inflow": C.supersonic_inflow(fs),
                                  "outflow": C.supersonic_outflow(),
                                  "wall": C.fixed_temperature_no_slip_wall(300.0)})


def plate_case(nx: int = 1024, ny: int = 512, **kw):
    fs = plate_freestream()
    return gridgen.flat_plate(nx, ny), plate_config(fs, **kw), fs


def hex_config(flux: str = "hanel", limiter: str = "venkatakrishnan") -> C.FVConfig:
    b = {t: C.supersonic_outflow() for t in ("bottom", "east", "north", "south", "top", "west")}
    return C.FVConfig(flux=flux, reconstruction_order=2, thermo_interpolator="rho_u",
                      limiter=limiter, epsilon=1e-25, venkat_K=5.0, boundaries=b)


def hex_case(n: int = 160, permute: bool = True, **kw):
    return gridgen.hex_box(n, permute=permute), hex_config(**kw)


def hex_prim(grid_io, n_total: int, seed: int = 5) -> np.ndarray:
    prim = np.zeros((7, n_total))
    cc = gridgen.cell_centres(grid_io)
    prim[:, :grid_io.num_cells] = gridgen.smooth_field(cc, seed, 3)
    return prim


def batch_instances(n: int = 64, seed: int = 7, n_min: int = 10_000, n_max: int = 400_000):
    """Seeded wedge/ramp instances of uneven size: cell count U[n_min, n_max]
    with nx:ny = 2:1, half wedges and half compression ramps, angle
    U[5, 20] degrees, Mach U[1.5, 4] at p = 101325 Pa, T = 300 K."""
    u = gridgen.uniform01(seed, 3 * n)
    out = []
    for i in range(n):
        cells = n_min + u[3 * i] * (n_max - n_min)
        ny = max(4, int(round(math.sqrt(cells / 2.0))))
        out.append({"kind": "wedge" if i % 2 == 0 else "ramp", "nx": 2 * ny, "ny": ny,
                    "angle": 5.0 + 15.0 * u[3 * i + 1], "mach": 1.5 + 2.5 * u[3 * i + 2]})
    return out


def batch_case(spec: dict):
    """(GridIO, FVConfig, FlowState) of one batch instance: AUSMDV, 2nd
    order, Barth-Jespersen, rho_u, FP64."""
    make = wedge_case if spec["kind"] == "wedge" else ramp_case
    return make(spec["nx"], spec["ny"], mach=spec["mach"], angle=spec["angle"])
