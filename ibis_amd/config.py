"""Residual configuration: gas model, transport properties, flux/limiter
choices and boundary conditions.

The JSON this module reads and writes uses the reference's config keys
(``convective_flux``, ``viscous_flux``, ``grid.boundaries``, ``gas_model``,
``transport_properties``; src/finite_volume/finite_volume/finite_volume.cpp:16-43,
convective_flux.cpp:61-82, src/gas/gas/gas_model.h:58-62,
src/gas/gas/transport_properties.h:88-110) and the boundary recipes of the
job DSL (src/ibis/ibis/commands/prep/prep.py:541-614), so a reference
``config/config.json`` can be loaded unchanged.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

FLUX = {"hanel": 0, "ausmdv": 1, "ldfss2": 2, "rusanov": 3}
THERMO = {"rho_p": 0, "rho_T": 1, "rho_u": 2, "p_T": 3}
LIMITER = {"unlimited": 0, "barth_jespersen": 1, "venkatakrishnan": 2}
# boundary action kinds (src/finite_volume/finite_volume/boundaries/boundary.cpp:449-492)
BC_KIND = {
    "flow_state_copy": 1,
    "internal_copy": 2,
    "internal_copy_reflect_normal": 3,
    "internal_vel_copy_reflect": 4,
    "fix_temperature": 5,
    "subsonic_inflow": 6,
    "subsonic_outflow": 7,
    "constant_flux": 8,
}
MAX_TAGS = 16
MAX_ACT = 4
NPAR = 8


@dataclasses.dataclass
class IdealGas:
    """Calorically perfect gas (src/gas/gas/gas_model.h:47-62)."""
    R: float = 287.0
    Cv: float = 2.5 * 287.0
    Cp: float = 3.5 * 287.0
    gamma: float = (3.5 * 287.0) / (2.5 * 287.0)

    @staticmethod
    def from_R(R: float) -> "IdealGas":
        Cv = 5.0 / 2.0 * R
        Cp = 7.0 / 2.0 * R
        return IdealGas(R, Cv, Cp, Cp / Cv)

    def speed_of_sound(self, T: float) -> float:
        return math.sqrt(self.gamma * self.R * T)

    def as_dict(self):
        return {"R": self.R, "Cv": self.Cv, "Cp": self.Cp, "gamma": self.gamma}


@dataclasses.dataclass
class TransportProperties:
    """Sutherland viscosity + constant Prandtl number conductivity
    (src/gas/gas/transport_properties.h:15-80; air values from
    share/species_database/air.json)."""
    mu0: float = 1.716e-5
    T0: float = 273.0
    Ts: float = 110.0
    Pr: float = 0.72

    def viscosity(self, T: float) -> float:
        return self.mu0 * math.pow(T / self.T0, 3.0 / 2.0) * (self.T0 + self.Ts) / (T + self.Ts)

    def as_dict(self):
        return {"viscosity": {"type": "sutherland", "mu_0": self.mu0, "T_0": self.T0,
                              "T_s": self.Ts},
                "thermal_conductivity": {"type": "constant_prandtl_number", "Pr": self.Pr}}


@dataclasses.dataclass
class FlowState:
    T: float
    p: float
    rho: float
    energy: float
    vx: float = 0.0
    vy: float = 0.0
    vz: float = 0.0

    @staticmethod
    def from_pT(p: float, T: float, vx: float = 0.0, vy: float = 0.0, vz: float = 0.0,
                gas: Optional[IdealGas] = None) -> "FlowState":
        gas = gas or IdealGas()
        rho = p / (gas.R * T)          # rho_from_pT, gas_model.h:11-14
        return FlowState(T, p, rho, gas.Cv * T, vx, vy, vz)

    def as_dict(self):
        return {"T": self.T, "p": self.p, "rho": self.rho, "energy": self.energy,
                "vx": self.vx, "vy": self.vy, "vz": self.vz}

    def params(self):
        return [self.T, self.p, self.rho, self.energy, self.vx, self.vy, self.vz]


@dataclasses.dataclass
class BoundaryCondition:
    pre_reconstruction: List[dict]
    post_convective_flux: List[dict] = dataclasses.field(default_factory=list)
    pre_viscous_grad: List[dict] = dataclasses.field(default_factory=list)
    ghost_cells: bool = True

    def as_dict(self):
        return {"ghost_cells": self.ghost_cells, "pre_reconstruction": self.pre_reconstruction,
                "post_convective_flux": self.post_convective_flux,
                "pre_viscous_grad": self.pre_viscous_grad}


def _fsc(fs: FlowState):
    return {"type": "flow_state_copy", "flow_state": fs.as_dict()}


# boundary recipes, src/ibis/ibis/commands/prep/prep.py:541-614
def supersonic_inflow(fs: FlowState) -> BoundaryCondition:
    return BoundaryCondition([_fsc(fs)])


def supersonic_outflow() -> BoundaryCondition:
    return BoundaryCondition([{"type": "internal_copy"}])


def slip_wall() -> BoundaryCondition:
    return BoundaryCondition([{"type": "internal_copy_reflect_normal"}])


def adiabatic_no_slip_wall() -> BoundaryCondition:
    return BoundaryCondition([{"type": "internal_copy_reflect_normal"}], [],
                             [{"type": "internal_vel_copy_reflect"}])


def fixed_temperature_no_slip_wall(temperature: float) -> BoundaryCondition:
    return BoundaryCondition([{"type": "internal_copy_reflect_normal"}], [],
                             [{"type": "internal_vel_copy_reflect"},
                              {"type": "fix_temperature", "temperature": temperature}])


def subsonic_inflow(fs: FlowState) -> BoundaryCondition:
    return BoundaryCondition([{"type": "subsonic_inflow", "flow_state": fs.as_dict()}])


def subsonic_outflow(pressure: float) -> BoundaryCondition:
    return BoundaryCondition([{"type": "subsonic_outflow", "pressure": pressure}])


def constant_flux(fs: FlowState) -> BoundaryCondition:
    return BoundaryCondition([_fsc(fs)], [{"type": "constant_flux", "flow_state": fs.as_dict()}])


def _action_params(a: dict) -> List[float]:
    kind = a["type"]
    if kind not in BC_KIND:
        raise ValueError(f"Unknown boundary action {kind}")
    if kind in ("flow_state_copy", "subsonic_inflow", "constant_flux"):
        f = a["flow_state"]
        return [f["T"], f["p"], f["rho"], f["energy"], f["vx"], f["vy"], f["vz"]]
    if kind == "fix_temperature":
        return [a["temperature"]]
    if kind == "subsonic_outflow":
        return [a["pressure"]]
    return []


@dataclasses.dataclass
class FVConfig:
    flux: str = "hanel"
    ldfss_delta: float = 2.0
    reconstruction_order: int = 2
    thermo_interpolator: str = "rho_u"
    limiter: str = "barth_jespersen"
    epsilon: float = 1e-25
    venkat_K: float = 5.0
    viscous: bool = False
    signal_factor: float = 4.0
    gas: IdealGas = dataclasses.field(default_factory=IdealGas)
    transport: TransportProperties = dataclasses.field(default_factory=TransportProperties)
    boundaries: Dict[str, BoundaryCondition] = dataclasses.field(default_factory=dict)

    def tags(self) -> List[str]:
        """Boundary tags in the order the reference iterates them: a JSON
        object is an ordered map, so sorted byte-wise (src/grid/grid/grid.h:414)."""
        return sorted(self.boundaries.keys(), key=lambda s: s.encode())

    # ---- reference JSON --------------------------------------------------
    def to_json(self) -> dict:
        fc = {"type": self.flux}
        if self.flux == "ldfss2":
            fc["delta"] = self.ldfss_delta
        lim = {"type": self.limiter}
        if self.limiter == "barth_jespersen":
            lim["epsilon"] = self.epsilon
        if self.limiter == "venkatakrishnan":
            lim["K"] = self.venkat_K
        return {
            "convective_flux": {"flux_calculator": fc,
                                "reconstruction_order": self.reconstruction_order,
                                "thermo_interpolator": self.thermo_interpolator,
                                "limiter": lim},
            "viscous_flux": {"enabled": self.viscous, "signal_factor": self.signal_factor},
            "gas_model": self.gas.as_dict(),
            "transport_properties": self.transport.as_dict(),
            "grid": {"boundaries": {k: v.as_dict() for k, v in self.boundaries.items()}},
        }

    @staticmethod
    def from_json(cfg: dict) -> "FVConfig":
        cf = cfg["convective_flux"]
        vf = cfg["viscous_flux"]
        fc = cf["flux_calculator"]
        out = FVConfig(flux=fc["type"], ldfss_delta=fc.get("delta", 2.0),
                       reconstruction_order=int(cf["reconstruction_order"]),
                       viscous=bool(vf["enabled"]), signal_factor=float(vf["signal_factor"]))
        if out.reconstruction_order > 1:
            out.thermo_interpolator = cf["thermo_interpolator"]
            lim = cf["limiter"]
            out.limiter = lim["type"]
            out.epsilon = lim.get("epsilon", 1e-25)
            out.venkat_K = lim.get("K", 5.0)
        gm = cfg.get("gas_model")
        if gm is not None:
            out.gas = IdealGas(gm["R"], gm["Cv"], gm["Cp"], gm["gamma"])
        tp = cfg.get("transport_properties")
        if tp is not None:
            v, k = tp["viscosity"], tp["thermal_conductivity"]
            if v["type"] != "sutherland":
                raise ValueError(f"Unknown viscosity model {v['type']}")
            if k["type"] != "constant_prandtl_number":
                raise ValueError(f"Unknown thermal conductivity model {k['type']}")
            out.transport = TransportProperties(v["mu_0"], v["T_0"], v["T_s"], k["Pr"])
        for tag, b in cfg["grid"]["boundaries"].items():
            out.boundaries[tag] = BoundaryCondition(list(b["pre_reconstruction"]),
                                                    list(b["post_convective_flux"]),
                                                    list(b["pre_viscous_grad"]),
                                                    bool(b.get("ghost_cells", True)))
        return out

    # ---- flat encoding shared by the native library and the oracle --------
    def validate(self):
        if self.flux not in FLUX:
            raise ValueError(f"Unknown flux calculator {self.flux}")
        if self.reconstruction_order not in (1, 2):
            raise ValueError(f"Invalid reconstruction order {self.reconstruction_order}")
        if self.thermo_interpolator not in THERMO:
            raise ValueError(f"Unknown thermo interpolator {self.thermo_interpolator}")
        if self.limiter not in LIMITER:
            raise ValueError(f"Unknown limiter {self.limiter}")
        if len(self.boundaries) > MAX_TAGS:
            raise ValueError(f"at most {MAX_TAGS} boundary tags are supported")

    def encode(self) -> dict:
        """Per-tag boundary action tables (kinds and parameters), tags sorted."""
        self.validate()
        tags = self.tags()
        tabs = {}
        for name, attr in (("pre", "pre_reconstruction"), ("visc", "pre_viscous_grad"),
                           ("post", "post_convective_flux")):
            kind = np.zeros(MAX_TAGS * MAX_ACT, np.int32)
            par = np.zeros(MAX_TAGS * MAX_ACT * NPAR, np.float64)
            for t, tag in enumerate(tags):
                acts = getattr(self.boundaries[tag], attr)
                if len(acts) > MAX_ACT:
                    raise ValueError(f"at most {MAX_ACT} actions per boundary list")
                for a, act in enumerate(acts):
                    pp = _action_params(act)
                    kind[t * MAX_ACT + a] = BC_KIND[act["type"]]
                    par[(t * MAX_ACT + a) * NPAR:(t * MAX_ACT + a) * NPAR + len(pp)] = pp
            tabs[name] = (kind, par)
        ghost = np.array([1 if self.boundaries[t].ghost_cells else 0 for t in tags], np.int32)
        return {"tags": tags, "ghost": ghost, **tabs}


def ssp_rk3():
    """ssp-rk3 Butcher tableau (src/ibis/ibis/commands/prep/prep.py:852-853)."""
    a = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.25, 0.0]])
    b = np.array([1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0])
    return a, b


def butcher(method: str):
    """Butcher tableaux of the job DSL (prep.py:845-859) as (a, b) with
    a[i, j] the coefficient of stage j in stage i."""
    if method == "ssp-rk3":
        return ssp_rk3()
    if method == "euler":
        return np.zeros((1, 1)), np.array([1.0])
    if method == "midpoint":
        return np.array([[0.0, 0.0], [0.5, 0.0]]), np.array([0.0, 1.0])
    if method == "rk3":
        return (np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [-1.0, 2.0, 0.0]]),
                np.array([1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0]))
    if method == "rk4":
        return (np.array([[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]], float),
                np.array([1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]))
    raise ValueError(f"Unknown method {method}")
