"""The pybind11 module python_api (CPU side): the reference's binding
surface (src/python/python/python_api.cpp:14-70) -- GasState, PyIdealGas and
its thermodynamic closures (src/gas/gas/gas_model.h:11-40), Vector3, the flux
calculator names -- and error reporting.  The GPU residual through it is in
test_gpu_python_api.py."""
import math

import pytest

from ibis_amd.pyapi import python_api


def test_ideal_gas_closures():
    m = python_api()
    g = m.PyIdealGas(287.0)
    assert (g.R(), g.Cv(), g.Cp(), g.gamma()) == (287.0, 2.5 * 287.0, 3.5 * 287.0, 1.4)
    s = m.GasState()
    s.p, s.T = 101325.0, 300.0
    g.update_thermo_from_pT(s)
    assert s.rho == 101325.0 / (287.0 * 300.0) and s.energy == 717.5 * 300.0
    assert g.speed_of_sound(s) == math.sqrt(1.4 * 287.0 * 300.0)
    s.rho, s.T = 1.2, 310.0
    g.update_thermo_from_rhoT(s)
    assert s.p == 1.2 * 287.0 * 310.0 and s.energy == 717.5 * 310.0
    s.rho, s.p = 0.9, 8.0e4
    g.update_thermo_from_rhop(s)
    assert s.T == 8.0e4 / (0.9 * 287.0) and s.energy == 717.5 * s.T
    s.rho, s.energy = 1.1, 2.0e5
    g.update_thermo_from_rhou(s)
    assert s.T == 2.0e5 / 717.5 and s.p == 1.1 * 287.0 * s.T


def test_vector3_and_flux_names():
    m = python_api()
    v = m.Vector3(1.0, 2.0)
    assert (v.x, v.y, v.z) == (1.0, 2.0, 0.0)
    v.z = 3.0
    assert m.Vector3(4.0, 5.0, 6.0).z == 6.0 and v.z == 3.0
    assert [c().name() for c in (m.PyHanel, m.PyAusmdv, m.PyLdfss2, m.PyRusanov)] == \
        ["hanel", "ausmdv", "ldfss2", "rusanov"]


def test_errors_are_python_exceptions(tmp_path):
    m = python_api()
    with pytest.raises(RuntimeError, match="cannot open"):
        m.GridBlock(str(tmp_path / "missing.su2"), '{"boundaries": {}}')
    with pytest.raises(RuntimeError, match="JSON"):
        m.GridBlock(str(tmp_path / "missing.su2"), '{"boundaries": ')
