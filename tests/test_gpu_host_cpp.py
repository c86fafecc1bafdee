"""The C++ host layer (ibis_amd/host: GridBlock, FiniteVolume, RungeKutta
over the C ABI) through its compiled driver ibis_amd/lib/ibis_fv_run, on a
reference-style job: the SU2 grid file and the config.json the reference
reads, against the oracle on the same grid and state.

The driver reads the files, builds the grid and the residual, and writes
dU/dt, the state after the boundary actions, dt, the bad-cell count, the
cell gradients after compute_convective_gradient and after
compute_viscous_gradient, and the state after five ssp-rk3 steps.  Bar:
bit for bit for the inviscid cases, 1e-12 relative per conserved variable
for the viscous plate (device and host pow may differ by an ulp).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_pair, rel_err
from ibis_amd import cases
from ibis_amd.config import ssp_rk3
from ibis_amd.gridio import write_su2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ibis_amd", "lib", "ibis_fv_run")
STEPS = 5


def _cases():
    g_w, c_w, f_w = cases.wedge_case(48, 24)
    g_p, c_p, f_p = cases.plate_case(32, 16)
    g_h, c_h = cases.hex_case(10, flux="hanel", limiter="venkatakrishnan")
    return {
        "wedge": (g_w, c_w, cases.perturbed_prim(g_w, f_w, 48 * 24 + 2 * 72, seed=51), True),
        "plate_viscous": (g_p, c_p, cases.perturbed_prim(g_p, f_p, 32 * 16 + 2 * 48, seed=52),
                          False),
        "hex10": (g_h, c_h, cases.hex_prim(g_h, 10 ** 3 + 6 * 100), True),
    }


@pytest.mark.parametrize("name", ["wedge", "plate_viscous", "hex10"])
def test_cpp_driver_matches_oracle(name, tmp_path):
    g_io, cfg, prim, exact = _cases()[name]
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is not built: run __graft_entry__.build()")
    su2, conf, pbin, out = (str(tmp_path / n) for n in ("grid.su2", "config.json", "prim.bin",
                                                       "out"))
    write_su2(g_io, su2)
    with open(conf, "w") as f:
        json.dump(cfg.to_json(), f)
    np.ascontiguousarray(prim, dtype=np.float64).tofile(pbin)
    print(f"{name}: ibis_fv_run", flush=True)
    r = subprocess.run([DRIVER, su2, conf, pbin, out, str(STEPS), "ssp-rk3", "0.5"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip(), flush=True)

    og, ofv = oracle_pair(g_io, cfg)
    n, nt, nc = og.N, og.Nt, og.dim + 2
    p_cpu = prim.copy()
    d_cpu = ofv.compute_dudt(p_cpu)
    dt_cpu = ofv.estimate_dt(p_cpu)
    d_gpu = np.fromfile(out + ".dudt.bin").reshape(nc, nt)
    p_gpu = np.fromfile(out + ".prim.bin").reshape(7, nt)
    lines = open(out + ".txt").read().split()
    if exact:
        assert np.array_equal(d_gpu[:, :n], d_cpu[:, :n])
        assert float(lines[0]) == dt_cpu
    else:
        assert np.all(rel_err(d_gpu, d_cpu, n) <= 1e-12)
        assert float(lines[0]) == pytest.approx(dt_cpu, rel=1e-12)
    assert np.array_equal(p_gpu[:, n:], p_cpu[:, n:])
    assert int(lines[1]) == 0

    # gradients: after compute_dudt, the convective then the viscous
    # post-processing gradients on the same FiniteVolume
    for step, fname in (("convective", ".cgrad.bin"), ("viscous", ".vgrad.bin")):
        getattr(ofv, f"compute_{step}_gradient")(prim.copy())
        want = ofv.intermediates()["grads"]
        got = np.fromfile(out + fname).reshape(7, 3, n)
        if exact:
            assert np.array_equal(got, want), step
        else:
            scale = np.abs(want).max(axis=2, keepdims=True)
            assert np.all(np.abs(got - want) <= 1e-12 * np.where(scale > 0, scale, 1.0)), step

    # five ssp-rk3 steps from the initial state
    a, b = ssp_rk3()
    c_cpu, pp = ofv.prim_to_cons(prim.copy()), prim.copy()
    dt_prev, t, dts = np.finfo(np.float64).max, 0.0, []
    for _ in range(STEPS):
        dt_prev = ofv.rk_step(a, b, 0.5, c_cpu, pp, dt_prev, t, 1e30)
        t += dt_prev
        dts.append(dt_prev)
    c_gpu = np.fromfile(out + ".cons.bin").reshape(nc, nt)
    got_dts = [float(x) for x in lines[2:2 + STEPS]]
    if exact:
        assert got_dts == dts
        assert np.array_equal(c_gpu[:, :n], c_cpu[:, :n])
    else:
        assert np.allclose(got_dts, dts, rtol=1e-12, atol=0)
        assert np.all(rel_err(c_gpu, c_cpu, n) <= 1e-12)
