"""The residual and its adjoint through the pybind11 module python_api
(ibis_amd.pyapi.PyFiniteVolume, built from an SU2 file and the job config),
against the oracle and against the ctypes path, bit for bit; and the
torch.autograd.Function wrappers over it."""
import numpy as np
import pytest

from helpers import gpu_objects, oracle_pair, rel_err, to_dev
from ibis_amd import cases
from ibis_amd.gridio import write_su2

pytestmark = pytest.mark.gpu


def _pyfv(g_io, cfg, tmp_path):
    from ibis_amd.pyapi import PyFiniteVolume
    path = str(tmp_path / "grid.su2")
    write_su2(g_io, path)
    return PyFiniteVolume(path, cfg)


@pytest.mark.parametrize("case", ["wedge", "plate_viscous"])
def test_python_api_residual_matches_oracle(case, tmp_path):
    import torch
    if case == "wedge":
        g_io, cfg, fs = cases.wedge_case(48, 24)
        prim = cases.perturbed_prim(g_io, fs, 48 * 24 + 2 * 72, seed=71)
    else:
        g_io, cfg, fs = cases.plate_case(32, 16)
        prim = cases.perturbed_prim(g_io, fs, 32 * 16 + 2 * 48, seed=72)
    fv = _pyfv(g_io, cfg, tmp_path)
    og, ofv = oracle_pair(g_io, cfg)
    p_cpu = prim.copy()
    d_cpu = ofv.compute_dudt(p_cpu)
    p = to_dev(fv.to_grid_order(prim))
    d = fv.compute_dudt(p)
    torch.cuda.synchronize()
    d_gpu = fv.to_reference_order(d.cpu().numpy())
    n = og.N
    if case == "wedge":
        assert np.array_equal(d_gpu[:, :n], d_cpu[:, :n])
        assert fv.estimate_dt(p) == ofv.estimate_dt(p_cpu)
    else:
        assert np.all(rel_err(d_gpu, d_cpu, n) <= 1e-12)
    assert np.array_equal(fv.to_reference_order(p.cpu().numpy())[:, n:], p_cpu[:, n:])
    assert fv.count_bad_cells(p) == 0


def test_python_api_autograd_equals_ctypes_path(tmp_path):
    """wall_force and residual as torch.autograd.Functions over the pybind
    FiniteVolume: values and gradients identical to the ctypes FiniteVolume's."""
    import torch
    from ibis_amd.autograd import residual, wall_force
    g_io, cfg, fs = cases.plate_case(32, 16)
    prim_ref = cases.perturbed_prim(g_io, fs, 32 * 16 + 2 * (32 + 16), seed=73)
    pyfv = _pyfv(g_io, cfg, tmp_path)
    gb, fv = gpu_objects(g_io, cfg)
    assert np.array_equal(pyfv.cell_new_of_old, gb.cell_new_of_old)
    prim0 = to_dev(gb.to_grid_order(prim_ref))
    grads = []
    for f in (pyfv, fv):
        x = prim0.clone().requires_grad_(True)
        J = wall_force(x, f, "wall", (0.3, 1.0, 0.0))
        J.backward()
        y = prim0.clone().requires_grad_(True)
        d = residual(y, f, limiter="frozen")
        w = torch.randn(d.shape, dtype=torch.float64, device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(9))
        w[:, gb.N:] = 0.0
        (d * w).sum().backward()
        grads.append((float(J.detach()), x.grad.clone(), y.grad.clone(), d.detach().clone()))
    a, b = grads
    assert a[0] == b[0]
    for k in (1, 2, 3):
        assert torch.equal(a[k], b[k])
    fv.close()
    gb.close()
