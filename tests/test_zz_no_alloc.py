"""ibis_rk_step_noalloc (stage buffers in the library's fixed scratch array)
against ibis_rk_step (stage buffers allocated on the first call), through
the C ABI: the same states, primitive rows and dt, bit for bit, on two
meshes in a row, the second smaller than the first, so that the second
call finds the first call's contents in the scratch.  A mesh whose stage
buffers exceed the scratch is refused.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import gpu_objects, to_dev
from ibis_amd import cases
from ibis_amd.config import butcher

pytestmark = pytest.mark.gpu


def _run(entry, nx, ny, method, nsteps, seed):
    import torch
    from ibis_amd.native import check, lib, ptr
    g_io, cfg, fs = cases.wedge_case(nx, ny)
    gb, fv = gpu_objects(g_io, cfg)
    prim = to_dev(gb.to_grid_order(cases.perturbed_prim(g_io, fs, gb.Nt, seed=seed)))
    cons = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    fv.prim_to_cons(prim, cons)
    a, b = butcher(method)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    dt_prev, taken, t, dts = C.c_double(np.finfo(np.float64).max), C.c_double(), 0.0, []
    fn = getattr(lib(), entry)
    print(f"{entry}: wedge {nx}x{ny}, {method}, {nsteps} steps", flush=True)
    for _ in range(nsteps):
        check(fn(fv._h, len(b), ptr(a), ptr(b), 0.5, ptr(cons), ptr(prim), C.byref(dt_prev), t,
                 1e30, -1.0, C.byref(taken), None), entry)
        t += taken.value
        dts.append(taken.value)
    torch.cuda.synchronize()
    out = (cons.cpu().numpy()[:, :gb.N], prim.cpu().numpy()[:, :gb.N], dts)
    fv.close()
    gb.close()
    return out


@pytest.mark.parametrize("nx,ny,method,nsteps,seed", [(96, 48, "ssp-rk3", 4, 21),
                                                     (40, 20, "rk4", 3, 22)])
def test_noalloc_rk_step_equals_rk_step(nx, ny, method, nsteps, seed):
    """Run in this order: the 96 x 48 wedge first (ssp-rk3), then the smaller
    40 x 20 wedge (rk4, four stage buffers)."""
    want = _run("ibis_rk_step", nx, ny, method, nsteps, seed)
    got = _run("ibis_rk_step_noalloc", nx, ny, method, nsteps, seed)
    assert got[2] == want[2]
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1])


def test_noalloc_rk_step_refuses_meshes_above_the_scratch():
    """The 1024 x 512 plate with ssp-rk3 needs 4 x 4 + 7 doubles per cell
    (12.1M doubles, 97 MB), above the 64 MiB scratch."""
    import torch
    from ibis_amd.native import IbisError, check, lib, ptr
    g_io, cfg, fs = cases.plate_case(1024, 512)
    gb, fv = gpu_objects(g_io, cfg)
    prim = torch.zeros((7, gb.Nt), dtype=torch.float64, device="cuda")
    cons = torch.zeros((4, gb.Nt), dtype=torch.float64, device="cuda")
    a, b = butcher("ssp-rk3")
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    dt_prev, taken = C.c_double(1.0), C.c_double()
    with pytest.raises(IbisError, match="fixed 67108864-byte scratch"):
        check(lib().ibis_rk_step_noalloc(fv._h, 3, ptr(a), ptr(b), 0.5, ptr(cons), ptr(prim),
                                         C.byref(dt_prev), 0.0, 1e30, -1.0, C.byref(taken),
                                         None), "ibis_rk_step_noalloc")
    fv.close()
    gb.close()
