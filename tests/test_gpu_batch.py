"""Batched stepping: every instance's state after K steps equals the state
it reaches when stepped alone with the plain path, bit for bit, with and
without the captured HIP graph."""
import numpy as np
import pytest

from helpers import to_dev
from ibis_amd import cases

pytestmark = pytest.mark.gpu


def _small_specs():
    return cases.batch_instances(n=6, seed=71, n_min=400, n_max=3000)


def _instances(specs):
    out = []
    for s in specs:
        g_io, cfg, fs = cases.batch_case(s)
        out.append((g_io, cfg, fs))
    return out


def _alone(g_io, cfg, fs, nsteps):
    """Plain path: one FiniteVolume + RungeKutta per instance."""
    import torch
    from ibis_amd.finite_volume import FiniteVolume, GridBlock, RungeKutta
    gb = GridBlock(g_io, cfg, reorder=True)
    fv = FiniteVolume(gb, cfg)
    prim = cases.uniform_prim(fs, gb.N, gb.Nt)
    d_prim = to_dev(prim)
    d_cons = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    fv.prim_to_cons(d_prim, d_cons)
    rk = RungeKutta(fv, "ssp-rk3", cfl=0.5, max_time=1e30)
    dts = [rk.take_step(d_cons, d_prim) for _ in range(nsteps)]
    torch.cuda.synchronize()
    return d_cons.cpu().numpy(), d_prim.cpu().numpy(), dts


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_instances_alone(use_graph):
    import torch
    from ibis_amd.batch import Batch
    specs = _small_specs()
    inst = _instances(specs)
    nsteps = 4
    ref = [_alone(g, c, f, nsteps) for g, c, f in inst]
    bt = Batch([(g, c) for g, c, _ in inst])
    prims = [cases.uniform_prim(f, b.N, b.Nt) for (g, c, f), b in zip(inst, bt.blocks)]
    d_prim = to_dev(bt.gather(prims))
    d_cons = torch.zeros((bt.ncons, bt.Nt), dtype=torch.float64, device="cuda")
    from ibis_amd.native import check, lib, ptr
    check(lib().ibis_prim_to_cons(bt.residual_handle(), ptr(d_prim), ptr(d_cons), None),
          "ibis_prim_to_cons")
    torch.cuda.synchronize()
    bt.rk_init(d_cons, d_prim, cfl=0.5, max_time=1e30)
    print(f"batch of {bt.n}: N={bt.N} graph={use_graph}", flush=True)
    bt.rk_steps(2, use_graph)
    bt.rk_steps(nsteps - 2, use_graph)
    torch.cuda.synchronize()
    t, dt = bt.times()
    cons = d_cons.cpu().numpy()
    for i, (c_ref, p_ref, dts) in enumerate(ref):
        n_i = bt.blocks[i].N
        got = bt.scatter(cons, i)
        assert np.array_equal(got[:, :n_i], c_ref[:, :n_i]), i
        assert dt[i] == dts[-1]
        assert t[i] == float(np.cumsum(dts)[-1])
