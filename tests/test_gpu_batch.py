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


def _batch_run(inst, prims, method, nsteps, use_graph, fuse, options=None):
    """Step a batch with the stage updates fused into the residual's face
    kernel (fuse 1) or as separate passes (fuse 0); returns cons, prim (all
    rows, ghosts included) and the per-instance times."""
    import torch
    from ibis_amd.batch import Batch
    from ibis_amd.native import check, lib, ptr
    bt = Batch([(g, c) for g, c in inst])
    d_prim = to_dev(bt.gather([p(b) for p, b in zip(prims, bt.blocks)]))
    d_cons = torch.zeros((bt.ncons, bt.Nt), dtype=torch.float64, device="cuda")
    check(lib().ibis_prim_to_cons(bt.residual_handle(), ptr(d_prim), ptr(d_cons), None),
          "ibis_prim_to_cons")
    torch.cuda.synchronize()
    bt.set_option("fuse_stages", fuse)
    for k, v in (options or {}).items():
        bt.set_option(k, v)
    bt.rk_init(d_cons, d_prim, cfl=0.5, max_time=1e30, method=method)
    bt.rk_steps(nsteps, use_graph)
    torch.cuda.synchronize()
    out = (d_cons.cpu().numpy(), d_prim.cpu().numpy(), bt.times())
    bt.close()
    return out


@pytest.mark.parametrize("method", ["ssp-rk3", "rk4", "rk3", "euler", "midpoint"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_fused_stage_epilogue_equals_separate_passes(method, use_graph):
    """The stage update and primitive conversion in the face kernel's
    epilogue (and the time step before the first residual) give the bits of
    the separate passes: every row of cons and prim, ghosts included, and
    every instance's t and dt, for tableaux with 1-4 stages (rk4's zero and
    rk3's negative a_ij are skipped by the stage sums, as the reference's
    loop skips a_ij < 1e-14)."""
    specs = _small_specs()[:4]
    inst = [(g, c) for g, c, _ in _instances(specs)]
    fss = [f for _, _, f in _instances(specs)]
    prims = [lambda b, f=f: cases.uniform_prim(f, b.N, b.Nt) for f in fss]
    print(f"fused stages: {method} graph={use_graph}", flush=True)
    fused = _batch_run(inst, prims, method, 3, use_graph, 1)
    sep = _batch_run(inst, prims, method, 3, use_graph, 0)
    assert fused[0].tobytes() == sep[0].tobytes()
    assert fused[1].tobytes() == sep[1].tobytes()
    for a, b in zip(fused[2], sep[2]):
        assert a.tobytes() == b.tobytes()


def _run_with_new_state(inst, prims, fuse, use_graph):
    """Two steps, then the caller writes the starting state back into the
    bound buffers, then two more steps; returns cons, prim and the times."""
    import torch
    from ibis_amd.batch import Batch
    from ibis_amd.native import check, lib, ptr
    bt = Batch([(g, c) for g, c in inst])
    start = to_dev(bt.gather([p(b) for p, b in zip(prims, bt.blocks)]))
    d_prim = start.clone()
    d_cons = torch.zeros((bt.ncons, bt.Nt), dtype=torch.float64, device="cuda")
    check(lib().ibis_prim_to_cons(bt.residual_handle(), ptr(d_prim), ptr(d_cons), None),
          "ibis_prim_to_cons")
    torch.cuda.synchronize()
    bt.set_option("fuse_stages", fuse)
    bt.rk_init(d_cons, d_prim, cfl=0.5, max_time=1e30)
    bt.rk_steps(2, use_graph)
    torch.cuda.synchronize()
    d_prim.copy_(start)
    check(lib().ibis_prim_to_cons(bt.residual_handle(), ptr(d_prim), ptr(d_cons), None),
          "ibis_prim_to_cons")
    torch.cuda.synchronize()
    bt.rk_steps(2, use_graph)
    torch.cuda.synchronize()
    out = (d_cons.cpu().numpy(), d_prim.cpu().numpy(), bt.times())
    bt.close()
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_fused_steps_take_the_time_step_of_a_state_the_caller_wrote(use_graph):
    """With the stage updates fused, a step's last stage writes the local
    time steps of its result for the next step; a state the caller writes
    between two ibis_batch_rk_steps calls must still set the next step's dt.
    Fused and separate passes (which evaluate the time step from the state
    at every step) agree bit for bit after 2 steps, a rewrite of the
    starting state, and 2 more steps."""
    specs = _small_specs()[:3]
    made = _instances(specs)
    inst = [(g, c) for g, c, _ in made]
    prims = [lambda b, f=f: cases.uniform_prim(f, b.N, b.Nt) for _, _, f in made]
    print(f"state rewritten between calls: graph={use_graph}", flush=True)
    fused = _run_with_new_state(inst, prims, 1, use_graph)
    sep = _run_with_new_state(inst, prims, 0, use_graph)
    assert fused[0].tobytes() == sep[0].tobytes()
    assert fused[1].tobytes() == sep[1].tobytes()
    for a, b in zip(fused[2], sep[2]):
        assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("tile", [256, 128, 512])
def test_fused_stage_epilogue_3d(tile):
    """Two hex boxes (6^3 and 9^3, jittered, permuted), Hanel with the
    Venkatakrishnan limiter, a smooth non-uniform state: fused stage
    updates (with tiles of 128, 256 or 512 cells) and separate ones (256)
    agree bit for bit after three ssp-rk3 steps."""
    inst = [cases.hex_case(n, flux="hanel", limiter="venkatakrishnan") for n in (6, 9)]

    def prim_of(g_io):
        def make(b):
            p = cases.hex_prim(g_io, b.Nt)
            p[:, b.N:] = p[:, :1]
            return b.to_grid_order(p)
        return make

    prims = [prim_of(g) for g, _ in inst]
    print(f"fused stages: 3D, tile {tile}", flush=True)
    fused = _batch_run(inst, prims, "ssp-rk3", 3, True, 1, {"tile": tile})
    sep = _batch_run(inst, prims, "ssp-rk3", 3, True, 0)
    assert fused[0].tobytes() == sep[0].tobytes()
    assert fused[1].tobytes() == sep[1].tobytes()
    for a, b in zip(fused[2], sep[2]):
        assert a.tobytes() == b.tobytes()
