"""Parity of the MI355X residual with the oracle (run on the GPU box).

Inviscid residuals compare bit for bit: the kernels and the oracle evaluate
the same IEEE expression trees with contraction disabled.  Viscous and
LDFSS2 residuals go through pow() (Sutherland, LDFSS2), where the device and
host math libraries may differ in the last bit, so they are compared at
1e-12 relative per conserved variable (max |gpu - cpu| / max |cpu|).
"""
import numpy as np
import pytest

from helpers import gpu_dudt, gpu_objects, oracle_pair, rel_err, to_dev
from ibis_amd import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _check(grid_io, cfg, prim_ref, exact, allow_reconstruction=True, options=None):
    og, ofv = oracle_pair(grid_io, cfg)
    p_cpu = prim_ref.copy()
    d_cpu = ofv.compute_dudt(p_cpu, allow_reconstruction=allow_reconstruction)
    gb, fv = gpu_objects(grid_io, cfg)
    for k, v in (options or {}).items():
        fv.set_option(k, v)
    d_gpu, p_gpu = gpu_dudt(gb, fv, prim_ref, allow_reconstruction)
    n = og.N
    print(f"case N={n} F={og.F} flux={cfg.flux} lim={cfg.limiter} visc={cfg.viscous}", flush=True)
    assert np.isfinite(d_cpu[:, :n]).all()
    if exact:
        assert np.array_equal(d_gpu[:, :n], d_cpu[:, :n]), rel_err(d_gpu, d_cpu, n)
    else:
        err = rel_err(d_gpu, d_cpu, n)
        assert (err <= TOL).all(), err
    # ghost rows written by the boundary actions (bit-exact in all cases)
    assert np.array_equal(p_gpu[:, n:], p_cpu[:, n:])
    # time step
    dt_cpu = ofv.estimate_dt(p_cpu)
    dt_gpu = fv.estimate_dt(to_dev(gb.to_grid_order(p_cpu)))
    if exact:
        assert dt_gpu == dt_cpu
    else:
        assert abs(dt_gpu - dt_cpu) <= TOL * dt_cpu
    return og, ofv, gb, fv


WEDGE_VARIANTS = [
    dict(flux="ausmdv", limiter="barth_jespersen", thermo="rho_u"),
    dict(flux="hanel", limiter="barth_jespersen", thermo="rho_u"),
    dict(flux="rusanov", limiter="barth_jespersen", thermo="rho_u"),
    dict(flux="ausmdv", limiter="unlimited", thermo="rho_u"),
    dict(flux="ausmdv", limiter="venkatakrishnan", thermo="rho_u"),
    dict(flux="ausmdv", limiter="barth_jespersen", thermo="rho_p"),
    dict(flux="ausmdv", limiter="barth_jespersen", thermo="rho_T"),
    dict(flux="hanel", limiter="barth_jespersen", thermo="p_T"),
    dict(flux="ausmdv", limiter="barth_jespersen", thermo="rho_u", order=1),
]


@pytest.mark.parametrize("kw", WEDGE_VARIANTS, ids=lambda k: "-".join(str(v) for v in k.values()))
def test_wedge_variants(kw):
    g_io, cfg, fs = cases.wedge_case(48, 24, **kw)
    prim = cases.perturbed_prim(g_io, fs, g_io.num_cells + 2 * (48 + 24), seed=11)
    _check(g_io, cfg, prim, exact=True)


@pytest.mark.parametrize("layout", [0, 2])
def test_wedge_subsonic_boundaries(layout):
    """Subsonic inflow / outflow and slip-wall actions on the second-order
    inviscid path: boundary actions fused with the ghost face values (layout
    0), and run in the cell kernel's staging of each ghost (layout 2, the
    option fold_boundary) or as a kernel of their own (fold_boundary 0)."""
    from ibis_amd import config as C
    g_io, cfg, fs = cases.wedge_case(48, 24, mach=0.5)
    cfg.boundaries = {"inflow": C.subsonic_inflow(fs), "outflow": C.subsonic_outflow(fs.p * 0.98),
                      "wall": C.slip_wall()}
    prim = cases.perturbed_prim(g_io, fs, g_io.num_cells + 2 * (48 + 24), seed=15)
    _check(g_io, cfg, prim, exact=True, options={"face_state_layout": layout})
    if layout == 2:
        _check(g_io, cfg, prim, exact=True, options={"fold_boundary": 0})


@pytest.mark.parametrize("zmode", ["flat", "uneven"])
def test_2d_mesh_off_the_z0_plane(zmode):
    """A 2D wedge lifted to z = 1.5 (every centroid at the same z: the 2D
    kernels take each z difference as +0 without loading it) and one whose
    vertices' z differ (they load it): the fused inviscid residual with and
    without precomputed offsets bit for bit, and the fused viscous one at the
    viscous bar, against the oracle."""
    from ibis_amd import config as C
    from ibis_amd import gridgen

    def lift(g_io):
        if zmode == "flat":
            g_io.xyz[:, 2] = 1.5
        else:
            g_io.xyz[:, 2] = 1.5 + 1e-3 * gridgen.uniform01(17, g_io.xyz.shape[0])
        return g_io

    g_io, cfg, fs = cases.wedge_case(40, 20)
    g_io = lift(g_io)
    prim = cases.perturbed_prim(g_io, fs, 40 * 20 + 2 * (40 + 20), seed=23)
    print(f"2D off z=0: {zmode}", flush=True)
    _check(g_io, cfg, prim, exact=True)
    _check(g_io, cfg, prim, exact=True, options={"precomputed_offsets": 0})
    g_io, cfg, fs = cases.wedge_case(40, 20, mach=0.5)
    g_io = lift(g_io)
    cfg.viscous = True
    cfg.boundaries = {"inflow": C.subsonic_inflow(fs), "outflow": C.subsonic_outflow(fs.p * 0.98),
                      "wall": C.fixed_temperature_no_slip_wall(300.0)}
    _check(g_io, cfg, cases.perturbed_prim(g_io, fs, 40 * 20 + 2 * 60, seed=24), exact=False)


def test_wedge_ldfss2_tolerance():
    g_io, cfg, fs = cases.wedge_case(48, 24, flux="ldfss2")
    prim = cases.perturbed_prim(g_io, fs, g_io.num_cells + 2 * (48 + 24), seed=12)
    _check(g_io, cfg, prim, exact=False)


def test_wedge_first_order_when_reconstruction_disallowed():
    g_io, cfg, fs = cases.wedge_case(48, 24)
    prim = cases.perturbed_prim(g_io, fs, g_io.num_cells + 2 * (48 + 24), seed=13)
    _check(g_io, cfg, prim, exact=True, allow_reconstruction=False)


def test_wedge_192x96_full_size():
    """192 x 96 wedge, AUSMDV / Barth-Jespersen / rho_u, uniform Mach 3 and
    perturbed states."""
    g_io, cfg, fs = cases.wedge_case(192, 96)
    nt = g_io.num_cells + 2 * (192 + 96)
    _check(g_io, cfg, cases.uniform_prim(fs, g_io.num_cells, nt), exact=True)
    _check(g_io, cfg, cases.perturbed_prim(g_io, fs, nt, seed=14), exact=True)


@pytest.mark.parametrize("flux,limiter", [("hanel", "venkatakrishnan"),
                                          ("ausmdv", "barth_jespersen"),
                                          ("rusanov", "unlimited")])
def test_hex_jittered_permuted(flux, limiter):
    g_io, cfg = cases.hex_case(16, flux=flux, limiter=limiter)
    prim = cases.hex_prim(g_io, 16 ** 3 + 6 * 16 ** 2)
    _check(g_io, cfg, prim, exact=True)


@pytest.mark.parametrize("layout,pack,fk,stage,pf,sc", [(0, 1, 0, 1, 0, 0), (1, 1, 0, 0, 0, 0),
                                                         (2, 1, 0, 1, 0, 0), (2, 0, 0, 1, 0, 0),
                                                         (2, 0, 0, 0, 1, 0), (2, 1, 64, 0, 0, 0),
                                                         (2, 0, 1, 1, 1, 0), (2, 1, 1, 0, 0, 0),
                                                         (2, 0, 128, 1, 1, 0), (2, 1, 256, 1, 0, 0),
                                                         (2, 0, 1, 1, 0, 1), (2, 0, 0, 0, 1, 1)])
@pytest.mark.parametrize("offsets", [0, 1])
def test_implementation_options_give_identical_bits(layout, pack, fk, stage, pf, sc, offsets):
    opts = {"face_state_layout": layout, "precomputed_offsets": offsets, "pack_cells": pack,
            "face_kernel": fk, "stage_offsets": stage, "cell_prefetch": pf, "stage_cells": sc}
    g_io, cfg = cases.hex_case(12, flux="hanel", limiter="venkatakrishnan")
    _check(g_io, cfg, cases.hex_prim(g_io, 12 ** 3 + 6 * 12 ** 2), exact=True, options=opts)
    g_io, cfg, fs = cases.wedge_case(40, 20)
    prim = cases.perturbed_prim(g_io, fs, 40 * 20 + 2 * (40 + 20), seed=21)
    _check(g_io, cfg, prim, exact=True, options=opts)


TILE_OPTIONS = [{"tile": 128}, {"tile": 512}, {"fold_boundary": 0}, {"fold_boundary": 1},
                {"tile": 512, "face_kernel": 0}, {"tile": 128, "stage_cells": 0, "cell_prefetch": 1},
                {"tile": 512, "face_state_layout": 0, "pack_cells": 1, "face_kernel": 0}]


@pytest.mark.parametrize("opts", TILE_OPTIONS, ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()))
def test_tile_sizes_give_identical_bits(opts):
    """Tiles of 128 and 512 cells (the tile kernels' block size follows)
    on the 12^3 hex box, the wedge, the first-order wedge and the
    hex/pyramid mesh, bit for bit against the oracle."""
    from ibis_amd import gridgen
    g_io, cfg = cases.hex_case(12, flux="hanel", limiter="venkatakrishnan")
    _check(g_io, cfg, cases.hex_prim(g_io, 12 ** 3 + 6 * 12 ** 2), exact=True, options=opts)
    g_io, cfg, fs = cases.wedge_case(40, 20)
    prim = cases.perturbed_prim(g_io, fs, 40 * 20 + 2 * (40 + 20), seed=21)
    _check(g_io, cfg, prim, exact=True, options=opts)
    g_io, cfg, fs = cases.wedge_case(40, 20, order=1)
    _check(g_io, cfg, prim, exact=True, options=opts)
    g_io, cfg = cases.hex_case(8, flux="hanel", limiter="venkatakrishnan")
    gh = gridgen.split_hexes(g_io, 3)
    _check(gh, cfg, cases.hex_prim(gh, gh.num_cells + 6 * 64), exact=True, options=opts)


def test_tile_option_switches_on_one_residual():
    """One residual switched between tile sizes (512, 128, 256, 512) and to
    the FP32-flux mode and back: every FP64 residual equals the oracle's bit
    for bit, so the per-entry face data follows each switch."""
    g_io, cfg = cases.hex_case(10, flux="ausmdv", limiter="barth_jespersen")
    prim = cases.hex_prim(g_io, 10 ** 3 + 6 * 10 ** 2)
    og, ofv = oracle_pair(g_io, cfg)
    d_cpu = ofv.compute_dudt(prim.copy())
    gb, fv = gpu_objects(g_io, cfg)
    for tile in (512, 128, 256, 512):
        fv.set_option("tile", tile)
        fv.set_precision("fp32_flux")
        gpu_dudt(gb, fv, prim)
        fv.set_precision("fp64")
        d_gpu, _ = gpu_dudt(gb, fv, prim)
        print(f"tile {tile}", flush=True)
        assert np.array_equal(d_gpu[:, :og.N], d_cpu[:, :og.N]), tile


MIXED_OPTIONS = [{}, {"face_state_layout": 0, "face_kernel": 0, "stage_cells": 0,
                      "precomputed_offsets": 0, "pack_cells": 1},
                 {"stage_cells": 0, "cell_prefetch": 1, "face_kernel": 0},
                 {"face_kernel": 128, "stage_cells": 0}]


@pytest.mark.parametrize("opts", MIXED_OPTIONS, ids=["default", "layout0", "prefetch", "slot128"])
def test_mixed_elements_ragged_stencils(opts):
    """Cells with different face counts: a 2D triangle / quad mesh (3 and 4
    faces) and a 3D hex / pyramid mesh (5 and 6 faces, triangle and quad
    faces), bit for bit against the oracle."""
    from ibis_amd import gridgen
    g_io, cfg, fs = cases.wedge_case(24, 12)
    gm = gridgen.split_quads(g_io, 2)
    prim = cases.perturbed_prim(gm, fs, gm.num_cells + 2 * (24 + 12), seed=31)
    _check(gm, cfg, prim, exact=True, options=opts)
    g_io, cfg = cases.hex_case(8, flux="hanel", limiter="venkatakrishnan")
    gh = gridgen.split_hexes(g_io, 3)
    _check(gh, cfg, cases.hex_prim(gh, gh.num_cells + 6 * 64), exact=True, options=opts)


@pytest.mark.parametrize("case", ["wedge", "hex"])
def test_stored_normals_entry_kernel_matches_tile_kernel(case):
    """A caller's mesh whose normals are not exactly cross(t1, t2) keeps its
    stored normals.  The per-entry face kernel then loads them with the
    entry geometry; the tile-list kernel reads them from the face geometry.
    The two kernels must agree bit for bit, and the stored normals must be
    used (the residual differs from the unperturbed mesh's)."""
    import ctypes as C
    from types import SimpleNamespace

    import torch

    from ibis_amd.batch import _concat_grids
    from ibis_amd.finite_volume import FiniteVolume, GridBlock
    from ibis_amd.native import MeshDescStruct, check, lib, ptr

    if case == "wedge":
        g_io, cfg, fs = cases.wedge_case(32, 16)
        prim_ref = cases.perturbed_prim(g_io, fs, 32 * 16 + 2 * (32 + 16), seed=41)
    else:
        g_io, cfg = cases.hex_case(10, flux="hanel", limiter="venkatakrishnan")
        prim_ref = cases.hex_prim(g_io, 10 ** 3 + 6 * 10 ** 2)
    gb = GridBlock(g_io, cfg)
    arr, N, G, F, _, _ = _concat_grids([gb])
    geom = arr["face_geom"]
    geom[0, ::3] = np.nextafter(geom[0, ::3], np.inf)  # n_x of every third face, one ulp
    d = MeshDescStruct()
    d.dim, d.ntags = gb.dim, len(arr["tag_off"]) - 1
    d.n_cells, d.n_ghost, d.n_faces = N, G, F
    for k, v in arr.items():
        setattr(d, k, ptr(v))
    m = C.c_void_p()
    check(lib().ibis_mesh_create(C.byref(d), C.byref(m)), "ibis_mesh_create")
    fv = ref_fv = None
    try:
        mesh = SimpleNamespace(_mesh=m, dim=gb.dim, Nt=gb.Nt)
        fv = FiniteVolume(mesh, cfg)
        prim0 = to_dev(gb.to_grid_order(prim_ref))
        out = {}
        for fk in (1, 0):
            fv.set_option("face_state_layout", 2)
            fv.set_option("face_kernel", fk)
            prim = prim0.clone()
            dudt = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
            fv.compute_dudt(prim, None, dudt)
            torch.cuda.synchronize()
            out[fk] = dudt.cpu().numpy()[:, :N]
        print(f"stored normals {case}: N={N} F={F}", flush=True)
        assert np.array_equal(out[1], out[0])
        ref_fv = FiniteVolume(gb, cfg)
        d_ref, _ = gpu_dudt(gb, ref_fv, prim_ref)
        assert not np.array_equal(out[1], gb.to_grid_order(d_ref)[:, :N])
    finally:
        for h in (fv, ref_fv):
            if h is not None:
                h.close()
        lib().ibis_mesh_destroy(m)
        gb.close()


def test_plate_viscous():
    g_io, cfg, fs = cases.plate_case(64, 32)
    prim = cases.perturbed_prim(g_io, fs, 64 * 32 + 2 * (64 + 32), seed=2)
    _check(g_io, cfg, prim, exact=False)


@pytest.mark.timeout(300)
def test_plate_viscous_1024x512_full_size():
    g_io, cfg, fs = cases.plate_case(1024, 512)
    prim = cases.perturbed_prim(g_io, fs, 1024 * 512 + 2 * (1024 + 512), seed=2)
    _check(g_io, cfg, prim, exact=False)


def test_count_bad_cells():
    g_io, cfg, fs = cases.wedge_case(48, 24)
    nt = g_io.num_cells + 2 * (48 + 24)
    prim = cases.uniform_prim(fs, g_io.num_cells, nt)
    prim[2, 5] = -1.0          # T < 0
    prim[0, 17] = np.nan       # rho NaN
    prim[0, 100] = np.inf      # rho Inf
    prim[0, 7] = -2.0          # rho < 0
    og, ofv = oracle_pair(g_io, cfg)
    gb, fv = gpu_objects(g_io, cfg)
    assert ofv.count_bad_cells(prim) == 4
    assert fv.count_bad_cells(to_dev(gb.to_grid_order(prim))) == 4


def test_rk_steps_match_oracle():
    """Five ssp-rk3 steps at CFL 0.5 on the wedge: identical state and dt."""
    import torch
    from ibis_amd.config import ssp_rk3
    from ibis_amd.finite_volume import RungeKutta
    g_io, cfg, fs = cases.wedge_case(48, 24)
    nt = g_io.num_cells + 2 * (48 + 24)
    prim = cases.perturbed_prim(g_io, fs, nt, seed=15)
    og, ofv = oracle_pair(g_io, cfg)
    cons = ofv.prim_to_cons(prim)
    gb, fv = gpu_objects(g_io, cfg)
    d_prim = to_dev(gb.to_grid_order(prim))
    d_cons = to_dev(gb.to_grid_order(cons))
    rk = RungeKutta(fv, "ssp-rk3", cfl=0.5, max_time=1.0)
    a, b = ssp_rk3()
    dt_prev, t = np.finfo(np.float64).max, 0.0
    p_cpu, c_cpu = prim.copy(), cons.copy()
    for step in range(5):
        dt_cpu = ofv.rk_step(a, b, 0.5, c_cpu, p_cpu, dt_prev, t, 1.0)
        dt_prev = dt_cpu
        t += dt_cpu
        dt_gpu = rk.take_step(d_cons, d_prim)
        assert dt_gpu == dt_cpu
    torch.cuda.synchronize()
    c_gpu = gb.to_reference_order(d_cons.cpu().numpy())
    n = og.N
    assert np.array_equal(c_gpu[:, :n], c_cpu[:, :n])


@pytest.mark.timeout(900)
def test_wedge_192x96_2000_steps():
    """The 192 x 96 wedge workload: uniform Mach-3 free
    stream, ssp-rk3 at CFL 0.5, 2000 steps.  The plain path (a host dt per
    step) equals the oracle's state after steps 1, 10, 100 and 2000 and its
    dt at every step; the device-resident path (a batch of one, each step
    one HIP graph replay) equals it after 2000 steps."""
    import torch
    from ibis_amd.batch import Batch
    from ibis_amd.config import ssp_rk3
    from ibis_amd.finite_volume import RungeKutta
    from ibis_amd.native import check, lib, ptr
    g_io, cfg, fs = cases.wedge_case(192, 96)
    n = g_io.num_cells
    prim = cases.uniform_prim(fs, n, n + 2 * (192 + 96))
    og, ofv = oracle_pair(g_io, cfg)
    cons = ofv.prim_to_cons(prim)
    gb, fv = gpu_objects(g_io, cfg)
    d_prim, d_cons = to_dev(gb.to_grid_order(prim)), to_dev(gb.to_grid_order(cons))
    rk = RungeKutta(fv, "ssp-rk3", cfl=0.5, max_time=1e30)
    a, b = ssp_rk3()
    dt_prev, t = np.finfo(np.float64).max, 0.0
    p_cpu, c_cpu = prim.copy(), cons.copy()
    for step in range(1, 2001):
        dt_cpu = ofv.rk_step(a, b, 0.5, c_cpu, p_cpu, dt_prev, t, 1e30)
        dt_prev = dt_cpu
        t += dt_cpu
        assert rk.take_step(d_cons, d_prim) == dt_cpu, step
        if step in (1, 10, 100, 2000):
            torch.cuda.synchronize()
            c_gpu = gb.to_reference_order(d_cons.cpu().numpy())
            assert np.array_equal(c_gpu[:, :n], c_cpu[:, :n]), step
            print(f"wedge 192x96: step {step}, t = {t:.6e}, state identical to the oracle", flush=True)
    bt = Batch([(g_io, cfg)])
    b0 = bt.blocks[0]
    bp = to_dev(bt.gather([b0.to_grid_order(prim)]))
    bc = torch.zeros((bt.ncons, bt.Nt), dtype=torch.float64, device="cuda")
    check(lib().ibis_prim_to_cons(bt.residual_handle(), ptr(bp), ptr(bc), None), "prim_to_cons")
    bt.rk_init(bc, bp, cfl=0.5, max_time=1e30)
    bt.rk_steps(2000, use_graph=True)
    c_bat = b0.to_reference_order(bt.scatter(bc.cpu().numpy(), 0))
    assert np.array_equal(c_bat[:, :n], c_cpu[:, :n])
    tb, _ = bt.times()
    assert tb[0] == t
    bt.close()


@pytest.mark.timeout(600)
def test_renumbering_invariance_hex160():
    """Config-3 size (160^3 jittered, cell-permuted hexes): the residual on
    the locality-renumbered grid equals the residual on the reference
    numbering, bit for bit."""
    import torch
    g_io, cfg = cases.hex_case(160)
    nt = 160 ** 3 + 6 * 160 ** 2
    prim = cases.hex_prim(g_io, nt)
    outs = []
    for reorder in (True, False):
        print(f"hex160 reorder={reorder}: building grid", flush=True)
        gb, fv = gpu_objects(g_io, cfg, reorder=reorder)
        print("  residual", flush=True)
        d, _ = gpu_dudt(gb, fv, prim)
        outs.append(d)
        del fv, gb
        torch.cuda.empty_cache()
    n = 160 ** 3
    assert np.array_equal(outs[0][:, :n], outs[1][:, :n])
    assert np.isfinite(outs[0][:, :n]).all()


def _visc_cases():
    """Viscous cases for the fused viscous path: the plate (AUSMDV,
    Barth-Jespersen, fixed-temperature no-slip wall), the same with Hanel and
    Venkatakrishnan and rho_p, a split-quad plate (triangles and quads),
    a jittered permuted 3D box with slip and no-slip walls, and a subsonic
    wedge with viscous walls."""
    from ibis_amd import config as C
    from ibis_amd import gridgen
    out = {}
    g_io, cfg, fs = cases.plate_case(64, 32)
    out["plate"] = (g_io, cfg, cases.perturbed_prim(g_io, fs, 64 * 32 + 2 * (64 + 32), seed=2))
    g_io, cfg, fs = cases.plate_case(48, 24, flux="hanel")
    cfg.limiter, cfg.thermo_interpolator = "venkatakrishnan", "rho_p"
    out["plate_hanel_vk"] = (g_io, cfg, cases.perturbed_prim(g_io, fs, 48 * 24 + 2 * 72, seed=3))
    gm = gridgen.split_quads(g_io, 2)
    out["plate_triquad"] = (gm, cfg, cases.perturbed_prim(gm, fs, gm.num_cells + 2 * 72, seed=4))
    g_io, cfg = cases.hex_case(10, flux="ausmdv", limiter="barth_jespersen")
    cfg.viscous = True
    tags = sorted(cfg.boundaries)
    cfg.boundaries[tags[0]] = C.fixed_temperature_no_slip_wall(320.0)
    cfg.boundaries[tags[1]] = C.slip_wall()
    prim = cases.hex_prim(g_io, 10 ** 3 + 6 * 10 ** 2)
    prim[:, 10 ** 3:] = prim[:, :1]
    out["hex_walls"] = (g_io, cfg, prim)
    g_io, cfg, fs = cases.wedge_case(48, 24, mach=0.5)
    cfg.viscous = True
    cfg.boundaries = {"inflow": C.subsonic_inflow(fs), "outflow": C.subsonic_outflow(fs.p * 0.98),
                      "wall": C.fixed_temperature_no_slip_wall(300.0)}
    out["wedge_subsonic"] = (g_io, cfg, cases.perturbed_prim(g_io, fs, 48 * 24 + 2 * 72, seed=5))
    return out


@pytest.mark.parametrize("name", ["plate", "plate_hanel_vk", "plate_triquad", "hex_walls",
                                  "wedge_subsonic"])
def test_fused_viscous_path_equals_general_path(name):
    """The four-kernel viscous path (boundary actions, cell kernel with the
    viscous gradients, viscous terms per face, face kernel with the viscous
    flux) against the
    reference-shaped general path on the same device: dU/dt, the ghost rows,
    estimate_dt (face density and temperature) and, with keep_gradients,
    every gradient and limiter slot, bit for bit; and against the oracle at
    the viscous bar."""
    import torch
    g_io, cfg, prim_ref = _visc_cases()[name]
    _check(g_io, cfg, prim_ref, exact=False)
    gb, fv = gpu_objects(g_io, cfg)
    outs = []
    # the fused path with its boundary passes folded into the cell kernel's
    # staging (the default) and as a kernel of their own, then the general path
    for fuse, fold in ((1, 1), (1, 0), (0, 1)):
        fv.set_option("fuse_viscous", fuse)
        fv.set_option("fold_boundary", fold)
        fv.set_option("keep_gradients", 1)
        d, p = gpu_dudt(gb, fv, prim_ref)
        dt = fv.estimate_dt(to_dev(gb.to_grid_order(p)))
        grads = fv.cell_gradients().cpu().numpy()
        torch.cuda.synchronize()
        outs.append((d, p, dt, grads))
    n = gb.N
    print(f"{name}: N={n}", flush=True)
    for o in outs[:2]:
        assert o[0][:, :n].tobytes() == outs[2][0][:, :n].tobytes()
        assert o[1].tobytes() == outs[2][1].tobytes()
        assert o[2] == outs[2][2]
        assert o[3].tobytes() == outs[2][3].tobytes()
