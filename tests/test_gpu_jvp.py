"""Parity of the tangent (JVP) residual with the oracle's Dual residual.

ibis_compute_dudt_jvp replaces FiniteVolume<Ibis::dual> under
SteadyStateLinearisation::matrix_vector_product
(src/solvers/solvers/steady_state.cpp:52-128).  Inviscid cases compare both
the value and the tangent bit for bit (same Dual expression trees,
contraction disabled on both sides); cases that go through pow() (viscous
Sutherland, LDFSS2) compare at 1e-12 relative per conserved variable.
"""
import numpy as np
import pytest

from helpers import gpu_objects, oracle_pair, rel_err, to_dev
from ibis_amd import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _tangent(shape, seed, scale):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) * scale[:, None]


def _jvp_case(grid_io, cfg, prim_ref, seed, exact, allow_reconstruction=True, nan_ok=False):
    import torch
    og, ofv = oracle_pair(grid_io, cfg)
    n, nt = og.N, og.N + og.G
    scale = np.abs(prim_ref[:, :n]).max(axis=1) * 1e-3 + 1e-3
    pdot_ref = _tangent(prim_ref.shape, seed, scale)
    p_cpu, pd_cpu = prim_ref.copy(), pdot_ref.copy()
    d_cpu, dd_cpu = ofv.compute_dudt_jvp(p_cpu, pd_cpu, allow_reconstruction)
    gb, fv = gpu_objects(grid_io, cfg)
    prim = to_dev(gb.to_grid_order(prim_ref))
    pdot = to_dev(gb.to_grid_order(pdot_ref))
    dudt = torch.zeros((gb.dim + 2, nt), dtype=torch.float64, device="cuda")
    ddot = torch.zeros_like(dudt)
    print(f"jvp case N={n} flux={cfg.flux} lim={cfg.limiter} visc={cfg.viscous} "
          f"recon={allow_reconstruction}", flush=True)
    fv.compute_dudt_jvp(prim, pdot, dudt, ddot, allow_reconstruction)
    torch.cuda.synchronize()
    d_gpu = gb.to_reference_order(dudt.cpu().numpy())
    dd_gpu = gb.to_reference_order(ddot.cpu().numpy())
    p_gpu = gb.to_reference_order(prim.cpu().numpy())
    pd_gpu = gb.to_reference_order(pdot.cpu().numpy())
    if nan_ok:
        # NaN tangents where the oracle has them (same positions), the rest
        # compared like any other tangent
        bad = ~np.isfinite(dd_cpu[:, :n])
        assert np.array_equal(bad, ~np.isfinite(dd_gpu[:, :n]))
        dd_cpu, dd_gpu = np.where(np.isfinite(dd_cpu), dd_cpu, 0.0), np.where(
            np.isfinite(dd_gpu), dd_gpu, 0.0)
    assert np.isfinite(dd_cpu[:, :n]).all() and np.abs(dd_cpu[:, :n]).max() > 0
    if exact:
        assert np.array_equal(d_gpu[:, :n], d_cpu[:, :n]), rel_err(d_gpu, d_cpu, n)
        assert np.array_equal(dd_gpu[:, :n], dd_cpu[:, :n]), rel_err(dd_gpu, dd_cpu, n)
    else:
        assert (rel_err(d_gpu, d_cpu, n) <= TOL).all(), rel_err(d_gpu, d_cpu, n)
        assert (rel_err(dd_gpu, dd_cpu, n) <= TOL).all(), rel_err(dd_gpu, dd_cpu, n)
    # ghost rows of the value and the tangent written by the boundary actions
    assert np.array_equal(p_gpu[:, n:], p_cpu[:, n:])
    assert np.array_equal(pd_gpu[:, n:], pd_cpu[:, n:])
    # the value part is the plain residual
    d_plain = ofv.compute_dudt(prim_ref.copy(), allow_reconstruction=allow_reconstruction)
    if exact:
        assert np.array_equal(d_gpu[:, :n], d_plain[:, :n])
    return og, ofv, gb, fv


@pytest.mark.parametrize("kw", [
    dict(flux="ausmdv", limiter="barth_jespersen", thermo="rho_u"),
    dict(flux="hanel", limiter="venkatakrishnan", thermo="rho_p"),
    dict(flux="rusanov", limiter="unlimited", thermo="p_T"),
], ids=lambda k: "-".join(k.values()))
def test_wedge_jvp(kw):
    g_io, cfg, fs = cases.wedge_case(48, 24, **kw)
    prim = cases.perturbed_prim(g_io, fs, 48 * 24 + 2 * (48 + 24), seed=31)
    _jvp_case(g_io, cfg, prim, seed=32, exact=True)


def test_wedge_jvp_first_order_preconditioner():
    """allow_reconstruction = false, as the JFNK preconditioner runs it
    (steady_state.cpp:47-50)."""
    g_io, cfg, fs = cases.wedge_case(48, 24)
    prim = cases.perturbed_prim(g_io, fs, 48 * 24 + 2 * (48 + 24), seed=33)
    _jvp_case(g_io, cfg, prim, seed=34, exact=True, allow_reconstruction=False)


def test_wedge_jvp_ldfss2():
    """The reference's Dual pow (UT/dual.h:239-245) adds power.dual *
    real * log(base): with a constant exponent that is 0 * log(base), NaN
    for the negative base LDFSS2 passes it on subsonic faces
    (ldfss2.cpp, (sqrt(0.5 (ML^2 + MR^2)) - 1)^2).  The tangent is NaN there
    in the reference, the oracle and here alike."""
    g_io, cfg, fs = cases.wedge_case(48, 24, flux="ldfss2")
    prim = cases.perturbed_prim(g_io, fs, 48 * 24 + 2 * (48 + 24), seed=35)
    _jvp_case(g_io, cfg, prim, seed=36, exact=False, nan_ok=True)


def test_hex_jvp():
    g_io, cfg = cases.hex_case(12, flux="hanel", limiter="venkatakrishnan")
    prim = cases.hex_prim(g_io, 12 ** 3 + 6 * 12 ** 2)
    _jvp_case(g_io, cfg, prim, seed=37, exact=True)


def test_mixed_elements_jvp():
    """Tangent residual on ragged stencils: triangle / quad and hex / pyramid."""
    from ibis_amd import gridgen
    g_io, cfg, fs = cases.wedge_case(24, 12)
    gm = gridgen.split_quads(g_io, 2)
    prim = cases.perturbed_prim(gm, fs, gm.num_cells + 2 * (24 + 12), seed=33)
    _jvp_case(gm, cfg, prim, seed=38, exact=True)
    g_io, cfg = cases.hex_case(8, flux="hanel", limiter="venkatakrishnan")
    gh = gridgen.split_hexes(g_io, 3)
    _jvp_case(gh, cfg, cases.hex_prim(gh, gh.num_cells + 6 * 64), seed=39, exact=True)


def test_plate_viscous_jvp():
    g_io, cfg, fs = cases.plate_case(64, 32)
    prim = cases.perturbed_prim(g_io, fs, 64 * 32 + 2 * (64 + 32), seed=2)
    _jvp_case(g_io, cfg, prim, seed=38, exact=False)


def test_cons_to_prim_jvp():
    import torch
    g_io, cfg, fs = cases.wedge_case(48, 24)
    nt = 48 * 24 + 2 * (48 + 24)
    prim = cases.perturbed_prim(g_io, fs, nt, seed=39)
    og, ofv = oracle_pair(g_io, cfg)
    cons = ofv.prim_to_cons(prim)
    cdot = np.random.default_rng(40).standard_normal(cons.shape)
    p_cpu, pd_cpu = ofv.cons_to_prim_jvp(cons, cdot)
    gb, fv = gpu_objects(g_io, cfg)
    p = torch.zeros((7, nt), dtype=torch.float64, device="cuda")
    pd = torch.zeros_like(p)
    fv.cons_to_prim_jvp(to_dev(gb.to_grid_order(cons)), to_dev(gb.to_grid_order(cdot)), p, pd)
    torch.cuda.synchronize()
    n = og.N
    assert np.array_equal(gb.to_reference_order(p.cpu().numpy())[:, :n], p_cpu[:, :n])
    assert np.array_equal(gb.to_reference_order(pd.cpu().numpy())[:, :n], pd_cpu[:, :n])
