"""FP32 reconstruction/flux with FP64 accumulation against the FP64 residual.

The FP64 residual equals the oracle bit for bit (test_gpu_residual.py), so
it is the reference for the reduced-precision one.  The drift is reported
per conserved variable as max |d32 - d64| / max |d64| and
rms(d32 - d64) / rms(d64) over the valid cells.  Bound: 1e-3 (float has a
2^-24 unit roundoff; the residual is a sum of O(1/h) larger face fluxes, so
the relative error grows like 2^-24 * L / h).
"""
import numpy as np
import pytest

from helpers import gpu_objects, to_dev
from ibis_amd import cases

pytestmark = pytest.mark.gpu

NAMES = ("mass", "momentum_x", "momentum_y", "momentum_z", "energy")


def drift(d32, d64, n):
    out = {}
    nc = d64.shape[0]
    names = NAMES if nc == 5 else ("mass", "momentum_x", "momentum_y", "energy")
    for q in range(nc):
        a, b = d32[q, :n], d64[q, :n]
        out[names[q]] = (np.abs(a - b).max() / np.abs(b).max(),
                         np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
    return out


def _pair(grid_io, cfg, prim_ref, mode="fp32_flux"):
    import torch
    gb, fv = gpu_objects(grid_io, cfg)
    prim = to_dev(gb.to_grid_order(prim_ref))
    d64 = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    d32 = torch.zeros_like(d64)
    fv.compute_dudt(prim, gb, d64)
    fv.set_precision(mode)
    fv.compute_dudt(prim, gb, d32)
    # the slot-major face kernel gives the same float results
    fv.set_option("face_kernel", 128)
    d32b = torch.zeros_like(d64)
    fv.compute_dudt(prim, gb, d32b)
    # and so does the per-entry tile kernel
    fv.set_option("face_kernel", 1)
    d32c = torch.zeros_like(d64)
    fv.compute_dudt(prim, gb, d32c)
    fv.set_option("face_kernel", 0)
    fv.set_precision("fp64")
    d64b = torch.zeros_like(d64)
    fv.compute_dudt(prim, gb, d64b)
    torch.cuda.synchronize()
    assert torch.equal(d64, d64b)  # switching back restores the FP64 bits
    assert torch.equal(d32, d32b)
    assert torch.equal(d32, d32c)
    return gb, d32.cpu().numpy(), d64.cpu().numpy()


@pytest.mark.parametrize("mode", ["fp32_flux", "fp32_recon"])
@pytest.mark.parametrize("flux", ["hanel", "ausmdv"])
def test_hex_fp32_drift(flux, mode):
    g_io, cfg = cases.hex_case(24, flux=flux, limiter="venkatakrishnan")
    prim = cases.hex_prim(g_io, 24 ** 3 + 6 * 24 ** 2)
    gb, d32, d64 = _pair(g_io, cfg, prim, mode)
    dr = drift(d32, d64, gb.N)
    print(f"hex24 {flux} {mode} drift (max_rel, rms_rel): {dr}", flush=True)
    for name, (mx, rms) in dr.items():
        assert 1e-12 < mx < 1e-3, (name, mx)
        assert 1e-12 < rms < 1e-3, (name, rms)


@pytest.mark.parametrize("mode", ["fp32_flux", "fp32_recon"])
def test_wedge_fp32_drift(mode):
    g_io, cfg, fs = cases.wedge_case(96, 48)
    prim = cases.perturbed_prim(g_io, fs, 96 * 48 + 2 * (96 + 48), seed=61)
    gb, d32, d64 = _pair(g_io, cfg, prim, mode)
    dr = drift(d32, d64, gb.N)
    print(f"wedge {mode} drift (max_rel, rms_rel): {dr}", flush=True)
    for name, (mx, rms) in dr.items():
        assert mx < 1e-3, (name, mx)


@pytest.mark.parametrize("mode", ["fp32_flux", "fp32_recon"])
def test_mixed_elements_fp32_drift(mode):
    from ibis_amd import gridgen
    g_io, cfg = cases.hex_case(8, flux="hanel", limiter="venkatakrishnan")
    gh = gridgen.split_hexes(g_io, 3)
    gb, d32, d64 = _pair(gh, cfg, cases.hex_prim(gh, gh.num_cells + 6 * 64), mode)
    dr = drift(d32, d64, gb.N)
    print(f"hex/pyramid {mode} drift (max_rel, rms_rel): {dr}", flush=True)
    for name, (mx, rms) in dr.items():
        assert mx < 1e-3, (name, mx)


def test_fp32_modes_rejected_where_unsupported():
    """The reduced-precision modes need a fused second-order path: the
    viscous plate on the general path (fuse_viscous 0) refuses both, a
    first-order residual refuses both; unknown mode names raise."""
    import torch
    from ibis_amd.native import IbisError
    g_io, cfg, fs = cases.plate_case(16, 8)
    gb, fv = gpu_objects(g_io, cfg)
    prim = to_dev(gb.to_grid_order(cases.perturbed_prim(g_io, fs, gb.Nt, seed=2)))
    d = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    fv.set_option("fuse_viscous", 0)
    for mode in ("fp32_flux", "fp32_recon"):
        fv.set_precision(mode)
        with pytest.raises(IbisError):
            fv.compute_dudt(prim.clone(), gb, d)
    fv.set_precision("fp64")
    g_io, cfg, fs = cases.wedge_case(24, 12, order=1)
    gb, fv = gpu_objects(g_io, cfg)
    prim = to_dev(gb.to_grid_order(cases.perturbed_prim(g_io, fs, gb.Nt, seed=3)))
    d = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    for mode in ("fp32_flux", "fp32_recon"):
        fv.set_precision(mode)
        with pytest.raises(IbisError):
            fv.compute_dudt(prim.clone(), gb, d)
    with pytest.raises(IbisError):
        fv.set_precision("bf16")


@pytest.mark.parametrize("name", ["plate", "hex_walls"])
def test_viscous_fp32_flux_drift(name):
    """The FP32-flux mode on the fused viscous path (float face values and
    convective flux; FP64 viscous terms, area products and sums) drifts
    from the FP64 residual by less than 1e-3 per conserved variable, and
    switching back restores the FP64 bits.  The FP32-reconstruction mode
    stays inviscid-only and is refused."""
    import torch
    from test_gpu_residual import _visc_cases
    from ibis_amd.native import IbisError
    g_io, cfg, prim_ref = _visc_cases()[name]
    gb, fv = gpu_objects(g_io, cfg)
    prim = to_dev(gb.to_grid_order(prim_ref))
    d64 = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    d32 = torch.zeros_like(d64)
    fv.compute_dudt(prim.clone(), gb, d64)
    fv.set_precision("fp32_flux")
    fv.compute_dudt(prim.clone(), gb, d32)
    fv.set_precision("fp64")
    d64b = torch.zeros_like(d64)
    fv.compute_dudt(prim.clone(), gb, d64b)
    torch.cuda.synchronize()
    assert torch.equal(d64, d64b)
    dr = drift(d32.cpu().numpy(), d64.cpu().numpy(), gb.N)
    print(f"{name} viscous fp32_flux drift {dr}", flush=True)
    for q, (mx, rms) in dr.items():
        assert mx < 1e-3 and rms < 1e-3, (q, mx, rms)
    assert not torch.equal(d32, d64)  # the float path did run
    fv.set_precision("fp32_recon")
    with pytest.raises(IbisError):
        fv.compute_dudt(prim.clone(), gb, d32)
    fv.set_precision("fp64")
