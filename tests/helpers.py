"""Shared helpers: run the oracle and the MI355X residual on the same inputs."""
import numpy as np

from oracle import OracleFV, OracleGrid


def oracle_pair(grid_io, cfg):
    enc = cfg.encode()
    og = OracleGrid(grid_io, enc["tags"], enc["ghost"])
    return og, OracleFV(og, cfg)


def gpu_objects(grid_io, cfg, reorder=True):
    from ibis_amd.finite_volume import FiniteVolume, GridBlock
    gb = GridBlock(grid_io, cfg, reorder=reorder)
    return gb, FiniteVolume(gb, cfg)


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def gpu_dudt(gb, fv, prim_ref, allow_reconstruction=True):
    """Residual on the GPU for a reference-ordered prim; returns (dudt, prim
    after the call) in reference order."""
    import torch
    prim = to_dev(gb.to_grid_order(prim_ref))
    dudt = torch.zeros((gb.dim + 2, gb.Nt), dtype=torch.float64, device="cuda")
    fv.compute_dudt(prim, gb, dudt, allow_reconstruction=allow_reconstruction)
    torch.cuda.synchronize()
    return gb.to_reference_order(dudt.cpu().numpy()), gb.to_reference_order(prim.cpu().numpy())


def rel_err(a, b, n):
    """Per conserved variable: max |a-b| over valid cells / max |b|."""
    out = []
    for q in range(a.shape[0]):
        scale = np.abs(b[q, :n]).max()
        out.append(np.abs(a[q, :n] - b[q, :n]).max() / (scale if scale > 0 else 1.0))
    return np.array(out)
