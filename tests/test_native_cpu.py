"""CPU-side checks of the product library (no GPU needed).

* libibis_gpu.so loads and exports every function include/ibis_gpu.h declares.
* The host grid builder (reference numbering) reproduces the oracle's grid
  bit for bit, and the locality renumbering is a pure permutation of it.
* Configuration errors are reported, as the reference's factories throw.
"""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import oracle_pair
from ibis_amd import cases, config, native
from ibis_amd.finite_volume import GridBlock


def test_library_exports_every_header_symbol():
    L = native.lib()
    syms = native.header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(L, name), name


CASES = {
    "wedge": lambda: cases.wedge_case(40, 20)[:2],
    "plate": lambda: cases.plate_case(32, 16)[:2],
    "hex": lambda: cases.hex_case(10),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("reorder", [False, True])
def test_grid_builder_matches_oracle(name, reorder):
    g_io, cfg = CASES[name]()
    og, _ = oracle_pair(g_io, cfg)
    gb = GridBlock(g_io, cfg, reorder=reorder, device=False)
    ex = gb.export()
    cn, fn = gb.cell_new_of_old, gb.face_new_of_old
    if not reorder:
        assert np.array_equal(cn, np.arange(og.Nt)) and np.array_equal(fn, np.arange(og.F))
    assert (gb.N, gb.G, gb.F) == (og.N, og.G, og.F)
    assert np.array_equal(ex["face_geom"][:, fn], og.face_geom)
    assert np.array_equal(ex["face_lr"][:, fn], cn[og.face_lr])
    assert np.array_equal(ex["centroid"][:, cn], og.centroid)
    assert np.array_equal(ex["volume"][cn], og.volume)
    assert np.array_equal(ex["lsq"][:, cn[:og.N]], og.lsq)
    assert np.array_equal(fn[og.bfaces], ex["bfaces"])
    assert np.array_equal(cn[og.ghosts], ex["ghosts"])
    for c in range(og.N):
        a, b = og.cell_off[c], og.cell_off[c + 1]
        k = cn[c]
        p, q = ex["cell_off"][k], ex["cell_off"][k + 1]
        assert np.array_equal(fn[og.cell_faces[a:b]], ex["cell_faces"][p:q])
        assert np.array_equal(cn[og.neighbours[a:b]], ex["neighbours"][p:q])
        assert np.array_equal(og.outsigns[a:b], ex["outsigns"][p:q])


def test_reorder_improves_locality_on_permuted_hex():
    g_io, cfg = cases.hex_case(16)
    span = {}
    for reorder in (False, True):
        gb = GridBlock(g_io, cfg, reorder=reorder, device=False)
        ex = gb.export()
        L, R = ex["face_lr"]
        inner = (L < gb.N) & (R < gb.N)
        span[reorder] = np.median(np.abs(L[inner] - R[inner]))
    assert span[True] < span[False] / 10


def test_grid_errors_are_reported():
    g_io, cfg = cases.wedge_case(8, 4)[:2]
    cfg.boundaries.pop("wall")  # wall faces then have no ghost cell
    with pytest.raises(native.IbisError, match="no cell on one side"):
        GridBlock(g_io, cfg, device=False)


def test_unknown_configuration_values_raise():
    with pytest.raises(ValueError, match="Unknown flux calculator"):
        config.FVConfig(flux="roe").encode()
    with pytest.raises(ValueError, match="Unknown limiter"):
        config.FVConfig(limiter="minmod").encode()
    with pytest.raises(ValueError, match="Unknown boundary action"):
        config.FVConfig(boundaries={"a": config.BoundaryCondition([{"type": "magic"}])}).encode()


def test_config_json_round_trip():
    fs = cases.freestream(3.0)
    cfg = cases.plate_config(fs)
    back = config.FVConfig.from_json(cfg.to_json())
    assert back.to_json() == cfg.to_json()
    assert back.tags() == ["inflow", "outflow", "wall"]


def _mesh_desc(gb, ex, tag_off=None):
    """An ibis_mesh_desc over a grid export (arrays kept alive by the caller)."""
    keep = dict(ex)
    keep["left"] = np.ascontiguousarray(ex["face_lr"][0])
    keep["right"] = np.ascontiguousarray(ex["face_lr"][1])
    if tag_off is not None:
        keep["tag_off"] = tag_off
    d = native.MeshDescStruct(gb.dim, gb.ntags, gb.N, gb.G, gb.F, native.ptr(keep["left"]),
                              native.ptr(keep["right"]), native.ptr(keep["face_geom"]),
                              native.ptr(keep["cell_off"]), native.ptr(keep["cell_faces"]),
                              native.ptr(keep["outsigns"]), native.ptr(keep["neighbours"]),
                              native.ptr(keep["centroid"]), native.ptr(keep["volume"]),
                              native.ptr(keep["lsq"]), native.ptr(keep["tag_off"]),
                              native.ptr(keep["bfaces"]))
    return d, keep


def test_mesh_create_rejects_a_ghost_on_no_tagged_face():
    """A descriptor whose last tag lists none of its boundary faces leaves
    those faces' ghosts on no tagged face: ibis_mesh_create refuses it on
    the host, before any device call (the boundary and adjoint kernels
    index their tables with the ghost's boundary face)."""
    g_io, cfg = cases.wedge_case(8, 4)[:2]
    gb = GridBlock(g_io, cfg, reorder=True, device=False)
    ex = gb.export()
    bad = ex["tag_off"].copy()
    bad[-1] = bad[-2]  # the last tag keeps no boundary faces
    d, keep = _mesh_desc(gb, ex, tag_off=bad)
    h = C.c_void_p()
    rc = native.lib().ibis_mesh_create(C.byref(d), C.byref(h))
    assert rc == 1  # IBIS_ERR_INVALID
    assert b"on no tagged boundary face" in native.lib().ibis_last_error()
    assert not h.value


def test_entry_kernel_index_range():
    """face_kernel 1 addresses the face values with 32-bit starts, up to
    maxf*5*N + 5*G; larger meshes are created with the tile-list kernel."""
    fit = native.lib().ibis_entry_indices_fit
    fit.restype = C.c_int
    fit.argtypes = [C.c_int64, C.c_int64, C.c_int64]
    assert fit(6, 4_096_000, 153_600) == 1  # the 160^3 box
    n = (2 ** 31 - 1) // 30  # largest hex mesh without ghosts that fits
    assert fit(6, n, 0) == 1
    assert fit(6, n + 1, 0) == 0
    assert fit(6, 72_000_000, 1_000_000) == 0  # ~72M hexes fit in HBM but not in int32 starts
    assert fit(4, 100_000_000, 0) == 1 and fit(4, 110_000_000, 0) == 0  # 2D quads


def test_cpp_driver_reports_input_errors(tmp_path):
    """ibis_fv_run (the C++ host layer's driver) refuses a missing argument
    list and a malformed config.json before it touches a GPU."""
    import subprocess
    drv = os.path.join(os.path.dirname(native.lib_path()), "ibis_fv_run")
    assert os.path.exists(drv), "ibis_fv_run is not built"
    r = subprocess.run([drv], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    bad = tmp_path / "config.json"
    bad.write_text('{"grid": {"boundaries": {"wall": [1, 2,]}}}')
    r = subprocess.run([drv, "g.su2", str(bad), "p.bin", str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "JSON" in r.stderr, r.stderr


@pytest.mark.parametrize("name", ["wedge", "hex"])
def test_su2_round_trip(name, tmp_path):
    """write_su2 then read_su2 gives back the same coordinates (bit for bit),
    elements and markers; the C++ host layer reads the same files."""
    from ibis_amd.gridio import read_su2, write_su2
    g = CASES[name]()[0]
    path = str(tmp_path / "g.su2")
    write_su2(g, path)
    h = read_su2(path)
    assert h.dim == g.dim
    assert np.array_equal(h.xyz, g.xyz)
    assert np.array_equal(h.ctype, g.ctype) and np.array_equal(h.cvtx, g.cvtx)
    assert sorted(h.markers) == sorted(g.markers)
    for t in g.markers:
        assert np.array_equal(h.markers[t][0], g.markers[t][0])
        assert np.array_equal(h.markers[t][1][:, :2], g.markers[t][1][:, :2])
