"""Native flow snapshots (src/io/io/native.cpp, io.cpp:32-158): file layout,
byte format, round trips, cell order."""
import json
import os

import numpy as np
import pytest

from ibis_amd import cases
from ibis_amd.finite_volume import GridBlock
from ibis_amd.snapshot import FVIO, pad_time_index


def test_pad_time_index():
    """io.cpp:32-37"""
    assert pad_time_index(0) == "0000"
    assert pad_time_index(37) == "0037"
    assert pad_time_index(123456) == "123456"


@pytest.fixture
def block():
    g_io, cfg, fs = cases.wedge_case(12, 6)
    gb = GridBlock(g_io, cfg, reorder=True, device=False)
    prim = cases.perturbed_prim(g_io, fs, gb.Nt, seed=91)
    return gb, gb.to_grid_order(prim), prim


def test_binary_layout_and_round_trip(tmp_path, block):
    gb, prim, prim_ref = block
    io = FVIO("native_binary", input_dir=str(tmp_path), output_dir=str(tmp_path))
    io.write(prim, gb, 0.25)
    io.write(prim, gb, 0.5)
    assert open(tmp_path / "flows").read() == "0000\n0001\n"
    assert json.load(open(tmp_path / "0001" / "meta_data.json")) == {"time": 0.5}
    n = gb.N
    assert sorted(os.listdir(tmp_path / "0000")) == ["T", "meta_data.json", "p", "vx", "vy"]
    # raw little-endian FP64 in the reference cell order
    raw = np.frombuffer(open(tmp_path / "0000" / "T", "rb").read(), dtype="<f8")
    assert raw.size == n and np.array_equal(raw, prim_ref[2, :n])
    back, meta = io.read(gb, 0)
    assert meta == {"time": 0.25}
    for row in (1, 2, 4, 5):  # p, T, vx, vy exactly
        assert np.array_equal(back[row, :n], prim[row, :n])
    # rho and e from p and T (update_thermo_from_pT)
    assert np.allclose(back[0, :n], prim[0, :n], rtol=1e-14)
    assert np.allclose(back[3, :n], prim[3, :n], rtol=1e-14)


def test_text_format(tmp_path, block):
    gb, prim, prim_ref = block
    io = FVIO("native_text", input_dir=str(tmp_path), output_dir=str(tmp_path))
    io.write(prim, gb, 1.0)
    lines = open(tmp_path / "0000" / "p").read().splitlines()
    assert len(lines) == gb.N
    assert lines[0] == f"{prim_ref[1, 0]:.16f}"
    back, _ = io.read(gb, 0)
    assert np.allclose(back[1, :gb.N], prim[1, :gb.N], rtol=1e-15)


def test_3d_writes_vz_and_rejects_short_files(tmp_path):
    g_io, cfg = cases.hex_case(4)
    gb = GridBlock(g_io, cfg, reorder=True, device=False)
    prim = gb.to_grid_order(cases.hex_prim(g_io, gb.Nt))
    io = FVIO("native_binary", input_dir=str(tmp_path), output_dir=str(tmp_path))
    io.write(prim, gb, 0.0)
    assert "vz" in os.listdir(tmp_path / "0000")
    np.zeros(3).tofile(tmp_path / "0000" / "T")
    with pytest.raises(ValueError):
        io.read(gb, 0)
    with pytest.raises(ValueError):
        FVIO("vtk_binary")
