"""The host grid builder (ibis_amd/csrc/grid.cpp) and the CPU oracle under
AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/driver.cpp is built with g++ -fsanitize=address,undefined
together with grid.cpp, status.cpp and oracle/ibis_oracle.cpp (host code
only; no GPU code is built or run here).  It builds each case's grid with
the library's builder (reference numbering and locality renumbering) and
with the oracle, checks the two agree bit for bit, and runs the oracle's
residual, estimate_dt, tangent residual and an ssp-rk3 step.  Any
sanitizer report fails the run.  Cases: a 2D wedge (quads), a mixed
triangle/quad mesh, a jittered permuted hex box and a hex/pyramid mesh.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ibis_amd import cases, gridgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("san") / "driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-ffp-contract=off", "-Wno-unknown-pragmas",
           os.path.join(ROOT, "tests", "sanitize", "driver.cpp"),
           os.path.join(ROOT, "oracle", "ibis_oracle.cpp"),
           os.path.join(ROOT, "ibis_amd", "csrc", "grid.cpp"),
           os.path.join(ROOT, "ibis_amd", "csrc", "status.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    return out


def _write_case(path, g_io, cfg, prim):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import fv_struct
    enc = cfg.encode()
    tags = enc["tags"]
    btag, btype, bvtx = g_io.flat_markers(tags)
    s = fv_struct(cfg)
    with open(path, "wb") as f:
        f.write(np.array([g_io.dim, len(tags)], np.int32).tobytes())
        f.write(np.array([g_io.xyz.shape[0]], np.int64).tobytes())
        f.write(np.ascontiguousarray(g_io.xyz, dtype=np.float64).tobytes())
        f.write(np.array([g_io.num_cells], np.int64).tobytes())
        f.write(np.ascontiguousarray(g_io.ctype, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(g_io.cvtx, dtype=np.int64).tobytes())
        f.write(np.ascontiguousarray(enc["ghost"], dtype=np.int32).tobytes())
        f.write(np.array([btag.shape[0]], np.int64).tobytes())
        f.write(btag.tobytes())
        f.write(btype.tobytes())
        f.write(np.ascontiguousarray(bvtx, dtype=np.int64).tobytes())
        raw = bytes(s)
        f.write(np.array([len(raw)], np.int64).tobytes())
        f.write(raw)
        f.write(np.array([prim.size], np.int64).tobytes())
        f.write(np.ascontiguousarray(prim, dtype=np.float64).tobytes())


def _cases():
    out = []
    g, cfg, fs = cases.wedge_case(24, 12)
    out.append(("wedge", g, cfg, cases.perturbed_prim(g, fs, 24 * 12 + 2 * 36, seed=5)))
    g, cfg, fs = cases.plate_case(16, 8)
    out.append(("plate_viscous", g, cfg, cases.perturbed_prim(g, fs, 16 * 8 + 2 * 24, seed=6)))
    g, cfg = cases.hex_case(7)
    out.append(("hex", g, cfg, cases.hex_prim(g, 7 ** 3 + 6 * 49)))
    g, cfg, fs = cases.wedge_case(12, 6)
    gm = gridgen.split_quads(g, 2)
    out.append(("tri_quad", gm, cfg, cases.perturbed_prim(gm, fs, gm.num_cells + 36, seed=31)))
    g, cfg = cases.hex_case(6)
    gh = gridgen.split_hexes(g, 3)
    out.append(("hex_pyramid", gh, cfg, cases.hex_prim(gh, gh.num_cells + 6 * 36)))
    return out


@pytest.mark.parametrize("idx", range(5), ids=["wedge", "plate_viscous", "hex", "tri_quad",
                                               "hex_pyramid"])
def test_grid_builder_and_oracle_clean_under_sanitizers(driver, tmp_path, idx):
    name, g, cfg, prim = _cases()[idx]
    path = str(tmp_path / f"{name}.bin")
    _write_case(path, g, cfg, prim)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([driver, path], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, f"{name}: rc {r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}"
    assert r.stdout.startswith("ok "), r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
