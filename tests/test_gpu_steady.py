"""Device GMRES / FGMRES, the steady-state linearisation and the JFNK step.

Known answers: the reference's GMRES and FGMRES 5x5 tests
(src/linear_algebra/linear_algebra/gmres.cpp:320-500).  The matvec compares
bit for bit with the oracle's Dual residual on an inviscid case; a JFNK step
compares with the oracle's GMRES on the same Dual matvec to the Krylov
tolerance (the two sum their inner products in different orders).
"""
import numpy as np
import pytest

from helpers import gpu_objects, oracle_pair, to_dev
from ibis_amd import cases
import steady as ost

pytestmark = pytest.mark.gpu

EXPECTED = np.array([1.0, 0.0, -1.0, 1.5, 0.5])


def _tridiag(n, lo, d, hi):
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = d
        if i + 1 < n:
            a[i, i + 1] = hi
            a[i + 1, i] = lo
    return a


def test_device_gmres_known_answer():
    import torch
    from ibis_amd.steady import DenseOperator, Gmres
    A = DenseOperator(to_dev(_tridiag(5, -1.0, 2.0, -0.5)))
    rhs = to_dev(np.array([2.0, -0.5, -2.75, 3.75, -0.5]))
    x = torch.zeros(5, dtype=torch.float64, device="cuda")
    res = Gmres(5, 5, 1e-14).solve(A, rhs, x)
    torch.cuda.synchronize()
    assert res.success and res.n_iters <= 5
    assert np.allclose(x.cpu().numpy(), EXPECTED, rtol=1e-12, atol=1e-12)


def test_device_fgmres_known_answer():
    import torch
    from ibis_amd.steady import DenseOperator, FGmres
    A = DenseOperator(to_dev(_tridiag(5, -1.0, 2.0, -1.0)))
    P = DenseOperator(to_dev(2.0 * np.eye(5)))
    rhs = to_dev(np.array([2.0, 0.0, -3.5, 3.5, -0.5]))
    x = torch.zeros(5, dtype=torch.float64, device="cuda")
    res = FGmres(5, 5, 1e-14, 2, 1e-1).solve(A, rhs, x, preconditioner=P)
    torch.cuda.synchronize()
    assert res.success
    assert np.allclose(x.cpu().numpy(), EXPECTED, rtol=1e-12, atol=1e-12)


def test_device_gmres_larger_system_matches_numpy():
    """A 300 x 300 diagonally dominant system: device GMRES (one Arnoldi
    step per host synchronisation) against the oracle's GMRES."""
    import torch
    from ibis_amd.steady import DenseOperator, Gmres
    rng = np.random.default_rng(50)
    n = 300
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    b = rng.standard_normal(n)
    x_cpu, ok, k, res = ost.gmres(lambda v: a @ v, b, np.zeros(n), 40, 1e-12)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    r = Gmres(n, 40, 1e-12).solve(DenseOperator(to_dev(a)), to_dev(b), x)
    torch.cuda.synchronize()
    assert r.success == ok and abs(r.n_iters - k) <= 1
    assert np.allclose(x.cpu().numpy(), x_cpu, rtol=1e-10, atol=1e-12)
    assert np.linalg.norm(a @ x.cpu().numpy() - b) <= 1e-11 * np.linalg.norm(b)


def _steady_case():
    g_io, cfg, fs = cases.wedge_case(24, 12)
    nt = 24 * 12 + 2 * (24 + 12)
    prim = cases.perturbed_prim(g_io, fs, nt, seed=51)
    og, ofv = oracle_pair(g_io, cfg)
    cons = ofv.prim_to_cons(prim)
    gb, fv = gpu_objects(g_io, cfg)
    return og, ofv, gb, fv, prim, cons


def _to_grid_vec(gb, vec_ref, N, nc):
    """cell-major vector in reference cell order -> grid cell order."""
    full = np.zeros((nc, gb.Nt))
    full[:, :N] = vec_ref.reshape(N, nc).T
    return np.ascontiguousarray(gb.to_grid_order(full)[:, :N].T).reshape(-1)


def _to_ref_vec(gb, vec_grid, N, nc):
    full = np.zeros((nc, gb.Nt))
    full[:, :N] = vec_grid.reshape(N, nc).T
    return np.ascontiguousarray(gb.to_reference_order(full)[:, :N].T).reshape(-1)


def test_steady_matvec_matches_oracle():
    import torch
    from ibis_amd.steady import SteadyStateLinearisation
    og, ofv, gb, fv, prim, cons = _steady_case()
    N, nc = og.N, ofv.ncons
    v = np.random.default_rng(52).standard_normal(N * nc)
    dt_star = 3e-6
    for recon in (True, False):
        want = ost.steady_matvec(ofv, cons, v, dt_star, recon)
        d_cons = to_dev(gb.to_grid_order(cons))
        sys_ = SteadyStateLinearisation(fv, d_cons, recon)
        sys_.set_pseudo_time_step(dt_star)
        out = torch.zeros(N * nc, dtype=torch.float64, device="cuda")
        sys_.matrix_vector_product(to_dev(_to_grid_vec(gb, v, N, nc)), out)
        torch.cuda.synchronize()
        got = _to_ref_vec(gb, out.cpu().numpy(), N, nc)
        print(f"steady matvec recon={recon}", flush=True)
        assert np.array_equal(got, want)


def test_jfnk_steps_match_oracle_gmres():
    import torch
    from ibis_amd.steady import Jfnk
    og, ofv, gb, fv, prim, cons = _steady_case()
    N, nc = og.N, ofv.ncons
    cfl = 3.0
    jf = Jfnk(fv, {"max_steps": 3, "tolerance": 1e-10,
                   "linear_solver": {"type": "gmres", "max_iters": 40, "tol": 1e-9},
                   "cfl": {"type": "constant", "value": cfl}})
    d_cons = to_dev(gb.to_grid_order(cons))
    d_prim = to_dev(gb.to_grid_order(prim))
    norms0 = jf.initialise(d_cons, d_prim)
    r0 = ofv.compute_dudt(prim.copy())
    assert norms0["global"] == pytest.approx(np.sqrt((r0[:, :N] ** 2).sum()), rel=1e-12)
    c_cpu, p_cpu = cons.copy(), prim.copy()
    for step in range(2):
        dt_star = cfl * ofv.estimate_dt(p_cpu)
        dU, c_cpu, p_cpu, (ok, k, res) = ost.jfnk_step(ofv, c_cpu, p_cpu, dt_star, 40, 1e-9)
        lr = jf.step(d_cons, d_prim, step)
        torch.cuda.synchronize()
        print(f"jfnk step {step}: gpu iters {lr.n_iters} cpu iters {k}", flush=True)
        assert jf.last_dt_star == pytest.approx(dt_star, rel=1e-14)
        assert lr.success and ok
        c_gpu = gb.to_reference_order(d_cons.cpu().numpy())
        scale = np.abs(c_cpu[:, :N]).max(axis=1)
        diff = np.abs(c_gpu[:, :N] - c_cpu[:, :N]).max(axis=1) / scale
        # dU is solved to 1e-9 relative by both; dU is a fraction of U
        assert (diff <= 1e-8).all(), diff
        c_cpu = c_gpu.copy()  # re-sync, so the steps stay comparable
        p_cpu = ofv.cons_to_prim(c_cpu)
