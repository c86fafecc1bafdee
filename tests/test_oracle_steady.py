"""The oracle's Krylov solvers, steady linearisation and CFL schedules
against the reference's own known answers, and the host side of the
device JFNK (configuration, CFL schedule through the C ABI)."""
import numpy as np
import pytest

from ibis_amd import cases
from helpers import oracle_pair
import steady as ost


def _tridiag(n, lo, d, hi):
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = d
        if i + 1 < n:
            a[i, i + 1] = hi
            a[i + 1, i] = lo
    return a


EXPECTED = np.array([1.0, 0.0, -1.0, 1.5, 0.5])


def test_gmres_known_answer():
    """gmres.cpp:320-400: Gmres{sys, 5, 1e-14} on the 5x5 system."""
    a = _tridiag(5, -1.0, 2.0, -0.5)
    rhs = np.array([2.0, -0.5, -2.75, 3.75, -0.5])
    x, ok, k, res = ost.gmres(lambda v: a @ v, rhs, np.zeros(5), 5, 1e-14)
    assert ok
    assert np.allclose(x, EXPECTED, rtol=1e-12, atol=1e-12)


def test_fgmres_known_answer():
    """gmres.cpp:403-500: FGmres{sys, 5, 1e-14, P = 2 I, 2, 1e-1}."""
    a = _tridiag(5, -1.0, 2.0, -1.0)
    p = 2.0 * np.eye(5)
    rhs = np.array([2.0, 0.0, -3.5, 3.5, -0.5])

    def precondition(v):
        z, *_ = ost.gmres(lambda u: p @ u, v, np.zeros(5), 2, 1e-1)
        return z

    x, ok, k, res = ost.gmres(lambda v: a @ v, rhs, np.zeros(5), 5, 1e-14, precondition)
    assert ok
    assert np.allclose(x, EXPECTED, rtol=1e-12, atol=1e-12)


def test_gmres_planted_error_is_caught():
    """A dropped Givens term must break the known answer (the check is not
    vacuous)."""
    a = _tridiag(5, -1.0, 2.0, -0.5)
    rhs = np.array([2.0, -0.5, -2.75, 3.75, -0.5])
    orig = ost._rotate

    def broken(H, g, cs, sn, j):
        orig(H, g, cs, sn, j)
        g[j + 1] = 0.0  # claims convergence immediately
    ost._rotate = broken
    try:
        x, *_ = ost.gmres(lambda v: a @ v, rhs, np.zeros(5), 5, 1e-14)
    finally:
        ost._rotate = orig
    assert not np.allclose(x, EXPECTED, rtol=1e-6)


def test_cfl_linear_schedule_known_answer_oracle_and_library():
    """cfl.cpp:86-96, through the oracle and through ibis_cfl_eval."""
    from ibis_amd.steady import CflSchedule
    times, cfls = [0.0, 1.0, 2.0], [0.1, 0.2, 0.5]
    s = CflSchedule({"type": "linear_interpolate", "times": times, "cfls": cfls})
    for t, want in ((0.0, 0.1), (0.5, 0.15), (1.5, 0.35), (2.5, 0.5)):
        assert ost.cfl_linear(times, cfls, t) == pytest.approx(want, rel=1e-14)
        assert s.eval(t) == ost.cfl_linear(times, cfls, t)


def test_cfl_residual_based_closed_form():
    """ResidualBasedCfl::eval (cfl.cpp:58-68): above the threshold the CFL
    holds and only the stored residual moves; below it the CFL grows by
    (r_prev / r)^power, clipped to [start, max]."""
    from ibis_amd.steady import CflSchedule
    s = CflSchedule({"type": "residual_based", "growth_threshold": 0.5, "power": 1.0,
                     "start_cfl": 2.0, "max_cfl": 50.0})
    assert s.eval(0.9) == 2.0           # above threshold: prev_res <- 0.9
    c = 2.0 * (0.9 / 0.3)
    assert s.eval(0.3) == c
    assert s.eval(0.03) == 50.0         # clipped at max
    c = max(2.0, min(50.0 * (0.03 / 0.3), 50.0))
    assert s.eval(0.3) == c             # shrinks with the residual ratio
    assert s.eval(3.0) == c             # above threshold: held
    assert s.eval(0.45) == max(2.0, min(c * (3.0 / 0.45), 50.0))


def test_cfl_constant_and_bad_configs():
    from ibis_amd.native import IbisError
    from ibis_amd.steady import CflSchedule, jfnk_desc
    assert CflSchedule({"type": "constant", "value": 7.5}).eval(123.0) == 7.5
    with pytest.raises(IbisError):
        CflSchedule({"type": "banana"})
    with pytest.raises(IbisError):
        jfnk_desc({"linear_solver": {"type": "cg"}, "cfl": {"type": "constant", "value": 1}})
    d = jfnk_desc({"linear_solver": {"type": "fgmres", "max_iters": 30, "tolerance": 1e-3,
                                     "max_preconditioner_iters": 10,
                                     "preconditioner_tolerance": 1e-2},
                   "cfl": {"type": "constant", "value": 4.0}})
    assert (d.linear_solver, d.max_iters, d.max_precond_iters) == (1, 30, 10)
    assert (d.tol, d.precond_tol, d.cfl_value) == (1e-3, 1e-2, 4.0)


def _small_steady():
    g_io, cfg, fs = cases.wedge_case(16, 8)
    nt = 16 * 8 + 2 * (16 + 8)
    prim = cases.perturbed_prim(g_io, fs, nt, seed=41)
    og, ofv = oracle_pair(g_io, cfg)
    cons = ofv.prim_to_cons(prim)
    return og, ofv, prim, cons


def test_oracle_steady_matvec_matches_finite_difference():
    """J v from the Dual residual against a central difference of the plain
    residual in the conserved variables (steady_state.cpp:52-111)."""
    og, ofv, prim, cons = _small_steady()
    N, nc = og.N, ofv.ncons
    rng = np.random.default_rng(42)
    v = rng.standard_normal(N * nc) * np.tile(np.abs(cons[:, :N]).max(axis=1), N) * 1e-3
    dt_star = 1e-4
    jv = ost.steady_matvec(ofv, cons, v, dt_star)

    def R(c):
        full = cons.copy()
        full[:, :N] = c
        return ost.cons_vector(ofv.compute_dudt(ofv.cons_to_prim(full)), N, nc)

    h = 1e-6
    base = cons[:, :N]
    dv = v.reshape(N, nc).T
    fd = (R(base + h * dv) - R(base - h * dv)) / (2 * h)
    want = 1 / dt_star * v - fd
    err = np.abs(jv - want).max() / np.abs(want).max()
    assert err < 1e-6, err


def test_oracle_jfnk_step_solves_the_linear_system():
    """The Newton update satisfies (I/dt* - J) dU = R to the GMRES tolerance."""
    og, ofv, prim, cons = _small_steady()
    N, nc = og.N, ofv.ncons
    dt_star = 5.0 * ofv.estimate_dt(prim)
    dU, new, p, (ok, k, res) = ost.jfnk_step(ofv, cons, prim, dt_star, 60, 1e-8)
    rhs = ost.cons_vector(ofv.compute_dudt(prim.copy()), N, nc)
    lin = ost.steady_matvec(ofv, cons, dU, dt_star) - rhs
    assert ok and k > 0 and np.isfinite(dU).all()
    assert np.linalg.norm(lin) <= 1e-7 * np.linalg.norm(rhs)
    assert np.array_equal(new[:, :N], cons[:, :N] + dU.reshape(N, nc).T)
