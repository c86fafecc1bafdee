"""CPU restatement of the reference's Krylov solvers and JFNK step.

TEST INFRASTRUCTURE ONLY: the checker for ibis_amd's device GMRES / JFNK.
Nothing in the product imports this module.

  gmres / fgmres    src/linear_algebra/linear_algebra/gmres.cpp:11-303
  steady matvec     src/solvers/solvers/steady_state.cpp:52-128
  jfnk_step         src/solvers/solvers/jfnk.cpp:39-93
  cfl schedules     src/solvers/solvers/cfl.cpp:9-68

Pinned by the reference's own GMRES / FGMRES 5x5 known answers
(gmres.cpp:320-500) and CFL schedule test (cfl.cpp:86-96); the matvec is
pinned through the oracle's Dual residual (itself checked against central
finite differences) by the identity J v = (R(U) - R(U - h v)) / h + O(h).
"""
from __future__ import annotations

import numpy as np


def _rotate(H, g, cs, sn, j):
    """apply_rotations_to_hessenberg_ (gmres.cpp:22-74) as sequential Givens
    rotations: previous rotations on column j, then a new one."""
    for i in range(j):
        a, b = H[i, j], H[i + 1, j]
        H[i, j] = cs[i] * a + sn[i] * b
        H[i + 1, j] = -sn[i] * a + cs[i] * b
    denom = np.sqrt(H[j, j] * H[j, j] + H[j + 1, j] * H[j + 1, j])
    si = H[j + 1, j] / denom
    ci = H[j, j] / denom
    cs[j], sn[j] = ci, si
    H[j, j] = ci * H[j, j] + si * H[j + 1, j]
    H[j + 1, j] = 0.0
    gj = g[j]
    g[j] = ci * gj
    g[j + 1] = -si * gj


def gmres(matvec, rhs, x0, max_iters, tol, precondition=None):
    """Gmres::solve / FGmres::solve.  ``precondition(v)`` returns z ~ P^-1 v
    (FGMRES); returns (x, success, n_iters, residual)."""
    n = rhs.size
    m = min(n, max_iters)
    x = x0.copy()
    r0 = rhs - matvec(x)
    beta = np.sqrt(np.dot(r0, r0))
    if beta < tol:
        return x, True, 0, beta
    V = np.zeros((m + 1, n))
    Z = np.zeros((m, n))
    H = np.zeros((m + 1, m))
    g = np.zeros(m + 1)
    cs, sn = np.zeros(m), np.zeros(m)
    g[0] = beta
    V[0] = r0 * (1.0 / beta)
    success, k, residual = False, 0, beta
    for j in range(m):
        if precondition is not None:
            Z[j] = precondition(V[j])
            w = matvec(Z[j])
        else:
            w = matvec(V[j])
        for i in range(j + 1):
            H[i, j] = np.dot(w, V[i])
            w = w + V[i] * (-H[i, j])
        H[j + 1, j] = np.sqrt(np.dot(w, w))
        # a breakdown (w == 0) makes this basis vector non-finite, as in
        # LA/gmres.cpp:119-182; the rotation then zeroes g[j + 1], so the
        # residual test below ends the loop before the vector is used
        with np.errstate(divide="ignore", invalid="ignore"):
            V[j + 1] = w * (1.0 / H[j + 1, j])
        _rotate(H, g, cs, sn, j)
        residual = abs(g[j + 1]) / beta
        k = j + 1
        if abs(g[j + 1]) < tol * beta:
            success = True
            break
    y = np.zeros(k)
    y[k - 1] = g[k - 1] / H[k - 1, k - 1]
    for i in range(k - 2, -1, -1):
        s = g[i]
        for c in range(i + 1, k):
            s -= H[i, c] * y[c]
        y[i] = s / H[i, i]
    basis = Z if precondition is not None else V
    x = x + (basis[:k].T @ y) * 1.0
    return x, success, k, residual


def cfl_linear(times, cfls, t):
    """LinearSchedule::eval (cfl.cpp:23-41)."""
    if t <= times[0]:
        return cfls[0]
    if t >= times[-1]:
        return cfls[-1]
    for i in range(len(times) - 1):
        if times[i] <= t < times[i + 1]:
            frac = (t - times[i]) / (times[i + 1] - times[i])
            return cfls[i] + frac * (cfls[i + 1] - cfls[i])
    raise RuntimeError("unreachable")


def cons_vector(cons, N, nc):
    """cell-major vector of the valid cells' conserved values."""
    return np.ascontiguousarray(cons[:nc, :N].T).reshape(-1)


def steady_matvec(ofv, cons, vec, dt_star, allow_reconstruction=True):
    """SteadyStateLinearisation::matrix_vector_product with the oracle's Dual
    residual: 1/dt* v - dual(R(U + eps v))."""
    g = ofv.grid
    N, nc, Nt = g.N, ofv.ncons, g.N + g.G
    cq = np.zeros((nc, Nt))
    cq[:, :N] = cons[:, :N]
    cd = np.zeros((nc, Nt))
    cd[:, :N] = vec.reshape(N, nc).T
    prim, pdot = ofv.cons_to_prim_jvp(cq, cd)
    _, ddot = ofv.compute_dudt_jvp(prim, pdot, allow_reconstruction)
    return 1 / dt_star * vec - cons_vector(ddot, N, nc)


def jfnk_step(ofv, cons, prim, dt_star, max_iters, tol):
    """One Jfnk::step with GMRES at a given dt*: returns (dU, cons, prim, rhs)."""
    g = ofv.grid
    N, nc = g.N, ofv.ncons
    dudt = ofv.compute_dudt(prim.copy())
    rhs = cons_vector(dudt, N, nc)
    dU, ok, k, res = gmres(lambda v: steady_matvec(ofv, cons, v, dt_star), rhs,
                           np.zeros_like(rhs), max_iters, tol)
    new = cons.copy()
    new[:, :N] += dU.reshape(N, nc).T
    p = ofv.cons_to_prim(new)
    return dU, new, p, (ok, k, res)
