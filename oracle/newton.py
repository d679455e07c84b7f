"""Oracle: Newton loop and linear solvers (test infrastructure only).

``newton`` restates solver.solve_newton / solve_damped_newton / damped_newton
(solver.py:740-948) with solve_linear (:544-659):
- P5: Dirichlet entries are overwritten with their prescribed values before
  every assembly; the linear solve returns Dirichlet *values* and free
  *increments*; free DOFs are incremented and Dirichlet DOFs are *set*.
- P6: the stopping test is absolute, ||r_free||_2 > tol; at least one
  iteration; divergence when ||r||/||r_old|| > 10 with itt > 1, or NaN/Inf;
  after ``maxiter`` the loop stops with diverged=True.  Returns
  (dofs, (n_iter, res_norm, diverged)).
- P4: the system is the reduced one, rhs = -r[free].

``pcg`` restates the preconditioned CG that solver.py:1116 calls
(jax.scipy.sparse.linalg.cg; third-party, its published algorithm): x0 = 0,
r0 = b - A x0, p0 = z0 = M r0, gamma = r.z, stop when r.r <= max(tol^2 b.b,
atol^2) or k = maxiter.  ``fcg`` is flexible CG (Notay 2000, Polak-Ribiere
beta) used by the learned preconditioner.

Pinned by: scipy SuperLU (the reference's own direct solver, solver.py:1519)
and the reference's 5x5 known-answer test; PCG is checked against SuperLU and
against scipy.sparse.linalg.cg with the same Jacobi M.
"""
import numpy as np
import scipy.sparse.linalg as spla


def pcg(A, b, M, tol=1e-5, atol=0.0, maxiter=None, x0=None, history=None):
    """Jacobi/any-M PCG; returns (x, iterations)."""
    n = b.shape[0]
    maxiter = 10 * n if maxiter is None else maxiter
    x = np.zeros(n) if x0 is None else x0.copy()
    atol2 = max(tol * tol * float(b @ b), atol * atol)
    r = b - A @ x
    z = M(r)
    p = z.copy()
    gamma = float(r @ z)
    k = 0
    while float(r @ r) > atol2 and k < maxiter:
        Ap = A @ p
        alpha = gamma / float(p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = M(r)
        gamma_new = float(r @ z)
        beta = gamma_new / gamma
        gamma = gamma_new
        p = z + beta * p
        k += 1
        if history is not None:
            history.append(np.sqrt(float(r @ r)))
    return x, k


def fcg(A, b, M, tol=1e-5, maxiter=1000, x0=None):
    """Flexible CG (Polak-Ribiere beta, one-term truncation, Notay 2000)."""
    n = b.shape[0]
    x = np.zeros(n) if x0 is None else x0.copy()
    r = b - A @ x
    bn = np.sqrt(float(b @ b))
    z = M(r)
    p = z.copy()
    gamma = float(r @ z)
    k = 0
    r_old = None
    while np.sqrt(float(r @ r)) > tol * bn and k < maxiter:
        Ap = A @ p
        alpha = gamma / float(p @ Ap)
        x = x + alpha * p
        r_old = r
        r = r - alpha * Ap
        k += 1
        if np.sqrt(float(r @ r)) <= tol * bn:
            break
        z_new = M(r)
        beta = float(z_new @ (r - r_old)) / gamma
        gamma = float(r @ z_new)
        p = z_new + beta * p
    return x, k


def linear_solve(A, b, kind="lapack", tol=1e-5, atol=0.0, maxiter=None):
    if kind == "lapack":
        return spla.spsolve(A.tocsc(), b), 0
    if kind == "cg":
        d = A.diagonal()
        return pcg(A, b, lambda v: v / d, tol=tol, atol=atol, maxiter=maxiter)
    raise ValueError(kind)


def solve_linear(dofs, residual_fn, tangent_fn, dirichlet_mask, dirichlet_values, lin_solve):
    """solve_linear (solver.py:544-659) on flat DOFs with nodal imposition: the Dirichlet entries
    of dofs are overwritten with their prescribed values (:584-604), the residual and reduced
    tangent are assembled there, A x = -r[free] is solved (P4), and the result holds the
    Dirichlet *values* at Dirichlet DOFs and the *increment* x at free DOFs (:646-656)."""
    dmask = np.asarray(dirichlet_mask, dtype=bool).ravel()
    free = ~dmask
    dvals = np.asarray(dirichlet_values, dtype=np.float64).ravel()
    imp = np.asarray(dofs, dtype=np.float64).ravel().copy()
    imp[dmask] = dvals[dmask]
    rhs = -residual_fn(imp)
    A = tangent_fn(imp)
    delta = imp.copy()
    delta[free] = lin_solve(A, rhs[free])
    return delta


def newton(u0, residual_fn, tangent_fn, dirichlet_mask, dirichlet_values, lin_solve,
           tol=1e-8, maxiter=30, damping=1.0, history=None):
    """damped_newton (solver.py:837-948) around solve_linear (:544-659) on flat DOFs.

    residual_fn(u) -> full residual; tangent_fn(u) -> reduced CSR;
    lin_solve(A, b) -> x (reduced).
    """
    dmask = np.asarray(dirichlet_mask, dtype=bool).ravel()
    free = ~dmask
    dvals = np.asarray(dirichlet_values, dtype=np.float64).ravel()
    dofs = np.asarray(u0, dtype=np.float64).ravel().copy()
    itt, res_old, not_stop, div, res = 0, 0.0, True, False, 0.0
    while not_stop:
        # solve_linear: impose Dirichlet values, assemble, solve, compose (P5)
        delta = solve_linear(dofs, residual_fn, tangent_fn, dmask, dvals, lin_solve)
        # damped update
        delta[free] *= damping
        dofs[free] += delta[free]
        dofs[dmask] = delta[dmask]
        r = residual_fn(dofs)
        rf = r[free]
        res = float(np.linalg.norm(rf))
        not_stop = res > tol
        if itt < maxiter:
            with np.errstate(divide="ignore", invalid="ignore"):
                div = bool((res / res_old > 10) and itt > 1) if res_old != 0 else False
            if np.any(~np.isfinite(rf)):
                div = True
            nxt = not div
        else:
            nxt, div = False, True
        if history is not None:
            history.append(res)
        itt += 1
        res_old = res
        not_stop = not_stop and nxt
    return dofs, (itt, res, div)


def adaptive_load_stepping(u0, solve_step, max_multiplier=1.0, min_increment=0.01, max_increment=1.0,
                           init_increment=0.2, target_num_newton_iter=7, log=None):
    """solver.adaptive_load_stepping without implicit differentiation (solver.py:296-379, the
    while_loop branch :455-457).  solve_step(u, multiplier) -> (u_new, (n_iter, res, diverged))
    is one Newton solve at that load multiplier from u.

    Loop while multiplier < max_multiplier and increment > min_increment:
      multiplier += increment; solve;
      diverged: multiplier -= increment, increment *= 0.5, keep the previous u;
      else: increment *= 1 + 0.5 (target - n_iter) / target;
      increment = min(increment, max_increment); if multiplier + increment > max_multiplier:
      increment = max_multiplier - multiplier.
    Returns (u, multiplier, increment, steps) with steps = [(multiplier, n_iter, diverged)] --
    the reference returns its carry (dofs, multiplier, increment, load_step, settings, res)."""
    u, multiplier, increment = np.array(u0, dtype=np.float64), 0.0, float(init_increment)
    steps = []
    while multiplier < max_multiplier and increment > min_increment:
        multiplier += increment
        u_new, (n_iter, res, diverged) = solve_step(u, multiplier)
        steps.append((multiplier, int(n_iter), bool(diverged)))
        if log:
            log(f"multiplier {multiplier:.6f}: newton {n_iter} res {res:.3e} diverged {diverged}")
        if diverged:
            multiplier -= increment
            increment *= 0.5
        else:
            u = u_new
            increment *= 1.0 + 0.5 * (target_num_newton_iter - n_iter) / target_num_newton_iter
        increment = min(increment, max_increment)
        if multiplier + increment > max_multiplier:
            increment = max_multiplier - multiplier
    return u, multiplier, increment, steps
