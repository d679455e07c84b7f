"""solver.solver with the 'hip' solver backend.

Same call and return as the reference (solver.py:41-137):
    sol, infos = solver(dofs, settings, static_settings, **kwargs)
with infos = (n_iter, res_norm, diverged) for 'newton' / 'damped newton'.

static_settings keys read here (reference keys unchanged):
  'solver type'            'newton' | 'damped newton' | 'linear'
  'solver backend'         must be 'hip'
  'solver'                 'cg' (Jacobi-PCG) -- the reference 'jax' backend's cg
  'type of preconditioner' 'jacobi' (default) | 'none' | 'learned'
  'verbose'                as in the reference
new keys of this backend:
  'hip precision'          'fp64' (default): FP64 tangent + FP64 PCG;
                           'mixed': FP32 tangent (Krylov matrix) for the inner
                           solves, the Newton residual and its norm stay FP64
                           (iterative refinement)
  'linear tolerance'       Krylov relative tolerance (default 1e-5, the
                           default of jax.scipy.sparse.linalg.cg that the
                           reference's newton mode uses, P7)
  'linear atol', 'linear maxiter'
  'newton tolerance mode'  'absolute' (reference, P6) | 'relative'
                           (tol * ||r_free(u0)||_2)
  'linear forcing'         True: Krylov tolerance max(linear tolerance,
                           0.5 tol/||r_k||) -- no inner solve tighter than the
                           Newton test needs (default False: fixed tolerance)
  'reliable delta'         mixed precision on models linear in u: one CG on the
                           FP32 matrix re-anchored to the FP64 residual each
                           time its residual drops by this factor (default 0.1;
                           0 turns it off and restarts CG at every Newton step)
  'hip tangent arithmetic' 'fp64' (default: in 'mixed' the FP64 tangent is
                           rounded once to FP32) | 'fp32' (element arithmetic of
                           the tangent in FP32)
  'hip krylov vectors'     'fp64' (default: with 'mixed', Jacobi-PCG keeps FP64
                           vectors and widens the FP32 matrix in the SpMV) |
                           'fp32' (FP32 vectors: an all-FP32 inner solve)
Newton semantics (P4-P6) are implemented in the device driver apx_newton
(solver.py:740-948).
"""
import ctypes as C

import numpy as np

from . import _native as N
from . import utility
from .backend import current_stream, get_problem


class NewtonInfos(tuple):
    """(n_iter, res_norm, diverged) plus .details (timings, Krylov counts, history)."""

    def __new__(cls, n_iter, res_norm, diverged, details):
        obj = super().__new__(cls, (n_iter, res_norm, diverged))
        obj.details = details
        return obj


def _opt(static_settings, key, default):
    try:
        return static_settings[key]
    except KeyError:
        return default


def _check_backend(static_settings):
    backend = static_settings["solver backend"]
    if backend != "hip":
        raise NotImplementedError(f"autopdex_amd provides the 'hip' solver backend only (got {backend!r})")
    solver = _opt(static_settings, "solver", "cg")
    if solver not in ("cg", "fcg"):
        raise NotImplementedError("the 'hip' backend solves with (flexible) preconditioned CG: 'solver': 'cg'")


def _precond(static_settings):
    p = _opt(static_settings, "type of preconditioner", "jacobi")
    table = {"jacobi": N.APX_PRECOND_JACOBI, "none": N.APX_PRECOND_NONE, None: N.APX_PRECOND_NONE,
             "learned": N.APX_PRECOND_LEARNED}
    if p not in table:
        raise NotImplementedError(f"preconditioner {p!r} is not available in the 'hip' backend")
    return table[p]


def newton_options(static_settings, newton_tol=1e-8, maxiter=30, damping=1.0, reuse_tangent=None):
    prec = _opt(static_settings, "hip precision", "fp64")
    if prec not in ("fp64", "mixed"):
        raise ValueError("'hip precision' must be 'fp64' or 'mixed'")
    o = N.NewtonOptions()
    o.newton_tol = float(newton_tol)
    o.relative = 1 if _opt(static_settings, "newton tolerance mode", "absolute") == "relative" else 0
    o.maxiter = int(maxiter)
    o.damping = float(damping)
    o.precision = N.APX_FP64 if prec == "fp64" else N.APX_FP32
    o.precond = _precond(static_settings)
    o.lin_tol = float(_opt(static_settings, "linear tolerance", 1e-5))
    o.lin_atol = float(_opt(static_settings, "linear atol", 0.0))
    o.lin_maxiter = int(_opt(static_settings, "linear maxiter", 0))
    if reuse_tangent is None:
        reuse_tangent = all(getattr(getattr(m, "integrand", m), "kernel", 0) in
                            (N.MODEL_POISSON_SOURCE, N.MODEL_POISSON_VARK, N.MODEL_LINE_TRACTION)
                            for m in static_settings["model"])
    o.reuse_tangent = 1 if reuse_tangent else 0
    o.forcing = 1 if _opt(static_settings, "linear forcing", False) else 0
    o.reliable_delta = float(_opt(static_settings, "reliable delta", 0.1)) if prec == "mixed" else 0.0
    arith = _opt(static_settings, "hip tangent arithmetic", "fp64")
    if arith not in ("fp64", "fp32"):
        raise ValueError("'hip tangent arithmetic' must be 'fp64' (FP64 tangent rounded to FP32 in 'mixed') or 'fp32'")
    o.fp32_element_arith = 1 if arith == "fp32" else 0
    vecs = _opt(static_settings, "hip krylov vectors", "fp64")
    if vecs not in ("fp64", "fp32"):
        raise ValueError("'hip krylov vectors' must be 'fp64' (FP32 matrix with FP64 Jacobi-PCG vectors) or 'fp32'")
    o.krylov_fp64_vectors = 1 if vecs == "fp64" else 0
    return o


def solver(dofs, settings, static_settings, **kwargs):
    _check_backend(static_settings)
    solver_type = static_settings["solver type"]
    match solver_type:
        case "newton":
            return solve_newton(dofs, settings, static_settings, **kwargs)
        case "damped newton":
            return solve_damped_newton(dofs, settings, static_settings, **kwargs)
        case "linear":
            return solve_linear(dofs, settings, static_settings, **kwargs), None
        case _:
            raise NotImplementedError(f"solver type {solver_type!r} is not available in the 'hip' backend")


def _load_multiplier(settings, multiplier):
    settings.update({"load multiplier": multiplier})
    return settings


def adaptive_load_stepping(dofs, settings, static_settings, multiplier_settings=_load_multiplier,
                           path_dependent=True, implicit_diff_mode=None, max_multiplier=1.0, min_increment=0.01,
                           max_increment=1.0, init_increment=0.2, max_load_steps=1000, target_num_newton_iter=7,
                           newton_tol=1e-10, **kwargs):
    """solver.adaptive_load_stepping (solver.py:139-457).

    implicit_diff_mode ('forward', 'backward' / 'reverse' or None) only decides, in the
    reference, which JAX derivative rule the returned root carries (:203-293); the root itself is
    the same.  Without JAX the derivatives of the returned root are computed explicitly by
    autopdex_amd.implicit_diff (root_jvp, root_vjp, root_jvp2) with the device linear solve, so
    the mode is validated and the solve runs unchanged.

    The increment control is the reference's (:296-379): multiplier += increment, one Newton
    solve; on divergence the step is undone and the increment halved, otherwise the increment
    is scaled by 1 + 0.5 (target - n_iter) / target; it is capped by max_increment and by what
    is left to max_multiplier; the loop runs while multiplier < max_multiplier and increment >
    min_increment.  The DOFs stay on the device between load steps; a diverged step is undone
    by a device copy.  Returns (dofs, multiplier, increment, load_step, settings, res_norm) like
    the reference's carry (its load_step is never incremented and stays 0)."""
    import torch
    if implicit_diff_mode not in (None, "forward", "backward", "reverse"):
        raise ValueError("implicit_diff_mode must be None, 'forward', 'backward' or 'reverse'")
    if implicit_diff_mode is not None and path_dependent:
        raise NotImplementedError("implicit differentiation of a path-dependent load history is not implemented "
                                  "(the reference differentiates the final root only with path_dependent=False)")
    _check_backend(static_settings)
    if static_settings["solver type"] not in ("newton", "damped newton"):
        raise NotImplementedError("adaptive load stepping works with solver types 'newton' and 'damped newton'")
    verbose = _opt(static_settings, "verbose", 0)
    damping = 1.0 if static_settings["solver type"] == "newton" else float(kwargs.get("damping_coefficient", 0.8))
    maxiter = int(kwargs.get("maxiter", 30))
    prob = get_problem(dofs, settings, static_settings)
    opts = newton_options(static_settings, newton_tol, maxiter, damping)
    u = prob.to_device(prob.flat(dofs))
    u_prev = torch.empty_like(u)
    multiplier, increment, res = 0.0, float(init_increment), 0.0
    steps = []
    while True:
        if multiplier < max_multiplier - min_increment and increment < min_increment:
            print("Adaptive load stepping could not converge; increment size below min_increment!")
        if not (multiplier < max_multiplier and increment > min_increment):
            break
        multiplier += increment
        if verbose > -1:
            print(f"Multiplier: {multiplier}")
        settings = multiplier_settings(settings, multiplier)
        prob.update(settings)
        dv = prob.to_device(np.ascontiguousarray(utility.dict_flatten(settings["dirichlet conditions"]),
                                                 dtype=np.float64))
        u_prev.copy_(u)
        st = N.NewtonStats()
        N.call("apx_newton", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(dv.data_ptr()), C.byref(opts),
               C.byref(st), current_stream())
        needed, res, diverged = st.iterations, st.residual_norm, bool(st.diverged)
        steps.append((multiplier, needed, diverged, st.krylov_iterations))
        if diverged:
            multiplier -= increment
            increment *= 0.5
            u.copy_(u_prev)
        else:
            increment *= 1.0 + 0.5 * (target_num_newton_iter - needed) / target_num_newton_iter
        increment = min(increment, max_increment)
        if multiplier + increment > max_multiplier:
            increment = max_multiplier - multiplier
    sol = utility.reshape_as(u.cpu().numpy(), dofs)
    adaptive_load_stepping.last_steps = steps  # (multiplier, Newton iterations, diverged, Krylov iterations)
    if implicit_diff_mode is not None:
        # the reference's differentiable branch returns (dofs, (multiplier, increment, load_step,
        # settings, res)) with settings at max_multiplier (:381-409)
        settings = multiplier_settings(settings, max_multiplier)
        return sol, (multiplier, increment, 0.0, settings, res)
    return sol, multiplier, increment, 0.0, settings, res


def solve_newton(dofs, settings, static_settings, newton_tol=1e-8, maxiter=30, **kwargs):
    """solver.py:740-768 (damping 1.0)."""
    return solve_damped_newton(dofs, settings, static_settings, newton_tol, 1.0, maxiter)


def solve_damped_newton(dofs, settings, static_settings, newton_tol=1e-8, damping_coefficient=0.8, maxiter=30,
                        **kwargs):
    """solver.py:770-948 on the device."""
    _check_backend(static_settings)
    if "nodal imposition" not in static_settings["solution structure"]:
        raise NotImplementedError("the 'hip' backend imposes Dirichlet conditions nodally ('nodal imposition')")
    prob = get_problem(dofs, settings, static_settings)
    opts = newton_options(static_settings, newton_tol, maxiter, damping_coefficient)
    u = prob.to_device(prob.flat(dofs))
    dv = prob.to_device(np.ascontiguousarray(utility.dict_flatten(settings["dirichlet conditions"]),
                                             dtype=np.float64))
    st = N.NewtonStats()
    N.call("apx_newton", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(dv.data_ptr()), C.byref(opts),
           C.byref(st), current_stream())
    sol = utility.reshape_as(u.cpu().numpy(), dofs)
    n = st.iterations
    details = {
        "krylov_iterations": st.krylov_iterations, "initial_residual": st.initial_residual,
        "seconds_total": st.seconds_total, "seconds_residual": st.seconds_residual,
        "seconds_tangent": st.seconds_tangent, "seconds_krylov": st.seconds_krylov,
        "fallbacks": st.fallbacks, "history": list(st.history[:min(n, 64)]),
    }
    verbose = _opt(static_settings, "verbose", 0)
    if verbose > 0:
        for i, r in enumerate(details["history"]):
            print(f"Residual after Newton iteration {i + 1}: {r}")
    if st.diverged and n > maxiter:
        print("Warning: Newton scheme could not converge!")
    return sol, NewtonInfos(n, st.residual_norm, bool(st.diverged), details)


def solve_linear(dofs, settings, static_settings, **kwargs):
    """solver.py:544-659: one linear solve at dofs; returns Dirichlet values at
    Dirichlet DOFs and the increment at free DOFs (P5)."""
    import torch
    _check_backend(static_settings)
    prob = get_problem(dofs, settings, static_settings)
    opts = newton_options(static_settings)
    dmask = utility.dict_flatten(settings["dirichlet dofs"]).astype(bool)
    dvals = np.asarray(utility.dict_flatten(settings["dirichlet conditions"]), dtype=np.float64)
    flat = prob.flat(dofs).copy()
    flat[dmask] = dvals[dmask]
    u = prob.to_device(flat)
    stream = current_stream()
    r = u.new_empty(prob.n_dofs)
    N.call("apx_assemble_residual", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(r.data_ptr()), stream)
    fp64 = opts.precision == N.APX_FP64
    dt = torch.float64 if fp64 else torch.float32
    vals = torch.empty(prob.nnz, dtype=dt, device=u.device)
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals.data_ptr()),
           opts.precision, stream)
    r2f = torch.from_numpy(prob.free_dofs().astype(np.int64)).to(u.device)
    b = (-r[r2f]).to(dt).contiguous()
    x = torch.empty_like(b)
    ks = N.KrylovStats()
    N.call("apx_pcg", prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
           opts.precision, opts.precond, opts.lin_tol, opts.lin_atol, opts.lin_maxiter, C.byref(ks), stream)
    out = u.clone()
    out[r2f] = x.to(torch.float64)
    return utility.reshape_as(out.cpu().numpy(), dofs)
