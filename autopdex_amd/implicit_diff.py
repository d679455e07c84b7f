"""Implicit differentiation of the 'hip' backend's roots (host + device), the counterpart of
autopdex.implicit_diff.custom_root (implicit_diff.py:487-542) with the device linear solve as
the 'sensitivity solver backend' (solver.py:203-293).

The reference makes its root solvers differentiable for JAX: custom_jvp / custom_vjp rules
around _root_jvp / _root_vjp (implicit_diff.py:97-337).  There is no JAX here, so the same rules
are explicit functions of a converged root u(theta) of r(u, theta) = 0, where theta are scalar
entries of settings that the element parameters are computed from (Young's modulus, Poisson's
ratio, the load multiplier, ...):

- root_jvp: u_dot = -K^-1 (dr/dtheta . theta_dot) on the free DOFs, Dirichlet DOFs fixed
  (_root_jvp :219-337);
- root_vjp: for a cotangent g, w = -K^-T g and d<g, u>/dtheta_k = w . dr/dtheta_k
  (_root_vjp :103-217);
- root_jvp2: the second derivative u_ij = d2u/dtheta_i dtheta_j that the reference gets by
  nesting forward rules, -K^-1 (r_uu[u_i, u_j] + r_u,i u_j + r_u,j u_i + r_ij) on the free DOFs.

Everything runs on the device through the C ABI: K is the FP64 reduced tangent at the root and
every solve is Jacobi-PCG (apx_pcg) to 'sensitivity linear tolerance' (default 1e-12).  The
device models are linear in their parameters (Neo-Hooke in (lam, mu), tractions in t), so
dr/dtheta . v and dK/dtheta . v are one residual or tangent assembly with the element parameters
replaced by their derivative along v.  That derivative of the settings -> parameters map (the
descriptors' own code, e.g. lam = E nu / ((1 + nu)(1 - 2 nu)) of a user function of settings)
is taken by complex step, exact for analytic maps.  Two difference quotients remain, both
Richardson-extrapolated central differences (truncation O(h^4)): r_uu[a, b] = d/de K(u + e b) a,
the third derivative of the energy, and the second derivative of the parameter map.
"""
import ctypes as C

import numpy as np

from . import _native as N
from . import utility
from .backend import current_stream, get_problem

_CS = 1e-30  # complex step


class _Root:
    """The root u and its reduced tangent K on the device, with parameter-swapped assemblies."""

    def __init__(self, dofs, settings, static_settings):
        import torch
        self.torch = torch
        self.prob = prob = get_problem(dofs, settings, static_settings)
        for i, m in enumerate(prob.models):
            if not getattr(m, "linear_in_params", False):
                raise NotImplementedError(f"set {i}: implicit differentiation needs a model that is linear in its "
                                          "parameters (Neo-Hooke, tractions)")
        self.dofs, self.settings = dofs, settings
        self.tol = float(static_settings.get("sensitivity linear tolerance", 1e-12))
        self.maxiter = int(static_settings.get("sensitivity linear maxiter", 0))
        dmask = utility.dict_flatten(settings["dirichlet dofs"]).astype(bool)
        flat = prob.flat(dofs).copy()
        flat[dmask] = np.asarray(utility.dict_flatten(settings["dirichlet conditions"]), dtype=np.float64)[dmask]
        self.u = prob.to_device(flat)
        self.r2f = torch.from_numpy(prob.free_dofs().astype(np.int64)).to(self.u.device)
        self.params0 = [prob.raw_params(i, settings) for i in range(len(prob.models))]
        self.K = self.tangent(self.u)

    # -- device primitives -------------------------------------------------------------------
    def residual(self, u):
        r = u.new_empty(self.prob.n_dofs)
        N.call("apx_assemble_residual", self.prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(r.data_ptr()),
               current_stream())
        return r

    def tangent(self, u):
        vals = u.new_empty(self.prob.nnz)
        N.call("apx_assemble_tangent", self.prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals.data_ptr()),
               N.APX_FP64, current_stream())
        return vals

    def spmv(self, vals, x):
        x = x.contiguous()
        y = x.new_empty(self.prob.n_free)
        N.call("apx_spmv", self.prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(x.data_ptr()),
               C.c_void_p(y.data_ptr()), N.APX_FP64, current_stream())
        return y

    def solve(self, b):
        """K x = b on the free DOFs (K symmetric: the potentials' tangents are)."""
        b = b.contiguous()
        x = b.new_empty(self.prob.n_free)
        ks = N.KrylovStats()
        N.call("apx_pcg", self.prob.plan, C.c_void_p(self.K.data_ptr()), C.c_void_p(b.data_ptr()),
               C.c_void_p(x.data_ptr()), N.APX_FP64, N.APX_PRECOND_JACOBI, self.tol, 0.0, self.maxiter,
               C.byref(ks), current_stream())
        if not ks.converged:
            raise RuntimeError(f"sensitivity solve did not converge ({ks.iterations} iterations, residual "
                               f"{ks.final_residual:.3e} from {ks.initial_residual:.3e})")
        return x

    def full(self, x_free):
        out = self.u.new_zeros(self.prob.n_dofs)
        out[self.r2f] = x_free
        return out

    # -- parameter derivatives ------------------------------------------------------------------
    def _with(self, params):
        for i, p in enumerate(params):
            self.prob.set_params(i, np.real(np.asarray(p, dtype=np.complex128)))

    def restore(self):
        self._with(self.params0)

    def dparams(self, tangents, shift=None):
        """d params / d theta . theta_dot by complex step (settings[k] += i h theta_dot[k]); shift
        moves the base point by a real offset {key: delta} first."""
        s = dict(self.settings)
        for k, dv in (shift or {}).items():
            s[k] = s[k] + dv
        for k, tv in tangents.items():
            s[k] = s[k] + 1j * _CS * tv
        return [np.imag(np.asarray(self.prob.raw_params(i, s), dtype=np.complex128)) / _CS
                for i in range(len(self.prob.models))]

    def d2params(self, t_i, t_j, rel=1e-3):
        """d2 params / d theta_i d theta_j along (t_i, t_j): central differences of the complex-step
        first derivative along t_j, Richardson-extrapolated."""
        d = rel * max(max(abs(complex(self.settings[k])) for k in t_j), 1.0)

        def diff(h):
            plus = self.dparams(t_i, {k: h * v for k, v in t_j.items()})
            minus = self.dparams(t_i, {k: -h * v for k, v in t_j.items()})
            return [(a - b) / (2 * h) for a, b in zip(plus, minus)]

        d1, d2 = diff(d), diff(2 * d)
        return [(4 * a - b) / 3 for a, b in zip(d1, d2)]

    def r_theta(self, dp):
        """dr/dtheta . v on the free DOFs: the residual at u with the parameters dp."""
        self._with(dp)
        try:
            return self.residual(self.u)[self.r2f]
        finally:
            self.restore()

    def K_theta(self, dp):
        self._with(dp)
        try:
            return self.tangent(self.u)
        finally:
            self.restore()

    def r_uu(self, a_full, b_full, rel=1e-4):
        """d/de K(u + e b) a at e = 0 (free rows), Richardson-extrapolated central differences."""
        nb = float(self.torch.linalg.norm(b_full))
        if nb == 0.0:
            return a_full.new_zeros(self.prob.n_free)
        h = rel * max(float(self.torch.linalg.norm(self.u)), 1.0) / nb
        a = a_full[self.r2f]

        def diff(e):
            kp = self.spmv(self.tangent(self.u + e * b_full), a)
            km = self.spmv(self.tangent(self.u - e * b_full), a)
            return (kp - km) / (2 * e)

        return (4 * diff(h) - diff(2 * h)) / 3


def _tangent_dict(t):
    if isinstance(t, str):
        return {t: 1.0}
    return dict(t)


def root_jvp(dofs, settings, static_settings, tangents):
    """Forward sensitivity of the root `dofs` (a converged solution for `settings`) along
    tangents = {settings key: tangent value} (or one key for a unit tangent).  Returns du with the
    structure of dofs (zero on Dirichlet DOFs).  implicit_diff._root_jvp (implicit_diff.py:219-337)."""
    R = _Root(dofs, settings, static_settings)
    du = R.full(R.solve(-R.r_theta(R.dparams(_tangent_dict(tangents)))))
    return utility.reshape_as(du.cpu().numpy(), dofs)


def root_vjp(dofs, settings, static_settings, cotangent, keys):
    """Reverse sensitivity: d<cotangent, u>/d settings[key] for each key, with one adjoint solve
    K^T w = -cotangent (free DOFs).  implicit_diff._root_vjp (implicit_diff.py:103-217)."""
    R = _Root(dofs, settings, static_settings)
    g = R.prob.to_device(np.ascontiguousarray(utility.dict_flatten(cotangent), dtype=np.float64))[R.r2f]
    w = R.solve(-g)
    return {k: float(w @ R.r_theta(R.dparams({k: 1.0}))) for k in ([keys] if isinstance(keys, str) else keys)}


def root_jvp2(dofs, settings, static_settings, tangents_i, tangents_j):
    """Second-order forward sensitivity u_ij = d2u / dtheta_i dtheta_j of the root along two
    tangents (forward over forward).  Returns (u_i, u_j, u_ij), each with the structure of dofs."""
    R = _Root(dofs, settings, static_settings)
    t_i, t_j = _tangent_dict(tangents_i), _tangent_dict(tangents_j)
    dp_i, dp_j = R.dparams(t_i), R.dparams(t_j)
    u_i = R.full(R.solve(-R.r_theta(dp_i)))
    u_j = u_i if t_i == t_j else R.full(R.solve(-R.r_theta(dp_j)))
    rhs = (R.r_uu(u_i, u_j) + R.spmv(R.K_theta(dp_i), u_j[R.r2f]) + R.spmv(R.K_theta(dp_j), u_i[R.r2f])
           + R.r_theta(R.d2params(t_i, t_j)))
    u_ij = R.full(R.solve(-rhs))
    return tuple(utility.reshape_as(x.cpu().numpy(), dofs) for x in (u_i, u_j, u_ij))
