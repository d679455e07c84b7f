"""Oracle: implicit differentiation of a root u(theta) of r(u, theta) = 0 (test infrastructure only).

Restates the linear-solve rules of implicit_diff._root_jvp / _root_vjp (implicit_diff.py:97-337)
on flat FP64 DOFs with Dirichlet DOFs held fixed:
  forward:  u_dot[free] = -K^-1 (dr/dtheta . theta_dot)[free]
  reverse:  w = -K^-T g[free],  (d<g, u>/dtheta) = w . (dr/dtheta)[free]
and, for the second order that the reference gets by nesting jacfwd (forward over forward),
  u_ij[free] = -K^-1 (r_uu[u_i, u_j] + r_u,theta_i u_j + r_u,theta_j u_i + r_theta_i,theta_j)[free].
Here r_uu[a, b] (the third derivative of the energy) is the central difference of the tangent
along b applied to a -- a CPU check with an O(eps^2) truncation error, stated where it is used.

Pinned by the reference's own values for Cook's membrane
(tests/test_user_elem_impl_diff_and_adaptive_load_step.py:151-172): d(u.u)/dE, d(u.u)/dnu and the
four second derivatives, plus reverse = forward for the first derivatives.
"""
import numpy as np
import scipy.sparse.linalg as spla


def solve(K, b):
    return spla.spsolve(K.tocsc(), b)


def root_jvp(K, b_v):
    """K: reduced tangent at the root; b_v: (dr/dtheta . v) on the free DOFs."""
    return -solve(K, b_v)


def root_vjp(K, g, b_list):
    """g: cotangent on the free DOFs; b_list: dr/dtheta_k on the free DOFs for each parameter."""
    w = -solve(K.T.tocsr(), g)
    return [float(w @ b) for b in b_list]


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def lame_derivatives(E, nu):
    """d(lam, mu)/d(E, nu) and the second derivatives, closed form."""
    a = (1 + nu) * (1 - 2 * nu)
    dlam = np.array([nu / a, E * (1 + 2 * nu ** 2) / a ** 2])
    dmu = np.array([1 / (2 * (1 + nu)), -E / (2 * (1 + nu) ** 2)])
    # d/dnu of (1 + 2 nu^2) / a^2 with a' = -1 - 4 nu
    dd = (4 * nu * a ** 2 - (1 + 2 * nu ** 2) * 2 * a * (-1 - 4 * nu)) / a ** 4
    d2lam = np.array([[0.0, (1 + 2 * nu ** 2) / a ** 2], [(1 + 2 * nu ** 2) / a ** 2, E * dd]])
    d2mu = np.array([[0.0, -1 / (2 * (1 + nu) ** 2)], [-1 / (2 * (1 + nu) ** 2), E / (1 + nu) ** 3]])
    return dlam, dmu, d2lam, d2mu
