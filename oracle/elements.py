"""Oracle: Q1 / Hex8 element kernels in FP64 numpy (test infrastructure only).

Shape functions restate autopdex/spaces.py:1913-1922 (Q1) and
:12057-12081 (Hex8); the physical gradient is (df/dxi) . (dX/dxi)^-1 as in the
custom JVP at spaces.py:12162-12191; the element potential is
sum_q w_q det(J_q) psi_q with the SIGNED determinant
(autopdex/models.py:1188-1269, det via utility.matrix_det :771-817).

Residuals and tangents are analytic derivatives of the potential (the
reference obtains them with jacrev / jacfwd(jacrev), assembler.py:1080-1140).
Pinned by: finite differences of ``potential`` (residual) and of ``residual``
(tangent), an independent torch float64 autograd of the same potential
(tests/test_oracle_elements.py), symmetry, zero row sums (Poisson), P(I)=0 and
the linear-elastic limit of the Neo-Hooke tangent, and the reference's 5x5
known-answer test end to end.

Element DOF order is node-major, component-fastest (a*d + i), as produced by
dict_flatten of the (npe, d) local DOF array (assembler.py:109-116).
"""
import numpy as np

# --- shape functions ------------------------------------------------------------


def q1_shape(xi):
    """spaces.py:1913-1922: nodes at (-1,-1), (1,-1), (1,1), (-1,1)."""
    x0 = 0.5 * xi[0]
    x1 = 0.5 - x0
    x2 = 0.5 * xi[1]
    x3 = 0.5 - x2
    x4 = x0 + 0.5
    x5 = x2 + 0.5
    N = np.array([x1 * x3, x3 * x4, x4 * x5, x1 * x5])
    dN = np.array([[-0.5 * x3, -0.5 * x1],
                   [0.5 * x3, -0.5 * x4],
                   [0.5 * x5, 0.5 * x4],
                   [-0.5 * x5, 0.5 * x1]])
    return N, dN


HEX8_SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                       [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)


def hex8_shape(xi):
    """spaces.py:12057-12081: N_a = (1+s0 xi0)(1+s1 xi1)(1+s2 xi2)/8, node order above."""
    f = 1.0 + HEX8_SIGNS * np.asarray(xi)[None, :]
    N = f[:, 0] * f[:, 1] * f[:, 2] / 8.0
    dN = np.stack([HEX8_SIGNS[:, 0] * f[:, 1] * f[:, 2],
                   f[:, 0] * HEX8_SIGNS[:, 1] * f[:, 2],
                   f[:, 0] * f[:, 1] * HEX8_SIGNS[:, 2]], axis=1) / 8.0
    return N, dN


def line2_shape(xi):
    """2-node line: N = ((1-xi)/2, (1+xi)/2)."""
    return np.array([(1 - xi) / 2, (1 + xi) / 2]), np.array([-0.5, 0.5])


def lagrange3(x):
    """Quadratic Lagrange polynomials on the nodes -1, 0, 1 and their derivatives."""
    return (np.array([x * (x - 1) / 2, 1 - x * x, x * (x + 1) / 2]),
            np.array([x - 0.5, -2 * x, x + 0.5]))


def line3_shape(xi):
    """3-node line of spaces.fem_iso_line_quad_brick (line branch, spaces.py:353-380): node
    order xi = -1, 1, 0 (ends first, then the interior node)."""
    L, dL = lagrange3(xi)
    order = [0, 2, 1]
    return L[order], dL[order]


# quadrilateral branch of spaces.fem_iso_line_quad_brick: node a is the tensor product
# L_i(xi_0) L_j(xi_1) with (i, j) from its square_indices ordering (corners
# counter-clockwise from (-1,-1), then midsides bottom, right, top, left, then the centre);
# i, j index the 1D nodes -1, 0, 1
Q2_IJ = np.array([[0, 0], [2, 0], [2, 2], [0, 2], [1, 0], [2, 1], [1, 2], [0, 1], [1, 1]])


def q2_shape(xi):
    """9-node biquadratic quad in the reference's node order (Q2_IJ)."""
    Lx, dLx = lagrange3(xi[0])
    Ly, dLy = lagrange3(xi[1])
    i, j = Q2_IJ[:, 0], Q2_IJ[:, 1]
    return Lx[i] * Ly[j], np.stack([dLx[i] * Ly[j], Lx[i] * dLy[j]], axis=1)


def shape(npe, dim, xi):
    if npe == 4 and dim == 2:
        return q1_shape(xi)
    if npe == 9 and dim == 2:
        return q2_shape(xi)
    if npe == 8 and dim == 3:
        return hex8_shape(xi)
    raise NotImplementedError((npe, dim))


def _gauss(dim, npe=None):
    """The rule each element type is integrated with: 2 points per direction (order 2) for
    Q1 / Hex8 / 2-node lines, 3 points (order 4 = 2 x order 2 as the reference's Q2 user
    elements use) for Q2 quads and 3-node lines."""
    from .mesh import gauss_legendre_nd
    return gauss_legendre_nd(dim, 4 if npe in (3, 9) else 2)


def mat_det(m):
    """utility.matrix_det (utility.py:771-817): closed-form determinant of (..., n, n), n = 1, 2, 3."""
    n = m.shape[-1]
    if n == 1:
        return m[..., 0, 0]
    if n == 2:
        return m[..., 0, 0] * m[..., 1, 1] - m[..., 0, 1] * m[..., 1, 0]
    a, b, c = m[..., 0, 0], m[..., 0, 1], m[..., 0, 2]
    d, e, f = m[..., 1, 0], m[..., 1, 1], m[..., 1, 2]
    g, h, i = m[..., 2, 0], m[..., 2, 1], m[..., 2, 2]
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)


def mat_inv(m):
    """utility.matrix_inv (utility.py:705-769): adjugate / determinant of (..., n, n), n = 1, 2, 3."""
    n = m.shape[-1]
    if n == 1:
        return 1.0 / m
    if n == 2:
        a, b, c, d = m[..., 0, 0], m[..., 0, 1], m[..., 1, 0], m[..., 1, 1]
        inv_det = 1.0 / (a * d - b * c)
        return inv_det[..., None, None] * np.stack([np.stack([d, -b], -1), np.stack([-c, a], -1)], -2)
    a, b, c = m[..., 0, 0], m[..., 0, 1], m[..., 0, 2]
    d, e, f = m[..., 1, 0], m[..., 1, 1], m[..., 1, 2]
    g, h, i = m[..., 2, 0], m[..., 2, 1], m[..., 2, 2]
    inv_det = 1.0 / mat_det(m)
    adj = np.stack([np.stack([e * i - f * h, c * h - b * i, b * f - c * e], -1),
                    np.stack([f * g - d * i, a * i - c * g, c * d - a * f], -1),
                    np.stack([d * h - e * g, b * g - a * h, a * e - b * d], -1)], -2)
    return inv_det[..., None, None] * adj


def geometry(xe, dN):
    """Per element: J_ij = sum_a x_ai dN_aj, signed det J, G = dN J^-1 (shape (E, npe, dim))."""
    J = np.swapaxes(xe, 1, 2) @ dN
    det = mat_det(J)
    Jinv = mat_inv(J)
    G = dN[None] @ Jinv
    return det, G


# --- models --------------------------------------------------------------------


class PoissonSource:
    """User integrand of examples/miscellaneous/short_example.py:23-33:
    psi = 1/2 grad(phi).grad(phi) - b phi,
    b = amp (sin(freq x.x) - cos(freq |x - c|^2)), (amp, freq, c) = (20, 10, (1, 0.5)).
    """
    d = 1

    def __init__(self, amp=20.0, freq=10.0, center=(1.0, 0.5)):
        self.amp, self.freq, self.center = amp, freq, np.asarray(center, dtype=np.float64)

    def source(self, x):
        x2 = x - self.center
        return self.amp * (np.sin(self.freq * np.sum(x * x, -1)) - np.cos(self.freq * np.sum(x2 * x2, -1)))

    def _loop(self, xe, ue, nodal=None):
        E, npe, dim = xe.shape
        xq, wq = _gauss(dim)
        for q in range(len(wq)):
            N, dN = shape(npe, dim, xq[q])
            det, G = geometry(xe, dN)
            x = N @ xe
            grad = (ue[:, None, :] @ G)[:, 0, :]
            yield wq[q] * det, N, G, x, grad

    def potential(self, xe, ue, nodal=None):
        out = np.zeros(xe.shape[0])
        for wd, N, G, x, grad in self._loop(xe, ue):
            phi = ue @ N
            out += wd * (0.5 * np.sum(grad * grad, -1) - self.source(x) * phi)
        return out

    def residual(self, xe, ue, nodal=None):
        out = np.zeros(ue.shape)
        for wd, N, G, x, grad in self._loop(xe, ue):
            out += wd[:, None] * ((G @ grad[:, :, None])[:, :, 0] - self.source(x)[:, None] * N[None, :])
        return out

    def tangent(self, xe, ue, nodal=None):
        E, npe = ue.shape
        out = np.zeros((E, npe, npe))
        for wd, N, G, x, grad in self._loop(xe, ue):
            out += wd[:, None, None] * np.einsum("eai,ebi->eab", G, G)
        return out


class PoissonVarK:
    """Variable-coefficient Poisson: psi = 1/2 k |grad phi|^2 - f phi with
    k = exp(sum_a N_a g_a) and f = sum_a N_a f_a at each Gauss point (nodal g, f).
    The reference counterpart is a user integrand reading nodal data through
    settings (models.py:1188-1269; cf. models.poisson_weak coefficient_fun :96-134).
    ``nodal`` = (g_e, f_e), each (E, npe).
    """
    d = 1

    def _loop(self, xe, ue, nodal):
        E, npe, dim = xe.shape
        ge, fe = nodal
        xq, wq = _gauss(dim)
        for q in range(len(wq)):
            N, dN = shape(npe, dim, xq[q])
            det, G = geometry(xe, dN)
            grad = np.einsum("ea,eai->ei", ue, G)
            k = np.exp(ge @ N)
            f = fe @ N
            yield wq[q] * det, N, G, grad, k, f

    def potential(self, xe, ue, nodal):
        out = np.zeros(xe.shape[0])
        for wd, N, G, grad, k, f in self._loop(xe, ue, nodal):
            out += wd * (0.5 * k * np.sum(grad * grad, -1) - f * (ue @ N))
        return out

    def residual(self, xe, ue, nodal):
        out = np.zeros(ue.shape)
        for wd, N, G, grad, k, f in self._loop(xe, ue, nodal):
            out += wd[:, None] * (k[:, None] * np.einsum("eai,ei->ea", G, grad) - f[:, None] * N[None, :])
        return out

    def tangent(self, xe, ue, nodal):
        E, npe = ue.shape
        out = np.zeros((E, npe, npe))
        for wd, N, G, grad, k, f in self._loop(xe, ue, nodal):
            out += (wd * k)[:, None, None] * np.einsum("eai,ebi->eab", G, G)
        return out


def neo_hooke_psi(F, lam, mu):
    """models.py:1122-1146 (Ciarlet): mu/2 (trC - 3 - 2 lnJ) + lam/4 (J^2 - 1 - 2 lnJ).

    F: (..., n, n) with n = 2 (plane strain, F33 = 1 so trC gains +1) or 3.
    """
    n = F.shape[-1]
    trC = np.einsum("...ij,...ij->...", F, F) + (1.0 if n == 2 else 0.0)
    J = np.linalg.det(F)
    lnJ = np.log(J)
    return (mu / 2) * (trC - 3 - 2 * lnJ) + (lam / 4) * (J ** 2 - 1 - 2 * lnJ)


def neo_hooke_stress(F, lam, mu):
    """First Piola P = mu (F - F^-T) + lam/2 (J^2 - 1) F^-T."""
    J = np.linalg.det(F)
    FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
    return mu * (F - FiT) + (lam / 2) * (J ** 2 - 1)[..., None, None] * FiT


def neo_hooke_moduli(F, lam, mu):
    """A_iJkL = dP_iJ/dF_kL = mu d_ik d_JL + (mu - lam/2 (J^2-1)) Fi_Jk Fi_Li + lam J^2 Fi_Ji Fi_Lk."""
    n = F.shape[-1]
    J = np.linalg.det(F)
    Fi = np.linalg.inv(F)
    I = np.eye(n)
    c1 = mu - (lam / 2) * (J ** 2 - 1)
    c2 = lam * J ** 2
    A = (mu * np.einsum("ik,JL->iJkL", I, I)[None]
         + c1[:, None, None, None, None] * np.einsum("eJk,eLi->eiJkL", Fi, Fi)
         + c2[:, None, None, None, None] * np.einsum("eJi,eLk->eiJkL", Fi, Fi))
    return A


class NeoHooke:
    """Hyperelastic user potential psi(F), F = I + du/dX (models.py:1122-1146).

    dim 2 = plane strain (F embedded in 3x3 with F33 = 1, as in
    models.hyperelastic_steady_state_weak :963-967); dim 3 = full 3D.
    param = (lam, mu), reference order.
    """

    def __init__(self, lam, mu):
        self.lam, self.mu = float(lam), float(mu)

    def _loop(self, xe, ue):
        E, npe, dim = xe.shape
        xq, wq = _gauss(dim, npe)
        for q in range(len(wq)):
            N, dN = shape(npe, dim, xq[q])
            det, G = geometry(xe, dN)
            F = np.eye(dim)[None] + np.einsum("eai,eaJ->eiJ", ue, G)
            yield wq[q] * det, G, F

    def potential(self, xe, ue, nodal=None):
        out = np.zeros(xe.shape[0])
        for wd, G, F in self._loop(xe, ue):
            out += wd * neo_hooke_psi(F, self.lam, self.mu)
        return out

    def residual(self, xe, ue, nodal=None):
        out = np.zeros(ue.shape)
        for wd, G, F in self._loop(xe, ue):
            P = neo_hooke_stress(F, self.lam, self.mu)
            out += wd[:, None, None] * np.einsum("eiJ,eaJ->eai", P, G)
        return out.reshape(out.shape[0], -1)

    def tangent(self, xe, ue, nodal=None):
        E, npe, d = ue.shape
        out = np.zeros((E, npe, d, npe, d))
        for wd, G, F in self._loop(xe, ue):
            A = neo_hooke_moduli(F, self.lam, self.mu)
            out += wd[:, None, None, None, None] * np.einsum("eaJ,eiJkL,ebL->eaibk", G, A, G)
        return out.reshape(E, npe * d, npe * d)


class LineTraction:
    """Dead-load traction on boundary lines in 2D:
    Pi = - sum_q w_q |dX/dxi| t . u(xi_q), t = traction vector; 2-node lines with the
    2-point rule, 3-node lines with the 3-point rule (isoparametric_surface_element_galerkin,
    models.py:1723-1850, with neumann_weak :744-780: -virt_u . t times w |G_1 x E_3|).
    Residual r_ai = - sum_q w_q |dX/dxi| N_a t_i; tangent zero.
    """

    def __init__(self, traction):
        self.t = np.asarray(traction, dtype=np.float64)

    def residual(self, xe, ue, nodal=None):
        E, npe, dim = xe.shape
        xq, wq = _gauss(1, npe)
        out = np.zeros((E, npe, dim))
        for q in range(len(wq)):
            N, dN = line2_shape(xq[q]) if npe == 2 else line3_shape(xq[q])
            t = np.einsum("a,eai->ei", dN, xe)
            L = np.sqrt(np.sum(t * t, -1))
            out -= (wq[q] * L)[:, None, None] * N[None, :, None] * self.t[None, None, :]
        return out.reshape(E, -1)

    def potential(self, xe, ue, nodal=None):
        return np.einsum("ea,ea->e", self.residual(xe, ue), ue.reshape(ue.shape[0], -1))

    def tangent(self, xe, ue, nodal=None):
        E = xe.shape[0]
        n = ue.reshape(E, -1).shape[1]
        return np.zeros((E, n, n))
