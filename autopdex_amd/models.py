"""Element models for the 'user potential' assembling mode (host descriptors).

The reference builds a user potential from a Python integrand with
models.mixed_reference_domain_potential (models.py:1188-1269) and lets JAX
differentiate it.  Device kernels cannot trace Python, so here the integrand
is a *descriptor* naming one of the hand-written element kernels plus its
parameters.  The descriptors below cover the integrands of the
assemble-and-solve path:

- poisson_source_integrand: the integrand of
  examples/miscellaneous/short_example.py:23-33 (and the reference's
  known-answer test tests/test_dicts_as_dofs_user_potential.py:24-34),
  psi = 1/2 |grad phi|^2 - b phi, b = 20 (sin(10 x.x) - cos(10 |x-(1,0.5)|^2));
- poisson_variable_coefficient_integrand: psi = 1/2 k |grad phi|^2 - f phi,
  k = exp(N.g), f = N.f from nodal arrays in settings;
- hyperelastic_integrand(neo_hooke, lam, mu): psi = neo_hooke(I + grad u),
  models.py:1122-1146, plane strain in 2D (F33 = 1, models.py:963-967);
- line_traction_potential: dead-load traction on 2-node boundary lines.

The 'user element' assembling mode (assembler.py:1309-1389) is covered for the
hyperelastic Galerkin elements of the reference: isoparametric_domain_element_galerkin
of hyperelastic_steady_state_weak(neo_hooke, E, nu, mode) on Q1 (2-point rule), Q2
(9 nodes, 3-point rule) or Hex8 elements, and isoparametric_surface_element_galerkin of
neumann_weak(t) on 2- or 3-node boundary lines.  E, nu and t are functions of (x) or
(x, settings) as in the reference; they are evaluated on the host per solver call and
must be uniform over their set.

Any other integrand raises NotImplementedError for the 'hip' backend -- there
is no silent fallback.
"""
from inspect import signature

import numpy as np

from . import _native as N
from . import seeder, spaces


def neo_hooke(F, param):
    """Ciarlet Neo-Hooke strain energy (models.py:1122-1146), param = (lam, mu)."""
    lam, mu = param
    F = np.asarray(F, dtype=np.float64)
    tr_C = np.einsum("ij,ij", F, F)
    J = np.linalg.det(F)
    ln_J = np.log(J)
    return (mu / 2) * (tr_C - 3 - 2 * ln_J) + (lam / 4) * (J ** 2 - 1 - 2 * ln_J)


def _scalar(v):
    """A real parameter as float; a complex one (complex-step differentiation of the parameter
    map, autopdex_amd.implicit_diff) is kept complex."""
    v = np.asarray(v)
    return complex(v) if np.iscomplexobj(v) else float(v)


class Integrand:
    """Descriptor of a device integrand."""
    kernel = None
    linear_in_params = False

    def dofs_per_node(self, n_dim):
        return 1

    def params(self, settings, n_dim):
        return []

    def nodal_fields(self, settings):
        return None

    def __call__(self, *args, **kwargs):
        raise NotImplementedError("autopdex_amd integrands are evaluated by device kernels only")


class _PoissonSource(Integrand):
    kernel = N.MODEL_POISSON_SOURCE

    def __init__(self, amplitude, frequency, center):
        self.amplitude, self.frequency = float(amplitude), float(frequency)
        self.center = tuple(float(c) for c in center)

    def params(self, settings, n_dim):
        return [self.amplitude, self.frequency, *self.center]


def poisson_source_integrand(amplitude=20.0, frequency=10.0, center=(1.0, 0.5)):
    return _PoissonSource(amplitude, frequency, center)


class _PoissonVarK(Integrand):
    kernel = N.MODEL_POISSON_VARK

    def __init__(self, log_coefficient_key, source_key):
        self.g_key, self.f_key = log_coefficient_key, source_key

    def nodal_fields(self, settings):
        return (np.asarray(settings[self.g_key], dtype=np.float64).ravel(),
                np.asarray(settings[self.f_key], dtype=np.float64).ravel())


def poisson_variable_coefficient_integrand(log_coefficient_key="log coefficient", source_key="source"):
    """k = exp(sum_a N_a g_a), f = sum_a N_a f_a with nodal g = settings[log_coefficient_key],
    f = settings[source_key]."""
    return _PoissonVarK(log_coefficient_key, source_key)


class _Hyperelastic(Integrand):
    kernel = N.MODEL_NEO_HOOKE

    def __init__(self, lam, mu):
        self.lam, self.mu = lam, mu

    def dofs_per_node(self, n_dim):
        return n_dim

    linear_in_params = True  # P = mu (F - F^-T) + lam/2 (J^2 - 1) F^-T

    def params(self, settings, n_dim):
        lam = settings[self.lam] if isinstance(self.lam, str) else self.lam
        mu = settings[self.mu] if isinstance(self.mu, str) else self.mu
        return [_scalar(lam), _scalar(mu)]


def hyperelastic_integrand(strain_energy_fun, lam, mu):
    """psi(I + grad u); lam / mu are numbers or settings keys.  Only neo_hooke is available."""
    if strain_energy_fun is not neo_hooke:
        raise NotImplementedError("the device kernels implement models.neo_hooke only")
    return _Hyperelastic(lam, mu)


class UserPotential:
    """Result of mixed_reference_domain_potential: a device element model."""

    def __init__(self, integrand, field_key, n_int):
        self.integrand, self.field_key, self.n_int = integrand, field_key, n_int

    @property
    def linear_in_params(self):
        return self.integrand.linear_in_params

    @property
    def kernel(self):
        return self.integrand.kernel

    def dofs_per_node(self, n_dim):
        return self.integrand.dofs_per_node(n_dim)

    def params(self, settings, n_dim):
        return self.integrand.params(settings, n_dim)

    def nodal_fields(self, settings):
        return self.integrand.nodal_fields(settings)

    def __call__(self, *args, **kwargs):
        raise NotImplementedError("autopdex_amd user potentials are evaluated by device kernels only")


def mixed_reference_domain_potential(integrand_fun, ansatz_fun, ref_int_coor, ref_int_weights, mapping_key):
    """Same call as models.mixed_reference_domain_potential (models.py:1188).

    integrand_fun must be an autopdex_amd integrand descriptor; ansatz_fun the
    dict {field: spaces.fem_iso_line_quad_brick}; the quadrature the 2-point
    Gauss-Legendre tensor rule.
    """
    if not isinstance(integrand_fun, Integrand):
        raise NotImplementedError(
            "the 'hip' solver backend runs hand-written element kernels; pass one of the autopdex_amd.models "
            "integrand descriptors (poisson_source_integrand, poisson_variable_coefficient_integrand, "
            "hyperelastic_integrand) instead of a Python integrand")
    if len(ansatz_fun) != 1 or any(f is not spaces.fem_iso_line_quad_brick for f in ansatz_fun.values()):
        raise NotImplementedError("only a single field with spaces.fem_iso_line_quad_brick is supported")
    x = np.asarray(ref_int_coor, dtype=np.float64)
    w = np.asarray(ref_int_weights, dtype=np.float64)
    dim = x.shape[1] if x.ndim == 2 else 1
    xr, wr = seeder.gauss_legendre_nd(dim, 2)
    if x.shape != xr.shape or not (np.allclose(x, xr, rtol=0, atol=1e-15) and np.allclose(w, wr, rtol=0, atol=1e-15)):
        raise NotImplementedError("the element kernels integrate with the 2-point Gauss-Legendre rule only")
    if mapping_key not in ansatz_fun:
        raise ValueError("mapping_key must be a key of ansatz_fun")
    return UserPotential(integrand_fun, mapping_key, x.shape[0])


class LineTractionPotential:
    """User potential of 2-node boundary lines: Pi = - sum_q w_q |dX/dxi| t . u.

    t = traction * settings[multiplier_key] (multiplier 1 when the key is absent).
    """
    kernel = N.MODEL_LINE_TRACTION
    linear_in_params = True

    def __init__(self, traction, multiplier_key="load multiplier"):
        self.traction = np.asarray(traction, dtype=np.float64)
        self.multiplier_key = multiplier_key

    def dofs_per_node(self, n_dim):
        return n_dim

    def params(self, settings, n_dim):
        m = _scalar(settings.get(self.multiplier_key, 1.0)) if self.multiplier_key else 1.0
        return [_scalar(v) for v in self.traction * m]

    def nodal_fields(self, settings):
        return None

    def __call__(self, *args, **kwargs):
        raise NotImplementedError("autopdex_amd user potentials are evaluated by device kernels only")


def line_traction_potential(traction, multiplier_key="load multiplier"):
    return LineTractionPotential(traction, multiplier_key)


# ---- 'user element' mode: hyperelastic Galerkin elements and Neumann lines ----------------

def _uniform(fun, sample_x, settings, what):
    """fun(x) or fun(x, settings) (the reference dispatches on the parameter count,
    models.py:937-944) at every point of sample_x (every distinct node of the element set);
    it must not vary over them, because the device kernels take it uniform.  A function that
    varies raises NotImplementedError."""
    if sample_x is None or len(sample_x) == 0:
        raise ValueError(f"{what}: sample points are needed to evaluate the user function")
    two = len(signature(fun).parameters) != 1

    def at(x):
        v = np.asarray(fun(np.asarray(x), settings) if two else fun(np.asarray(x)))
        return v.astype(np.complex128 if np.iscomplexobj(v) else np.float64)

    first = at(sample_x[0])
    for x in sample_x[1:]:
        v = at(x)
        if v.shape != first.shape or not np.allclose(v, first, rtol=1e-14, atol=0.0):
            raise NotImplementedError(f"{what} varies over the element set (at x = {np.asarray(x).tolist()}); "
                                      "the device kernels take it uniform")
    return first


class _HyperelasticWeak:
    """hyperelastic_steady_state_weak (models.py:917-1000) for neo_hooke: the virtual work
    P : virt_F with lam = E nu / ((1 + nu)(1 - 2 nu)), mu = E / (2 (1 + nu)); 'plain strain'
    embeds F in 3 x 3 with F33 = 1 (:963-967)."""
    kernel = N.MODEL_NEO_HOOKE

    def __init__(self, youngs_mod_fun, poisson_ratio_fun, mode):
        self.youngs_mod_fun, self.poisson_ratio_fun, self.mode = youngs_mod_fun, poisson_ratio_fun, mode

    def dofs_per_node(self, n_dim):
        return n_dim

    def lame(self, settings, sample_x):
        Em = _scalar(_uniform(self.youngs_mod_fun, sample_x, settings, "Young's modulus"))
        nu = _scalar(_uniform(self.poisson_ratio_fun, sample_x, settings, "Poisson's ratio"))
        lam = Em * nu / ((1 + nu) * (1 - 2 * nu))
        mu = Em / (2 * (1 + nu))
        return lam, mu


def hyperelastic_steady_state_weak(strain_energy_fun, youngs_mod_fun, poisson_ratio_fun, mode,
                                   volume_load_fun=None):
    """Same call as models.hyperelastic_steady_state_weak (models.py:917-1000)."""
    if strain_energy_fun is not neo_hooke:
        raise NotImplementedError("the device kernels implement models.neo_hooke only")
    if volume_load_fun is not None:
        raise NotImplementedError("volume loads are not implemented by the device kernels")
    assert mode in ("plain strain", "3d"), "Hyperelastic model supports only 'plain strain' and '3d' modes."
    return _HyperelasticWeak(youngs_mod_fun, poisson_ratio_fun, mode)


class _NeumannWeak:
    """neumann_weak (models.py:744-780): - virt_u . t(x[, settings])."""

    def __init__(self, neumann_fun):
        self.neumann_fun = neumann_fun


def neumann_weak(neumann_fun):
    """Same call as models.neumann_weak (models.py:744-780)."""
    return _NeumannWeak(neumann_fun)


def _points_per_direction(ref_int_coor, ref_int_weights):
    """2 or 3 when the rule is seeder.gauss_legendre_nd's 2- or 3-point tensor rule, else None."""
    x = np.asarray(ref_int_coor, dtype=np.float64)
    w = np.asarray(ref_int_weights, dtype=np.float64)
    dim = x.shape[1] if x.ndim == 2 else 1
    for order, n in ((2, 2), (4, 3)):
        xr, wr = seeder.gauss_legendre_nd(dim, order)
        if x.shape == xr.shape and np.allclose(x, xr, rtol=0, atol=1e-15) and np.allclose(w, wr, rtol=0, atol=1e-15):
            return n
    return None


class UserElement:
    """Result of isoparametric_domain_element_galerkin: a device element model.

    Q1 quads with the 2-point rule, Q2 (9-node) quads with the 3-point rule and Hex8 bricks
    with the 2-point rule map to the Neo-Hooke kernels."""
    user_element = True
    linear_in_params = True

    def __init__(self, weak, n_1d, dim):
        self.weak, self.n_1d, self.dim = weak, n_1d, dim

    def kernel_for(self, npe, n_dim):
        ok = {(4, 2): 2, (9, 2): 3, (8, 3): 2}
        if ok.get((npe, n_dim)) != self.n_1d or n_dim != self.dim:
            raise NotImplementedError(
                f"no device kernel for {npe}-node elements in {n_dim}D with a {self.n_1d}-point rule "
                "(Q1 / Hex8 with 2 points, Q2 with 3 points per direction)")
        if (n_dim == 2) != (self.weak.mode == "plain strain"):
            raise NotImplementedError("'plain strain' needs 2D elements and '3d' needs 3D elements")
        return self.weak.kernel

    def dofs_per_node(self, n_dim):
        return n_dim

    def params(self, settings, n_dim, sample_x=None):
        return list(self.weak.lame(settings, sample_x))

    def nodal_fields(self, settings):
        return None

    def __call__(self, *args, **kwargs):
        raise NotImplementedError("autopdex_amd user elements are evaluated by device kernels only")


def isoparametric_domain_element_galerkin(weak_form_fun, ansatz_fun, ref_int_coor, ref_int_weights,
                                          initial_config=True):
    """Same call as models.isoparametric_domain_element_galerkin (models.py:1616-1720)."""
    if not isinstance(weak_form_fun, _HyperelasticWeak):
        raise NotImplementedError("the device kernels implement hyperelastic_steady_state_weak(neo_hooke, ...)")
    if ansatz_fun is not spaces.fem_iso_line_quad_brick:
        raise NotImplementedError("only spaces.fem_iso_line_quad_brick is supported")
    if not initial_config:
        raise NotImplementedError("the device kernels map with the initial configuration")
    n = _points_per_direction(ref_int_coor, ref_int_weights)
    if n is None:
        raise NotImplementedError("the element kernels integrate with the 2- or 3-point Gauss-Legendre rule")
    x = np.asarray(ref_int_coor)
    return UserElement(weak_form_fun, n, x.shape[1] if x.ndim == 2 else 1)


class UserSurfaceElement:
    """Result of isoparametric_surface_element_galerkin with neumann_weak: dead-load
    traction on 2-node (2-point rule) or 3-node (3-point rule) boundary lines in 2D."""
    user_element = True
    linear_in_params = True
    kernel = N.MODEL_LINE_TRACTION

    def __init__(self, weak, n_1d):
        self.weak, self.n_1d = weak, n_1d

    def kernel_for(self, npe, n_dim):
        if n_dim != 2 or {2: 2, 3: 3}.get(npe) != self.n_1d:
            raise NotImplementedError("surface elements: 2-node lines with 2 points or 3-node lines with 3 "
                                      "points, in 2D")
        return self.kernel

    def dofs_per_node(self, n_dim):
        return n_dim

    def params(self, settings, n_dim, sample_x=None):
        t = _uniform(self.weak.neumann_fun, sample_x, settings, "traction")
        if t.shape != (n_dim,):
            raise ValueError("the traction function must return one value per displacement component")
        return [_scalar(v) for v in t]

    def nodal_fields(self, settings):
        return None

    def __call__(self, *args, **kwargs):
        raise NotImplementedError("autopdex_amd user elements are evaluated by device kernels only")


def isoparametric_surface_element_galerkin(weak_form_fun, ansatz_fun, ref_int_coor, ref_int_weights,
                                           tangent_contributions, initial_config=True):
    """Same call as models.isoparametric_surface_element_galerkin (models.py:1723-1850).  The
    Neumann weak form does not depend on u, so the tangent is zero whatever
    tangent_contributions says."""
    if not isinstance(weak_form_fun, _NeumannWeak):
        raise NotImplementedError("the device kernels implement surface elements of neumann_weak only")
    if ansatz_fun is not spaces.fem_iso_line_quad_brick:
        raise NotImplementedError("only spaces.fem_iso_line_quad_brick is supported")
    if not initial_config:
        raise NotImplementedError("the device kernels map with the initial configuration")
    n = _points_per_direction(ref_int_coor, ref_int_weights)
    if n is None:
        raise NotImplementedError("surface elements integrate with the 2- or 3-point Gauss-Legendre rule")
    return UserSurfaceElement(weak_form_fun, n)
