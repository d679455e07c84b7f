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

Any other integrand raises NotImplementedError for the 'hip' backend -- there
is no silent fallback.
"""
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


class Integrand:
    """Descriptor of a device integrand."""
    kernel = None

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

    def params(self, settings, n_dim):
        lam = settings[self.lam] if isinstance(self.lam, str) else self.lam
        mu = settings[self.mu] if isinstance(self.mu, str) else self.mu
        return [float(lam), float(mu)]


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

    def __init__(self, traction, multiplier_key="load multiplier"):
        self.traction = np.asarray(traction, dtype=np.float64)
        self.multiplier_key = multiplier_key

    def dofs_per_node(self, n_dim):
        return n_dim

    def params(self, settings, n_dim):
        m = float(settings.get(self.multiplier_key, 1.0)) if self.multiplier_key else 1.0
        return list(self.traction * m)

    def nodal_fields(self, settings):
        return None

    def __call__(self, *args, **kwargs):
        raise NotImplementedError("autopdex_amd user potentials are evaluated by device kernels only")


def line_traction_potential(traction, multiplier_key="load multiplier"):
    return LineTractionPotential(traction, multiplier_key)
