"""Host-side logic of the product package (no GPU): mesher, geometry, seeder,
utility and model descriptors agree with the oracle / reference semantics, and
unsupported inputs raise instead of falling back."""
import numpy as np
import pytest

from autopdex_amd import geometry, mesher, models, seeder, spaces, utility
from oracle import mesh as OM
from tests import problems as P


@pytest.mark.parametrize("n,verts,kind", [((7, 5), P.TAPERED, "quad"), ((3, 4, 2), P.UNIT_CUBE, "brick")])
def test_mesher_matches_oracle(n, verts, kind):
    c, e = mesher.structured_mesh(n, verts, kind)
    co, eo = OM.structured_mesh(n, verts, kind)
    assert np.array_equal(e, eo) and e.dtype == np.int64
    assert np.allclose(c, co, rtol=0, atol=1e-14)


def test_masks_match_oracle_bitwise():
    c, _ = mesher.structured_mesh((64, 64), P.UNIT_SQUARE, "quad")
    m = geometry.in_sdfs(c, lambda x: geometry.psdf_polygon(x, P.UNIT_SQUARE))
    mo = OM.in_sdfs(c, lambda x: OM.psdf_polygon(x, P.UNIT_SQUARE))
    assert np.array_equal(m, mo)
    # single point call as in the reference (vmap over points)
    assert geometry.in_sdf(c[1], lambda x: geometry.psdf_polygon(x, P.UNIT_SQUARE)) == mo[1]
    assert np.array_equal(geometry.in_planes(c, [0, 0], [1, 0]), OM.in_planes(c, [0, 0], [1, 0]))


def test_gauss_and_utility():
    for dim in (1, 2, 3):
        x, w = seeder.gauss_legendre_nd(dim, 2)
        xo, wo = OM.gauss_legendre_nd(dim, 2)
        assert np.array_equal(x, xo) and np.array_equal(w, wo)
    d = {"a": np.arange(4.0), "b": np.arange(6.0).reshape(3, 2)}
    flat = utility.dict_flatten(d)
    assert flat.size == 10 and np.array_equal(utility.reshape_as(flat, d)["b"], d["b"])
    sel = {"a": np.array([1, 0, 0, 1], bool), "b": np.zeros((3, 2), bool)}
    out = utility.mask_op(d, sel, {"a": np.full(4, 9.0), "b": np.zeros((3, 2))}, "set")
    assert list(out["a"]) == [9, 1, 2, 9]
    assert np.array_equal(utility.dof_select(np.array([True, False]), [True, False]), [[True, False], [False, False]])


def test_space_evaluation():
    xI = np.array([[0, 0], [2, 0], [2, 1], [0, 1]], float)
    fI = np.array([1.0, 2.0, 3.0, 4.0])
    # xi=(-1,-1) is node 0
    assert spaces.fem_iso_line_quad_brick(np.array([-1.0, -1.0]), xI, fI) == 1.0
    assert abs(spaces.fem_iso_line_quad_brick(np.zeros(2), xI, fI) - 2.5) < 1e-15


def test_model_descriptors_reject_unsupported():
    with pytest.raises(NotImplementedError):
        models.mixed_reference_domain_potential(lambda *a: 0.0, {"phi": spaces.fem_iso_line_quad_brick},
                                                *seeder.gauss_legendre_nd(2, 2), "phi")
    with pytest.raises(NotImplementedError):
        models.mixed_reference_domain_potential(models.poisson_source_integrand(),
                                                {"phi": spaces.fem_iso_line_quad_brick},
                                                *seeder.gauss_legendre_nd(2, 4), "phi")
    with pytest.raises(NotImplementedError):
        models.hyperelastic_integrand(lambda F, p: 0.0, 1.0, 1.0)
    pot = models.mixed_reference_domain_potential(models.hyperelastic_integrand(models.neo_hooke, "lam", 0.4),
                                                  {"u": spaces.fem_iso_line_quad_brick},
                                                  *seeder.gauss_legendre_nd(2, 2), "u")
    assert pot.params({"lam": 40.0}, 2) == [40.0, 0.4] and pot.dofs_per_node(2) == 2
    with pytest.raises(NotImplementedError):
        pot()


def test_solver_rejects_other_backends_and_types():
    from autopdex_amd import solver
    dofs, settings, static, o = P.poisson_source(4)
    with pytest.raises(NotImplementedError):
        solver.solver(dofs, settings, dict(static, **{"solver backend": "scipy"}))
    with pytest.raises(NotImplementedError):
        solver.solver(dofs, settings, dict(static, **{"solver type": "minimize"}))
    o64 = solver.newton_options(static)
    assert o64.reuse_tangent == 1 and o64.precision == 0 and abs(o64.lin_tol - 1e-5) < 1e-20
    dofs, settings, static, o = P.neo_hooke_cube(2)
    assert solver.newton_options(static).reuse_tangent == 0


def test_random_field_statistics():
    """learned.random_field: zero mean, unit variance, squared-exponential correlation
    exp(-d^2 / (2 l^2)) at a lag d (averaged over seeds; CPU)."""
    from autopdex_amd import learned
    H, length = 65, 0.1
    lag = 8  # nodes; d = lag / (H - 1)
    d = lag / (H - 1)
    vals, corr = [], []
    for s in range(40):
        f = learned.random_field(H, H, length, 1.0, seed=s, device="cpu").numpy().reshape(H, H)
        vals.append(f.ravel())
        corr.append(np.mean(f[:, :-lag] * f[:, lag:]))
    v = np.concatenate(vals)
    assert abs(v.mean()) < 0.1 and abs(v.var() - 1.0) < 0.15
    assert abs(np.mean(corr) - np.exp(-d ** 2 / (2 * length ** 2))) < 0.1


def test_default_levels_reach_a_3x3_grid():
    from autopdex_amd import learned
    assert learned.default_levels(2049, 2049) == 11   # 2049 -> 1025 -> ... -> 5 -> 3
    assert learned.default_levels(1025, 1025) == 10
    assert learned.default_levels(257, 257) == 8
    assert learned.default_levels(33, 33) == 5
    assert learned.default_levels(513, 129) == 7      # limited by the shorter side
    assert learned.default_levels(3, 3) == 1


@pytest.mark.parametrize("frac", [1.0, -0.1, 1.5])
def test_fit_rejects_self_fraction_outside_unit_interval(frac):
    """The flexible-CG pair phase takes a fraction in [0, 1) of the steps; anything else is
    rejected before any device work (no problem object is needed to reach the check)."""
    from autopdex_amd import learned
    with pytest.raises(ValueError, match="self_fraction"):
        learned.fit(None, 9, 9, self_fraction=frac)
