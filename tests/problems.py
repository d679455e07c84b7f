"""Problem builders shared by the tests: each returns the reference-style
(dofs, settings, static_settings) for the autopdex_amd 'hip' backend and the
matching oracle description (coords, sets, mask, values, d)."""
import numpy as np

from autopdex_amd import geometry, mesher, models, seeder, spaces, utility
from oracle import elements as OE
from oracle import mesh as OM

UNIT_SQUARE = [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]
TAPERED = [[0.0, 0.0], [48.0, 44.0], [48.0, 60.0], [0.0, 44.0]]
UNIT_CUBE = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]


def golden(name):
    """A known answer from tests/golden/reference_goldens.json (the reference's own test values)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")) as f:
        return json.load(f)[name]["value"]


def static(model_list, **extra):
    s = {
        "assembling mode": tuple("user potential" for _ in model_list),
        "solution structure": ("nodal imposition",),
        "model": tuple(model_list),
        "solver type": "newton",
        "solver backend": "hip",
        "solver": "cg",
        "type of preconditioner": "jacobi",
        "verbose": -1,
    }
    s.update(extra)
    return s


def poisson_source(n, dict_dofs=True):
    """Config-1/2 family: unit square, n x n Q1, psdf-polygon Dirichlet mask (corners free)."""
    coords, elems = mesher.structured_mesh((n, n), UNIT_SQUARE, "quad")
    mask = geometry.in_sdfs(coords, lambda x: geometry.psdf_polygon(x, UNIT_SQUARE))
    pot = models.mixed_reference_domain_potential(
        models.poisson_source_integrand(), {"phi": spaces.fem_iso_line_quad_brick},
        *seeder.gauss_legendre_nd(dimension=2, order=2), "phi")
    if dict_dofs:
        settings = {"connectivity": ({"phi": elems},), "dirichlet dofs": {"phi": mask},
                    "node coordinates": {"phi": coords},
                    "dirichlet conditions": utility.dict_zeros_like({"phi": mask})}
        dofs = utility.dict_zeros_like({"phi": mask})
    else:
        settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                    "dirichlet conditions": np.zeros(mask.shape)}
        dofs = np.zeros(mask.shape)
    oracle = dict(coords=coords, sets=[(OE.PoissonSource(), elems, None)], mask=mask.ravel(),
                  values=np.zeros(mask.size), d=1)
    return dofs, settings, static([pot]), oracle


def poisson_rect(nx, ny, verts=((0.0, 0.0), (2.0, 0.0), (2.0, 0.7), (0.0, 1.0))):
    """Poisson with source on an nx x ny Q1 quadrilateral (non-square grid, skewed
    elements), psdf-polygon Dirichlet mask with free corners."""
    verts = [list(v) for v in verts]
    coords, elems = mesher.structured_mesh((nx, ny), verts, "quad")
    mask = geometry.in_sdfs(coords, lambda x: geometry.psdf_polygon(x, verts))
    pot = models.mixed_reference_domain_potential(
        models.poisson_source_integrand(), {"phi": spaces.fem_iso_line_quad_brick},
        *seeder.gauss_legendre_nd(dimension=2, order=2), "phi")
    settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": np.zeros(mask.shape)}
    oracle = dict(coords=coords, sets=[(OE.PoissonSource(), elems, None)], mask=mask.ravel(),
                  values=np.zeros(mask.size), d=1)
    return np.zeros(mask.shape), settings, static([pot]), oracle


def poisson_vark_trapezoid(n, seed=0):
    """poisson_vark on a trapezoid: no element is a parallelogram."""
    return poisson_vark(n, seed, verts=((0.0, 0.0), (2.0, 0.0), (2.0, 0.7), (0.0, 1.0)))


def poisson_kinked(n):
    """Poisson with source on the unit square with a few interior nodes moved off the grid: the
    grid mixes parallelograms with general quadrilaterals."""
    dofs, settings, st, oracle = poisson_source(n, dict_dofs=False)
    coords = settings["node coordinates"].copy()
    h = 1.0 / n
    for k, (i, j) in enumerate(((2, 3), (n // 2, n // 3), (n - 3, n - 2))):
        coords[i * (n + 1) + j] += h * np.array([0.2 - 0.1 * k, 0.15])
    settings["node coordinates"] = coords
    oracle["coords"] = coords
    return dofs, settings, st, oracle


def grf(n_nodes_1d, seed, length=0.05, sigma=1.0):
    """Squared-exponential Gaussian random field on an (m x m) node grid of [0,1]^2
    by circulant embedding (periodic FFT on a doubled grid), seeded PCG64."""
    m = n_nodes_1d
    h = 1.0 / (m - 1)
    M2 = 2 * m
    k = np.minimum(np.arange(M2), M2 - np.arange(M2)) * h
    cov = sigma ** 2 * np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) / (2 * length ** 2))
    lam = np.maximum(np.fft.fft2(cov).real, 0.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    xi = rng.standard_normal((M2, M2)) + 1j * rng.standard_normal((M2, M2))
    f = np.fft.fft2(np.sqrt(lam / (M2 * M2)) * xi)
    return f.real[:m, :m].ravel()


def poisson_vark(n, seed=0, verts=None):
    """Config-4 family: -div(k grad u) = f, k = exp(g), g GRF, f ~ N(0,1) per node (on the unit
    square, or on the quadrilateral verts)."""
    verts = UNIT_SQUARE if verts is None else [list(v) for v in verts]
    coords, elems = mesher.structured_mesh((n, n), verts, "quad")
    mask = geometry.in_sdfs(coords, lambda x: geometry.psdf_polygon(x, verts))
    g = grf(n + 1, seed)
    f = np.random.Generator(np.random.PCG64(seed + 1_000_000)).standard_normal(coords.shape[0])
    pot = models.mixed_reference_domain_potential(
        models.poisson_variable_coefficient_integrand("log coefficient", "source"),
        {"u": spaces.fem_iso_line_quad_brick}, *seeder.gauss_legendre_nd(2, 2), "u")
    settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": np.zeros(mask.shape), "log coefficient": g, "source": f}
    oracle = dict(coords=coords, sets=[(OE.PoissonVarK(), elems, (g, f))], mask=mask.ravel(),
                  values=np.zeros(mask.size), d=1)
    return np.zeros(mask.shape), settings, static([pot]), oracle


def neo_hooke_cube(n, lam=10.0, mu=1.0, disp=0.1):
    """Config-3 family: unit cube n^3 Hex8, x=0 clamped, x=1 u_x = disp (y, z free)."""
    coords, elems = mesher.structured_mesh((n, n, n), UNIT_CUBE, "brick")
    left = geometry.in_planes(coords, [0, 0, 0], [1, 0, 0])
    right = geometry.in_planes(coords, [1, 0, 0], [1, 0, 0])
    mask = utility.dof_select(left, [True, True, True]) | utility.dof_select(right, [True, False, False])
    values = np.zeros(mask.shape)
    values[right, 0] = disp
    pot = models.mixed_reference_domain_potential(
        models.hyperelastic_integrand(models.neo_hooke, lam, mu), {"u": spaces.fem_iso_line_quad_brick},
        *seeder.gauss_legendre_nd(3, 2), "u")
    settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": values}
    oracle = dict(coords=coords, sets=[(OE.NeoHooke(lam, mu), elems, None)], mask=mask.ravel(),
                  values=values.ravel(), d=3)
    return np.zeros(mask.shape), settings, static([pot]), oracle


def tapered_panel(n, lam=40.0, mu=0.4, total_traction=1.0, multiplier=1.0, verts=None):
    """Config-5 family: tapered panel n x n Q1, x=0 clamped, vertical traction on x=48
    (verts: another quadrilateral with its right edge at x = 48 and length 16)."""
    coords, elems = mesher.structured_mesh((n, n), TAPERED if verts is None else verts, "quad")
    left = geometry.in_planes(coords, [0, 0], [1, 0])
    mask = utility.dof_select(left, [True, True])
    lines = mesher.boundary_lines((n, n), "right")
    t = np.array([0.0, total_traction / 16.0])  # right edge length 16
    pot = models.mixed_reference_domain_potential(
        models.hyperelastic_integrand(models.neo_hooke, lam, mu), {"u": spaces.fem_iso_line_quad_brick},
        *seeder.gauss_legendre_nd(2, 2), "u")
    trac = models.line_traction_potential(t, "load multiplier")
    settings = {"connectivity": (elems, lines), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": np.zeros(mask.shape), "load multiplier": multiplier}
    oracle = dict(coords=coords, sets=[(OE.NeoHooke(lam, mu), elems, None), (OE.LineTraction(t * multiplier), lines, None)],
                  mask=mask.ravel(), values=np.zeros(mask.size), d=2)
    return np.zeros(mask.shape), settings, static([pot, trac]), oracle


def cook_membrane_fixture():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "cook_membrane_q2.json")) as f:
        return json.load(f)


def cook_membrane_q2(multiplier=1.0):
    """The reference's Q2 user-element test problem (Cook's membrane, 12 Q2 elements, 63 nodes,
    tests/test_user_elem_impl_diff_and_adaptive_load_step.py:29-105): plane-strain Neo-Hooke
    with E = 100, nu = 0.3 through isoparametric_domain_element_galerkin, the x = 0 edge
    clamped, traction (0, load multiplier) on the 3-node lines of x = 48 through
    isoparametric_surface_element_galerkin, 3-point rules (order 2 * 2).  'load multiplier'
    is multiplier * q_0 as multiplier_settings of that test sets it."""
    fx = cook_membrane_fixture()
    x = np.asarray(fx["x_nodes"], dtype=np.float64)
    elems = np.asarray(fx["elements"], dtype=np.int64)
    surf = np.asarray(fx["surface_elements"], dtype=np.int64)
    pts = fx["pts"]
    left = geometry.in_planes(x, pts[0], [1.0, 0.0])
    mask = utility.dof_select(left, [True, True])
    neumann = surf[geometry.select_elements_in_plane(x, surf, pts[1], [1.0, 0.0])]
    weak1 = models.hyperelastic_steady_state_weak(models.neo_hooke, lambda x, s: s["youngs modulus"],
                                                  lambda x, s: s["poisson ratio"], "plain strain")
    elem1 = models.isoparametric_domain_element_galerkin(weak1, spaces.fem_iso_line_quad_brick,
                                                         *seeder.gauss_legendre_nd(dimension=2, order=4))
    weak2 = models.neumann_weak(lambda x, s: np.asarray([0.0, s["load multiplier"]]))
    elem2 = models.isoparametric_surface_element_galerkin(weak2, spaces.fem_iso_line_quad_brick,
                                                          *seeder.gauss_legendre_nd(dimension=1, order=4),
                                                          tangent_contributions=False)
    settings = {"dirichlet dofs": mask, "connectivity": (elems, neumann),
                "load multiplier": multiplier * fx["q_0"], "node coordinates": x,
                "dirichlet conditions": np.zeros(mask.shape),
                "youngs modulus": fx["youngs_modulus"], "poisson ratio": fx["poisson_ratio"]}
    st = static([elem1, elem2])
    st["assembling mode"] = ("user element", "user element")
    E, nu = fx["youngs_modulus"], fx["poisson_ratio"]
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    oracle = dict(coords=x, sets=[(OE.NeoHooke(lam, mu), elems, None),
                                  (OE.LineTraction([0.0, multiplier * fx["q_0"]]), neumann, None)],
                  mask=mask.ravel(), values=np.zeros(mask.size), d=2)
    return np.zeros(mask.shape), settings, st, oracle


def oracle_cook_membrane(E=None, nu=None, newton_tol=None, log=None):
    """The oracle's adaptive load stepping on cook_membrane_q2 with the reference's defaults
    (init_increment 0.2, target 7 Newton iterations) and a direct (LAPACK) linear solve, as the
    reference test runs it.  Returns (u, steps)."""
    from oracle import assembly as OA
    from oracle import newton as ON
    fx = cook_membrane_fixture()
    E = fx["youngs_modulus"] if E is None else E
    nu = fx["poisson_ratio"] if nu is None else nu
    tol = fx["newton_tol"] if newton_tol is None else newton_tol
    _, _, _, o = cook_membrane_q2()
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    (_, elems, _), (_, lines, _) = o["sets"]
    coords, mask = o["coords"], o["mask"]
    free = ~mask

    def step(u, m):
        sets = [(OE.NeoHooke(lam, mu), elems, None), (OE.LineTraction([0.0, m * fx["q_0"]]), lines, None)]
        return ON.newton(u, lambda v: OA.residual(sets, coords, v, 2),
                         lambda v: OA.reduced_tangent(sets, coords, v, 2, free), mask, o["values"],
                         lambda A, b: ON.linear_solve(A, b, "lapack")[0], tol=tol, maxiter=30)

    u, m, inc, steps = ON.adaptive_load_stepping(np.zeros(mask.size), step, log=log)
    return u, steps


def oracle_cook_sensitivities(u, eps=1e-6):
    """d(u.u)/d(E, nu) (forward and reverse) and the second derivatives of u.u at the converged
    Cook's-membrane solution u, by the oracle's implicit-differentiation rules (oracle/implicit_diff.py).
    The residual is linear in (lam, mu), so dr/dlam and dr/dmu are residuals with unit
    parameters; r_uu uses a central difference of the tangent with step eps * |u| / |u_j|."""
    from oracle import assembly as OA
    from oracle import implicit_diff as OI
    fx = cook_membrane_fixture()
    E, nu = fx["youngs_modulus"], fx["poisson_ratio"]
    _, _, _, o = cook_membrane_q2()
    (_, elems, _), line = o["sets"]
    coords, mask = o["coords"], o["mask"]
    free = ~mask
    lam, mu = OI.lame(E, nu)
    sets = [(OE.NeoHooke(lam, mu), elems, None), line]
    K = OA.reduced_tangent(sets, coords, u, 2, free)

    def unit(a, b):
        return [(OE.NeoHooke(a, b), elems, None)]

    r_l = OA.residual(unit(1.0, 0.0), coords, u, 2)[free]
    r_m = OA.residual(unit(0.0, 1.0), coords, u, 2)[free]
    K_l = OA.reduced_tangent(unit(1.0, 0.0), coords, u, 2, free)
    K_m = OA.reduced_tangent(unit(0.0, 1.0), coords, u, 2, free)
    dlam, dmu, d2lam, d2mu = OI.lame_derivatives(E, nu)
    uf = u[free]
    du = []
    for i in range(2):
        x = np.zeros(u.size)
        x[free] = OI.root_jvp(K, dlam[i] * r_l + dmu[i] * r_m)
        du.append(x)
    first = [2 * uf @ d[free] for d in du]
    first_rev = OI.root_vjp(K, 2 * uf, [dlam[i] * r_l + dmu[i] * r_m for i in range(2)])
    second = np.zeros((2, 2))
    for i in range(2):
        for j in range(2):
            h = eps * np.linalg.norm(u) / np.linalg.norm(du[j])
            Kp = OA.reduced_tangent(sets, coords, u + h * du[j], 2, free)
            Km = OA.reduced_tangent(sets, coords, u - h * du[j], 2, free)
            r_uu = (Kp @ du[i][free] - Km @ du[i][free]) / (2 * h)
            K_i = dlam[i] * K_l + dmu[i] * K_m
            K_j = dlam[j] * K_l + dmu[j] * K_m
            rhs = r_uu + K_i @ du[j][free] + K_j @ du[i][free] + d2lam[i, j] * r_l + d2mu[i, j] * r_m
            u_ij = OI.root_jvp(K, rhs)
            second[i, j] = 2 * (du[i][free] @ du[j][free] + uf @ u_ij)
    return np.array(first), np.array(first_rev), second


def oracle_newton(o, tol=1e-8, maxiter=30, kind="lapack", lin_tol=1e-10, u0=None):
    from oracle import assembly as OA
    from oracle import newton as ON
    coords, sets, mask, d = o["coords"], o["sets"], o["mask"], o["d"]
    free = ~mask
    u0 = np.zeros(mask.size) if u0 is None else u0
    return ON.newton(u0, lambda u: OA.residual(sets, coords, u, d),
                     lambda u: OA.reduced_tangent(sets, coords, u, d, free), mask, o["values"],
                     lambda A, b: ON.linear_solve(A, b, kind, tol=lin_tol)[0], tol=tol, maxiter=maxiter)
