"""GPU parity: HIP kernels (through the C ABI) against the FP64 oracle.

Tolerances: element residuals / tangents / diagonals are FP64 on both sides
and differ only by summation order -> max-abs error <= 1e-12 of the max-abs
value; the reduced CSR pattern is compared bit for bit; solutions reach the
reference's FP64-mode bar (relative solution error 1e-8, relative residual
1e-10).
"""
import ctypes as C

import numpy as np
import pytest

from tests import problems as P

pytestmark = pytest.mark.gpu


def _setup(builder, *a, **k):
    from autopdex_amd import backend
    dofs, settings, static, o = builder(*a, **k)
    prob = backend.get_problem(dofs, settings, static)
    return dofs, settings, static, o, prob


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


CASES = [
    ("poisson_q1", P.poisson_source, (12,)),
    ("poisson_vark_q1", P.poisson_vark, (12, 3)),
    ("neo_hooke_hex8", P.neo_hooke_cube, (4,)),
    ("neo_hooke_plane_strain_traction", P.tapered_panel, (8,)),
    # several owner tiles in each direction (tiles own 4 x 63 nodes), partial edge tiles
    ("poisson_q1_tiles", P.poisson_source, (130,)),
    ("poisson_rect_tiles", P.poisson_rect, (9, 200)),
    ("poisson_rect_tiles_t", P.poisson_rect, (70, 5)),
    ("poisson_vark_q1_tiles", P.poisson_vark, (70, 5)),
    ("neo_hooke_plane_strain_tiles", P.tapered_panel, (66,)),
    # 'user element' mode: Q2 Neo-Hooke and 3-node traction lines on an unstructured mesh
    ("cook_membrane_q2", P.cook_membrane_q2, ()),
    # owner tiles on grids that are not all parallelograms: the general per-point element path
    ("poisson_vark_trapezoid_tiles", P.poisson_vark_trapezoid, (70, 7)),
    ("poisson_kinked_tiles", P.poisson_kinked, (70,)),
]
# bit 0: owner-tile residual kernels in use; bit 1: owner-tile tangent too
STRUCTURED = {"poisson_q1": 3, "poisson_vark_q1": 3, "neo_hooke_hex8": 0, "neo_hooke_plane_strain_traction": 1,
              "poisson_q1_tiles": 3, "poisson_rect_tiles": 3, "poisson_rect_tiles_t": 3, "poisson_vark_q1_tiles": 3,
              "neo_hooke_plane_strain_tiles": 1, "cook_membrane_q2": 0, "poisson_vark_trapezoid_tiles": 3,
              "poisson_kinked_tiles": 3}


@pytest.mark.parametrize("name,builder,args", CASES, ids=[c[0] for c in CASES])
def test_element_assembly_matches_oracle(torch_gpu, name, builder, args):
    from autopdex_amd import _native as N
    from autopdex_amd.backend import current_stream
    from oracle import assembly as OA
    torch = torch_gpu
    print(f"[case] {name} {args}", flush=True)
    dofs, settings, static, o, prob = _setup(builder, *args)
    assert prob.info.structured == STRUCTURED[name]
    rng = np.random.default_rng(7)
    u = rng.standard_normal(prob.n_dofs) * (0.02 if o["d"] > 1 else 1.0)
    ud = torch.from_numpy(u).cuda()
    st = current_stream()
    # residual
    r = torch.empty_like(ud)
    N.call("apx_assemble_residual", prob.plan, C.c_void_p(ud.data_ptr()), C.c_void_p(r.data_ptr()), st)
    r_ref = OA.residual(o["sets"], o["coords"], u, o["d"])
    assert _rel(r.cpu().numpy(), r_ref) < 1e-12
    # reduced CSR: pattern bit-exact, values FP64 and FP32
    A_ref = OA.reduced_tangent(o["sets"], o["coords"], u, o["d"], ~o["mask"])
    indptr, indices = prob.csr_pattern()
    assert np.array_equal(indptr, A_ref.indptr) and np.array_equal(indices, A_ref.indices)
    v64 = torch.empty(prob.nnz, dtype=torch.float64, device="cuda")
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(ud.data_ptr()), C.c_void_p(v64.data_ptr()), N.APX_FP64, st)
    assert _rel(v64.cpu().numpy(), A_ref.data) < 1e-12
    v32 = torch.empty(prob.nnz, dtype=torch.float32, device="cuda")
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(ud.data_ptr()), C.c_void_p(v32.data_ptr()), N.APX_FP32, st)
    assert _rel(v32.cpu().numpy().astype(np.float64), A_ref.data) < 1e-5
    # the mixed method's matrix: FP64 arithmetic rounded once, bit for bit the FP64 values as float
    vr = torch.empty(prob.nnz, dtype=torch.float32, device="cuda")
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(ud.data_ptr()), C.c_void_p(vr.data_ptr()),
           N.APX_FP64_ROUNDED, st)
    assert np.array_equal(vr.cpu().numpy().view(np.uint32), v64.cpu().numpy().astype(np.float32).view(np.uint32))
    # diagonal (full length)
    g = torch.empty_like(ud)
    N.call("apx_assemble_diagonal", prob.plan, C.c_void_p(ud.data_ptr()), C.c_void_p(g.data_ptr()), st)
    g_ref = OA.diagonal(o["sets"], o["coords"], u, o["d"])
    assert _rel(g.cpu().numpy(), g_ref) < 1e-12


def test_assembly_is_bitwise_deterministic(torch_gpu):
    from autopdex_amd import _native as N
    from autopdex_amd.backend import current_stream
    torch = torch_gpu
    dofs, settings, static, o, prob = _setup(P.poisson_source, 64)
    u = torch.from_numpy(np.random.default_rng(1).standard_normal(prob.n_dofs)).cuda()
    st = current_stream()
    outs = []
    for _ in range(3):
        r = torch.empty_like(u)
        v = torch.empty(prob.nnz, dtype=torch.float64, device="cuda")
        N.call("apx_assemble_residual", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(r.data_ptr()), st)
        N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()), N.APX_FP64, st)
        outs.append((r.cpu().numpy().tobytes(), v.cpu().numpy().tobytes()))
    assert outs[0] == outs[1] == outs[2]


# The 40 x 40 grid's SELL columns are stored as int16 offsets from their rows; the rows of the
# 3 x 40000 strip (unit-square elements) reach 40,000 columns away, past int16, so it takes the
# int32 columns.
STRIP = ((0.0, 0.0), (3.0, 0.0), (3.0, 40000.0), (0.0, 40000.0))


@pytest.mark.parametrize("builder,args,int16_cols",
                         [(P.poisson_source, (40,), True), (P.poisson_rect, (3, 40000, STRIP), False)],
                         ids=["grid40_int16_columns", "strip3x40000_int32_columns"])
def test_spmv_and_pcg_match_oracle(torch_gpu, builder, args, int16_cols):
    from autopdex_amd import _native as N
    from autopdex_amd.backend import current_stream
    from oracle import assembly as OA
    from oracle import newton as ON
    torch = torch_gpu
    dofs, settings, static, o, prob = _setup(builder, *args)
    u = np.zeros(prob.n_dofs)
    A = OA.reduced_tangent(o["sets"], o["coords"], u, 1, ~o["mask"])
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    assert (np.abs(A.indices - rows).max() <= 32767) == int16_cols
    st = current_stream()
    vals = torch.from_numpy(A.data.copy()).cuda()
    x = np.random.default_rng(3).standard_normal(prob.n_free)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    N.call("apx_spmv", prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(xd.data_ptr()), C.c_void_p(y.data_ptr()),
           N.APX_FP64, st)
    assert _rel(y.cpu().numpy(), A @ x) < 1e-13
    b = A @ x
    bd = torch.from_numpy(b).cuda()
    sol = torch.empty_like(bd)
    ks = N.KrylovStats()
    N.call("apx_pcg", prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(bd.data_ptr()), C.c_void_p(sol.data_ptr()),
           N.APX_FP64, N.APX_PRECOND_JACOBI, 1e-10, 0.0, 0, C.byref(ks), st)
    d = A.diagonal()
    x_ref, k_ref = ON.pcg(A, b, lambda v: v / d, tol=1e-10)
    assert ks.converged == 1
    assert abs(ks.iterations - k_ref) <= max(2, int(0.02 * k_ref))
    assert np.linalg.norm(sol.cpu().numpy() - x) / np.linalg.norm(x) < 1e-8
    assert np.linalg.norm(b - A @ sol.cpu().numpy()) / np.linalg.norm(b) < 1.5e-10


def test_reference_known_answer_5x5(torch_gpu):
    """tests/test_dicts_as_dofs_user_potential.py:63 through solver.solver('hip')."""
    from autopdex_amd import solver
    dofs, settings, static, o = P.poisson_source(5)
    sol, infos = solver.solver(dofs, settings, static)
    assert np.isclose(sol["phi"].sum(), P.golden("short_example_5x5_sum_phi"), rtol=1e-10, atol=0)
    assert infos[2] is False and infos[1] <= 1e-8


def test_config1_200x200_matches_superlu(torch_gpu):
    from autopdex_amd import solver
    dofs, settings, static, o = P.poisson_source(200)
    static = dict(static, **{"linear tolerance": 1e-12, "newton tolerance mode": "relative"})
    sol, infos = solver.solver(dofs, settings, static, newton_tol=1e-10)
    ref, _ = P.oracle_newton(o)
    u = sol["phi"].ravel()
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-8
    assert infos[1] / infos.details["initial_residual"] <= 1e-10


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
def test_mixed_precision_refinement_reaches_fp64(torch_gpu, arith):
    """FP32 PCG with FP64 reliable updates reaches the FP64 solution, with the FP32 tangent
    either the FP64 one rounded (default) or assembled with FP32 element arithmetic."""
    from autopdex_amd import solver
    dofs, settings, static, o = P.poisson_source(128, dict_dofs=False)
    s64 = dict(static, **{"linear tolerance": 1e-12, "newton tolerance mode": "relative"})
    s32 = dict(static, **{"linear tolerance": 1e-4, "newton tolerance mode": "relative", "hip precision": "mixed",
                          "hip tangent arithmetic": arith})
    u64, i64 = solver.solver(dofs, settings, s64, newton_tol=1e-10)
    u32, i32 = solver.solver(dofs, settings, s32, newton_tol=1e-10)
    assert i32[2] is False
    assert i32[1] / i32.details["initial_residual"] <= 1e-10
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-8


def test_neo_hooke_cube_load_steps_match_oracle(torch_gpu):
    from autopdex_amd import solver
    n, steps = 4, 3
    u_dev = None
    u_ref = None
    for k in range(1, steps + 1):
        dofs, settings, static, o = P.neo_hooke_cube(n, disp=0.1 * k / steps)
        static = dict(static, **{"linear tolerance": 1e-12})
        if u_dev is None:
            u_dev, u_ref = dofs, np.zeros(o["mask"].size)
        sol, infos = solver.solver(u_dev, settings, static, newton_tol=1e-9)
        ref, rinfo = P.oracle_newton(o, tol=1e-9, u0=u_ref)
        assert infos[2] is False and rinfo[2] is False
        assert abs(infos[0] - rinfo[0]) <= 1
        assert np.linalg.norm(sol.ravel() - ref) / np.linalg.norm(ref) < 1e-8
        u_dev, u_ref = sol, ref


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tapered_panel_traction_matches_oracle(torch_gpu, precision):
    """Plane-strain Neo-Hooke with traction, Newton to 1e-9 (absolute).  'mixed' re-assembles the
    tangent every Newton step in FP64, rounds it to FP32 and solves in FP32 to 1e-5; the FP64
    Newton residual decides, so the solution meets the same bar."""
    from autopdex_amd import solver
    dofs, settings, static, o = P.tapered_panel(8, multiplier=1.0)
    if precision == "fp64":
        static = dict(static, **{"linear tolerance": 1e-12})
    else:
        static = dict(static, **{"linear tolerance": 1e-5, "hip precision": "mixed"})
    sol, infos = solver.solver(dofs, settings, static, newton_tol=1e-9)
    ref, rinfo = P.oracle_newton(o, tol=1e-9)
    assert infos[2] is False and rinfo[2] is False
    assert np.linalg.norm(sol.ravel() - ref) / np.linalg.norm(ref) < 1e-8


def test_linear_solver_type_matches_oracle_solve_linear(torch_gpu):
    """'solver type': 'linear' (solver.py:544-659) from a nonzero starting point with nonzero
    Dirichlet values: one linear solve at the imposed DOFs; the result holds the Dirichlet values
    at Dirichlet DOFs and the increment at free DOFs (P5), and infos is None (solver.py:41-137).
    Against the oracle's solve_linear with SuperLU, to 1e-8."""
    from autopdex_amd import solver
    from oracle import assembly as OA
    from oracle import newton as ON
    dofs, settings, static, o = P.poisson_source(24, dict_dofs=False)
    mask = o["mask"]
    rng = np.random.default_rng(12)
    values = np.where(mask, rng.standard_normal(mask.size), 0.0)
    u0 = 0.1 * rng.standard_normal(mask.size)
    settings = dict(settings, **{"dirichlet conditions": values})
    st = dict(static, **{"solver type": "linear", "linear tolerance": 1e-13})
    sol, infos = solver.solver(u0, settings, st)
    assert infos is None
    ref = ON.solve_linear(u0, lambda u: OA.residual(o["sets"], o["coords"], u, 1),
                          lambda u: OA.reduced_tangent(o["sets"], o["coords"], u, 1, ~mask), mask, values,
                          lambda A, b: ON.linear_solve(A, b, "lapack")[0])
    assert np.array_equal(sol.ravel()[mask], values[mask])
    assert np.linalg.norm(sol.ravel() - ref) / np.linalg.norm(ref) < 1e-8


def test_damped_newton_solver_type_matches_oracle(torch_gpu):
    """'solver type': 'damped newton' (solver.py:770-835, damped_newton :837-948) with the
    reference's default damping 0.8 on a Neo-Hooke cube: the same Newton count and solution as the
    oracle's damped Newton."""
    from autopdex_amd import solver
    dofs, settings, static, o = P.neo_hooke_cube(4, disp=0.05)
    st = dict(static, **{"solver type": "damped newton", "linear tolerance": 1e-12})
    sol, infos = solver.solver(dofs, settings, st, newton_tol=1e-9)
    from oracle import assembly as OA
    from oracle import newton as ON
    free = ~o["mask"]
    ref, rinfo = ON.newton(np.zeros(o["mask"].size), lambda u: OA.residual(o["sets"], o["coords"], u, 3),
                           lambda u: OA.reduced_tangent(o["sets"], o["coords"], u, 3, free), o["mask"], o["values"],
                           lambda A, b: ON.linear_solve(A, b, "lapack")[0], tol=1e-9, damping=0.8)
    print(f"[case] damped newton: device {tuple(infos)} oracle {rinfo}", flush=True)
    assert infos[2] is False and rinfo[2] is False
    assert infos[0] == rinfo[0] and infos[0] > 3  # damping 0.8 converges linearly: more steps than Newton
    assert np.linalg.norm(sol.ravel() - ref) / np.linalg.norm(ref) < 1e-8


def test_adaptive_load_stepping_matches_oracle(torch_gpu):
    """solver.adaptive_load_stepping on the device against the oracle's restatement: a Neo-Hooke
    cube pulled to u_x = 2 with a first increment of 1 (which diverges) -- same multiplier
    sequence, same Newton counts, same final solution."""
    from autopdex_amd import solver
    from oracle import newton as ON
    n, total = 3, 2.0
    dofs, settings, static, o = P.neo_hooke_cube(n, disp=total)
    full = settings["dirichlet conditions"].copy()
    static = dict(static, **{"linear tolerance": 1e-12, "verbose": -1})

    def scale(s, m):
        s["dirichlet conditions"] = full * m
        return s

    sol, m, inc, _, _, res = solver.adaptive_load_stepping(dofs, settings, static, multiplier_settings=scale,
                                                           init_increment=1.0, newton_tol=1e-10)
    dev_steps = solver.adaptive_load_stepping.last_steps

    def oracle_step(u, mult):
        oo = dict(o, values=o["values"] * mult)
        return P.oracle_newton(oo, tol=1e-10, u0=u)

    u_ref, m_ref, _, ref_steps = ON.adaptive_load_stepping(np.zeros(o["mask"].size), oracle_step, init_increment=1.0)
    print(f"[case] device steps {[(round(a, 4), b, c) for a, b, c, _ in dev_steps]}", flush=True)
    assert any(s[2] for s in ref_steps), "the first increment should diverge"
    assert [(round(a, 12), b, c) for a, b, c, _ in dev_steps] == [(round(a, 12), b, c) for a, b, c in ref_steps]
    assert m == m_ref == 1.0
    assert np.linalg.norm(sol.ravel() - u_ref) / np.linalg.norm(u_ref) < 1e-8


def test_cook_membrane_q2_user_elements_reach_reference_golden(torch_gpu):
    """The reference's Q2 user-element test (tests/test_user_elem_impl_diff_and_adaptive_load_step.py):
    Cook's membrane, 12 Q2 elements, plane-strain Neo-Hooke E = 100, nu = 0.3, traction
    (0, 20 m) on 3-node lines, adaptive load stepping to m = 1 with newton_tol 1e-8.  The
    reference asserts allclose(dofs . dofs, 19390.35027108); the device run must meet that, agree
    with the value to 1e-10, and follow the oracle's multiplier sequence and Newton counts
    (the reference solves each Newton step directly; here Jacobi-PCG to 1e-12)."""
    from autopdex_amd import solver
    fx = P.cook_membrane_fixture()
    golden = fx["golden"]["dofs_dot_dofs"]
    dofs, settings, static, o = P.cook_membrane_q2()
    static = dict(static, **{"linear tolerance": 1e-12})

    def multiplier_settings(s, m):
        s["load multiplier"] = m * fx["q_0"]
        return s

    sol, m, inc, _, _, res = solver.adaptive_load_stepping(dofs, settings, static, multiplier_settings, False, None,
                                                           newton_tol=fx["newton_tol"])
    steps = solver.adaptive_load_stepping.last_steps
    u_ref, ref_steps = P.oracle_cook_membrane()
    val = float(sol.ravel() @ sol.ravel())
    print(f"[case] cook membrane dofs.dofs {val!r} golden {golden!r} steps {[(round(a, 4), b) for a, b, c, _ in steps]}",
          flush=True)
    assert np.isclose(val, golden, rtol=fx["golden_rtol"], atol=fx["golden_atol"])
    assert abs(val / golden - 1.0) < 1e-10
    assert m == 1.0
    assert [(round(a, 12), b, c) for a, b, c, _ in steps] == [(round(a, 12), b, c) for a, b, c in ref_steps]
    assert np.linalg.norm(sol.ravel() - u_ref) / np.linalg.norm(u_ref) < 1e-9


def test_cook_membrane_sensitivities_match_reference(torch_gpu):
    """autopdex_amd.implicit_diff on the device for the reference's sensitivities of u.u
    (tests/test_user_elem_impl_diff_and_adaptive_load_step.py:151-172): first order in forward
    (root_jvp) and reverse (root_vjp) mode to 1e-9 of the reference's values, second order
    (root_jvp2, forward over forward) to 1e-7; and against the oracle's rules."""
    from autopdex_amd import implicit_diff, solver
    fx = P.cook_membrane_fixture()
    g = fx["golden"]
    dofs, settings, static, o = P.cook_membrane_q2()
    static = dict(static, **{"linear tolerance": 1e-12, "sensitivity solver backend": "hip"})

    def multiplier_settings(s, m):
        s["load multiplier"] = m * fx["q_0"]
        return s

    sol, (m, inc, _, settings, res) = solver.adaptive_load_stepping(
        dofs, settings, static, multiplier_settings, False, "forward", newton_tol=fx["newton_tol"])
    u = sol.ravel()
    keys = ["youngs modulus", "poisson ratio"]
    first = [2 * u @ implicit_diff.root_jvp(sol, settings, static, k).ravel() for k in keys]
    rev = implicit_diff.root_vjp(sol, settings, static, 2 * sol, keys)
    ref1 = [g["d_dE"], g["d_dnu"]]
    print(f"[case] cook sensitivities forward {first} reverse {rev}", flush=True)
    for k, key in enumerate(keys):
        assert abs(first[k] / ref1[k] - 1) < 1e-9 and abs(rev[key] / ref1[k] - 1) < 1e-9
    second = {}
    for i, j, name in [(0, 0, "d2_dE2"), (0, 1, "d2_dE_dnu"), (1, 1, "d2_dnu2")]:
        u_i, u_j, u_ij = implicit_diff.root_jvp2(sol, settings, static, keys[i], keys[j])
        second[name] = 2 * (u_i.ravel() @ u_j.ravel() + u @ u_ij.ravel())
    print(f"[case] cook second order {second}", flush=True)
    for name, v in second.items():
        assert np.isclose(v, g[name], rtol=1e-5, atol=1e-8)  # the reference's allclose
        assert abs(v / g[name] - 1) < 1e-7
    u_ref, _ = P.oracle_cook_membrane()
    o1, _, o2 = P.oracle_cook_sensitivities(u_ref, eps=1e-5)
    assert np.allclose(first, o1, rtol=1e-9, atol=0)
    assert np.allclose([second["d2_dE2"], second["d2_dE_dnu"], second["d2_dnu2"]], [o2[0, 0], o2[0, 1], o2[1, 1]],
                       rtol=1e-7, atol=0)


@pytest.mark.parametrize("name,builder,args", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_tangent_bcoo_export_matches_reference_order(torch_gpu, name, builder, args):
    """assembler.assemble_tangent(form='bcoo'): the full matrix with duplicates, entry for entry in
    the reference's order (per set element-major row-major blocks, sets concatenated)."""
    from autopdex_amd import assembler
    from oracle import assembly as OA
    dofs, settings, static, o, prob = _setup(builder, *args)
    rng = np.random.default_rng(9)
    u = rng.standard_normal(prob.n_dofs) * (0.02 if o["d"] > 1 else 1.0)
    M = assembler.assemble_tangent(utility_reshape(u, dofs), settings, static, form="bcoo")
    rows, cols, data = [], [], []
    for model, conn, nodal in o["sets"]:
        r_, c_, d_ = OA.tangent_coo(model, o["coords"], conn, u, o["d"], nodal)
        rows.append(r_), cols.append(c_), data.append(d_)
    rows, cols, data = np.concatenate(rows), np.concatenate(cols), np.concatenate(data)
    assert M.shape == (prob.n_dofs, prob.n_dofs)
    assert np.array_equal(M.row, rows) and np.array_equal(M.col, cols)
    assert _rel(M.data, data) < 1e-12
    M32 = assembler.assemble_tangent(utility_reshape(u, dofs), settings, static, precision="fp32", form="bcoo")
    assert _rel(M32.data.astype(np.float64), data) < 1e-5


def utility_reshape(u, dofs):
    from autopdex_amd import utility
    return utility.reshape_as(u, dofs)
