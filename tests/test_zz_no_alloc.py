"""apx_pcg_noalloc (the allocation-free Jacobi-PCG) against apx_pcg through the C ABI.

The two entry points run the same kernels with the same launch shapes, so the
solutions must agree bit for bit and take the same number of iterations; the
solution also meets the oracle's SuperLU solution.  Two calls in a row on one
device, the second on a smaller system (the shared scratch array keeps the first
call's contents), in FP64 and FP32, and a system too large for the scratch
array is rejected with APX_ERR_ARG before any work is queued.
"""
import ctypes as C

import numpy as np
import pytest

from tests import problems as P

pytestmark = pytest.mark.gpu


def _solve(name, prob, vals, b, precision, tol):
    import torch
    from autopdex_amd import _native as N
    from autopdex_amd.backend import current_stream
    x = torch.full_like(b, float("nan"))
    ks = N.KrylovStats()
    N.call(name, prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
           precision, N.APX_PRECOND_JACOBI, tol, 0.0, 0, C.byref(ks), current_stream())
    torch.cuda.synchronize()
    return x.cpu().numpy(), ks


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pcg_noalloc_matches_pcg_on_two_calls(torch_gpu, precision):
    torch = torch_gpu
    import scipy.sparse.linalg as spla
    from autopdex_amd import _native as N
    from autopdex_amd import backend
    from oracle import assembly as OA
    prec = N.APX_FP64 if precision == "fp64" else N.APX_FP32
    dt = torch.float64 if precision == "fp64" else torch.float32
    tol = 1e-10 if precision == "fp64" else 1e-5
    for n in (96, 40):  # the second call is smaller than the first
        print(f"[case] pcg_noalloc {precision} n={n}", flush=True)
        dofs, settings, static, o = P.poisson_source(n, dict_dofs=False)
        prob = backend.get_problem(dofs, settings, static)
        A = OA.reduced_tangent(o["sets"], o["coords"], np.zeros(o["mask"].size), 1, ~o["mask"])
        vals = torch.from_numpy(A.data.copy()).to(dt).cuda()
        rhs = np.random.default_rng(n).standard_normal(prob.n_free)
        b = torch.from_numpy(rhs).to(dt).cuda()
        x_na, ks_na = _solve("apx_pcg_noalloc", prob, vals, b, prec, tol)
        x_ref, ks_ref = _solve("apx_pcg", prob, vals, b, prec, tol)
        print(f"[case] iterations noalloc {ks_na.iterations} pcg {ks_ref.iterations}", flush=True)
        assert ks_na.converged == 1 and ks_na.iterations == ks_ref.iterations
        assert ks_na.final_residual == ks_ref.final_residual
        assert np.array_equal(x_na.view(np.uint8), x_ref.view(np.uint8))
        if precision == "fp64":
            x_lu = spla.spsolve(A.tocsc(), rhs)
            assert np.linalg.norm(x_na - x_lu) / np.linalg.norm(x_lu) < 1e-8


def test_pcg_noalloc_rejects_a_system_beyond_its_scratch(torch_gpu):
    torch = torch_gpu
    from autopdex_amd import _native as N
    from autopdex_amd import backend
    dofs, settings, static, o = P.poisson_source(700, dict_dofs=False)
    prob = backend.get_problem(dofs, settings, static)
    vals = torch.zeros(prob.nnz, dtype=torch.float64, device="cuda")
    b = torch.zeros(prob.n_free, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    ks = N.KrylovStats()
    rc = N.lib().apx_pcg_noalloc(prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(b.data_ptr()),
                                 C.c_void_p(x.data_ptr()), N.APX_FP64, N.APX_PRECOND_JACOBI, 1e-8, 0.0, 0,
                                 C.byref(ks), None)
    msg = N.lib().apx_last_error().decode()
    print(f"[case] oversize: rc {rc} {msg}", flush=True)
    assert rc == -1 and "scratch" in msg
