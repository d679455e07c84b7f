"""CPU tests of the C ABI: the library loads, exports every declared symbol, and
its host-side plan builder reproduces scipy's canonical reduced CSR pattern
(solver.py:1180-1222) bit for bit.  No device calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tests import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "autopdex_amd.h")).read()
    return sorted(set(re.findall(r"\b(apx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from autopdex_amd import _native as N
    L = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) <= set(N._SIGS) | {"apx_plan_build_pattern"}
    assert L.apx_version() >= 100


def _host_pattern(coords, conns, npes, models, mask, d):
    from autopdex_amd import _native as N
    plan = C.c_void_p()
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    N.call("apx_plan_create", C.byref(plan), coords.shape[1], coords.shape[0], coords.ctypes.data, d, m.ctypes.data)
    keep = []
    for conn, model in zip(conns, models):
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        keep.append(conn)
        p = np.zeros(2)
        N.call("apx_plan_add_set", plan, conn.ctypes.data, conn.shape[0], conn.shape[1], model, p.ctypes.data, 2, None)
    N.call("apx_plan_build_pattern", plan)
    n_free = int((m == 0).sum())
    indptr = np.empty(n_free + 1, dtype=np.int32)
    info = N.PlanInfo()
    N.call("apx_plan_get_info", plan, C.byref(info))
    indices = np.empty(info.nnz, dtype=np.int32)
    N.call("apx_plan_csr_pattern", plan, indptr.ctypes.data, indices.ctypes.data)
    r2f = np.empty(n_free, dtype=np.int32)
    N.call("apx_plan_free_dofs", plan, r2f.ctypes.data)
    N.call("apx_plan_destroy", plan)
    return indptr, indices, r2f


@pytest.mark.parametrize("name,builder,args", [
    ("poisson_q1_5x5", P.poisson_source, (5,)),
    ("poisson_q1_200", P.poisson_source, (200,)),
    ("hex8_4", P.neo_hooke_cube, (4,)),
    ("panel_traction_6", P.tapered_panel, (6,)),
])
def test_host_pattern_equals_scipy_reduced_csr(name, builder, args):
    from autopdex_amd import _native as N
    from oracle import assembly as OA
    dofs, settings, static, o = builder(*args)
    conns = [c for _, c, _ in o["sets"]]
    models = [N.MODEL_NEO_HOOKE if o["d"] > 1 else N.MODEL_POISSON_SOURCE for _ in conns]
    indptr, indices, r2f = _host_pattern(o["coords"], conns, None, models, o["mask"], o["d"])
    u = np.zeros(o["mask"].size)
    A = OA.reduced_tangent(o["sets"], o["coords"], u, o["d"], ~o["mask"])
    assert np.array_equal(indptr, A.indptr)
    assert np.array_equal(indices, A.indices)
    assert np.array_equal(r2f, np.flatnonzero(~o["mask"]))


def test_appendix_sizes_reduced_nnz():
    """Reduced nnz of the 200x200 problem (config 1): 354,037 with corners free."""
    from autopdex_amd import _native as N
    dofs, settings, static, o = P.poisson_source(200)
    indptr, indices, r2f = _host_pattern(o["coords"], [o["sets"][0][1]], None, [N.MODEL_POISSON_SOURCE], o["mask"], 1)
    assert indptr[-1] == 354037 and len(r2f) == 39605


def test_errors_are_reported_not_raised():
    from autopdex_amd import _native as N
    plan = C.c_void_p()
    coords = np.zeros((4, 2))
    m = np.zeros(4, np.uint8)
    rc = N.lib().apx_plan_create(C.byref(plan), 5, 4, coords.ctypes.data, 1, m.ctypes.data)
    assert rc == -1 and b"dim" in N.lib().apx_last_error()
    N.call("apx_plan_create", C.byref(plan), 2, 4, coords.ctypes.data, 1, m.ctypes.data)
    bad = np.array([[0, 1, 2, 9]], dtype=np.int64)
    N.call("apx_plan_add_set", plan, bad.ctypes.data, 1, 4, N.MODEL_POISSON_SOURCE, None, 0, None)
    rc = N.lib().apx_plan_build_pattern(plan)
    assert rc == -1 and b"outside" in N.lib().apx_last_error()
    rc = N.lib().apx_assemble_residual(plan, None, None, None)
    assert rc == -3  # not finalized
    N.call("apx_plan_destroy", plan)
