"""Assembly entry points (device), as autopdex.assembler for 'user potential'.

assemble_residual (assembler.py:587-637), assemble_tangent (:682-777) and
assemble_tangent_diagonal (:639-680) run the element kernels of
libautopdex_amd.  Residual and diagonal come back with the structure of
dofs.  The tangent comes back either as the *reduced* scipy CSR matrix
(form="csr": Dirichlet rows and columns removed -- the matrix the
reference's scipy backend solves with, solver.py:1180-1222), or as the
reference returns it (form="bcoo": the full n_dofs x n_dofs matrix with
duplicates, assembler.py:682-777 / :1115-1149, as a scipy COO matrix whose
entries keep the reference's order: per set, element-major row-major blocks,
rows repeat(gdofs), columns tile(gdofs), _get_indices assembler.py:47-141).
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _native as N
from . import utility
from .backend import current_stream, get_problem


def _upload_dofs(prob, dofs):
    return prob.to_device(prob.flat(dofs))


def assemble_residual(dofs, settings, static_settings):
    prob = get_problem(dofs, settings, static_settings)
    u = _upload_dofs(prob, dofs)
    r = u.new_empty(prob.n_dofs)
    N.call("apx_assemble_residual", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(r.data_ptr()), current_stream())
    return utility.reshape_as(r.cpu().numpy(), dofs)


def assemble_tangent_diagonal(dofs, settings, static_settings):
    prob = get_problem(dofs, settings, static_settings)
    u = _upload_dofs(prob, dofs)
    g = u.new_empty(prob.n_dofs)
    N.call("apx_assemble_diagonal", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(g.data_ptr()), current_stream())
    return utility.reshape_as(g.cpu().numpy(), dofs)


def _coo_indices(prob, settings):
    """Row and column indices of the reference's BCOO (int64): per set, element-major,
    row = repeat(gdofs, n), col = tile(gdofs, n), gdofs = node * d + comp."""
    d = prob.n_dofs // prob.n_nodes
    rows, cols = [], []
    for conn_s in settings["connectivity"]:
        conn = np.asarray(conn_s[prob.key] if isinstance(conn_s, dict) else conn_s, dtype=np.int64)
        g = (conn[:, :, None] * d + np.arange(d)).reshape(conn.shape[0], -1)
        n = g.shape[1]
        rows.append(np.repeat(g, n, axis=1).ravel())
        cols.append(np.tile(g, (1, n)).ravel())
    return np.concatenate(rows), np.concatenate(cols)


def assemble_tangent(dofs, settings, static_settings, precision="fp64", form="csr"):
    import torch
    prob = get_problem(dofs, settings, static_settings)
    u = _upload_dofs(prob, dofs)
    p = N.APX_FP64 if precision == "fp64" else N.APX_FP32
    if form == "bcoo":
        cnt = C.c_int64()
        N.call("apx_tangent_coo_entries", prob.plan, C.byref(cnt))
        data = torch.empty(cnt.value, dtype=torch.float64 if p == N.APX_FP64 else torch.float32, device=u.device)
        N.call("apx_assemble_tangent_coo", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(data.data_ptr()), p,
               current_stream())
        rows, cols = _coo_indices(prob, settings)
        return sp.coo_matrix((data.cpu().numpy(), (rows, cols)), shape=(prob.n_dofs, prob.n_dofs))
    if form != "csr":
        raise ValueError("form must be 'csr' (reduced) or 'bcoo' (the reference's full matrix with duplicates)")
    vals = torch.empty(prob.nnz, dtype=torch.float64 if p == N.APX_FP64 else torch.float32, device=u.device)
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals.data_ptr()), p,
           current_stream())
    indptr, indices = prob.csr_pattern()
    return sp.csr_matrix((vals.cpu().numpy(), indices, indptr), shape=(prob.n_free, prob.n_free))
