"""Assembly entry points (device), as autopdex.assembler for 'user potential'.

assemble_residual (assembler.py:587-637), assemble_tangent (:682-777) and
assemble_tangent_diagonal (:639-680) run the element kernels of
libautopdex_amd.  Residual and diagonal come back with the structure of
dofs.  The tangent comes back as the *reduced* scipy CSR matrix (Dirichlet
rows and columns removed) -- the matrix the reference's scipy backend
actually solves with (solver.py:1180-1222); the full BCOO with duplicates is
not formed on the device.
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


def assemble_tangent(dofs, settings, static_settings, precision="fp64"):
    import torch
    prob = get_problem(dofs, settings, static_settings)
    u = _upload_dofs(prob, dofs)
    p = N.APX_FP64 if precision == "fp64" else N.APX_FP32
    vals = torch.empty(prob.nnz, dtype=torch.float64 if p == N.APX_FP64 else torch.float32, device=u.device)
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals.data_ptr()), p,
           current_stream())
    indptr, indices = prob.csr_pattern()
    return sp.csr_matrix((vals.cpu().numpy(), indices, indptr), shape=(prob.n_free, prob.n_free))
