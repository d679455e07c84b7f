"""Oracle: global assembly and the reduced CSR system (test infrastructure only).

Restates
- the element gather (assembler.py:199-207) and the residual scatter-add
  ``.at[idx].add`` (assembler.py:331-436),
- the BCOO tangent with duplicates in element-major, row-major block order:
  row = repeat(gdofs, n), col = tile(gdofs, n) (assembler.py:47-141,
  :1115-1149),
- scipy COO -> canonical CSR (duplicates summed, sorted columns, explicit zeros
  kept) and deletion of Dirichlet rows and columns ``csr[:, free][free]``
  (solver.py:1180-1222),
- the tangent diagonal (assembler.py:219-329).
Global DOF g = node*d + comp (P3, assembler.py:129-137).

Pinned by brute force: the CSR is checked against a dense assembly built with
an explicit python loop over elements on small meshes, and end to end by the
reference's 5x5 known-answer test.
"""
import numpy as np
import scipy.sparse as sp


def gdofs(conn, d):
    """(E, npe) node ids -> (E, npe*d) global DOFs, component fastest."""
    conn = np.asarray(conn, dtype=np.int64)
    return (conn[:, :, None] * d + np.arange(d)[None, None, :]).reshape(conn.shape[0], -1)


def gather(coords, conn, u, d, nodal=None):
    conn = np.asarray(conn, dtype=np.int64)
    xe = np.asarray(coords)[conn]
    uf = np.asarray(u, dtype=np.float64).reshape(-1, d)
    ue = uf[conn]
    if d == 1:
        ue = ue[..., 0]
    ne = None if nodal is None else tuple(np.asarray(f)[conn] for f in nodal)
    return xe, ue, ne


def assemble_residual(model, coords, conn, u, d, nodal=None):
    xe, ue, ne = gather(coords, conn, u, d, nodal)
    re = model.residual(xe, ue, ne).reshape(len(conn), -1)
    r = np.zeros(np.asarray(u).size)
    np.add.at(r, gdofs(conn, d).ravel(), re.ravel())
    return r


def tangent_coo(model, coords, conn, u, d, nodal=None):
    """BCOO entries with duplicates (P11): element-major, row-major blocks."""
    xe, ue, ne = gather(coords, conn, u, d, nodal)
    ke = model.tangent(xe, ue, ne)
    g = gdofs(conn, d)
    n = g.shape[1]
    rows = np.repeat(g, n, axis=1).ravel()
    cols = np.tile(g, (1, n)).ravel()
    return rows, cols, ke.reshape(-1)


def assemble_diagonal(model, coords, conn, u, d, nodal=None):
    xe, ue, ne = gather(coords, conn, u, d, nodal)
    ke = model.tangent(xe, ue, ne)
    diag = np.zeros(np.asarray(u).size)
    np.add.at(diag, gdofs(conn, d).ravel(), np.einsum("eaa->ea", ke).ravel())
    return diag


def scipy_reduce(rows, cols, data, n, free):
    """solver.py:1180-1222: coo_matrix -> csr_matrix -> [:, free][free]."""
    A = sp.csr_matrix(sp.coo_matrix((data, (rows, cols)), shape=(n, n)))
    if free is not None:
        A = A[:, free]
        A = A[free]
    return A


def reduced_tangent(sets, coords, u, d, free):
    """Sum over sets (assembler.py:682-777) then reduce; sets = [(model, conn, nodal)]."""
    n = np.asarray(u).size
    R, C, D = [], [], []
    for model, conn, nodal in sets:
        r, c, v = tangent_coo(model, coords, conn, u, d, nodal)
        R.append(r), C.append(c), D.append(v)
    return scipy_reduce(np.concatenate(R), np.concatenate(C), np.concatenate(D), n, free)


def residual(sets, coords, u, d):
    r = np.zeros(np.asarray(u).size)
    for model, conn, nodal in sets:
        r += assemble_residual(model, coords, conn, u, d, nodal)
    return r


def diagonal(sets, coords, u, d):
    g = np.zeros(np.asarray(u).size)
    for model, conn, nodal in sets:
        g += assemble_diagonal(model, coords, conn, u, d, nodal)
    return g


def dense_bruteforce(model, coords, conn, u, d, nodal=None):
    """Explicit per-element python loop into a dense matrix (small meshes only)."""
    n = np.asarray(u).size
    K = np.zeros((n, n))
    xe, ue, ne = gather(coords, conn, u, d, nodal)
    ke = model.tangent(xe, ue, ne)
    g = gdofs(conn, d)
    for e in range(len(conn)):
        for a in range(g.shape[1]):
            for b in range(g.shape[1]):
                K[g[e, a], g[e, b]] += ke[e, a, b]
    return K
