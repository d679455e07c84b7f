"""Structured meshes (host, numpy) -- same call and output as autopdex.mesher.

structured_mesh follows autopdex/mesher.py:29-206: the domain is the
bilinear (2D) / trilinear (3D) image of a reference grid linspace(-1, 1);
node (i, j[, k]) gets index i*(ny+1)+j (2D) or i*(ny+1)*(nz+1)+j*(nz+1)+k
(3D); quads are [(i,j), (i,j+1), (i+1,j+1), (i+1,j)] and bricks list the
k-face then the k+1-face counter-clockwise.  Indices are int64.
"""
import numpy as np


def structured_mesh(n_elements, vertices, element_type, order=1):
    v = np.asarray(vertices, dtype=np.float64)
    dim = v.shape[1]
    if order != 1:
        raise NotImplementedError("Only order==1 is implemented at this moment.")
    if dim == 2:
        if element_type not in ("quad", "tri"):
            raise NotImplementedError("For 2D, element_type must be either 'quad' or 'tri'.")
        nx, ny = n_elements
        S, T = np.meshgrid(np.linspace(-1, 1, nx + 1), np.linspace(-1, 1, ny + 1), indexing="ij")
        s, t = S.ravel()[:, None], T.ravel()[:, None]
        coords = ((1 - s) * (1 - t) * v[0] + (1 + s) * (1 - t) * v[1] + (1 + s) * (1 + t) * v[2]
                  + (1 - s) * (1 + t) * v[3]) / 4
        I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        b = (I * (ny + 1) + J).ravel()
        quads = np.stack([b, b + 1, b + ny + 2, b + ny + 1], axis=1).astype(np.int64)
        if element_type == "quad":
            return coords, quads
        return coords, np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]], axis=0)
    if dim == 3:
        if element_type not in ("brick", "tet"):
            raise NotImplementedError("For 3D, element_type must be either 'brick' or 'tet'.")
        nx, ny, nz = n_elements
        S, T, U = np.meshgrid(np.linspace(-1, 1, nx + 1), np.linspace(-1, 1, ny + 1),
                              np.linspace(-1, 1, nz + 1), indexing="ij")
        s, t, u = S.ravel()[:, None], T.ravel()[:, None], U.ravel()[:, None]
        coords = ((1 - s) * (1 - t) * (1 - u) * v[0] + (1 + s) * (1 - t) * (1 - u) * v[1]
                  + (1 + s) * (1 + t) * (1 - u) * v[2] + (1 - s) * (1 + t) * (1 - u) * v[3]
                  + (1 - s) * (1 - t) * (1 + u) * v[4] + (1 + s) * (1 - t) * (1 + u) * v[5]
                  + (1 + s) * (1 + t) * (1 + u) * v[6] + (1 - s) * (1 + t) * (1 + u) * v[7]) / 8
        sx, sy = (ny + 1) * (nz + 1), nz + 1
        I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        b = (I * sx + J * sy + K).ravel()
        bricks = np.stack([b, b + sx, b + sx + sy, b + sy, b + 1, b + sx + 1, b + sx + sy + 1, b + sy + 1],
                          axis=1).astype(np.int64)
        if element_type == "brick":
            return coords, bricks
        n0, n1, n2, n3, n4, n5, n6, n7 = bricks.T
        tets = np.stack([np.stack(t_, axis=1) for t_ in (
            (n0, n1, n2, n6), (n0, n2, n3, n6), (n0, n3, n7, n6),
            (n0, n7, n4, n6), (n0, n4, n5, n6), (n0, n5, n1, n6))], axis=1)
        return coords, tets.reshape(-1, 4)
    raise ValueError("Unsupported dimension: vertices must have 2 or 3 columns.")


def boundary_lines(n_elements, edge):
    """2-node line connectivity along one edge of a structured (nx, ny) quad mesh.

    edge: 'left' (i = 0), 'right' (i = nx), 'bottom' (j = 0), 'top' (j = ny);
    lines are ordered by increasing free index.  Used for Neumann line sets.
    """
    nx, ny = n_elements
    if edge in ("left", "right"):
        i = 0 if edge == "left" else nx
        n = i * (ny + 1) + np.arange(ny + 1)
    elif edge in ("bottom", "top"):
        j = 0 if edge == "bottom" else ny
        n = np.arange(nx + 1) * (ny + 1) + j
    else:
        raise ValueError(edge)
    return np.stack([n[:-1], n[1:]], axis=1).astype(np.int64)
