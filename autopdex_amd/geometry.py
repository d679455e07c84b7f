"""Distance functions and node selection (host, numpy), as autopdex.geometry.

The functions accept a single point (shape (2,)) or a batch (shape (N, 2));
batches replace the reference's jax.vmap.  The arithmetic follows the
reference operation by operation, so masks built from them match the
reference's, including the polygon corners: psdf_polygon folds the edge
distances with r_equivalence o1*o2/(o1+o2) (geometry.py:33-45, :672-694),
which is 0/0 = NaN at a corner, and isclose(NaN, 0) is False
(geometry.py:840-871) -- corner nodes are therefore NOT selected.
"""
import numpy as np


def r_equivalence(o1, o2):
    return o1 * o2 / (o1 + o2)


def r_conjunction(o1, o2):
    return o1 + o2 - np.sqrt(o1 ** 2 + o2 ** 2)


def r_disjunction(o1, o2):
    return o1 + o2 + np.sqrt(o1 ** 2 + o2 ** 2)


def r_trimming(o, t):
    return np.sqrt(o ** 2 + (np.sqrt(t ** 2 + o ** 4) - t) ** 2 / 4)


def sdf_infinite_line(x, x_p1, x_p2):
    x = np.asarray(x, dtype=np.float64)
    p1 = np.asarray(x_p1, dtype=np.float64)
    p2 = np.asarray(x_p2, dtype=np.float64)
    line_length = np.linalg.norm(p2 - p1)
    return ((x[..., 0] - p1[0]) * (p2[1] - p1[1]) - (x[..., 1] - p1[1]) * (p2[0] - p1[0])) / line_length


def sdf_nd_sphere(x, xc, r):
    dxc = np.asarray(x, dtype=np.float64) - np.asarray(xc, dtype=np.float64)
    return (r ** 2 - np.sum(dxc * dxc, axis=-1)) / (2 * r)


def sdf_normal_to_direction(x, x_p1, normal):
    return (np.asarray(x, dtype=np.float64) - np.asarray(x_p1, dtype=np.float64)) @ np.asarray(normal, dtype=np.float64)


def centroid(x_p1, x_p2):
    return (np.asarray(x_p1, dtype=np.float64) + np.asarray(x_p2, dtype=np.float64)) / 2


def psdf_trimmed_line(x, x_p1, x_p2):
    p1 = np.asarray(x_p1, dtype=np.float64)
    p2 = np.asarray(x_p2, dtype=np.float64)
    omega = sdf_infinite_line(x, p1, p2)
    xc = centroid(p1, p2)
    length = np.linalg.norm(p2 - p1)
    trim = sdf_nd_sphere(x, xc, length / 2)
    return r_trimming(omega, trim)


def psdf_polygon(x, x_p_list):
    pts = np.asarray(x_p_list, dtype=np.float64)
    n_segments = pts.shape[0]
    with np.errstate(invalid="ignore", divide="ignore"):
        psdf_d = psdf_trimmed_line(x, pts[0], pts[1])
        for i in range(n_segments - 1):
            psdf_d = r_equivalence(psdf_d, psdf_trimmed_line(x, pts[-i], pts[-(i + 1)]))
    return psdf_d


def in_sdf(x, sdf_fun, tol=1e-6):
    return np.isclose(sdf_fun(np.asarray(x, dtype=np.float64)), 0, atol=tol)


def in_sdfs(x, sdf_fun, tol=1e-6):
    """Boolean per point: |sdf| <= tol (NaN -> False)."""
    return in_sdf(x, sdf_fun, tol)


def in_plane(x, x_p1, normal):
    return in_sdf(x, lambda y: sdf_normal_to_direction(y, x_p1, normal))


def in_planes(x_q, x_p1, normal):
    return in_plane(x_q, x_p1, normal)


def elem_in_plane(x_nodes, x_p1, normal):
    """geometry.py:1030-1045: every node of the element lies in the plane."""
    return bool(np.all(in_planes(np.asarray(x_nodes), x_p1, normal)))


def select_elements_in_plane(x_nodes, surface_elements, x_p1, normal):
    """geometry.py:1048-1067: boolean mask of the surface elements lying in the plane."""
    x_nodes = np.asarray(x_nodes, dtype=np.float64)
    se = np.asarray(surface_elements)
    inside = np.asarray(in_planes(x_nodes[se].reshape(-1, x_nodes.shape[1]), x_p1, normal))
    return inside.reshape(se.shape).all(axis=1)
