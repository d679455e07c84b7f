"""Oracle: meshes, Dirichlet masks and Gauss rules (test infrastructure only).

Pinned by: closed-form node/connectivity arithmetic, the reference's
known-answer test (5x5 Poisson, tests/test_dicts_as_dofs_user_potential.py:63)
and the closed form of the 2-point Gauss rule (+-1/sqrt(3), weights 1).
"""
import numpy as np


def structured_mesh(n_elements, vertices, element_type="quad"):
    """Restates autopdex/mesher.py:29-175 (quad / brick branches).

    2D: node (i, j) -> index i*(ny+1)+j at the bilinear image of
    (s_i, t_j) = linspace(-1, 1); quad = [(i,j), (i,j+1), (i+1,j+1), (i+1,j)]
    (mesher.py:88-113).  3D: node (i, j, k) -> i*(ny+1)*(nz+1)+j*(nz+1)+k,
    brick order per mesher.py:150-168.
    """
    v = np.asarray(vertices, dtype=np.float64)
    dim = v.shape[1]
    if dim == 2:
        assert element_type == "quad"
        nx, ny = n_elements
        s = np.linspace(-1.0, 1.0, nx + 1)
        t = np.linspace(-1.0, 1.0, ny + 1)
        coords = np.empty(((nx + 1) * (ny + 1), 2))
        for i in range(nx + 1):
            S = s[i]
            T = t
            # (1-s)(1-t) v0 + (1+s)(1-t) v1 + (1+s)(1+t) v2 + (1-s)(1+t) v3, /4
            blk = (((1 - S) * (1 - T))[:, None] * v[0] + ((1 + S) * (1 - T))[:, None] * v[1]
                   + ((1 + S) * (1 + T))[:, None] * v[2] + ((1 - S) * (1 + T))[:, None] * v[3]) / 4
            coords[i * (ny + 1):(i + 1) * (ny + 1)] = blk
        elems = np.empty((nx * ny, 4), dtype=np.int64)
        e = 0
        for i in range(nx):
            for j in range(ny):
                elems[e] = [i * (ny + 1) + j, i * (ny + 1) + j + 1,
                            (i + 1) * (ny + 1) + j + 1, (i + 1) * (ny + 1) + j]
                e += 1
        return coords, elems
    if dim == 3:
        assert element_type == "brick"
        nx, ny, nz = n_elements
        s = np.linspace(-1.0, 1.0, nx + 1)
        t = np.linspace(-1.0, 1.0, ny + 1)
        u = np.linspace(-1.0, 1.0, nz + 1)
        S, T, U = np.meshgrid(s, t, u, indexing="ij")
        S, T, U = S.ravel(), T.ravel(), U.ravel()
        coords = (((1 - S) * (1 - T) * (1 - U))[:, None] * v[0] + ((1 + S) * (1 - T) * (1 - U))[:, None] * v[1]
                  + ((1 + S) * (1 + T) * (1 - U))[:, None] * v[2] + ((1 - S) * (1 + T) * (1 - U))[:, None] * v[3]
                  + ((1 - S) * (1 - T) * (1 + U))[:, None] * v[4] + ((1 + S) * (1 - T) * (1 + U))[:, None] * v[5]
                  + ((1 + S) * (1 + T) * (1 + U))[:, None] * v[6] + ((1 - S) * (1 + T) * (1 + U))[:, None] * v[7]) / 8
        sy, sx = nz + 1, (ny + 1) * (nz + 1)
        elems = []
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    b = i * sx + j * sy + k
                    elems.append([b, b + sx, b + sx + sy, b + sy,
                                  b + 1, b + sx + 1, b + sx + sy + 1, b + sy + 1])
        return coords, np.asarray(elems, dtype=np.int64)
    raise ValueError("vertices must have 2 or 3 columns")


# --- R-function distance fields (autopdex/geometry.py) -----------------------

def r_equivalence(o1, o2):
    """geometry.py:33-45: o1*o2/(o1+o2)."""
    return o1 * o2 / (o1 + o2)


def r_trimming(o, t):
    """geometry.py:78-91: sqrt(o^2 + (sqrt(t^2 + o^4) - t)^2 / 4)."""
    return np.sqrt(o ** 2 + (np.sqrt(t ** 2 + o ** 4) - t) ** 2 / 4)


def sdf_infinite_line(x, p1, p2):
    """geometry.py:123-140."""
    p1 = np.asarray(p1, dtype=np.float64)
    p2 = np.asarray(p2, dtype=np.float64)
    length = np.linalg.norm(p2 - p1)
    return ((x[..., 0] - p1[0]) * (p2[1] - p1[1]) - (x[..., 1] - p1[1]) * (p2[0] - p1[0])) / length


def sdf_nd_sphere(x, xc, r):
    """geometry.py:143-157: (r^2 - |x-xc|^2) / (2r)."""
    d = x - xc
    return (r ** 2 - np.sum(d * d, axis=-1)) / (2 * r)


def psdf_trimmed_line(x, p1, p2):
    """geometry.py:462-484."""
    p1 = np.asarray(p1, dtype=np.float64)
    p2 = np.asarray(p2, dtype=np.float64)
    omega = sdf_infinite_line(x, p1, p2)
    xc = (p1 + p2) / 2
    length = np.linalg.norm(p2 - p1)
    trim = sdf_nd_sphere(x, xc, length / 2)
    return r_trimming(omega, trim)


def psdf_polygon(x, pts):
    """geometry.py:672-694: fold order (v0,v1), then (v0,v-1), (v-1,v-2), ..."""
    pts = np.asarray(pts, dtype=np.float64)
    n = pts.shape[0]
    with np.errstate(invalid="ignore", divide="ignore"):
        d = psdf_trimmed_line(x, pts[0], pts[1])
        for i in range(n - 1):
            d = r_equivalence(d, psdf_trimmed_line(x, pts[-i], pts[-(i + 1)]))
    return d


def in_sdfs(x, sdf_fun, tol=1e-6):
    """geometry.py:840-871: isclose(sdf, 0, atol=tol) -- NaN (polygon corners) is False."""
    v = sdf_fun(np.asarray(x, dtype=np.float64))
    return np.isclose(v, 0.0, atol=tol)


def in_planes(x, p1, normal, tol=1e-6):
    """geometry.py:891-921 with sdf_normal_to_direction (:185-202): isclose((x-p1).n, 0)."""
    v = (np.asarray(x) - np.asarray(p1, dtype=np.float64)) @ np.asarray(normal, dtype=np.float64)
    return np.isclose(v, 0.0, atol=tol)


def dof_select(nodes_mask, fields):
    """utility.py:448-462: outer product of a node mask and a per-component selector."""
    return np.outer(nodes_mask, np.asarray(fields, dtype=bool))


# --- Gauss-Legendre (autopdex/seeder.py) --------------------------------------

def gauss_legendre_nd(dimension, order=2):
    """seeder.py:1046-1062 + :370-393 + tensor_product_rule :324-367 (orders 2..5).

    Order 2|3: points 2*p-1 with p in {0.21132486540518713, 0.7886751345948129},
    weights 2*0.5.  Order 4|5: p in {0.1127016653792583, 0.5, 0.8872983346207417},
    weights 2*{0.2777777777777777, 0.44444444444444453, 0.2777777777777777} (the
    reference's digits).  xi_0 varies fastest, then xi_1, then xi_2.
    """
    if order in (2, 3):
        r = 2 * np.array([0.21132486540518713, 0.7886751345948129]) - 1
        w = 2 * np.array([0.5, 0.5])
    elif order in (4, 5):
        r = 2 * np.array([0.1127016653792583, 0.5, 0.8872983346207417]) - 1
        w = 2 * np.array([0.2777777777777777, 0.44444444444444453, 0.2777777777777777])
    else:
        raise ValueError(order)
    n = r.size
    if dimension == 1:
        return r, w
    if dimension == 2:
        xi = np.tile(r, n)
        yi = np.repeat(r, n)
        return np.stack([xi, yi], axis=1), np.outer(w, w).ravel()
    if dimension == 3:
        xi = np.tile(r, n * n)
        yi = np.tile(np.repeat(r, n), n)
        zi = np.repeat(r, n * n)
        return np.stack([xi, yi, zi], axis=1), np.outer(np.outer(w, w).ravel(), w).ravel()
    raise ValueError(dimension)
