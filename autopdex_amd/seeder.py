"""Gauss-Legendre rules (host), as autopdex.seeder (seeder.py:301-385, :1046-1062).

gauss_legendre_nd(dimension, order) returns points on [-1, 1]^dim (2 p - 1)
and weights 2 w, tensor ordered with xi_0 fastest (the order-4|5 weights carry the
reference's digits, seeder.py:387-393).  The device kernels implement the 2-point
rule (order 2 or 3) and, for Q2 quads and 3-node lines, the 3-point rule (order 4
or 5); models reject other rules.
"""
import numpy as np

_GL01 = {
    1: ([0.5], [1.0]),
    2: ([0.21132486540518713, 0.7886751345948129], [0.5, 0.5]),
    3: ([0.21132486540518713, 0.7886751345948129], [0.5, 0.5]),
    4: ([0.1127016653792583, 0.5, 0.8872983346207417], [0.2777777777777777, 0.44444444444444453, 0.2777777777777777]),
    5: ([0.1127016653792583, 0.5, 0.8872983346207417], [0.2777777777777777, 0.44444444444444453, 0.2777777777777777]),
}


def gauss_legendre_1d(order):
    """Points and weights on [0, 1]."""
    if order not in _GL01:
        p, w = np.polynomial.legendre.leggauss(int(np.ceil((order + 1) / 2)))
        return (p + 1) / 2, w / 2
    p, w = _GL01[order]
    return np.asarray(p), np.asarray(w)


def tensor_product_rule(roots_1d, weights_1d, dim):
    r, w = np.asarray(roots_1d), np.asarray(weights_1d)
    n = r.size
    if dim == 2:
        x_int = np.stack([np.tile(r, n), np.repeat(r, n)], axis=1)
        return x_int, np.outer(w, w).ravel()
    if dim == 3:
        x_int = np.stack([np.tile(r, n * n), np.tile(np.repeat(r, n), n), np.repeat(r, n * n)], axis=1)
        return x_int, np.outer(np.outer(w, w).ravel(), w).ravel()
    raise NotImplementedError("Not implemented for this dimensionality!")


def gauss_legendre_nd(dimension, order):
    p, w = gauss_legendre_1d(order)
    p, w = 2 * p - 1, 2 * w
    if dimension == 1:
        return p, w
    return tensor_product_rule(p, w, dimension)
