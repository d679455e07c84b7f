"""Pins for the network oracle (CPU): finite differences of its loss, exact
bilinear reproduction of the prolongation, R = P^T, and the conv adjoint."""
import numpy as np
import pytest

from oracle import unet as OU


def test_prolongation_reproduces_bilinear_and_restriction_is_adjoint():
    H, W = 9, 7
    Hc, Wc = 5, 4
    I, J = np.meshgrid(np.arange(Hc), np.arange(Wc), indexing="ij")
    c = (0.3 + 1.5 * I - 0.7 * J + 0.2 * I * J)[None, :, :, None]
    f = OU.prolong(c, H, W)
    i, j = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    exact = 0.3 + 1.5 * i / 2 - 0.7 * j / 2 + 0.2 * (i / 2) * (j / 2)
    # bilinear data is reproduced exactly away from the clipped last coarse row/col (W even)
    assert np.allclose(f[0, :, :W - 1, 0], exact[:, :W - 1])
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, Hc, Wc, 3))
    b = rng.standard_normal((2, H, W, 3))
    assert abs(np.sum(OU.prolong(a, H, W) * b) - np.sum(a * OU.restrict(b))) < 1e-12


def test_conv_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 6, 5, 3))
    g = rng.standard_normal((1, 6, 5, 4))
    W = rng.standard_normal((4, 9, 3))
    assert abs(np.sum(OU.conv3x3(x, W, None) * g) - np.sum(x * OU.conv3x3_dgrad(g, W))) < 1e-10
    dW, db = OU.conv3x3_wgrad(x, g, 3, 4)
    # d/dW <conv(x, W), g> = dW
    E = np.zeros_like(W)
    E[2, 4, 1] = 1.0
    assert abs(np.sum(OU.conv3x3(x, E, None) * g) - dW[2, 4, 1]) < 1e-10


@pytest.mark.parametrize("relu", [0, 1])
def test_backward_matches_finite_differences(relu):
    L, ch, H, W = 3, [16, 32, 64], 9, 11
    net = OU.UNetOracle(L, ch, 1, 1, relu, H, W, bf16_storage=False)
    theta = OU.init_params(L, ch, 1, 1, relu, seed=3).astype(np.float64)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, H, W, 1))
    t = rng.standard_normal((2, H, W, 1))

    def loss(th):
        net.forward(th, x)
        return net.loss_backward(t)

    l0, g = loss(theta)
    h = 1e-6
    for name in ("in", "e1", "down0", "b", "up1", "d0", "out"):
        w_off, b_off, ci, co, cop, wk = net.layout[name]
        for idx in (w_off, w_off + 3 * wk + 5 if co > 3 else w_off + 7, b_off):
            tp, tm = theta.copy(), theta.copy()
            tp[idx] += h
            tm[idx] -= h
            fd = (loss(tp)[0] - loss(tm)[0]) / (2 * h)
            if idx == b_off and not relu:
                # the linear network keeps its biases frozen at zero: no gradient is reported
                assert g[idx] == 0.0
                continue
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * np.abs(g).max(), (name, idx, fd, g[idx])


def _train_problem(d):
    """A small reduced system on a structured node grid (Poisson for d = 1, plane-strain
    Neo-Hooke at a deformed state for d = 2) with its reduced -> full DOF map."""
    from oracle import assembly as OA
    from tests import problems as P
    if d == 1:
        dofs, settings, static, o = P.poisson_source(8)
        u = np.zeros(o["mask"].size)
        H, W = 9, 9
    else:
        dofs, settings, static, o = P.tapered_panel(6)
        u = np.random.default_rng(2).standard_normal(o["mask"].size) * 0.05
        H, W = 7, 7
    free = ~o["mask"]
    A = OA.reduced_tangent(o["sets"], o["coords"], u, o["d"], free)
    return A, np.flatnonzero(free), H, W


@pytest.mark.parametrize("kind,d,w_e,skip", [("anorm", 1, 0.0, False), ("anorm", 2, 0.3, False),
                                            ("energy", 1, 0.0, False), ("residual", 2, 0.0, True)])
def test_train_step_loss_gradient_matches_finite_differences(kind, d, w_e, skip):
    """The training losses of the learned preconditioner (relative A-norm error, energy,
    residual; through the Jacobi scaling and the RMS normalisation of z = M(r)) and their
    parameter gradients: central differences of the FP64 loss."""
    import scipy.sparse.linalg as spla
    A, red2full, H, W = _train_problem(d)
    L, ch = 2, [16, 32]
    net = OU.UNetOracle(L, ch, d, d, 0, H, W, bf16_storage=False)
    theta = OU.init_params(L, ch, d, d, 0, seed=5).astype(np.float64)
    rng = np.random.default_rng(6)
    r_list = [rng.standard_normal(A.shape[0]) for _ in range(2)]
    e_list = [spla.spsolve(A.tocsc(), r) for r in r_list]

    def loss(th):
        return OU.train_step(net, th, A, red2full, r_list, e_list, kind=kind, w_e=w_e, jacobi_skip=skip,
                             quant=False)[:2]

    l0, g = loss(theta)
    assert np.isfinite(l0)
    h = 1e-5
    checked = 0
    for name in ("in", "e0", "down0", "b", "up0", "d0", "out"):
        w_off, b_off, ci, co, cop, wk = net.layout[name]
        for idx in (w_off, w_off + 2 * wk + 3 if co > 2 else w_off + 5):
            tp, tm = theta.copy(), theta.copy()
            tp[idx] += h
            tm[idx] -= h
            fd = (loss(tp)[0] - loss(tm)[0]) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-5 * abs(fd) + 1e-7 * np.abs(g).max(), (name, idx, fd, g[idx])
            checked += 1
    assert checked == 14


def test_anorm_loss_is_the_relative_a_norm_error():
    """With A e = r the A-norm loss equals (z - e).|A|(z - e) / e.|A|e computed directly from z
    (a plausible slip such as a dropped e.res term or a wrong matrix sign fails this)."""
    import scipy.sparse.linalg as spla
    A, red2full, H, W = _train_problem(1)
    net = OU.UNetOracle(2, [16, 32], 1, 1, 0, H, W, bf16_storage=False)
    theta = OU.init_params(2, [16, 32], 1, 1, 0, seed=1).astype(np.float64)
    r = np.random.default_rng(0).standard_normal(A.shape[0])
    e = spla.spsolve(A.tocsc(), r)
    l, _, (z,) = OU.train_step(net, theta, A, red2full, [r], [e], kind="anorm", quant=False)
    Ah = -A  # negative definite Q1 tangent: |A| = -A
    direct = float((z - e) @ (Ah @ (z - e))) / float(e @ (Ah @ e))
    assert abs(l - direct) < 1e-12 * max(1.0, abs(direct))
    assert l > 0
