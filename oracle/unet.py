"""Oracle: the multigrid-structured U-Net preconditioner in numpy (test infrastructure only).

No reference counterpart (the learned preconditioner is new).  Forward and
backward in FP64, optionally rounding every stored activation / activation
gradient to bf16 as the device kernels do.  Pinned by finite differences of
its own loss (tests/test_oracle_unet.py) and by the closed forms of the grid
transfers (P reproduces bilinear functions exactly; R = P^T).

Layer order and parameter layout (per layer: weights W[co_pad][k_pad] with
k = t*ci + c, t = 3*(dy+1) + (dx+1) for the row offset dy and column offset
dx, zero-padded; then biases[co]):
  in, (e_l, down_l) for l < L-1, bottom, (up_l, d_l) for l = L-2 .. 0, out.
"""
import numpy as np


def round16(c):
    return (c + 15) // 16 * 16


def kblocks(ci):
    return 9 * (ci // 16) if ci >= 16 else (9 * ci + 15) // 16


def layer_specs(levels, channels, c_in, c_out, relu):
    """[(name, ci, co, relu, level)] in parameter order."""
    ch = list(channels)
    L = levels
    specs = [("in", c_in, ch[0], relu, 0)]
    for l in range(L - 1):
        specs.append((f"e{l}", ch[l], ch[l], relu, l))
        specs.append((f"down{l}", ch[l], ch[l + 1], relu, l + 1))
    specs.append(("b", ch[L - 1], ch[L - 1], relu, L - 1))
    for l in range(L - 2, -1, -1):
        specs.append((f"up{l}", ch[l + 1], ch[l], 0, l + 1))
        specs.append((f"d{l}", ch[l], ch[l], relu, l))
    specs.append(("out", ch[0], c_out, 0, 0))
    return specs


def param_layout(specs):
    out, off = {}, 0
    for name, ci, co, r, lv in specs:
        cop, wk = round16(co), kblocks(ci) * 16
        out[name] = (off, off + cop * wk, ci, co, cop, wk)
        off += cop * wk + co
    return out, off


def bf16(a):
    """Round to bf16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta, layout, name):
    w_off, b_off, ci, co, cop, wk = layout[name]
    W = theta[w_off:w_off + cop * wk].reshape(cop, wk)[:co, :9 * ci].reshape(co, 9, ci)
    return W, theta[b_off:b_off + co]


def conv3x3(x, W, b):
    """x (B,H,Wd,ci); W (co,9,ci); y[b,i,j,o] = sum_t sum_c x[b,i+dy,j+dx,c] W[o,t,c] + b[o]."""
    B, H, Wd, ci = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((B, H, Wd, W.shape[0]))
    for t in range(9):
        dy, dx = t // 3, t % 3
        y += xp[:, dy:dy + H, dx:dx + Wd] @ W[:, t, :].T
    return y + (0 if b is None else b)


def conv3x3_wgrad(x, g, ci, co):
    B, H, Wd, _ = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    dW = np.zeros((co, 9, ci))
    for t in range(9):
        dy, dx = t // 3, t % 3
        dW[:, t, :] = np.einsum("bijo,bijc->oc", g, xp[:, dy:dy + H, dx:dx + Wd])
    return dW, g.sum((0, 1, 2))


def conv3x3_dgrad(g, W):
    """Adjoint of conv3x3 in x: flip taps, transpose channels."""
    Wf = W[:, ::-1, :].transpose(2, 1, 0)  # (ci, 9, co)
    return conv3x3(g, Wf, None)


def _w(d):
    return 1.0 if d == 0 else 0.5


def prolong(c, H, W):
    """Bilinear interpolation from the (ceil(H/2), ceil(W/2)) node grid to (H, W)."""
    B, Hc, Wc, C = c.shape
    out = np.zeros((B, H, W, C))
    for i in range(H):
        for j in range(W):
            I0, J0 = i // 2, j // 2
            I1, J1 = (I0 + 1 if i % 2 else I0), (J0 + 1 if j % 2 else J0)
            acc = 0
            for I in {I0, I1}:
                for J in {J0, J1}:
                    if I < Hc and J < Wc:
                        acc = acc + _w(i - 2 * I) * _w(j - 2 * J) * c[:, I, J]
            out[:, i, j] = acc
    return out


def restrict(f):
    """R = P^T: coarse(I, J) = sum w(di) w(dj) fine(2I+di, 2J+dj)."""
    B, H, W, C = f.shape
    Hc, Wc = (H + 1) // 2, (W + 1) // 2
    out = np.zeros((B, Hc, Wc, C))
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for I in range(Hc):
                i = 2 * I + di
                if i < 0 or i >= H:
                    continue
                for J in range(Wc):
                    j = 2 * J + dj
                    if 0 <= j < W:
                        out[:, I, J] += _w(di) * _w(dj) * f[:, i, j]
    return out


class UNetOracle:
    def __init__(self, levels, channels, c_in, c_out, relu, H, W, bf16_storage=True):
        self.L, self.ch, self.c_in, self.c_out, self.relu = levels, list(channels), c_in, c_out, relu
        self.specs = layer_specs(levels, channels, c_in, c_out, relu)
        self.layout, self.n_params = param_layout(self.specs)
        self.H, self.W = [H], [W]
        for _ in range(1, levels):
            self.H.append((self.H[-1] + 1) // 2)
            self.W.append((self.W[-1] + 1) // 2)
        self.q = bf16 if bf16_storage else (lambda a: np.asarray(a, dtype=np.float64))

    def _act(self, name, y):
        r = dict((s[0], s[3]) for s in self.specs)[name]
        return np.maximum(y, 0) if r else y

    def forward(self, theta, x0):
        q, L = self.q, self.L
        theta = np.asarray(theta, dtype=np.float64).copy()
        if self.q is bf16:  # the device multiplies with bf16 weights, adds FP32 biases
            for name in self.layout:
                w_off, b_off = self.layout[name][:2]
                theta[w_off:b_off] = bf16(theta[w_off:b_off])
        self.theta_used = theta
        A = {"x0": q(x0)}
        c = lambda name, x: self._act(name, conv3x3(x, *unpack(theta, self.layout, name)))
        A["a0"] = q(c("in", A["x0"]))
        for l in range(L - 1):
            A[f"d{l}"] = q(c(f"e{l}", A[f"a{l}"]))
            A[f"rd{l}"] = q(restrict(A[f"d{l}"]))
            A[f"a{l + 1}"] = q(c(f"down{l}", A[f"rd{l}"]))
        A[f"u{L - 1}"] = q(c("b", A[f"a{L - 1}"]))
        for l in range(L - 2, -1, -1):
            A[f"t{l}"] = q(c(f"up{l}", A[f"u{l + 1}"]))
            A[f"s{l}"] = q(A[f"d{l}"] + prolong(A[f"t{l}"], self.H[l], self.W[l]))
            A[f"u{l}"] = q(c(f"d{l}", A[f"s{l}"]))
        A["y"] = conv3x3(A["u0"], *unpack(theta, self.layout, "out"))
        self.A = A
        return A["y"]

    def loss_backward(self, target):
        """loss = mean_b ||y_b - t_b||^2 / ||t_b||^2; returns (loss, grad) with the device layout."""
        q, L, A, theta = self.q, self.L, self.A, self.theta_used
        y = A["y"]
        B = y.shape[0]
        t2 = (target ** 2).reshape(B, -1).sum(1)
        e2 = ((y - target) ** 2).reshape(B, -1).sum(1)
        loss = float(np.mean(e2 / t2))
        gy = q(2 * (y - target) / (t2[:, None, None, None] * B))
        grad = np.zeros(self.n_params)

        def put(name, dW, db):
            w_off, b_off, ci, co, cop, wk = self.layout[name]
            G = np.zeros((cop, wk))
            G[:co, :9 * ci] = dW.reshape(co, 9 * ci)
            grad[w_off:w_off + cop * wk] = G.ravel()
            grad[b_off:b_off + co] = db

        def back(name, x, g_out, out):
            ci, co = self.layout[name][2], self.layout[name][3]
            r = dict((s[0], s[3]) for s in self.specs)[name]
            g = g_out * (out > 0) if r else g_out
            put(name, *conv3x3_wgrad(x, g, ci, co))
            W, _ = unpack(theta, self.layout, name)
            return q(conv3x3_dgrad(g, W))

        gU = {0: back("out", A["u0"], gy, y)}
        gS = {}
        for l in range(L - 1):
            gS[l] = back(f"d{l}", A[f"s{l}"], gU[l], A[f"u{l}"])
            gT = q(restrict(gS[l]))
            gU[l + 1] = back(f"up{l}", A[f"u{l + 1}"], gT, A[f"t{l}"])
        gA = {L - 1: back("b", A[f"a{L - 1}"], gU[L - 1], A[f"u{L - 1}"])}
        for l in range(L - 2, -1, -1):
            gR = back(f"down{l}", A[f"rd{l}"], gA[l + 1], A[f"a{l + 1}"])
            gD = q(gS[l] + prolong(gR, self.H[l], self.W[l]))
            gA[l] = back(f"e{l}", A[f"a{l}"], gD, A[f"d{l}"])
        # input layer: weights only
        ci, co = self.layout["in"][2], self.layout["in"][3]
        g = gA[0] * (A["a0"] > 0) if self.relu else gA[0]
        put("in", *conv3x3_wgrad(A["x0"], g, ci, co))
        if not self.relu:
            # the linear network trains no biases (they stay at their initial zero, so the map is linear)
            for name in self.layout:
                b_off, co = self.layout[name][1], self.layout[name][3]
                grad[b_off:b_off + co] = 0.0
        return loss, grad


def init_params(levels, channels, c_in, c_out, relu, seed=0, gain=None):
    """He-normal weights (std sqrt(2 / fan_in) with ReLU, sqrt(1 / fan_in) without), zero biases."""
    specs = layer_specs(levels, channels, c_in, c_out, relu)
    layout, n = param_layout(specs)
    rng = np.random.default_rng(seed)
    theta = np.zeros(n, dtype=np.float32)
    for name, ci, co, r, lv in specs:
        w_off, b_off, ci, co, cop, wk = layout[name]
        std = np.sqrt((2.0 if r else 1.0) / (9 * ci)) if gain is None else gain
        W = np.zeros((cop, wk), dtype=np.float32)
        W[:co, :9 * ci] = rng.standard_normal((co, 9 * ci)) * std
        theta[w_off:w_off + cop * wk] = W.ravel()
    return theta
