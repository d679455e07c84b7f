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
        y = self.A["y"]
        B = y.shape[0]
        t2 = (target ** 2).reshape(B, -1).sum(1)
        e2 = ((y - target) ** 2).reshape(B, -1).sum(1)
        loss = float(np.mean(e2 / t2))
        return loss, self.backward(2 * (y - target) / (t2[:, None, None, None] * B))

    def backward(self, gy):
        """Parameter gradient for the output gradient gy = dL/dy (B, H, W, c_out) of the last
        forward pass (stored in bf16 like the device's when bf16_storage)."""
        q, L, A, theta = self.q, self.L, self.A, self.theta_used
        y = A["y"]
        gy = q(gy)
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
        return grad


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


# ---- the preconditioner map and the training losses of apx_unet_train_step ---------------------

LOSS_KINDS = {"residual": 0, "energy": 1, "anorm": 2}


def precondition(orc, theta, A, red2full, r_list, g_list=None, jacobi_skip=False, quant=True):
    """The learned preconditioner's map z = M(r) per image (apx_unet_apply; the forward half of
    apx_unet_train_step), with the definitions listed under train_step: Jacobi scaling dsq,
    RMS scale s_b, FP32 / bf16 rounding points.  Returns (z_list, y, dsq, scales, sA)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    red2full = np.asarray(red2full, dtype=np.int64)
    H, W = orc.H[0], orc.W[0]
    d = orc.c_out
    B = len(r_list)
    diag = A.diagonal()
    ad = np.abs(diag)
    dsq = np.where(ad > 0, 1.0 / np.sqrt(np.where(ad > 0, ad, 1.0)), 1.0)
    sA = -1.0 if diag[0] < 0 else 1.0
    f32 = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if quant else (lambda a: np.asarray(a, dtype=np.float64))
    x0 = np.zeros((B, H, W, orc.c_in))
    scales = []
    for b, r in enumerate(r_list):
        v = r * dsq
        rms = np.sqrt(np.sum(v * v) / max(n, 1))
        sb = float(f32(rms)) if rms > 0 else 1.0
        scales.append(sb)
        inv = float(f32(1.0 / sb)) if quant else 1.0 / sb
        xin = f32(f32(v) * inv)
        flat = np.zeros(H * W * d)
        flat[red2full] = xin
        x0[b, :, :, :d] = flat.reshape(H, W, d)
        if orc.c_in > d:
            x0[b, :, :, d] = f32(np.asarray(g_list[b])).reshape(H, W)
    y = orc.forward(theta, x0)
    zs = []
    for b, r in enumerate(r_list):
        yb = y[b].reshape(-1)[red2full]
        z = f32(scales[b] * f32(yb)) * dsq
        if jacobi_skip:
            z = z + r * dsq * dsq
        zs.append(z)
    return zs, y, dsq, scales, sA


def train_step(orc, theta, A, red2full, r_list, e_list=None, g_list=None, kind="anorm", w_e=0.0,
               jacobi_skip=False, quant=True):
    """Loss and parameter gradient of one training step of the learned preconditioner
    (apx_unet_train_step), restated from the device's definitions:

    - D = |diag A|, dsq = D^-1/2 per reduced DOF (1 where the diagonal is 0); the matrix sign
      sA = sign(A_00) (definite matrices only; the Q1 tangents of the structured mesher are
      negative definite);
    - per image b: s_b = RMS(dsq r_b) rounded to FP32; the network input is dsq r_b / s_b at the
      free DOFs of the node grid (node-major, component fastest, zero at Dirichlet DOFs), FP32
      then bf16, plus the nodal field g_b as the last channel when the network has d + 1 inputs;
    - z_b = dsq s_b y_b (s_b y_b formed in FP32) on the free DOFs, plus dsq^2 r_b with the Jacobi
      skip; w = A z_b, res = r_b - w;
    - kind 'residual': L_b = ||res||^2 / ||r_b||^2;  'energy': L_b = sA (z.w / 2 - z.r_b) / (s_b^2 n);
      'anorm': L_b = (z_b - e_b).sA A (z_b - e_b) / e_b.sA A e_b written as
      -sA (z_b - e_b).res / (sA e_b.r_b) (exact when A e_b = r_b);  each + w_e ||z - e||^2 / ||e||^2
      when e is given;  loss = mean_b L_b;
    - dL/dy_b = s_b dsq dL_b/dz_b / B, stored in bf16, then the network's backward pass.

    quant=False keeps every value in FP64 (no FP32 / bf16 rounding), so the loss is a smooth
    function of theta and its gradient can be checked by finite differences.
    Returns (loss, grad, z_list)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    H, W = orc.H[0], orc.W[0]
    d = orc.c_out
    B = len(r_list)
    f32 = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if quant else (lambda a: np.asarray(a, dtype=np.float64))
    zs, y, dsq, scales, sA = precondition(orc, theta, A, red2full, r_list, g_list, jacobi_skip, quant)
    red2full = np.asarray(red2full, dtype=np.int64)
    gy = np.zeros_like(y)
    loss = 0.0
    for b, r in enumerate(r_list):
        sb = scales[b]
        z = zs[b]
        w = A @ z
        res = r - w
        e = None if e_list is None else e_list[b]
        ee = max(float(e @ e), 1e-300) if e is not None else 1.0
        if kind == "residual":
            rr = max(float(r @ r), 1e-300)
            Lb = float(res @ res) / rr
            gz = -2.0 * (A.T @ res) / rr
        elif kind == "energy":
            norm = sb * sb * n
            Lb = sA * (0.5 * float(z @ w) - float(z @ r)) / norm
            gz = -sA * res / norm
        elif kind == "anorm":
            if e is None:
                raise ValueError("the A-norm loss needs the exact errors e")
            E = max(sA * float(e @ r), 1e-300)
            Lb = -sA * (float(z @ res) - float(e @ res)) / E
            gz = -2.0 * sA * res / E
        else:
            raise ValueError(kind)
        if e is not None and (w_e > 0 or kind == "anorm"):
            Lb += w_e * float((z - e) @ (z - e)) / ee
            gz = gz + 2.0 * w_e * (z - e) / ee
        loss += Lb
        g = np.zeros(H * W * d)
        g[red2full] = f32(sb * gz * dsq / B)
        gy[b] = g.reshape(H, W, d)
    grad = orc.backward(gy)
    return loss / B, grad, zs
