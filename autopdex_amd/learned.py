"""Learned preconditioner: a multigrid-structured U-Net trained on the GPU.

The network approximates A^-1 r on the structured node grid and is used
inside flexible CG (static_settings['type of preconditioner'] = 'learned').
Training data are (residual, error) pairs that the plain solver itself
produces: iterates x_k of solves A x = b give r_k = b - A x_k and
e_k = x* - x_k, with x* from Jacobi-PCG; the last part of fit() adds the
iterates of flexible CG with the network being trained.  Forward, backward
and Adam run in libautopdex_amd's HIP kernels (bf16 MFMA convolutions, FP32
master weights).

No counterpart in the reference; it plugs into the reference's Newton /
linear-solve flow through the 'hip' backend's preconditioner option.
"""
import ctypes as C
import hashlib
import os
import time

import numpy as np

from . import _native as N
from .backend import current_stream, torch_cuda

def _call(name, *args):
    N.call(name, *args)


def _lib():
    return N.lib()


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class UNetPreconditioner:
    """Network on an (height x width) node grid; levels with channels[l] in {16, 32, 64}."""

    def __init__(self, height, width, levels=4, channels=(16, 32, 64, 64), c_in=1, c_out=1, relu=False, batch=1,
                 seed=0, jacobi_skip=False):
        torch_cuda()
        self.height, self.width, self.levels = int(height), int(width), int(levels)
        self.channels = tuple(int(c) for c in channels)[:levels]
        self.c_in, self.c_out, self.relu, self.batch = int(c_in), int(c_out), bool(relu), int(batch)
        ch = (C.c_int * len(self.channels))(*self.channels)
        h = C.c_void_p()
        _call("apx_unet_create", C.byref(h), self.levels, ch, self.c_in, self.c_out, int(self.relu), self.batch,
              self.height, self.width)
        self.handle = h
        self.jacobi_skip = bool(jacobi_skip)
        _call("apx_unet_set_jacobi_skip", h, int(self.jacobi_skip))
        n, nl = C.c_int64(), C.c_int()
        _call("apx_unet_num_params", h, C.byref(n), C.byref(nl))
        self.n_params, self.n_layers = n.value, nl.value
        self.layers = []
        for i in range(self.n_layers):
            w, b = C.c_int64(), C.c_int64()
            ci, co, cop, wk, r, lv = (C.c_int() for _ in range(6))
            _call("apx_unet_layer", h, i, C.byref(w), C.byref(b), C.byref(ci), C.byref(co), C.byref(cop),
                  C.byref(wk), C.byref(r), C.byref(lv))
            self.layers.append(dict(w_off=w.value, b_off=b.value, ci=ci.value, co=co.value, co_pad=cop.value,
                                    k_pad=wk.value, relu=r.value, level=lv.value))
        self.set_params(self.initial_params(seed))
        self.step = 0

    def up_layers(self):
        """Indices of the coarse-to-fine ('up') convolutions; layer order: in, (e_l, down_l) for
        l = 0..L-2, b, (up_l, d_l) for l = L-2..0, out."""
        L = self.levels
        return [2 * L + 2 * k for k in range(L - 1)]

    def initial_params(self, seed, up_scale=1.0):
        """He-normal weights: std sqrt(gain / (9 ci)), gain 2 with ReLU else 1; zero biases.  The
        output layer is scaled down so the untrained network starts close to zero; up_scale < 1
        starts the coarse branches small (full weighting sums 4 smooth values per level, so an
        untrained deep branch amplifies smooth inputs by about 4 per level)."""
        rng = np.random.default_rng(seed)
        theta = np.zeros(self.n_params, dtype=np.float32)
        ups = set(self.up_layers())
        for i, ly in enumerate(self.layers):
            std = np.sqrt((2.0 if ly["relu"] else 1.0) / (9 * ly["ci"]))
            if i in ups:
                std *= up_scale
            if i == self.n_layers - 1:
                std *= 0.1
            W = np.zeros((ly["co_pad"], ly["k_pad"]), dtype=np.float32)
            W[:ly["co"], :9 * ly["ci"]] = rng.standard_normal((ly["co"], 9 * ly["ci"])) * std
            theta[ly["w_off"]:ly["w_off"] + W.size] = W.ravel()
        return theta

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        assert theta.size == self.n_params
        _call("apx_unet_set_params", self.handle, theta.ctypes.data)

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        _call("apx_unet_get_params", self.handle, out.ctypes.data)
        return out

    def get_grads(self):
        out = np.empty(self.n_params, dtype=np.float32)
        _call("apx_unet_get_grads", self.handle, out.ctypes.data)
        return out

    def buffers(self):
        x, y, s = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _call("apx_unet_buffers", self.handle, C.byref(x), C.byref(y), C.byref(s))
        return x.value, y.value, s.value

    def set_fused_tail(self, enable):
        """Coarse levels in one launch (True, the default) or one launch per layer; returns the
        level whose restriction starts the fused part (-1: none)."""
        level = C.c_int()
        _call("apx_unet_set_fused_tail", self.handle, int(bool(enable)), C.byref(level))
        return level.value

    def forward(self):
        _call("apx_unet_forward", self.handle, current_stream())

    def loss_backward(self, target):
        lval = C.c_double()
        _call("apx_unet_loss_backward", self.handle, _vp(target), C.byref(lval), current_stream())
        return lval.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _call("apx_unet_adam", self.handle, lr, beta1, beta2, eps, self.step, current_stream())

    def load_pair(self, problem, r, e=None, target=None, g=None, image=0):
        _call("apx_unet_load_pair", self.handle, problem.plan, _vp(r), _vp(g), _vp(e), _vp(target), image,
              current_stream())

    def apply(self, problem, r, g=None):
        z = r.new_empty(r.shape)
        _call("apx_unet_apply", self.handle, problem.plan, _vp(r), _vp(g), _vp(z), current_stream())
        return z

    def attach(self, problem, g=None):
        """Use this network as the plan's preconditioner; g = nodal input channel (device FP64)."""
        _call("apx_plan_attach_unet", problem.plan, self.handle, _vp(g))
        problem._attached_net = (self, g)

    def spec(self):
        return dict(height=self.height, width=self.width, levels=self.levels, channels=list(self.channels),
                    c_in=self.c_in, c_out=self.c_out, relu=self.relu, jacobi_skip=self.jacobi_skip)

    def save(self, path):
        theta = self.get_params()
        digest = hashlib.sha256(theta.tobytes()).hexdigest()
        np.savez(path, theta=theta, spec=np.array(repr(self.spec())), sha256=np.array(digest))
        return digest

    @classmethod
    def load(cls, path, height=None, width=None, batch=1):
        d = np.load(path, allow_pickle=False)
        import ast
        spec = ast.literal_eval(str(d["spec"]))
        theta = d["theta"]
        if hashlib.sha256(theta.tobytes()).hexdigest() != str(d["sha256"]):
            raise ValueError(f"{path}: parameter hash mismatch")
        net = cls(height or spec["height"], width or spec["width"], spec["levels"], spec["channels"], spec["c_in"],
                  spec["c_out"], spec["relu"], batch=batch, jacobi_skip=spec.get("jacobi_skip", False))
        net.set_params(theta)
        return net

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib().apx_unet_destroy(self.handle)
        except Exception:
            pass


def fcg(problem, vals, b, tol=1e-10, atol=0.0, maxiter=2000, precision="fp64"):
    """One flexible-CG solve with the attached network (x0 = 0); vals in CSR order."""
    x = b.new_empty(b.shape)
    ks = N.KrylovStats()
    p = N.APX_FP64 if precision == "fp64" else N.APX_FP32
    _call("apx_fcg", problem.plan, _vp(vals), _vp(b), _vp(x), p, tol, atol, maxiter, C.byref(ks), current_stream())
    return x, ks


def pcg(problem, vals, b, tol=1e-10, maxiter=0, precision="fp64"):
    x = b.new_empty(b.shape)
    ks = N.KrylovStats()
    p = N.APX_FP64 if precision == "fp64" else N.APX_FP32
    N.call("apx_pcg", problem.plan, _vp(vals), _vp(b), _vp(x), p, N.APX_PRECOND_JACOBI, tol, 0.0, maxiter,
           C.byref(ks), current_stream())
    return x, ks


def spmv(problem, vals, x):
    y = x.new_empty(x.shape)
    N.call("apx_spmv", problem.plan, _vp(vals), _vp(x), _vp(y), N.APX_FP64, current_stream())
    return y


def tangent(problem, u=None):
    torch = torch_cuda()
    if u is None:
        u = torch.zeros(problem.n_dofs, dtype=torch.float64, device=problem.device)
    vals = torch.empty(problem.nnz, dtype=torch.float64, device=problem.device)
    N.call("apx_assemble_tangent", problem.plan, _vp(u), _vp(vals), N.APX_FP64, current_stream())
    return vals


def random_field(height, width, length=0.05, sigma=1.0, seed=0, device="cuda"):
    """Squared-exponential Gaussian random field on a (height x width) node grid of the
    unit square, by circulant embedding (FFT on the doubled periodic grid), generated on
    the device; returns a flat FP64 tensor in node order i * width + j."""
    import torch
    h = 1.0 / (max(height, width) - 1)
    M0, M1 = 2 * height, 2 * width
    k0 = torch.minimum(torch.arange(M0, device=device), M0 - torch.arange(M0, device=device)).double() * h
    k1 = torch.minimum(torch.arange(M1, device=device), M1 - torch.arange(M1, device=device)).double() * h
    cov = sigma ** 2 * torch.exp(-(k0[:, None] ** 2 + k1[None, :] ** 2) / (2 * length ** 2))
    lam = torch.clamp(torch.fft.fft2(cov).real, min=0.0)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    xi = torch.complex(torch.randn(M0, M1, generator=gen, device=device, dtype=torch.float64),
                       torch.randn(M0, M1, generator=gen, device=device, dtype=torch.float64))
    f = torch.fft.fft2(torch.sqrt(lam / (M0 * M1)) * xi)
    return f.real[:height, :width].contiguous().view(-1)


def default_levels(height, width, coarsest=3):
    """Levels of node-centred coarsening (H -> (H+1)/2) until the coarsest grid has about
    `coarsest` nodes per side: the coarse levels must carry the smooth part of A^-1, so the
    network reaches as deep as a multigrid hierarchy would (at most 12 levels)."""
    L, h, w = 1, height, width
    while L < 12 and min((h + 1) // 2, (w + 1) // 2) >= coarsest:
        h, w, L = (h + 1) // 2, (w + 1) // 2, L + 1
    return L


def fit(problem, height, width, vals=None, levels=None, channels=None, n_noise=16, n_smooth=16,
        pair_tol=1e-7, steps=1500, batch=4, lr=1e-4, up_scale=0.05, seed=0, validate_every=250,
        log=None, time_budget=None, jacobi_skip=False, self_fraction=1.0 / 3.0, self_lr=None,
        self_iterates=(4, 8, 16, 32), bootstrap=None):
    """Train a linear U-Net preconditioner for `problem` (d DOFs per node on a structured
    height x width node grid: d input and output channels) and attach it.

    Training pairs come from the plain solver: Jacobi-PCG solves of A e = r, to a relative
    residual of pair_tol, for white-noise and smooth random right-hand sides r
    (squared-exponential fields of lengths 0.2 ... 0.005).  At 2048^2, pair_tol 1e-7 gave
    28 / 28 / 38 flexible-CG iterations in the benchmark's learned_mixed solve for seeds 0 / 1 / 2,
    against 31 / 32 / 40 with 1e-9, and 3-5 s less pair generation.
    The loss is the relative A-norm error, the quantity CG minimizes.  Every `validate_every`
    steps the parameters are scored by flexible CG (to 1e-6) on a held-out right-hand side and
    the best ones are kept.

    The last `self_fraction` of the steps also train on the residuals that flexible CG with the
    network meets: the best network so far runs `self_iterates` iterations of flexible CG on
    every training right-hand side b, and each iterate x_k adds the pair (b - A x_k, x* - x_k)
    (x* from the plain solve).  These steps run at `self_lr` (default 0.15 lr), decaying to a
    third of it.  At 2048^2 (pair_tol 1e-9) this took the benchmark's learned_mixed Newton solve
    from 48 to 31 flexible-CG iterations at the same number of steps (seed 1: 45 to 32); Jacobi-PCG
    iterates as extra pairs made the network fail instead.  self_fraction = 0 trains on the
    plain solver's pairs only.

    bootstrap = (n, k): only n right-hand sides (half white noise, half smooth) are solved by
    Jacobi-PCG up front; after k steps the best network so far solves the others by flexible CG
    to pair_tol (Jacobi-PCG finishes a solve that stagnates), which is far cheaper than plain
    solves, and they join the pairs.  At 2048^2, (8, 500) cut the setup from 48 to 32 s for one
    more flexible-CG iteration in learned_mixed (29 against 28, seed 0); (8, 250): 29 / 29 / 40
    iterations for seeds 0 / 1 / 2.  None (the default) solves every right-hand side up front.
    Returns (network with batch 1, attached; info dict)."""
    if not 0.0 <= self_fraction < 1.0:
        raise ValueError("self_fraction must be in [0, 1)")
    torch = torch_cuda()
    t0 = time.perf_counter()
    if vals is None:
        vals = tangent(problem)
    L = levels or default_levels(height, width)
    ch = tuple(channels or ((16, 32) + (64,) * 10)[:L])
    free = torch.from_numpy(problem.free_dofs().astype(np.int64)).to(problem.device)
    gen = torch.Generator(device=problem.device)
    gen.manual_seed(int(seed))
    lengths = (0.2, 0.05, 0.02, 0.005)

    def noise():
        return torch.randn(problem.n_free, generator=gen, device=problem.device, dtype=torch.float64)

    d = problem.n_dofs // (height * width)  # DOFs per node (fields interleaved node-major, P3)

    def smooth(s):
        comps = [random_field(height, width, lengths[s % len(lengths)], seed=seed * 1000 + d * s + c)
                 for c in range(d)]
        return torch.stack(comps, dim=1).reshape(-1)[free].contiguous()

    rhs = [noise() for _ in range(n_noise)] + [smooth(s) for s in range(n_smooth)]
    if bootstrap:
        nb, boot_steps = int(bootstrap[0]), int(bootstrap[1])
        first = list(range(min(nb // 2, n_noise))) + list(range(n_noise, n_noise + min(nb - nb // 2, n_smooth)))
    else:
        boot_steps, first = -1, list(range(len(rhs)))
    later = [i for i in range(len(rhs)) if i not in first]
    base = []  # (b, x*): the pair of the iterate x_0 = 0 of each right-hand side solved so far
    for i in first:
        base += solver_pairs(problem, vals, rhs[i], checkpoints=(0,), tol=pair_tol)
    pairs = list(base)
    b_val = noise() + smooth(n_smooth + 1) * 10.0  # held out
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    if log:
        log(f"[fit] {len(pairs)} (r, A^-1 r) pairs from {len(first)} plain solves in {t_pairs:.1f} s; levels {L}, "
            f"channels {ch}")
    net = UNetPreconditioner(height, width, L, ch, d, d, relu=False, batch=batch, seed=seed + 1,
                             jacobi_skip=jacobi_skip)
    net.set_params(net.initial_params(seed + 1, up_scale=up_scale))
    inf = UNetPreconditioner(height, width, L, ch, d, d, relu=False, batch=1, jacobi_skip=jacobi_skip)
    inf.attach(problem)
    t1 = time.perf_counter()
    best, history, done = None, [], 0
    seg = validate_every if validate_every else steps
    n_self = int(round(steps * self_fraction)) if self_iterates else 0
    n_first = steps - n_self
    lr_self = 0.15 * lr if self_lr is None else self_lr
    n_self_pairs, t_self, n_boot, t_boot = 0, 0.0, len(later), 0.0
    while done < steps:
        if done == boot_steps and later:
            # the remaining right-hand sides, solved by flexible CG with the best network so far
            tb = time.perf_counter()
            inf.set_params(best[1] if best is not None else net.get_params())
            for i in later:
                b = rhs[i]
                x, ks = fcg(problem, vals, b, tol=pair_tol, maxiter=2000)
                if not bool(torch.isfinite(x).all()):
                    x = torch.zeros_like(b)
                if not ks.converged:  # stagnation: Jacobi-PCG finishes from the iterate
                    r2 = b - spmv(problem, vals, x)
                    rel = pair_tol * float(b.norm()) / max(float(r2.norm()), 1e-300)
                    x2, _ = pcg(problem, vals, r2, tol=min(rel, 1.0))
                    x = x + x2
                base.append((b, x))
                pairs.append((b, x))
            later = []
            torch.cuda.synchronize()
            t_boot = time.perf_counter() - tb
            if log:
                log(f"[fit] {n_boot} pairs from flexible-CG solves in {t_boot:.1f} s")
        if done == n_first and n_self > 0:
            # pairs from the flexible-CG iterates of the best network so far
            ts = time.perf_counter()
            theta = best[1] if best is not None else net.get_params()
            inf.set_params(theta)
            net.set_params(theta)
            # (x* solves to pair_tol: an iterate whose residual is near that carries no signal)
            for b, xs in base:
                for k_it in self_iterates:
                    xk, _ = fcg(problem, vals, b, tol=0.0, maxiter=k_it)
                    rk = b - spmv(problem, vals, xk)
                    if bool(torch.isfinite(xk).all()) and float(rk.norm()) > 100.0 * pair_tol * float(b.norm()):
                        pairs.append((rk, xs - xk))
                        n_self_pairs += 1
            torch.cuda.synchronize()
            t_self = time.perf_counter() - ts
            if log:
                log(f"[fit] {n_self_pairs} pairs from flexible-CG iterates in {t_self:.1f} s")
        if done < n_first:
            k = min(seg, (boot_steps if 0 <= done < boot_steps else n_first) - done)
            lr0 = lr * (0.1 ** (done / steps))
            lr1 = lr * (0.1 ** ((done + k) / steps))
        else:
            k = min(seg, steps - done)
            lr0 = lr_self * (1.0 / 3.0) ** ((done - n_first) / n_self)
            lr1 = lr_self * (1.0 / 3.0) ** ((done + k - n_first) / n_self)
        history += train(net, problem, pairs, vals, steps=k, lr=lr0, lr_final=lr1,
                         warmup=50 if done == 0 else 0, seed=seed + done, loss="anorm", log=None)
        done += k
        theta = net.get_params()
        inf.set_params(theta)
        _, ks = fcg(problem, vals, b_val, tol=1e-6, maxiter=400)
        score = ks.iterations if ks.converged else 10 ** 6
        if log:
            log(f"[fit] step {done}: loss {np.mean(history[-50:]):.3e}, validation FCG {ks.iterations} it"
                f"{'' if ks.converged else ' (not converged)'}")
        if best is None or score <= best[0]:
            best = (score, theta, done)
        if time_budget and time.perf_counter() - t1 > time_budget:
            break
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    inf.set_params(best[1])
    info = dict(pairs=len(base) - n_boot, boot_pairs=n_boot, boot_pair_seconds=t_boot, self_pairs=n_self_pairs,
                pair_seconds=t_pairs, self_pair_seconds=t_self,
                train_seconds=t_train, steps=done, self_steps=max(0, done - n_first), best_step=best[2],
                validation_iterations=best[0], levels=L, channels=list(ch), batch=batch,
                final_loss=float(np.mean(history[-50:])), n_params=int(inf.n_params))
    return inf, info


def fit_fields(problem, height, width, fields, set_index=0, levels=None, channels=None, n_noise=2, n_smooth=2,
               pair_tol=1e-9, steps=1500, batch=4, lr=1e-4, up_scale=0.05, seed=0, validate_every=250,
               validation_field=None, log=None, jacobi_skip=False):
    """Train one network for a family of coefficient fields of a variable-coefficient model
    (models.poisson_variable_coefficient_integrand): the nodal log-coefficient g is the
    network's extra input channel, so it can follow the field it is applied to.

    fields: list of (g, f) device tensors (FP64, one value per node).  For each field the
    plan's tangent is assembled and the plain solver (Jacobi-PCG) produces (r, A^-1 r) pairs for
    noise and smooth right-hand sides.  Every training step draws one field and a batch of its
    pairs, so the A-norm loss uses that field's matrix.  validation_field (g, f), held out,
    scores the parameters by flexible-CG iterations (to 1e-6) every `validate_every` steps; the
    best are kept.  Returns (network with batch 1, attached with the validation field's g when
    given, else the last field's; info dict).  Leaves the last bound field on the plan."""
    torch = torch_cuda()
    t0 = time.perf_counter()
    L = levels or default_levels(height, width)
    ch = tuple(channels or ((16, 32) + (64,) * 10)[:L])
    free = torch.from_numpy(problem.free_dofs().astype(np.int64)).to(problem.device)
    gen = torch.Generator(device=problem.device)
    gen.manual_seed(int(seed))
    lengths = (0.2, 0.05, 0.02, 0.005)
    data = []
    for fi, (g, f) in enumerate(fields):
        problem.bind_nodal_device(set_index, g, f)
        vals = tangent(problem)
        rhs = [torch.randn(problem.n_free, generator=gen, device=problem.device, dtype=torch.float64)
               for _ in range(n_noise)]
        rhs += [random_field(height, width, lengths[(fi + s) % len(lengths)], seed=seed * 100003 + fi * 101 + s)[free]
                .contiguous() for s in range(n_smooth)]
        prs = []
        for b in rhs:
            prs += solver_pairs(problem, vals, b, checkpoints=(0,), tol=pair_tol)
        data.append((vals, g, prs))
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    if log:
        log(f"[fit_fields] {sum(len(d[2]) for d in data)} pairs on {len(data)} fields in {t_pairs:.1f} s; "
            f"levels {L}, channels {ch}")
    d = problem.n_dofs // (height * width)
    net = UNetPreconditioner(height, width, L, ch, d + 1, d, relu=False, batch=batch, seed=seed + 1,
                             jacobi_skip=jacobi_skip)
    net.set_params(net.initial_params(seed + 1, up_scale=up_scale))
    inf = UNetPreconditioner(height, width, L, ch, d + 1, d, relu=False, batch=1, jacobi_skip=jacobi_skip)
    val = None
    if validation_field is not None:
        gv, fv = validation_field
        problem.bind_nodal_device(set_index, gv, fv)
        val = (tangent(problem), gv,
               torch.randn(problem.n_free, generator=gen, device=problem.device, dtype=torch.float64))
    rng = np.random.default_rng(seed)
    B = batch
    history, best, done = [], None, 0
    t1 = time.perf_counter()
    seg = validate_every if (validate_every and val is not None) else steps
    while done < steps:
        k = min(seg, steps - done)
        for it in range(done, done + k):
            vals, g, prs = data[rng.integers(len(data))]
            idx = rng.integers(0, len(prs), size=B)
            r_list = (C.c_void_p * B)(*[prs[i][0].data_ptr() for i in idx])
            e_list = (C.c_void_p * B)(*[prs[i][1].data_ptr() for i in idx])
            g_list = (C.c_void_p * B)(*([g.data_ptr()] * B))
            lval = C.c_double()
            _call("apx_unet_train_step", net.handle, problem.plan, _vp(vals), r_list, g_list, e_list,
                  _LOSS_KINDS["anorm"], 0.0, C.byref(lval), current_stream())
            warm = min(1.0, (it + 1) / 50)
            net.adam(warm * lr * 0.1 ** (it / max(steps - 1, 1)))
            history.append(lval.value)
        done += k
        theta = net.get_params()
        if val is not None:
            inf.set_params(theta)
            inf.attach(problem, val[1])
            _, ks = fcg(problem, val[0], val[2], tol=1e-6, maxiter=400)
            score = ks.iterations if ks.converged else 10 ** 6
            if log:
                log(f"[fit_fields] step {done}: loss {np.mean(history[-50:]):.3e}, held-out field FCG {ks.iterations} it")
        else:
            score = 0
        if best is None or score <= best[0]:
            best = (score, theta, done)
    torch.cuda.synchronize()
    inf.set_params(best[1])
    inf.attach(problem, val[1] if val is not None else data[-1][1])
    info = dict(fields=len(data), pairs=sum(len(x[2]) for x in data), pair_seconds=t_pairs,
                train_seconds=time.perf_counter() - t1, steps=done, best_step=best[2],
                validation_iterations=best[0] if val is not None else None, levels=L, channels=list(ch),
                final_loss=float(np.mean(history[-50:])), n_params=int(inf.n_params))
    return inf, info


def fit_field_family(problem, height, width, field, train_seeds, set_index=0, levels=None, channels=None,
                     pair_tol=1e-6, steps=1500, lr=1e-4, up_scale=0.05, seed=0, validate_every=250,
                     validation_seed=None, log=None):
    """fit_fields for a large family of coefficient fields, streaming: field(seed) -> (g, f)
    device tensors (FP64, one value per node) is called once per training seed to make two
    (r, A^-1 r) pairs by plain Jacobi-PCG (to pair_tol) -- one white-noise and one smooth
    right-hand side -- which are kept in FP32 with g; the field's tangent is re-assembled at
    every training step that draws it (a few hundred microseconds at 1025 x 1025) instead of
    being stored.  Each step trains on the two pairs of one field (batch 2, relative A-norm
    loss with that field's matrix).  Returns (network with batch 1, attached with the validation
    field's g when given, else the last field's; info dict)."""
    torch = torch_cuda()
    t0 = time.perf_counter()
    L = levels or default_levels(height, width)
    ch = tuple(channels or ((16, 32) + (64,) * 10)[:L])
    free = torch.from_numpy(problem.free_dofs().astype(np.int64)).to(problem.device)
    gen = torch.Generator(device=problem.device)
    gen.manual_seed(int(seed))
    lengths = (0.2, 0.05, 0.02, 0.005)
    data = []
    for k, s in enumerate(train_seeds):
        g, f = field(s)
        problem.bind_nodal_device(set_index, g, f)
        vals = tangent(problem)
        rhs = [torch.randn(problem.n_free, generator=gen, device=problem.device, dtype=torch.float64),
               random_field(height, width, lengths[k % len(lengths)], seed=seed * 100003 + int(s))[free].contiguous()]
        prs = []
        for b in rhs:
            e, _ = pcg(problem, vals, b, tol=pair_tol)
            prs.append((b.float(), e.float()))
        data.append((g, prs))
        if log and (k + 1) % 250 == 0:
            log(f"[fit_field_family] pairs of {k + 1}/{len(train_seeds)} fields, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    d = problem.n_dofs // (height * width)
    net = UNetPreconditioner(height, width, L, ch, d + 1, d, relu=False, batch=2, seed=seed + 1)
    net.set_params(net.initial_params(seed + 1, up_scale=up_scale))
    inf = UNetPreconditioner(height, width, L, ch, d + 1, d, relu=False, batch=1)
    val = None
    if validation_seed is not None:
        gv, fv = field(validation_seed)
        problem.bind_nodal_device(set_index, gv, fv)
        val = (tangent(problem), gv, gv.new_empty(problem.n_free).normal_(generator=gen))
    rng = np.random.default_rng(seed)
    history, best, done = [], None, 0
    t1 = time.perf_counter()
    seg = validate_every if (validate_every and val is not None) else steps
    r64 = [torch.empty(problem.n_free, dtype=torch.float64, device=problem.device) for _ in range(2)]
    e64 = [torch.empty_like(r64[0]) for _ in range(2)]
    while done < steps:
        k = min(seg, steps - done)
        for it in range(done, done + k):
            g, prs = data[rng.integers(len(data))]
            problem.bind_nodal_device(set_index, g, g)  # the tangent does not read f
            vals = tangent(problem)
            for j in range(2):
                r64[j].copy_(prs[j][0])
                e64[j].copy_(prs[j][1])
            r_list = (C.c_void_p * 2)(*[t.data_ptr() for t in r64])
            e_list = (C.c_void_p * 2)(*[t.data_ptr() for t in e64])
            g_list = (C.c_void_p * 2)(g.data_ptr(), g.data_ptr())
            lval = C.c_double()
            _call("apx_unet_train_step", net.handle, problem.plan, _vp(vals), r_list, g_list, e_list,
                  _LOSS_KINDS["anorm"], 0.0, C.byref(lval), current_stream())
            warm = min(1.0, (it + 1) / 50)
            net.adam(warm * lr * 0.1 ** (it / max(steps - 1, 1)))
            history.append(lval.value)
        done += k
        theta = net.get_params()
        if val is not None:
            inf.set_params(theta)
            inf.attach(problem, val[1])
            _, ks = fcg(problem, val[0], val[2], tol=1e-6, maxiter=400)
            score = ks.iterations if ks.converged else 10 ** 6
            if log:
                log(f"[fit_field_family] step {done}: loss {np.mean(history[-50:]):.3e}, held-out field FCG "
                    f"{ks.iterations} it")
        else:
            score = 0
        if best is None or score <= best[0]:
            best = (score, theta, done)
    torch.cuda.synchronize()
    inf.set_params(best[1])
    inf.attach(problem, val[1] if val is not None else data[-1][0])
    info = dict(fields=len(data), pairs=2 * len(data), pair_tol=pair_tol, pair_seconds=t_pairs,
                train_seconds=time.perf_counter() - t1, steps=done, best_step=best[2],
                validation_iterations=best[0] if val is not None else None, levels=L, channels=list(ch),
                final_loss=float(np.mean(history[-50:])), n_params=int(inf.n_params))
    return inf, info


def solver_pairs(problem, vals, rhs, checkpoints=(0, 4, 16, 48, 128, 320, 800, 2000), tol=1e-10):
    """(r_k, e_k) pairs from Jacobi-PCG iterates of A x = b: e_k = x* - x_k, r_k = b - A x_k.
    Restarted PCG runs with maxiter = k reproduce the iterates x_k of the full solve."""
    torch = torch_cuda()
    xs, ks = pcg(problem, vals, rhs, tol=tol)
    out = []
    for k in checkpoints:
        if k >= ks.iterations:
            break
        if k == 0:
            xk = torch.zeros_like(rhs)
        else:
            xk, _ = pcg(problem, vals, rhs, tol=0.0, maxiter=k)
        out.append((rhs - spmv(problem, vals, xk), xs - xk))
    return out


_LOSS_KINDS = {"residual": 0, "energy": 1, "anorm": 2}


def train(net, problem, pairs, vals, steps=500, lr=2e-3, lr_final=2e-4, w_e=0.0, g=None, seed=0, log_every=50,
          log=print, time_budget=None, warmup=0, loss="anorm"):
    """Adam over random batches of (r, e[, g]) pairs; vals: FP64 tangent (CSR).  loss="anorm":
    relative A-norm error (z-e).|A|(z-e) / e.|A|e of z = M(r), e = A^-1 r (the quantity CG
    minimizes); "energy": sA (z.Az/2 - z.r) / r.r without e; "residual": ||r - A z||^2 / ||r||^2;
    optionally + w_e x relative error ||z - e||^2 / ||e||^2.
    Learning rate: linear warm-up over `warmup` steps, then geometric decay from lr to lr_final."""
    torch_cuda()
    loss_kind = loss
    if loss_kind not in _LOSS_KINDS:
        raise ValueError(f"loss must be one of {sorted(_LOSS_KINDS)}")
    rng = np.random.default_rng(seed)
    B = net.batch
    history = []
    t0 = time.perf_counter()
    for it in range(steps):
        idx = rng.integers(0, len(pairs), size=B)
        r_list = (C.c_void_p * B)(*[pairs[i][0].data_ptr() for i in idx])
        e_list = ((C.c_void_p * B)(*[pairs[i][1].data_ptr() for i in idx])
                  if (w_e > 0 or loss_kind == "anorm") else None)
        gs = [(pairs[i][2] if len(pairs[i]) > 2 else g) for i in idx]
        g_list = (C.c_void_p * B)(*[x.data_ptr() for x in gs]) if gs[0] is not None else None
        lval = C.c_double()
        _call("apx_unet_train_step", net.handle, problem.plan, _vp(vals), r_list, g_list, e_list,
              _LOSS_KINDS[loss_kind], float(w_e), C.byref(lval), current_stream())
        frac = it / max(steps - 1, 1)
        warm = min(1.0, (it + 1) / max(1, warmup))
        net.adam(warm * lr * (lr_final / lr) ** frac)
        history.append(lval.value)
        if log and (it % log_every == 0 or it == steps - 1):
            log(f"[train] step {it} loss {lval.value:.4e} ({time.perf_counter() - t0:.1f} s)")
        if time_budget and time.perf_counter() - t0 > time_budget:
            break
    return history
