"""Benchmark: time-to-solution of the 2048x2048 Q1 Poisson problem (4.2M DOFs).

Workload (one "step"): a full Newton solve on one GPU to a 1e-10 FP64
relative residual ||r_free(u)||/||r_free(u0)||, inputs resident in HBM, for
each method:
  plain         = FP64 assembly + FP64 Jacobi-PCG (the drop-in method)
  mixed         = FP64 tangent rounded once to FP32 + one FP32 Jacobi-PCG
                  re-anchored to the FP64 residual (reliable-update refinement)
  learned       = FP64 tangent + flexible CG with the U-Net preconditioner
                  (bf16 convolutions on MFMA, FP32 accumulation)
  learned_mixed = FP64 tangent rounded once to FP32 + FP32-matrix flexible CG
                  with the U-Net, each inner solve re-anchored by the FP64
                  Newton residual
  learned_mixed_fp32_elements = learned_mixed with the tangent's element
                  arithmetic in FP32 (FP32 accumulation) instead of rounding
The network is trained on the GPU before the timed steps (learned.fit: pairs
from plain Jacobi-PCG solves of random right-hand sides, A-norm loss); its
setup time is reported in `learned_setup` and is not part of any step.
The headline `value` is the time-to-solution (seconds, lower is better) of
the --headline method (default learned_mixed: learned + mixed precision);
every method's time, Krylov iterations, residual and error against plain are
reported beside it, with assembly throughput and a roofline for the heaviest
kernel of the headline method (the level-0 convolution for the learned
methods, the fused PCG SpMV otherwise) measured with HIP events on the launch
stream during the timed steps.

The CPU baseline times the oracle (FP64 numpy/scipy restatement) on the box's
host: residual + tangent assembly of a 1/16 element slab (scaled x16) plus a
fixed number of scipy Jacobi-PCG iterations on the full-size reduced matrix
(scaled to the plain method's iteration count).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 2048]
                [--methods plain,mixed,learned,learned_mixed,learned_mixed_fp32_elements]
                [--headline learned_mixed] [--train-steps 1500] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-solution at 1e-8 FP64 residual: learned+mixed vs plain; assembly GB/s"
HBM_PEAK = 8.0e12
# HBM bytes per executing launch at n = 2048, from `rocprofv3 --pmc FETCH_SIZE` and
# `--pmc WRITE_SIZE` runs of `python bench.py --probe`, scaled by the counters'
# reading of a 512 MiB device copy (FETCH_SIZE x2.000, WRITE_SIZE x1.000).
PMC_TRAFFIC_BYTES = {("k_pcg_spmv<double>", 2048): 660.2e6, ("k_pcg_spmv<float>", 2048): 408.5e6,
                     ("k_conv3x3_ws<16, 16, 4, 0>", 2048): 292.1e6}


def build_problem(n):
    from autopdex_amd import backend, geometry, mesher, models, seeder, spaces
    pts = [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]
    coords, elems = mesher.structured_mesh((n, n), pts, "quad")
    mask = geometry.in_sdfs(coords, lambda x: geometry.psdf_polygon(x, pts))
    pot = models.mixed_reference_domain_potential(
        models.poisson_source_integrand(), {"phi": spaces.fem_iso_line_quad_brick},
        *seeder.gauss_legendre_nd(2, 2), "phi")
    static = {"assembling mode": ("user potential",), "solution structure": ("nodal imposition",),
              "model": (pot,), "solver type": "newton", "solver backend": "hip", "solver": "cg",
              "type of preconditioner": "jacobi", "verbose": -1}
    settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": np.zeros(mask.shape)}
    dofs = np.zeros(mask.shape)
    t0 = time.perf_counter()
    prob = backend.get_problem(dofs, settings, static)
    plan_s = time.perf_counter() - t0
    return prob, settings, static, coords, elems, mask, plan_s


def options(static, method):
    from autopdex_amd import solver
    s = dict(static)
    s["newton tolerance mode"] = "relative"
    s["linear forcing"] = True
    s["linear maxiter"] = 200000
    if method == "plain":
        s["linear tolerance"] = 1e-10
    elif method == "mixed":
        s["hip precision"] = "mixed"
        s["reliable delta"] = float(os.environ.get("APX_RELIABLE_DELTA", "0.1"))
    elif method in ("learned", "learned_mixed", "learned_mixed_fp32_elements"):
        s["type of preconditioner"] = "learned"
        s["linear tolerance"] = 1e-10
        s["linear maxiter"] = 3000
        if method != "learned":
            s["hip precision"] = "mixed"
            s["reliable delta"] = float(os.environ.get("APX_LEARNED_DELTA", "1e-2"))
        if method == "learned_mixed_fp32_elements":
            s["hip tangent arithmetic"] = "fp32"
    else:
        raise ValueError(method)
    return solver.newton_options(s, newton_tol=1e-10, maxiter=100)


def run(prob, opts, u, dv, torch):
    from autopdex_amd import _native as N
    from autopdex_amd.backend import current_stream
    u.zero_()
    st = N.NewtonStats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    N.call("apx_newton", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(dv.data_ptr()), C.byref(opts),
           C.byref(st), current_stream())
    torch.cuda.synchronize()
    return time.perf_counter() - t0, st


def timer(prob, name):
    from autopdex_amd import _native as N
    ms, cnt, by = C.c_double(), C.c_int64(), C.c_int64()
    N.call("apx_get_timing", prob.plan, name.encode(), C.byref(ms), C.byref(cnt), C.byref(by))
    return ms.value, cnt.value, by.value


def cpu_baseline(prob, coords, elems, mask, plain_iters, vals64_cpu):
    """Oracle timed on the host: bounded sample, scaled to the full workload."""
    import scipy.sparse as sp
    from oracle import assembly as OA
    from oracle import elements as OE
    from oracle import newton as ON
    free = ~mask.ravel()
    u = np.zeros(mask.size)
    E = elems.shape[0]
    slab = elems[: E // 16]
    sets = [(OE.PoissonSource(), slab, None)]
    t0 = time.perf_counter()
    OA.residual(sets, coords, u, 1)
    t_res = (time.perf_counter() - t0) * 16
    t0 = time.perf_counter()
    OA.reduced_tangent(sets, coords, u, 1, free)
    t_tan = (time.perf_counter() - t0) * 16
    indptr, indices = prob.csr_pattern()
    A = sp.csr_matrix((vals64_cpu, indices, indptr), shape=(prob.n_free, prob.n_free))
    b = -OA.residual([(OE.PoissonSource(), elems[: E // 64], None)], coords, u, 1)[free]
    k_sample = 60
    d = A.diagonal()
    t0 = time.perf_counter()
    ON.pcg(A, b, lambda v: v / d, tol=0.0, maxiter=k_sample)
    t_it = (time.perf_counter() - t0) / k_sample
    total = 2 * t_res + t_tan + t_it * plain_iters
    return {"value": total, "unit": "s", "cores": 1, "kind": "port",
            "sample": (f"oracle FP64 residual + tangent/CSR assembly of a 1/16 element slab (x16: residual "
                       f"{t_res:.2f} s, tangent {t_tan:.2f} s) + {k_sample} scipy Jacobi-PCG iterations on the full "
                       f"{prob.n_free}-row reduced matrix ({t_it * 1e3:.1f} ms/it) x {plain_iters} iterations of the "
                       f"plain method")}


def kernel_table(prob):
    out = {}
    for name in ("pcg_spmv_f64", "pcg_update_f64", "pcg_spmv_f32", "pcg_update_f32", "residual", "tangent_f64",
                 "tangent_f64_to_f32", "tangent_f32", "fcg_spmv_f64", "fcg_spmv_f32", "fcg_precond",
                 "unet_conv_d0", "unet_forward"):
        ms, cnt, by = timer(prob, name)
        if cnt:
            out[name] = {"avg_ms": ms, "count": cnt, "bytes": by, "GBps": by / (ms * 1e-3) / 1e9 if by else None,
                         "frac_hbm": by / (ms * 1e-3) / HBM_PEAK if by else None}
    return out


def probe(args):
    """Bounded kernel probe for hardware-counter runs: 3 residual and 3 tangent
    assemblies, 20 fixed PCG iterations per precision, 3 network applications,
    and a 512 MiB device copy (the calibration kernel for FETCH_SIZE / WRITE_SIZE)."""
    import torch
    from autopdex_amd import _native as N
    from autopdex_amd.backend import current_stream
    prob, settings, static, coords, elems, mask, plan_s = build_problem(args.n)
    dev, st = prob.device, current_stream()
    u = torch.zeros(prob.n_dofs, dtype=torch.float64, device=dev)
    r = torch.empty_like(u)
    for _ in range(3):
        N.call("apx_assemble_residual", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(r.data_ptr()), st)
    r2f = torch.from_numpy(prob.free_dofs().astype(np.int64)).to(dev)
    for prec, dt in ((N.APX_FP64, torch.float64), (N.APX_FP32, torch.float32)):
        vals = torch.empty(prob.nnz, dtype=dt, device=dev)
        for _ in range(3):
            N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals.data_ptr()), prec, st)
        b = (-r[r2f]).to(dt).contiguous()
        x = torch.empty_like(b)
        ks = N.KrylovStats()
        N.call("apx_pcg", prob.plan, C.c_void_p(vals.data_ptr()), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
               prec, N.APX_PRECOND_JACOBI, 0.0, 0.0, 20, C.byref(ks), st)
    # three preconditioner applications of an (untrained) network of the benchmark's shape
    from autopdex_amd import learned
    H = args.n + 1
    L = learned.default_levels(H, H)
    net = learned.UNetPreconditioner(H, H, L, ((16, 32) + (64,) * 10)[:L], 1, 1, relu=False, batch=1)
    net.attach(prob)
    rb = (-r[r2f]).contiguous()
    for _ in range(3):
        net.apply(prob, rb)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "ok", "copy_bytes": src.numel() * 8, "n_free": prob.n_free, "nnz": prob.nnz}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--methods", default="plain,mixed,learned,learned_mixed,learned_mixed_fp32_elements")
    ap.add_argument("--headline", default="learned_mixed")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--seed", type=int, default=0, help="seed of the preconditioner's training data")
    ap.add_argument("--save-net", default=None, help="write the trained network (.npz)")
    ap.add_argument("--load-net", default=None, help="use a network written by --save-net instead of training")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true", help="bounded kernel probe for counter runs")
    args = ap.parse_args()
    if args.probe:
        return probe(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the assemble-and-solve path runs on a single MI355X "
              "(multi-GPU sharding is a design, not implemented). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from autopdex_amd import _native as N
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    methods = args.methods.split(",")
    assert "plain" in methods and args.headline in methods
    prob, settings, static, coords, elems, mask, plan_s = build_problem(args.n)
    dev = prob.device
    u = torch.zeros(prob.n_dofs, dtype=torch.float64, device=dev)
    dv = torch.zeros(prob.n_dofs, dtype=torch.float64, device=dev)
    opts = {m: options(static, m) for m in methods}
    print(f"[bench] plan built in {plan_s:.2f} s: n_free={prob.n_free} nnz={prob.nnz}", file=sys.stderr, flush=True)
    learned_info = None
    if any(m.startswith("learned") for m in methods):
        from autopdex_amd import learned
        if args.load_net:
            t0 = time.perf_counter()
            net = learned.UNetPreconditioner.load(args.load_net)
            net.attach(prob)
            learned_info = {"loaded": args.load_net, "load_seconds": time.perf_counter() - t0, "levels": net.levels,
                            "channels": list(net.channels), "n_params": int(net.n_params)}
        else:
            net, learned_info = learned.fit(prob, args.n + 1, args.n + 1, steps=args.train_steps, seed=args.seed,
                                            log=lambda m: print(m, file=sys.stderr, flush=True))
        if args.save_net:
            net.save(args.save_net)
        print(f"[bench] learned preconditioner: {learned_info}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        for m in methods:
            run(prob, opts[m], u, dv, torch)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    N.call("apx_set_timing", prob.plan, 1)
    times = {m: [] for m in methods}
    stats, errs = {}, {}
    for k in range(args.steps):
        u_plain = None
        for m in methods:
            t, st = run(prob, opts[m], u, dv, torch)
            times[m].append(t)
            stats[m] = st
            if m == "plain":
                u_plain = u.clone()
            else:
                errs[m] = float(torch.linalg.norm(u - u_plain) / torch.linalg.norm(u_plain))
        print(f"[bench] step {k + 1}/{args.steps}: " + ", ".join(
            f"{m} {times[m][-1]:.3f} s ({stats[m].krylov_iterations} it)" for m in methods),
            file=sys.stderr, flush=True)
    T = {m: float(np.median(times[m])) for m in methods}
    kernels = kernel_table(prob)
    H = args.headline
    if H.startswith("learned"):
        # the heaviest single kernel of flexible CG with the network: the level-0 16 -> 16 convolution
        dom = kernels["unet_conv_d0"]
        dom_name = "k_conv3x3_ws<16, 16, 4, 0>"
    else:
        prec_tag = "f64" if H == "plain" else "f32"
        dom = kernels[f"pcg_spmv_{prec_tag}"]
        dom_name = f"k_pcg_spmv<{'double' if prec_tag == 'f64' else 'float'}>"
    res = kernels.get("residual", {})
    dtypes = {"plain": "f64", "mixed": "f32+f64", "learned": "bf16+f64", "learned_mixed": "bf16+f32+f64",
              "learned_mixed_fp32_elements": "bf16+f32+f64"}
    out = {
        "metric": METRIC, "value": T[H], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": T[H] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": dtypes[H], "data": "synthetic",
        "config": {"workload": f"2D Poisson Q1 {args.n}x{args.n} (short_example source, psdf Dirichlet), "
                               f"Newton to 1e-10 FP64 relative residual, method '{H}'",
                   "n_dofs": int(prob.n_dofs), "n_free": int(prob.n_free), "nnz": int(prob.nnz),
                   "parallelism": "single GPU"},
        "methods": {m: {"time_s": T[m], "speedup_vs_plain": T["plain"] / T[m],
                        "krylov_iterations": stats[m].krylov_iterations, "newton_iterations": stats[m].iterations,
                        "relative_residual": stats[m].residual_norm / stats[m].initial_residual,
                        "rel_error_vs_plain": errs.get(m, 0.0)} for m in methods},
        "plan_build_s": plan_s,
        "learned_setup": learned_info,
        "assembly": {"residual_GBps": res.get("GBps"), "residual_ms": res.get("avg_ms"),
                     "residual_frac_hbm": res.get("frac_hbm"),
                     "dofs_assembled_per_s": prob.n_dofs / (res["avg_ms"] * 1e-3) if res else None,
                     "tangent_f64_ms": kernels.get("tangent_f64", {}).get("avg_ms"),
                     "tangent_f64_to_f32_ms": kernels.get("tangent_f64_to_f32", {}).get("avg_ms"),
                     "tangent_f32_ms": kernels.get("tangent_f32", {}).get("avg_ms")},
        "kernels": kernels,
        "network": ({"forward_ms": kernels["unet_forward"]["avg_ms"], "forward_GBps": kernels["unet_forward"]["GBps"],
                     "forward_bytes": kernels["unet_forward"]["bytes"]} if "unet_forward" in kernels else None),
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": dom["GBps"], "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": dom["frac_hbm"], "traffic": PMC_TRAFFIC_BYTES.get((dom_name, args.n)),
                     "algorithmic_bytes_per_launch": dom["bytes"], "avg_ms": dom["avg_ms"]},
    }
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        run(prob, opts["plain"], u, dv, torch)
        vals = torch.empty(prob.nnz, dtype=torch.float64, device=dev)
        N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals.data_ptr()),
               N.APX_FP64, None)
        torch.cuda.synchronize()
        out["cpu_baseline"] = cpu_baseline(prob, coords, elems, mask, stats["plain"].krylov_iterations,
                                           vals.cpu().numpy())
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
