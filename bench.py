"""Benchmark: time-to-solution of the 2048x2048 Q1 Poisson problem (4.2M DOFs).

Workload (one "step"): a full Newton solve on one GPU to a 1e-10 FP64
relative residual ||r_free(u)||/||r_free(u0)||, inputs resident in HBM, for
each method:
  plain         = FP64 assembly + FP64 Jacobi-PCG (the drop-in method)
  mixed         = FP64 tangent rounded once to FP32 + Jacobi-PCG with FP64 vectors
                  on that matrix, re-anchored to the FP64 residual (reliable
                  updates)
  mixed_fp32    = FP32 element arithmetic for the tangent + an all-FP32 Jacobi-PCG
                  (FP32 vectors) under the same FP64 refinement
  learned       = FP64 tangent + flexible CG with the U-Net preconditioner
                  (bf16 convolutions on MFMA, FP32 accumulation)
  learned_mixed = FP64 tangent rounded once to FP32 + FP32-matrix flexible CG
                  with the U-Net, each inner solve re-anchored by the FP64
                  Newton residual
  learned_mixed_fp32_elements = learned_mixed with the tangent's element
                  arithmetic in FP32 (FP32 accumulation) instead of rounding
  learned_family_mixed = learned_mixed with a network trained on the
                  variable-coefficient family (--family-net, written by
                  --workload c4 --n 2048 --save-net), applied with k = 1: a
                  network that never saw this operator (not in the default list)
The network is trained on the GPU before the timed steps (learned.fit: pairs
from plain Jacobi-PCG solves of random right-hand sides, then also from the
flexible-CG iterates of the network being trained, A-norm loss); its
setup time is reported in `learned_setup` (with the time-to-solution including
it and the number of solves after which the training has paid for itself)
and is not part of any step.
The headline `value` is the time-to-solution (seconds, lower is better) of
the --headline method (default learned_mixed: learned + mixed precision),
timed over --steps steps; every other method is timed over --aux-steps steps
and reported beside it (time, Krylov iterations, residual, error against
plain).  The timed steps run without timers.  The roofline is that of the
kernel with the largest total device time in the headline method's solve: a
warmup solve with every network kernel timed per launch picks it, and
untimed solves after the timed steps record that kernel's launches with HIP
events that take each dispatch's own start and end timestamps
(hipExtLaunchKernel), as rocprofv3's kernel trace does.

The CPU baseline times the oracle (FP64 numpy/scipy restatement) on the box's
host: residual + tangent assembly of a 1/16 element slab (scaled x16) plus a
fixed number of scipy Jacobi-PCG iterations on the full-size reduced matrix
(scaled to the plain method's iteration count).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--aux-steps 3] [--n 2048]
                [--methods plain,mixed,mixed_fp32,learned,learned_mixed,learned_mixed_fp32_elements]
                [--headline learned_mixed] [--train-steps 1500] [--no-cpu]
python bench.py --workload c4 [--fields 2000] [--heldout 200] [--train-steps 8000]
                (the learned preconditioner on held-out variable-coefficient fields at 1024^2)
python bench.py --probe [--workload c3|train]
                (bounded kernel probes for rocprofv3 runs: 2048^2 Poisson, 64^3 Hex8, 60 training
                steps of the 2048^2 network)
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
# The convolution's figure is the average over the launches of one network application (two at
# 513 x 513, four at 257 x 257), the mix its timer averages over.
PMC_TRAFFIC_BYTES = {("k_pcg_spmv<double, double, 4, short>", 2048): 584.8e6,
                     ("k_pcg_spmv<float, double, 4, short>", 2048): 434.1e6,
                     ("k_pcg_spmv<float, float, 4, short>", 2048): 333.1e6, ("k_fcg_spmv<float, short>", 2048): 467.7e6,
                     ("k_conv3x3<64, 64, false, 4, 2, false>", 2048): 37.33e6,
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
    elif method in ("mixed", "mixed_fp32"):
        s["hip precision"] = "mixed"
        s["reliable delta"] = float(os.environ.get("APX_RELIABLE_DELTA", "0.1"))
        if method == "mixed_fp32":
            s["hip tangent arithmetic"] = "fp32"
            s["hip krylov vectors"] = "fp32"
    elif method in ("learned", "learned_mixed", "learned_mixed_fp32_elements", "learned_family_mixed"):
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


def all_timers(prob):
    """{name: (avg_ms, launches, bytes per launch, flops per launch)} of every plan timer."""
    from autopdex_amd import _native as N
    out, i = {}, 0
    buf = C.create_string_buffer(256)
    ms, cnt, by, fl = C.c_double(), C.c_int64(), C.c_int64(), C.c_int64()
    while N.lib().apx_get_timer(prob.plan, i, buf, 256, C.byref(ms), C.byref(cnt), C.byref(by), C.byref(fl)) == 0:
        out[buf.value.decode()] = (ms.value, cnt.value, by.value, fl.value)
        i += 1
    return out


# single-kernel solver timers and the kernel each one times (rocprofv3 names)
# (the SpMVs of the 2048^2 grid read int16 column offsets: template argument "short")
SOLVER_KERNELS = {"fcg_spmv_f32": "k_fcg_spmv<float, short>", "fcg_spmv_f64": "k_fcg_spmv<double, short>",
                  "pcg_spmv_f64": "k_pcg_spmv<double, double, 4, short>",
                  "pcg_spmv_f32": "k_pcg_spmv<float, float, 4, short>",
                  "pcg_spmv_f32m_f64v": "k_pcg_spmv<float, double, 4, short>", "pcg_update_f64": "k_pcg_update<double>",
                  "pcg_update_f32": "k_pcg_update<float>"}


def dominant_kernel(timers):
    """The single kernel with the largest total device time among the per-launch network kernel
    timers (names "k_...") and the single-kernel solver timers."""
    best = None
    for name, (ms, cnt, by, fl) in timers.items():
        if not (name.startswith("k_") or name in SOLVER_KERNELS) or not cnt:
            continue
        if best is None or ms * cnt > best[1]:
            best = (name, ms * cnt)
    return best[0] if best else None


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


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
    return {"value": total, "unit": "s", "cores": 1, "kind": "port", "extrapolated": True, "host": host_info(),
            "sample": (f"oracle FP64 residual + tangent/CSR assembly of a 1/16 element slab (x16: residual "
                       f"{t_res:.2f} s, tangent {t_tan:.2f} s) + {k_sample} scipy Jacobi-PCG iterations on the full "
                       f"{prob.n_free}-row reduced matrix ({t_it * 1e3:.1f} ms/it) x {plain_iters} iterations of the "
                       f"plain method; single-threaded numpy/scipy (1 core of the host)")}


def kernel_table(prob):
    out = {}
    for name, (ms, cnt, by, fl) in all_timers(prob).items():
        if not cnt or name.endswith("_call"):
            continue
        out[name] = {"avg_ms": ms, "count": cnt, "total_ms": ms * cnt, "bytes": by,
                     "GBps": by / (ms * 1e-3) / 1e9 if by else None,
                     "frac_hbm": by / (ms * 1e-3) / HBM_PEAK if by else None}
        if fl:
            out[name]["TFLOPs"] = fl / (ms * 1e-3) / 1e12
    return out


MFMA_PEAK = 2.5e15  # bf16 dense


def roofline(name, ms, by, fl, traffic):
    """Roofline of one kernel from its algorithmic bytes and flops per launch: attainable =
    min(bf16 MFMA peak, flops/bytes x HBM peak); reported in the bound's unit."""
    gbps = by / (ms * 1e-3) / 1e9
    r = {"kernel": name, "avg_ms": ms, "algorithmic_bytes_per_launch": by, "traffic": traffic}
    if fl:
        ai = fl / by
        tflops = fl / (ms * 1e-3) / 1e12
        if ai * HBM_PEAK < MFMA_PEAK:
            r.update(bound="hbm", achieved=gbps, peak=HBM_PEAK / 1e9, unit="GB/s", frac=gbps * 1e9 / HBM_PEAK)
        else:
            r.update(bound="mfma", achieved=tflops, peak=MFMA_PEAK / 1e12, unit="TFLOP/s",
                     frac=tflops * 1e12 / MFMA_PEAK)
        r.update(flops_per_launch=fl, achieved_TFLOPs=tflops, arithmetic_intensity=ai,
                 attainable_TFLOPs=min(MFMA_PEAK, ai * HBM_PEAK) / 1e12,
                 frac_of_attainable=tflops * 1e12 / min(MFMA_PEAK, ai * HBM_PEAK))
    else:
        r.update(bound="hbm", achieved=gbps, peak=HBM_PEAK / 1e9, unit="GB/s", frac=gbps * 1e9 / HBM_PEAK)
    return r


def probe(args):
    """Bounded kernel probe for hardware-counter runs: 3 residual and 3 tangent
    assemblies, 20 fixed PCG iterations per precision (FP64, FP32, and the FP32
    matrix with FP64 vectors), 3 network applications, 5 flexible-CG iterations on
    the FP32 matrix, and a 512 MiB device copy (the calibration kernel for
    FETCH_SIZE / WRITE_SIZE)."""
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
    vals32 = torch.empty(prob.nnz, dtype=torch.float32, device=dev)
    N.call("apx_assemble_tangent", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(vals32.data_ptr()), N.APX_FP32, st)
    learned.fcg(prob, vals32, rb, tol=0.0, maxiter=5, precision="fp32")
    # 20 iterations of the mixed method's PCG (FP32 matrix, FP64 vectors): one Newton step
    from autopdex_amd import solver
    s = dict(static, **{"hip precision": "mixed", "reliable delta": 0.0, "linear maxiter": 20,
                        "linear tolerance": 0.0})
    opts = solver.newton_options(s, newton_tol=0.0, maxiter=1)
    dv = torch.zeros_like(u)
    u.zero_()
    stt = N.NewtonStats()
    N.call("apx_newton", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(dv.data_ptr()), C.byref(opts),
           C.byref(stt), st)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "ok", "copy_bytes": src.numel() * 8, "n_free": prob.n_free, "nnz": prob.nnz}))


def build_cube_problem(n, lam=10.0, mu=1.0, disp=0.02):
    """Compressible Neo-Hooke on the unit cube, n^3 Hex8: x = 0 clamped, u_x = disp on x = 1."""
    from autopdex_amd import backend, geometry, mesher, models, seeder, spaces, utility
    cube = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]
    coords, elems = mesher.structured_mesh((n, n, n), cube, "brick")
    left = geometry.in_planes(coords, [0, 0, 0], [1, 0, 0])
    right = geometry.in_planes(coords, [1, 0, 0], [1, 0, 0])
    mask = utility.dof_select(left, [True, True, True]) | utility.dof_select(right, [True, False, False])
    values = np.zeros(mask.shape)
    values[right, 0] = disp
    pot = models.mixed_reference_domain_potential(
        models.hyperelastic_integrand(models.neo_hooke, lam, mu), {"u": spaces.fem_iso_line_quad_brick},
        *seeder.gauss_legendre_nd(3, 2), "u")
    static = {"assembling mode": ("user potential",), "solution structure": ("nodal imposition",),
              "model": (pot,), "solver type": "newton", "solver backend": "hip", "solver": "cg",
              "type of preconditioner": "jacobi", "verbose": -1}
    settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": values}
    return backend.get_problem(np.zeros(mask.shape), settings, static), static, values


def probe_train(args):
    """Bounded training probe for kernel traces: the benchmark's network (batch 4) on 4 pairs
    (white-noise right-hand sides, plain Jacobi-PCG to 1e-3), 5 untimed and 60 timed
    learned.train steps at a fixed learning rate; prints the wall time per step."""
    import torch
    from autopdex_amd import learned
    prob = build_problem(args.n)[0]
    vals = learned.tangent(prob)
    gen = torch.Generator(device=prob.device)
    gen.manual_seed(args.seed)
    pairs = []
    for _ in range(4):
        b = torch.randn(prob.n_free, generator=gen, device=prob.device, dtype=torch.float64)
        pairs += learned.solver_pairs(prob, vals, b, checkpoints=(0,), tol=1e-3)
    H = args.n + 1
    L = learned.default_levels(H, H)
    net = learned.UNetPreconditioner(H, H, L, ((16, 32) + (64,) * 10)[:L], 1, 1, relu=False, batch=4, seed=1)
    net.set_params(net.initial_params(1, up_scale=0.05))
    learned.train(net, prob, pairs, vals, steps=5, lr=1e-5, lr_final=1e-5, log=None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    learned.train(net, prob, pairs, vals, steps=60, lr=1e-5, lr_final=1e-5, log=None)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / 60 * 1e3
    print(json.dumps({"probe": "train", "n": args.n, "batch": 4, "steps": 60, "ms_per_step": ms}))


def probe_cube(args):
    """The 64^3 Hex8 Neo-Hooke configuration's kernels (colour-pass assembly, 3-DOF SpMV), for
    kernel-trace and counter runs: one Newton iteration of 20 Jacobi-PCG iterations from u = 0
    with the first load step's Dirichlet values (imposed by apx_newton before its first
    residual), FP64, then the mixed method (FP32 element arithmetic, FP64 Krylov vectors).
    Prints the plan's kernel timers (HIP events) with their algorithmic bytes."""
    import torch
    from autopdex_amd import _native as N
    from autopdex_amd import solver
    n = args.n
    t0 = time.perf_counter()
    prob, static, values = build_cube_problem(n)
    plan_s = time.perf_counter() - t0
    dvals = torch.from_numpy(values.ravel().copy()).to(prob.device)  # Dirichlet values, imposed by Newton
    u = torch.zeros_like(dvals)

    def newton(opts):
        from autopdex_amd.backend import current_stream
        u.zero_()
        st = N.NewtonStats()
        N.call("apx_newton", prob.plan, C.c_void_p(u.data_ptr()), C.c_void_p(dvals.data_ptr()), C.byref(opts),
               C.byref(st), current_stream())
        torch.cuda.synchronize()
        return st
    out = {"probe": "cube", "n": n, "n_dofs": prob.n_dofs, "n_free": prob.n_free, "nnz": prob.nnz,
           "plan_build_s": plan_s, "methods": {}}
    for name, extra in (("fp64", {}), ("mixed_fp32_elements", {"hip precision": "mixed", "hip tangent arithmetic": "fp32",
                                                                 "reliable delta": 0.0})):
        s = dict(static, **extra, **{"linear maxiter": 20, "linear tolerance": 0.0})
        opts = solver.newton_options(s, newton_tol=0.0, maxiter=1)
        N.call("apx_set_timing", prob.plan, 0)
        newton(opts)  # warm-up
        N.call("apx_set_timing", prob.plan, 1)
        for _ in range(3):
            st = newton(opts)
        out["methods"][name] = kernel_table(prob)
        out["methods"][name]["krylov_iterations_per_call"] = st.krylov_iterations
        N.call("apx_set_timing", prob.plan, 0)  # level 0 clears the timers
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=prob.device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(out))


def build_vark_problem(n):
    """-div(exp(g) grad u) = f on the unit square, n x n Q1, psdf Dirichlet mask (corners free)."""
    from autopdex_amd import backend, geometry, mesher, models, seeder, spaces
    pts = [[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]
    coords, elems = mesher.structured_mesh((n, n), pts, "quad")
    mask = geometry.in_sdfs(coords, lambda x: geometry.psdf_polygon(x, pts))
    pot = models.mixed_reference_domain_potential(
        models.poisson_variable_coefficient_integrand("log coefficient", "source"),
        {"u": spaces.fem_iso_line_quad_brick}, *seeder.gauss_legendre_nd(2, 2), "u")
    static = {"assembling mode": ("user potential",), "solution structure": ("nodal imposition",),
              "model": (pot,), "solver type": "newton", "solver backend": "hip", "solver": "cg",
              "type of preconditioner": "jacobi", "verbose": -1}
    zeros = np.zeros(coords.shape[0])
    settings = {"connectivity": (elems,), "dirichlet dofs": mask, "node coordinates": coords,
                "dirichlet conditions": np.zeros(mask.shape), "log coefficient": zeros, "source": zeros}
    return backend.get_problem(np.zeros(mask.shape), settings, static)


def c4_workload(args):
    """The learned preconditioner on a family of variable-coefficient problems at 1024 x 1024:
    k = exp(g), g a squared-exponential Gaussian random field (length 0.05, sigma 1) sampled on
    the nodes by circulant embedding, f ~ N(0, 1) per node; train seeds 0 .. F-1, held-out
    seeds 10000 .. 10000 + M - 1.  One network (g as input channel) is trained on two plain
    Jacobi-PCG (r, A^-1 r) pairs per training field; every held-out field is then solved from
    u = 0 by flexible CG with the network and by FP64 Jacobi-PCG, both to a 1e-8 relative
    residual (the FP64 residual b - A x of each result is checked on the device).  A flexible-CG
    solve that stagnates is finished by Jacobi-PCG from its iterate, as the Newton driver does,
    and both parts count for the learned method."""
    import torch
    from autopdex_amd import _native as N
    from autopdex_amd import learned
    from autopdex_amd.backend import current_stream
    n = args.n
    H = n + 1
    t0 = time.perf_counter()
    prob = build_vark_problem(n)
    plan_s = time.perf_counter() - t0
    log = lambda m: print(m, file=sys.stderr, flush=True)
    log(f"[c4] plan {plan_s:.2f} s, n_free {prob.n_free}, nnz {prob.nnz}")

    def field(seed):
        g = learned.random_field(H, H, 0.05, 1.0, seed=seed).contiguous()
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed + 1_000_000)
        return g, torch.randn(H * H, generator=gen, device="cuda", dtype=torch.float64)

    net, info = learned.fit_field_family(prob, H, H, field, list(range(args.fields)), steps=args.train_steps,
                                         validation_seed=9999, log=log)
    log(f"[c4] trained: {info}")
    if args.save_net:
        net.save(args.save_net)
    r2f = torch.from_numpy(prob.free_dofs().astype(np.int64)).cuda()
    u0 = torch.zeros(prob.n_dofs, dtype=torch.float64, device="cuda")
    r = torch.empty_like(u0)
    rows = []
    for k, s in enumerate(range(10000, 10000 + args.heldout)):
        g, f = field(s)
        prob.bind_nodal_device(0, g, f)
        vals = learned.tangent(prob)
        N.call("apx_assemble_residual", prob.plan, C.c_void_p(u0.data_ptr()), C.c_void_p(r.data_ptr()),
               current_stream())
        b = (-r[r2f]).contiguous()
        net.attach(prob, g)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, kf = learned.fcg(prob, vals, b, tol=1e-8, maxiter=3000)
        fell_back = not kf.converged
        its_f = kf.iterations
        if fell_back:
            # as the Newton driver does: keep flexible CG's progress, finish with Jacobi-PCG (its
            # time and iterations count for the learned method)
            b2 = b - learned.spmv(prob, vals, x)
            tol2 = 1e-8 * float(torch.linalg.norm(b)) / max(float(torch.linalg.norm(b2)), 1e-300)
            x2, k2 = learned.pcg(prob, vals, b2, tol=min(tol2, 1.0))
            x = x + x2
            its_f += k2.iterations
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        xp, kp = learned.pcg(prob, vals, b, tol=1e-8)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        res_f = float(torch.linalg.norm(b - learned.spmv(prob, vals, x)) / torch.linalg.norm(b))
        res_p = float(torch.linalg.norm(b - learned.spmv(prob, vals, xp)) / torch.linalg.norm(b))
        rows.append((its_f, kp.iterations, t2 - t1, t3 - t2, res_f, res_p, not fell_back))
        if (k + 1) % 20 == 0:
            log(f"[c4] held-out {k + 1}/{args.heldout}: FCG {its_f} it {t2 - t1:.3f} s, "
                f"PCG {kp.iterations} it {t3 - t2:.3f} s")
    a = np.array([r_[:6] for r_ in rows], dtype=np.float64)
    out = {
        "metric": "C4 held-out fields: time to a 1e-8 relative residual, flexible CG with the learned "
                  "preconditioner vs FP64 Jacobi-PCG",
        "value": float(np.median(a[:, 3] / a[:, 2])), "unit": "x (median speedup per field)",
        "higher_is_better": True, "n_gpus": 1, "data": "synthetic",
        "config": {"workload": f"variable-coefficient Poisson {n}x{n} Q1, k = exp(g), g squared-exponential GRF "
                               f"(length 0.05, sigma 1), f ~ N(0,1)", "n_free": int(prob.n_free),
                   "nnz": int(prob.nnz), "train_fields": args.fields, "heldout_fields": args.heldout},
        "training": info,
        "heldout": {"fcg_iterations_median": float(np.median(a[:, 0])), "fcg_iterations_max": float(a[:, 0].max()),
                    "pcg_iterations_median": float(np.median(a[:, 1])),
                    "fcg_seconds_median": float(np.median(a[:, 2])), "pcg_seconds_median": float(np.median(a[:, 3])),
                    "fcg_seconds_total": float(a[:, 2].sum()), "pcg_seconds_total": float(a[:, 3].sum()),
                    "fcg_converged": int(sum(r_[6] for r_ in rows)),
                    "fallbacks_to_jacobi_pcg": int(sum(not r_[6] for r_ in rows)),
                    "fcg_true_relative_residual_max": float(a[:, 4].max()),
                    "pcg_true_relative_residual_max": float(a[:, 5].max())},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--aux-steps", type=int, default=3, help="timed steps of the methods other than the headline")
    ap.add_argument("--n", type=int, default=None,
                    help="elements per direction (default: 2048; 1024 for --workload c4; 64 for the c3 probe)")
    ap.add_argument("--methods", default="plain,mixed,mixed_fp32,learned,learned_mixed,learned_mixed_fp32_elements")
    ap.add_argument("--headline", default="learned_mixed")
    ap.add_argument("--train-steps", type=int, default=None,
                    help="training steps of the network (default: 1500; 8000 for --workload c4)")
    ap.add_argument("--seed", type=int, default=0, help="seed of the preconditioner's training data")
    ap.add_argument("--save-net", default=None, help="write the trained network (.npz)")
    ap.add_argument("--load-net", default=None, help="use a network written by --save-net instead of training")
    ap.add_argument("--family-net", default=None,
                    help="method learned_family_mixed: a variable-coefficient family network (bench.py --workload c4 "
                         "--save-net) applied with k = 1")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true", help="bounded kernel probe for counter runs")
    ap.add_argument("--workload", default="poisson", choices=("poisson", "c3", "c4", "train"),
                    help="poisson: the 2048^2 time-to-solution benchmark; c4: the learned preconditioner on "
                         "held-out variable-coefficient fields at 1024^2; c3 (with --probe): the 64^3 Hex8 "
                         "Neo-Hooke kernels; train (with --probe): 60 training steps of the 2048^2 network")
    ap.add_argument("--fields", type=int, default=2000, help="c4: training fields")
    ap.add_argument("--heldout", type=int, default=200, help="c4: held-out fields")
    args = ap.parse_args()
    if args.n is None:
        args.n = {"c3": 64, "c4": 1024}.get(args.workload, 2048)
    if args.train_steps is None:
        args.train_steps = 8000 if args.workload == "c4" else 1500
    if args.probe:
        return {"c3": probe_cube, "train": probe_train}.get(args.workload, probe)(args)
    if args.workload == "train":
        raise SystemExit("--workload train is a probe: use it with --probe")
    if args.workload == "c4":
        return c4_workload(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the assemble-and-solve path runs on a single MI355X "
              "(multi-GPU sharding is a design, not implemented). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from autopdex_amd import _native as N
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    methods = args.methods.split(",")
    H = args.headline
    assert "plain" in methods and H in methods
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
    # a network trained on the variable-coefficient family (never on this operator), applied with
    # the log-coefficient channel g = 0, i.e. k = 1
    family = None
    if "learned_family_mixed" in methods:
        from autopdex_amd import learned
        if not args.family_net:
            raise SystemExit("method learned_family_mixed needs --family-net (bench.py --workload c4 --save-net)")
        fnet = learned.UNetPreconditioner.load(args.family_net)
        family = (fnet, torch.zeros(prob.n_dofs, dtype=torch.float64, device=dev))
        learned_info = dict(learned_info or {}, family_net=args.family_net)

    def solve(m):
        if m.startswith("learned"):
            if m == "learned_family_mixed":
                family[0].attach(prob, family[1])
            else:
                net.attach(prob)
        return run(prob, opts[m], u, dv, torch)

    for w in range(args.warmup):
        for m in methods:
            if m == H or w == 0:
                solve(m)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    # one more untimed solve of the headline method with every network kernel timed per launch:
    # the kernel with the largest total device time is the one the roofline reports
    N.call("apx_set_timing", prob.plan, 2)
    solve(H)
    profile = all_timers(prob)
    dom = dominant_kernel(profile)
    N.call("apx_set_timing", prob.plan, 0)
    N.call("apx_set_timing_filter", prob.plan, (dom if dom and dom.startswith("k_") else "").encode())
    print(f"[bench] dominant kernel of '{H}': {dom}", file=sys.stderr, flush=True)
    # timed steps: no timers (no events between the launches)
    times = {m: [] for m in methods}
    stats, errs, u_plain = {}, {}, None
    for k in range(args.steps):
        for m in methods:
            if m != H and k >= args.aux_steps:
                continue
            t, st = solve(m)
            times[m].append(t)
            stats[m] = st
            if m == "plain" and u_plain is None:
                u_plain = u.clone()
            elif m != "plain" and u_plain is not None:
                errs[m] = float(torch.linalg.norm(u - u_plain) / torch.linalg.norm(u_plain))
        print(f"[bench] step {k + 1}/{args.steps}: " + ", ".join(
            f"{m} {times[m][-1]:.3f} s ({stats[m].krylov_iterations} it)" for m in methods if len(times[m]) == k + 1),
            file=sys.stderr, flush=True)
    T = {m: float(np.median(times[m])) for m in methods}
    # instrumented solves, untimed: the kernel-family timers of every method, and the dominant
    # kernel per launch with its own start and end timestamps
    N.call("apx_set_timing", prob.plan, 1)
    for m in methods:
        for _ in range(max(1, args.steps) if m == H else 1):
            solve(m)
    kernels = kernel_table(prob)
    timers = all_timers(prob)
    N.call("apx_set_timing", prob.plan, 0)
    if dom in timers and timers[dom][1]:
        ms, cnt, by, fl = timers[dom]
        rocname = SOLVER_KERNELS.get(dom, dom)
        roof = roofline(rocname, ms, by, fl, PMC_TRAFFIC_BYTES.get((rocname, args.n)))
        roof["launches_timed"] = cnt
        roof["share_of_headline_device_time"] = profile[dom][0] * profile[dom][1] / max(
            1e-12, sum(v[0] * v[1] for k_, v in profile.items() if k_.startswith("k_") or k_ in SOLVER_KERNELS))
    else:
        roof = None
    # the same for the headline's Krylov SpMV (the HBM-bound kernel of every method)
    spmv_name = {"plain": "pcg_spmv_f64", "mixed": "pcg_spmv_f32m_f64v", "mixed_fp32": "pcg_spmv_f32",
                 "learned": "fcg_spmv_f64"}.get(H, "fcg_spmv_f32")
    spmv_roof = None
    if spmv_name in timers and timers[spmv_name][1] and spmv_name != dom:
        ms, cnt, by, fl = timers[spmv_name]
        rocname = SOLVER_KERNELS[spmv_name]
        spmv_roof = roofline(rocname, ms, by, 0, PMC_TRAFFIC_BYTES.get((rocname, args.n)))
    res = kernels.get("residual", {})
    dtypes = {"plain": "f64", "mixed": "f32+f64", "mixed_fp32": "f32+f64", "learned": "bf16+f64",
              "learned_mixed": "bf16+f32+f64", "learned_mixed_fp32_elements": "bf16+f32+f64",
              "learned_family_mixed": "bf16+f32+f64"}
    if learned_info and "train_seconds" in learned_info:
        setup = learned_info["pair_seconds"] + learned_info["train_seconds"]
        learned_info["setup_seconds"] = setup
        if H.startswith("learned"):
            learned_info["time_to_solution_incl_setup_s"] = T[H] + setup
            saved = T["plain"] - T[H]
            learned_info["break_even_solves"] = setup / saved if saved > 0 else None
    out = {
        "metric": METRIC, "value": T[H], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": T[H] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": dtypes[H], "data": "synthetic",
        "config": {"workload": f"2D Poisson Q1 {args.n}x{args.n} (short_example source, psdf Dirichlet), "
                               f"Newton to 1e-10 FP64 relative residual, method '{H}'",
                   "n_dofs": int(prob.n_dofs), "n_free": int(prob.n_free), "nnz": int(prob.nnz),
                   "parallelism": "single GPU"},
        "methods": {m: {"time_s": T[m], "timed_steps": len(times[m]), "speedup_vs_plain": T["plain"] / T[m],
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
        "headline_profile": {k_: {"avg_ms": v[0], "launches": v[1], "total_ms": v[0] * v[1]}
                             for k_, v in sorted(profile.items(), key=lambda kv: -kv[1][0] * kv[1][1])
                             if (k_.startswith("k_") or k_ in SOLVER_KERNELS) and v[1]},
        "network": ({"forward_ms": kernels["unet_forward"]["avg_ms"], "forward_GBps": kernels["unet_forward"]["GBps"],
                     "forward_bytes": kernels["unet_forward"]["bytes"]} if "unet_forward" in kernels else None),
        "roofline": roof,
        "roofline_spmv": spmv_roof,
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
