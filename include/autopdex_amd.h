/*
 * autopdex_amd -- MI355X (gfx950) assemble-and-solve backend for AutoPDEx.
 *
 * Plain C ABI.  All pointers named *_dev are device pointers (HBM) owned by
 * the caller; *_host pointers are host memory.  Every function returns 0 on
 * success or a negative apx_status; apx_last_error() gives the message
 * (thread-local).  No exceptions cross this boundary.  A plan is not
 * thread-safe; each call enqueues its work on the given stream (hipStream_t
 * passed as void*, NULL = default stream).
 *
 * The reference has no native interface: its backend plug point is the
 * match on static_settings['solver backend'] in solve_linear
 * (autopdex/solver.py:621-640) and the host-solver callback contract
 * fn(mat, rhs, free_dofs) -> sol (solver.py:643-659).  Each entry point below
 * names the reference function it replaces.
 */
#ifndef AUTOPDEX_AMD_H
#define AUTOPDEX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct apx_plan apx_plan;

enum apx_status {
  APX_OK = 0,
  APX_ERR_ARG = -1,      /* invalid argument / shape                       */
  APX_ERR_HIP = -2,      /* HIP runtime error                              */
  APX_ERR_STATE = -3,    /* call order (e.g. plan not finalized)           */
  APX_ERR_UNSUPPORTED = -4, /* model / element / option not implemented     */
  APX_ERR_INTERNAL = -5
};

/* Element models (the kernels behind static_settings['model'][set]). */
enum apx_model {
  APX_MODEL_POISSON_SOURCE = 1, /* 1/2|grad u|^2 - b u, b = p0 (sin(p1 x.x) - cos(p1 |x-(p2,p3)|^2))
                                   examples/miscellaneous/short_example.py:23-33 */
  APX_MODEL_POISSON_VARK = 2,   /* 1/2 exp(N.g)|grad u|^2 - (N.f) u, nodal fields g, f       */
  APX_MODEL_NEO_HOOKE = 3,      /* Ciarlet Neo-Hooke psi(I + grad u), params (lambda, mu),
                                   models.py:1122-1146; 2D = plane strain (F33 = 1).
                                   Q1 quads / Hex8 bricks ('user potential',
                                   models.py:1188-1269, 2-point rule) or Q2 9-node quads
                                   ('user element' P : virt F of hyperelastic_steady_state_weak
                                   :917-1000 in isoparametric_domain_element_galerkin
                                   :1616-1720, 3 x 3 rule)                                    */
  APX_MODEL_LINE_TRACTION = 4   /* - t . u on boundary lines in 2D, params (t_x, t_y): 2-node
                                   lines (2-point rule) or 3-node lines (nodes xi = -1, 1, 0,
                                   3-point rule; isoparametric_surface_element_galerkin
                                   :1723-1850 with neumann_weak :744-780)                     */
};

/* APX_FP64_ROUNDED (tangent assembly only): FP64 element arithmetic, values rounded once to
 * FP32 -- the FP32 matrix of the mixed method, bit for bit the FP64 values cast to float.     */
enum apx_precision { APX_FP64 = 0, APX_FP32 = 1, APX_FP64_ROUNDED = 2 };
enum apx_precond { APX_PRECOND_JACOBI = 0, APX_PRECOND_NONE = 1, APX_PRECOND_LEARNED = 2 };

/* ---- plan (mesh + DOF layout + Dirichlet mask + CSR pattern) ------------- */

/* Replaces the per-call host COO->CSR of scipy_assembling (solver.py:1180-1222)
 * and _get_indices (assembler.py:47-141): the reduced pattern is built once.
 * coords_host: n_nodes x dim row-major (settings['node coordinates']).
 * dirichlet_mask_host: n_nodes*dofs_per_node bytes, nonzero = Dirichlet DOF
 * (settings['dirichlet dofs'] flattened, P3 numbering node*d + comp).       */
int apx_plan_create(apx_plan** out, int dim, int64_t n_nodes, const double* coords_host,
                    int dofs_per_node, const uint8_t* dirichlet_mask_host);

/* One element set = one entry of settings['connectivity'] / static_settings['model']
 * (assembler.py:587-637 sums over sets).  conn_host: n_elems x nodes_per_elem
 * int64 row-major, nodes_per_elem 2, 3 (lines), 4, 9 (quads) or 8 (bricks).
 * Returns the set index through *set_out.                                   */
int apx_plan_add_set(apx_plan* plan, const int64_t* conn_host, int64_t n_elems, int nodes_per_elem,
                     int model, const double* params, int n_params, int* set_out);

/* Update a set's scalar parameters (e.g. load multiplier x traction).       */
int apx_plan_set_params(apx_plan* plan, int set, const double* params, int n_params);

/* Bind nodal field `which` (0 = g, 1 = f for APX_MODEL_POISSON_VARK) of a set
 * to caller-owned device memory of n_nodes doubles.  The pointer is stored, not
 * copied: it must stay valid while the plan assembles.                      */
int apx_plan_bind_nodal_field(apx_plan* plan, int set, int which, const double* field_dev);

/* Build the reduced CSR pattern (bit-identical to scipy's canonical CSR of the
 * reduced system, P4), element colours and slot maps; upload to the device. */
int apx_plan_finalize(apx_plan* plan);
/* Host-only part of apx_plan_finalize (pattern, colours, slot maps; no device
 * calls), so the pattern can be built and inspected without a GPU.         */
int apx_plan_build_pattern(apx_plan* plan);
int apx_plan_destroy(apx_plan* plan);

/* structured: bit 0 = a set is a structured Q1 grid (mesher.structured_mesh
 * connectivity) and uses the owner-tile residual kernels; bit 1 = the tangent
 * is assembled by the owner-tile kernels too (one DOF per node, stencil
 * pattern).                                                                  */
typedef struct apx_plan_info {
  int64_t n_nodes, n_dofs, n_free, nnz;
  int dim, dofs_per_node, n_sets, structured;
  int n_colours[8];
  double build_seconds;
  int64_t device_bytes;
} apx_plan_info;
int apx_plan_get_info(const apx_plan* plan, apx_plan_info* info);

/* Copy the reduced CSR pattern to the host: indptr n_free+1, indices nnz.   */
int apx_plan_csr_pattern(const apx_plan* plan, int32_t* indptr_host, int32_t* indices_host);
/* Reduced index -> full DOF index (n_free int32, host).                     */
int apx_plan_free_dofs(const apx_plan* plan, int32_t* red2full_host);

/* ---- assembly ------------------------------------------------------------ */

/* assembler.assemble_residual (assembler.py:587-637, user potential
 * :1064-1088): r_dev (n_dofs, FP64, full length incl. Dirichlet rows)
 * = sum over sets of the scattered element residuals at u_dev (n_dofs).     */
int apx_assemble_residual(apx_plan* plan, const double* u_dev, double* r_dev, void* stream);

/* assembler.assemble_tangent (assembler.py:682-777, :1115-1149) followed by the
 * reduction of scipy_assembling: values of the reduced CSR (nnz) in FP64, FP32
 * or FP32 rounded from FP64 arithmetic (precision = apx_precision), pattern
 * from apx_plan_csr_pattern.                                                */
int apx_assemble_tangent(apx_plan* plan, const double* u_dev, void* vals_dev, int precision, void* stream);

/* assembler.assemble_tangent_diagonal (assembler.py:639-680, :1090-1113):
 * diag_dev (n_dofs, FP64, full length).                                     */
int apx_assemble_diagonal(apx_plan* plan, const double* u_dev, double* diag_dev, void* stream);

/* assembler.assemble_tangent as the reference returns it (assembler.py:682-777,
 * user_potential_assemble_tangent :1115-1149, indices _get_indices :47-141):
 * the full n_dofs x n_dofs tangent as BCOO with duplicates.  data_dev receives
 * the values in the reference's order -- per set, element-major row-major
 * blocks, rows repeat(gdofs, n), columns tile(gdofs, n) with gdofs = node*d +
 * comp (P11); sets concatenated -- apx_tangent_coo_entries() values in all
 * (precision = apx_precision).  The indices depend only on the connectivity
 * and are built by the caller.                                              */
int apx_tangent_coo_entries(const apx_plan* plan, int64_t* count);
int apx_assemble_tangent_coo(apx_plan* plan, const double* u_dev, void* data_dev, int precision, void* stream);

/* ---- linear algebra on the reduced system -------------------------------- */

/* y = A x with A the reduced CSR of the plan (vals in `precision`).         */
int apx_spmv(apx_plan* plan, const void* vals_dev, const void* x_dev, void* y_dev, int precision,
             void* stream);

typedef struct apx_krylov_stats {
  int iterations;
  int converged;
  double initial_residual;  /* ||b - A x0||_2                                 */
  double final_residual;    /* ||r_k||_2 (recursive residual)                 */
  double seconds;
} apx_krylov_stats;

/* Jacobi-PCG on the reduced system (replaces jax.scipy.sparse.linalg.cg with
 * M = 1/diag, solver.py:1093-1116): stops when r.r <= max(tol^2 b.b, atol^2)
 * or after maxiter.  x0 = 0.  vals/b/x in `precision`.                      */
int apx_pcg(apx_plan* plan, const void* vals_dev, const void* b_dev, void* x_dev, int precision,
            int precond, double tol, double atol, int maxiter, apx_krylov_stats* stats, void* stream);

/* Allocation-free apx_pcg.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels, iterates and
 * results as apx_pcg, but everything apx_pcg allocates (the SELL copy of the
 * values, the inverse diagonal, six Krylov vectors, reduction partials, device
 * scalars) lives in one 32 MiB device array per GPU, held until the process ends;
 * the stop flag is polled without events.  A solve needs
 * 256 + 262144 + round256(sell_entries * T) + round256(n_free * T) +
 * round256(6 n_free * T) bytes (T = 8 FP64, 4 FP32); a larger one returns
 * APX_ERR_ARG.  All calls on one GPU share its array: calls that overlap on one
 * GPU race.                                                                  */
int apx_pcg_noalloc(apx_plan* plan, const void* vals_dev, const void* b_dev, void* x_dev, int precision,
                    int precond, double tol, double atol, int maxiter, apx_krylov_stats* stats, void* stream);

/* ---- Newton ---------------------------------------------------------------- */

typedef struct apx_newton_options {
  double newton_tol;        /* absolute tolerance on ||r_free||_2 (P6)                */
  int relative;             /* 1: threshold = newton_tol * ||r_free(u0)||_2           */
  int maxiter;              /* reference default 30                                  */
  double damping;           /* 1.0 = solve_newton                                    */
  int precision;            /* tangent + Krylov precision (residual is always FP64)  */
  int precond;              /* apx_precond                                           */
  double lin_tol;           /* Krylov relative tolerance                             */
  double lin_atol;
  int lin_maxiter;          /* <= 0: 10 * n_free (jax cg default)                    */
  int reuse_tangent;        /* 1: model is linear in u, assemble the tangent once    */
  int forcing;              /* 1: Krylov tolerance max(lin_tol, 0.5 thr/||r_k||), so no
                               solve is tighter than the Newton test needs           */
  double reliable_delta;    /* > 0 with precision FP32 and reuse_tangent: one FP32 CG
                               whose residual is re-anchored to the FP64 residual
                               whenever it drops by this factor (reliable updates);
                               each anchor is a Newton iteration                      */
  int fp32_element_arith;   /* precision FP32 only.  0: the tangent is assembled in FP64
                               and rounded once to FP32 (default); 1: the element
                               arithmetic of the tangent runs in FP32 (FP32 accumulation)  */
  int krylov_fp64_vectors;  /* precision FP32 only.  1: Jacobi-PCG keeps FP64 vectors and
                               scalars with the FP32 matrix (each value widened in the
                               SpMV); 0: FP32 vectors.  Flexible CG always uses FP64
                               vectors.                                                    */
} apx_newton_options;

typedef struct apx_newton_stats {
  int iterations;
  int diverged;
  double residual_norm;     /* final ||r_free||_2                                     */
  double initial_residual;  /* ||r_free(u0)||_2                                       */
  int krylov_iterations;    /* summed over Newton iterations                          */
  int fallbacks;            /* learned-preconditioner stagnation -> Jacobi restarts   */
  double seconds_total, seconds_residual, seconds_tangent, seconds_krylov;
  double history[64];       /* ||r_free|| after each Newton iteration                 */
} apx_newton_stats;

/* solver.solve_newton / damped_newton (solver.py:740-948) with solve_linear
 * (:544-659), device resident.  u_dev (n_dofs, FP64): initial guess in,
 * solution out; dirichlet_values_dev (n_dofs) gives the prescribed values at
 * Dirichlet DOFs (P5: imposed absolutely at every iteration).               */
int apx_newton(apx_plan* plan, double* u_dev, const double* dirichlet_values_dev,
               const apx_newton_options* opts, apx_newton_stats* stats, void* stream);

/* ---- learned preconditioner (no counterpart in the reference) ---------------
 * A multigrid-structured U-Net on the structured node grid (height x width =
 * the plan's nodes, node index = row * width + col as produced by
 * mesher.structured_mesh).  Levels l = 0..levels-1 with channels[l] in
 * {16, 32, 64} (channels[0] = 16); input channels 1 (residual) or 2 (residual +
 * nodal log-coefficient, or two displacement residual components); output
 * channels = DOFs per node.  Activations bf16, bf16 MFMA with FP32
 * accumulation; FP32 master parameters and Adam moments.                    */
typedef struct apx_unet apx_unet;
int apx_unet_create(apx_unet** out, int levels, const int* channels, int c_in, int c_out, int relu, int batch,
                    int height, int width);
int apx_unet_destroy(apx_unet* net);
/* Parameter vector (FP32, host): per layer [co_pad x k_pad weights][co biases];
 * apx_unet_layer describes layer i (weight W[co][t*ci + c], t = 3(dy+1)+(dx+1)). */
int apx_unet_num_params(const apx_unet* net, int64_t* n_params, int* n_layers);
int apx_unet_layer(const apx_unet* net, int layer, int64_t* w_off, int64_t* b_off, int* ci, int* co, int* co_pad,
                   int* k_pad, int* relu, int* level);
int apx_unet_set_params(apx_unet* net, const float* params_host);
int apx_unet_get_params(const apx_unet* net, float* params_host);
int apx_unet_get_grads(const apx_unet* net, float* grads_host);
/* Raw buffers (device): input x (batch x H x W x c_in bf16), output y (FP32),
 * per-image input scale (FP32).                                             */
int apx_unet_buffers(apx_unet* net, void** x, void** y, void** scale);
int apx_unet_forward(apx_unet* net, void* stream);
/* loss = mean_b ||y_b - t_b||^2 / ||t_b||^2 for target t (device, FP32, layout of y);
 * backward into the parameter gradients.                                    */
int apx_unet_loss_backward(apx_unet* net, const float* target_dev, double* loss, void* stream);
int apx_unet_adam(apx_unet* net, float lr, float beta1, float beta2, float eps, int step, void* stream);
/* Network input of image `image` from a reduced FP64 residual r (n_free):
 * x = r / s at free DOFs (0 at Dirichlet DOFs), s = RMS(r), plus the nodal
 * field g_nodal (device, optional) as the last channel.  With e given, also the
 * training target t = e / s (device FP32, layout of y).                    */
int apx_unet_load_pair(apx_unet* net, apx_plan* plan, const double* r_dev, const double* g_nodal_dev,
                       const double* e_dev, float* target_dev, int image, void* stream);
/* One training step's loss and gradient, z_b = s_b net(r_b / s_b):
 *   loss_kind 0 (residual): L = mean_b ||r_b - A z_b||^2 / ||r_b||^2
 *   loss_kind 1 (energy):   L = mean_b sA (z_b.A z_b / 2 - z_b.r_b) / (s_b^2 n),
 *                           sA = sign of A, minimizer z_b = A^-1 r_b
 *   loss_kind 2 (A-norm):   L = mean_b (z_b-e_b).sA A (z_b-e_b) / e_b.sA A e_b,
 *                           e_b = A^-1 r_b from e_list (required)
 * plus w_e ||z_b - e_b||^2 / ||e_b||^2 when errors are given.  Loads the batch
 * inputs from r_list[b] (device FP64 residuals, host array of `batch`
 * pointers; g_list / e_list optional), runs forward and backward.  vals_dev:
 * FP64 tangent in CSR order.  Follow with apx_unet_adam.                   */
int apx_unet_train_step(apx_unet* net, apx_plan* plan, const void* vals_dev, const double* const* r_list,
                        const double* const* g_list, const double* const* e_list, int loss_kind, double w_e,
                        double* loss, void* stream);
/* z = D^-1/2 s net(D^-1/2 r / s) on the plan's free DOFs (FP64 in and out),
 * s = RMS of D^-1/2 r, D = |diag A| of the matrix of the plan's latest
 * flexible-CG solve or training step (none before the first: D = I).       */
int apx_unet_apply(apx_unet* net, apx_plan* plan, const double* r_dev, const double* g_nodal_dev, double* z_dev,
                   void* stream);
/* Make the network a correction to Jacobi: z = D^-1 r + D^-1/2 s net(D^-1/2 r / s)
 * (enable = 1) instead of z = D^-1/2 s net(D^-1/2 r / s) (0, the default).  Applies to
 * inference, flexible CG and training alike.                                */
int apx_unet_set_jacobi_skip(apx_unet* net, int enable);
/* Fused launches (enable = 1, the default) or one launch per layer (0); the two
 * give bit-identical outputs.  Fused: the network's coarse levels (from the first
 * restriction into a run of 64-channel levels of at most 9 x 9 nodes down to the
 * matching prolongation) run as one launch, one workgroup per image; in inference
 * (apx_unet_apply, flexible CG) the input layer is formed inside the level-0
 * conv_e launch and its output is not stored.  tail_level (optional) receives the
 * level whose restriction starts the fused coarse levels (-1: none).  No
 * counterpart in the reference (a launch and traffic optimisation of this
 * implementation).                                                          */
int apx_unet_set_fused_tail(apx_unet* net, int enable, int* tail_level);

/* Use `net` as preconditioner of the plan (apx_newton with precond
 * APX_PRECOND_LEARNED): flexible CG, Jacobi-PCG fallback on stagnation.     */
int apx_plan_attach_unet(apx_plan* plan, apx_unet* net, const double* g_nodal_dev);
/* One flexible-CG solve with the attached network (x0 = 0; vals in CSR order). */
int apx_fcg(apx_plan* plan, const void* vals_dev, const double* b_dev, double* x_dev, int precision, double tol,
            double atol, int maxiter, apx_krylov_stats* stats, void* stream);

/* ---- misc ------------------------------------------------------------------ */
const char* apx_last_error(void);
int apx_version(void);
/* Average device time in ms of the most recent launches of a named kernel
 * family ("spmv", "residual", "tangent", "pcg_update") recorded with HIP
 * events on the caller's stream (0 when never timed).  enabled: 0 off (clears
 * the timers), 1 the solvers' kernel families plus, per launch, the network
 * kernels named in the filter, 2 also every network kernel per launch (timer
 * name = the kernel's template as rocprofv3 shows it, e.g.
 * "k_conv3x3<64, 64, false, 4>").  No counterpart in the reference
 * (measurement of this implementation).                                      */
int apx_set_timing(apx_plan* plan, int enabled);
int apx_get_timing(apx_plan* plan, const char* name, double* avg_ms, int64_t* count, int64_t* bytes);
/* Kernel names (separated by ';') timed per launch at timing level 1.      */
int apx_set_timing_filter(apx_plan* plan, const char* names);
/* Timer `index` (0-based, name order): name, average ms per launch, launches,
 * algorithmic bytes and flops per launch; APX_ERR_ARG past the last timer.   */
int apx_get_timer(const apx_plan* plan, int index, char* name, int name_len, double* avg_ms, int64_t* count,
                  int64_t* bytes, int64_t* flops);

#ifdef __cplusplus
}
#endif
#endif
