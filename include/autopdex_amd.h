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
                                   models.py:1122-1146; 2D = plane strain (F33 = 1)          */
  APX_MODEL_LINE_TRACTION = 4   /* - t . u on 2-node boundary lines, params (t_x, t_y)       */
};

enum apx_precision { APX_FP64 = 0, APX_FP32 = 1 };
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
 * int64 row-major.  Returns the set index through *set_out.                 */
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
 * reduction of scipy_assembling: values of the reduced CSR (nnz) in FP64 or
 * FP32 (precision = apx_precision), pattern from apx_plan_csr_pattern.      */
int apx_assemble_tangent(apx_plan* plan, const double* u_dev, void* vals_dev, int precision, void* stream);

/* assembler.assemble_tangent_diagonal (assembler.py:639-680, :1090-1113):
 * diag_dev (n_dofs, FP64, full length).                                     */
int apx_assemble_diagonal(apx_plan* plan, const double* u_dev, double* diag_dev, void* stream);

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

/* ---- misc ------------------------------------------------------------------ */
const char* apx_last_error(void);
int apx_version(void);
/* Average device time in ms of the most recent launches of a named kernel
 * family ("spmv", "residual", "tangent", "pcg_update") recorded with HIP
 * events on the caller's stream (0 when never timed).                       */
int apx_set_timing(apx_plan* plan, int enabled);
int apx_get_timing(apx_plan* plan, const char* name, double* avg_ms, int64_t* count, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
