// Device-resident Newton driver.
//
// Restates solver.solve_newton / solve_damped_newton / damped_newton
// (solver.py:740-948) around solve_linear (:544-659) for the 'hip' backend:
//  - Dirichlet values are imposed absolutely (P5): u[D] = values before the
//    first assembly; the update adds damping * increment to free DOFs only.
//  - stop when ||r_free||_2 <= tol (absolute, P6; optional relative form
//    tol * ||r_free(u0)||), at least one iteration, divergence when
//    ||r||/||r_old|| > 10 with itt > 1 or a non-finite norm, and diverged =
//    true after maxiter iterations.
//  - the linear system is the reduced one with rhs -r_free (P4).
// The residual after an update is reused as the next iteration's right-hand
// side (the reference assembles the same vector twice).  Only one scalar
// (the residual norm) comes back to the host per Newton iteration.
#include <chrono>
#include <cmath>

#include "assembly.hpp"

namespace apx {

void newton(Plan& P, double* u, const double* dvals, const apx_newton_options& o, apx_newton_stats& s,
            hipStream_t st) {
  APX_REQUIRE(o.precond == APX_PRECOND_JACOBI || o.precond == APX_PRECOND_NONE, APX_ERR_UNSUPPORTED,
              "apx_newton: preconditioner not available in this entry point");
  const auto t_start = std::chrono::steady_clock::now();
  auto secs = [](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
  };
  const int prec = o.precision;
  const size_t tsz = prec == APX_FP64 ? 8 : 4;
  if (P.r_full.n < (size_t)P.n_dofs) P.r_full.alloc(P.n_dofs);
  // vectors: inv_diag, b, x (n_free each) in the Krylov precision
  DevBuf<char> vecs;
  vecs.alloc(3 * P.n_free * tsz);
  void* inv = vecs.p;
  void* b = vecs.p + P.n_free * tsz;
  void* x = vecs.p + 2 * P.n_free * tsz;
  void* vals;
  if (prec == APX_FP64) {
    if (P.vals64.n < (size_t)P.nnz) P.vals64.alloc(P.nnz);
    vals = P.vals64.p;
  } else {
    if (P.vals32.n < (size_t)P.nnz) P.vals32.alloc(P.nnz);
    vals = P.vals32.p;
  }
  double* r = P.r_full.p;

  s = apx_newton_stats{};
  impose_dirichlet(P, u, dvals, st);
  auto t0 = std::chrono::steady_clock::now();
  assemble_residual(P, u, r, st);
  const double res0 = free_norm(P, r, st);
  s.seconds_residual += secs(t0);
  s.initial_residual = res0;
  const double thr = o.relative ? o.newton_tol * res0 : o.newton_tol;

  int itt = 0;
  double res_old = 0.0, res = res0;
  bool not_stop = true, diverged = false, have_tangent = false;
  while (not_stop) {
    if (!have_tangent || !o.reuse_tangent) {
      t0 = std::chrono::steady_clock::now();
      assemble_tangent(P, u, vals, prec, st);
      if (o.precond == APX_PRECOND_JACOBI) {
        jacobi_inverse(P, vals, inv, prec, st);
      } else if (prec == APX_FP64) {
        fill<double>((double*)inv, P.n_free, 1.0, st);
      } else {
        fill<float>((float*)inv, P.n_free, 1.0f, st);
      }
      APX_HIP(hipStreamSynchronize(st));
      s.seconds_tangent += secs(t0);
      have_tangent = true;
    }
    gather_rhs(P, r, b, prec, st);
    t0 = std::chrono::steady_clock::now();
    apx_krylov_stats ks = pcg(P, vals, inv, b, x, prec, o.lin_tol, o.lin_atol, o.lin_maxiter, st);
    s.seconds_krylov += secs(t0);
    s.krylov_iterations += ks.iterations;
    scatter_update(P, u, x, o.damping, prec, st);
    t0 = std::chrono::steady_clock::now();
    assemble_residual(P, u, r, st);
    res = free_norm(P, r, st);
    s.seconds_residual += secs(t0);
    not_stop = res > thr;
    bool next;
    if (itt < o.maxiter) {
      diverged = (itt > 1 && res / res_old > 10.0) || !std::isfinite(res);
      next = !diverged;
    } else {
      next = false;
      diverged = true;
    }
    if (itt < 64) s.history[itt] = res;
    ++itt;
    res_old = res;
    not_stop = not_stop && next;
  }
  APX_HIP(hipStreamSynchronize(st));
  s.iterations = itt;
  s.diverged = diverged ? 1 : 0;
  s.residual_norm = res;
  s.seconds_total = secs(t_start);
}

}  // namespace apx
