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
//
// Mixed precision for models linear in u (reliable_delta > 0): the tangent
// is assembled once (in FP64, rounded to FP32) and a single Jacobi-PCG on the
// FP32 matrix runs across Newton iterations; whenever its recursive residual
// has dropped by reliable_delta, the correction is added to the FP64 DOFs, the
// FP64 residual is assembled (the Newton test) and the CG residual is replaced
// by it while the search direction is kept.  That is FP64 iterative
// refinement without discarding the Krylov space at each refinement step.
// The Krylov vectors are FP64 (krylov_fp64_vectors, each FP32 matrix value
// widened in the SpMV) or FP32.  With FP64 vectors only the matrix is
// perturbed (by its rounding, a relative 6e-8 per entry), so CG keeps the FP64
// iteration count while the SpMV moves 4 bytes less per stored entry; FP32
// vectors also lose orthogonality at FP32 rounding and need more iterations.
#include <chrono>
#include <cmath>

#include "assembly.hpp"
#include "unet.hpp"

namespace apx {

namespace {
struct NewtonCtx {
  Plan& P;
  double* u;
  double* r;
  hipStream_t st;
  std::vector<EventPair> ev_res, ev_tan;
  apx_newton_stats& s;

  void residual() {
    auto t0 = std::chrono::steady_clock::now();
    EventPair e = timing_begin(P, st);
    assemble_residual(P, u, r, st);
    timing_end(P, e, st);
    if (P.timing) ev_res.push_back(e);
    s.seconds_residual += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  double norm() {
    auto t0 = std::chrono::steady_clock::now();
    double v = free_norm(P, r, st);
    s.seconds_residual += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return v;
  }
  // The FP32 Krylov matrix is by default the FP64 tangent rounded once (the owner-tile kernels
  // store the rounded values directly).  Assembled with FP32 element arithmetic through the
  // general per-point path, the element sums carry accumulation errors of the size of rounding but
  // without its structure (row sums, symmetry): on the 2048^2 Poisson problem flexible CG with
  // the network needed 76 iterations over 6 Newton steps on such a matrix and 50 over 5 on the
  // rounded one.  The parallelogram path's FP32 arithmetic (constant tables) keeps the structure
  // and needs the same 50.
  bool fp32_arith = false;  // FP32 tangent: element arithmetic in FP32 instead of rounding the FP64 one
  // Tangent values of type M into vals and the preconditioner's inverse diagonal of type V
  // (an FP32 matrix may run with FP64 Krylov vectors: its diagonal is widened exactly).
  template <class M, class V>
  void tangent(M* vals, V* inv, int precond) {
    const int prec = sizeof(M) == 8 ? APX_FP64 : APX_FP32;
    auto t0 = std::chrono::steady_clock::now();
    EventPair e = timing_begin(P, st);
    if (prec == APX_FP32 && !fp32_arith) {
      assemble_tangent_rounded(P, u, (float*)vals, st);
    } else {
      assemble_tangent(P, u, vals, prec, st);
    }
    timing_end(P, e, st);
    if (P.timing) ev_tan.push_back(e);
    if (precond == APX_PRECOND_JACOBI) {
      if constexpr (sizeof(M) == sizeof(V))
        jacobi_inverse(P, vals, inv, prec, st);
      else
        jacobi_inverse_f32_to_f64(P, vals, inv, st);
    } else {
      fill<V>(inv, P.n_free, V(1), st);
    }
    APX_HIP(hipStreamSynchronize(st));
    s.seconds_tangent += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void finish(int prec) {
    APX_HIP(hipStreamSynchronize(st));
    timing_collect(P, "residual", ev_res, residual_bytes(P), ev_res.size());
    if (prec == APX_FP64)
      timing_collect(P, "tangent_f64", ev_tan, tangent_bytes(P, APX_FP64), ev_tan.size());
    else if (fp32_arith)
      timing_collect(P, "tangent_f32", ev_tan, tangent_bytes(P, APX_FP32), ev_tan.size());
    else
      timing_collect(P, "tangent_f64_to_f32", ev_tan,
                     tangent_by_tiles(P) ? tangent_bytes(P, APX_FP32) : tangent_bytes(P, APX_FP64) + P.sell_total * 12,
                     ev_tan.size());
  }
};

// Newton bookkeeping shared by both drivers (solver.py:898-930).
struct NewtonTest {
  int itt = 0;
  double res_old = 0.0;
  bool diverged = false;
  // returns true when the loop continues
  bool step(double res, double thr, int maxiter, apx_newton_stats& s) {
    const bool not_stop = res > thr;
    bool next;
    if (itt < maxiter) {
      diverged = (itt > 1 && res / res_old > 10.0) || !std::isfinite(res);
      next = !diverged;
    } else {
      next = false;
      diverged = true;
    }
    if (itt < 64) s.history[itt] = res;
    ++itt;
    res_old = res;
    return not_stop && next;
  }
};
}  // namespace

// The plan's SELL value buffer of type M.
template <class M>
static M* matrix_buffer(Plan& P) {
  if constexpr (sizeof(M) == 8) {
    if (P.vals64.n < (size_t)P.sell_total) P.vals64.alloc(P.sell_total);
    return P.vals64.p;
  } else {
    if (P.vals32.n < (size_t)P.sell_total) P.vals32.alloc(P.sell_total);
    return P.vals32.p;
  }
}

// Newton with one Krylov solve per iteration.  M: tangent (matrix) type, V: Krylov vector type
// of Jacobi-PCG (the learned path's flexible CG keeps FP64 vectors whatever M is).
template <class M, class V>
static void newton_standard(NewtonCtx& C, const double* dvals, const apx_newton_options& o) {
  Plan& P = C.P;
  apx_newton_stats& s = C.s;
  const int vprec = sizeof(V) == 8 ? APX_FP64 : APX_FP32;
  DevBuf<V> vecs;
  vecs.alloc(2 * P.n_free);
  V* inv = vecs.p;
  V* b = vecs.p + P.n_free;
  M* vals = matrix_buffer<M>(P);
  impose_dirichlet(P, C.u, dvals, C.st);
  C.residual();
  const double res0 = C.norm();
  s.initial_residual = res0;
  const double thr = o.relative ? o.newton_tol * res0 : o.newton_tol;

  NewtonTest test;
  double res = res0;
  bool have_tangent = false, go = true;
  bool learned = o.precond == APX_PRECOND_LEARNED, fell_back = false;
  PcgRun<M, V> pcg(P, vals, inv, C.st);
  DevBuf<double> b64, x64;
  if (learned) {
    APX_REQUIRE(P.net != nullptr, APX_ERR_STATE, "learned preconditioner requested but no network is attached");
    b64.alloc(P.n_free);
    x64.alloc(P.n_free);
  }
  while (go) {
    if (!have_tangent || !o.reuse_tangent) {
      C.tangent(vals, inv, learned ? APX_PRECOND_JACOBI : o.precond);
      have_tangent = true;
    }
    double eta = o.lin_tol;
    if (o.forcing && res > 0.0) eta = std::max(eta, std::min(0.5, 0.5 * thr / res));
    // an FP32 matrix bounds the true progress of one inner solve (about eps32 x cond(A)):
    // stop each FP32 flexible-CG solve after a 'reliable delta' reduction and let the FP64
    // Newton residual re-anchor the next one (flexible CG(1) keeps one direction, so the
    // restart costs little)
    if (learned && sizeof(M) == 4 && o.reliable_delta > 0.0) eta = std::max(eta, o.reliable_delta);
    auto t0 = std::chrono::steady_clock::now();
    if (learned) {
      gather_rhs(P, C.r, b64.p, APX_FP64, C.st);
      FcgResult fr = fcg_solve<M>(P, *P.net, vals, P.net_g, b64.p, x64.p, eta, o.lin_atol, o.lin_maxiter, C.st);
      s.krylov_iterations += fr.iterations;
      if (fr.stagnated) {
        // the network stopped reducing the residual: keep its progress, continue with Jacobi-PCG
        s.fallbacks += 1;
        learned = false;
        fell_back = true;
      }
      scatter_update(P, C.u, x64.p, o.damping, APX_FP64, C.st);
    } else {
      gather_rhs(P, C.r, b, vprec, C.st);
      // after a fallback the iteration budget is Jacobi-PCG's own (at least jax cg's default
      // 10 n_free): 'linear maxiter' was sized for flexible CG with the network
      const int fb_max = (int)std::min<int64_t>(10 * P.n_free, 2000000000);
      const int maxit = fell_back && o.lin_maxiter > 0 ? std::max(o.lin_maxiter, fb_max) : o.lin_maxiter;
      pcg.init(b, nullptr, eta, o.lin_atol, maxit, 0.0);
      pcg.iterations = 0;
      pcg.counted = 0;
      KState hs = pcg.run();
      pcg.collect_timing();
      s.krylov_iterations += hs.k;
      scatter_update(P, C.u, pcg.x, o.damping, vprec, C.st);
    }
    s.seconds_krylov += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    C.residual();
    res = C.norm();
    go = test.step(res, thr, o.maxiter, s);
  }
  s.iterations = test.itt;
  s.diverged = test.diverged ? 1 : 0;
  s.residual_norm = res;
}

// V: Krylov vector type (FP32, or FP64 with the FP32 matrix widened in the SpMV).
template <class V>
static void newton_reliable(NewtonCtx& C, const double* dvals, const apx_newton_options& o) {
  Plan& P = C.P;
  apx_newton_stats& s = C.s;
  const int vprec = sizeof(V) == 8 ? APX_FP64 : APX_FP32;
  DevBuf<V> vecs;
  vecs.alloc(2 * P.n_free);
  V* inv = vecs.p;
  V* b = vecs.p + P.n_free;
  float* vals = matrix_buffer<float>(P);
  impose_dirichlet(P, C.u, dvals, C.st);
  C.residual();
  const double res0 = C.norm();
  s.initial_residual = res0;
  const double thr = o.relative ? o.newton_tol * res0 : o.newton_tol;
  C.tangent(vals, inv, o.precond);
  gather_rhs(P, C.r, b, vprec, C.st);
  auto t0 = std::chrono::steady_clock::now();
  PcgRun<float, V> pcg(P, vals, inv, C.st);
  // the FP32 recursion stops on its own at half the Newton threshold; the
  // FP64 residual decides at every anchor
  pcg.init(b, nullptr, 0.0, 0.5 * thr, o.lin_maxiter, o.reliable_delta);
  NewtonTest test;
  double res = res0;
  bool go = true;
  while (go) {
    KState hs = pcg.run();
    pcg.collect_timing();
    scatter_update(P, C.u, pcg.x, 1.0, vprec, C.st);
    C.residual();
    res = C.norm();
    go = test.step(res, thr, o.maxiter, s);
    if (go && hs.k >= pcg.maxiter) {
      test.diverged = true;
      go = false;
    }
    if (go) pcg.reanchor(C.r);
  }
  s.seconds_krylov += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  s.krylov_iterations = pcg.iterations;
  s.iterations = test.itt;
  s.diverged = test.diverged ? 1 : 0;
  s.residual_norm = res;
}

void newton(Plan& P, double* u, const double* dvals, const apx_newton_options& o, apx_newton_stats& s,
            hipStream_t st) {
  APX_REQUIRE(o.precond == APX_PRECOND_JACOBI || o.precond == APX_PRECOND_NONE || o.precond == APX_PRECOND_LEARNED,
              APX_ERR_UNSUPPORTED, "apx_newton: unknown preconditioner");
  const auto t_start = std::chrono::steady_clock::now();
  if (P.r_full.n < (size_t)P.n_dofs) P.r_full.alloc(P.n_dofs);
  s = apx_newton_stats{};
  NewtonCtx C{P, u, P.r_full.p, st, {}, {}, s};
  C.fp32_arith = o.fp32_element_arith != 0;
  APX_REQUIRE(o.precision == APX_FP64 || o.precision == APX_FP32, APX_ERR_ARG, "apx_newton: unknown precision");
  const bool v64 = o.krylov_fp64_vectors != 0;
  if (o.precision == APX_FP32 && o.reuse_tangent && o.reliable_delta > 0.0 && o.precond != APX_PRECOND_LEARNED) {
    APX_REQUIRE(o.damping == 1.0, APX_ERR_ARG, "reliable-update refinement needs damping 1");
    if (v64)
      newton_reliable<double>(C, dvals, o);
    else
      newton_reliable<float>(C, dvals, o);
  } else if (o.precision == APX_FP64) {
    newton_standard<double, double>(C, dvals, o);
  } else if (v64) {
    newton_standard<float, double>(C, dvals, o);
  } else {
    newton_standard<float, float>(C, dvals, o);
  }
  C.finish(o.precision);
  s.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

}  // namespace apx
