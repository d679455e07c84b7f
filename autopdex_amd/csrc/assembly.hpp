// Internal interfaces shared by the library's translation units.
#pragma once
#include "plan.hpp"

namespace apx {

template <class T>
void fill(T* p, int64_t n, T v, hipStream_t s);

void assemble_residual(Plan& P, const double* u, double* r, hipStream_t st);
void assemble_tangent(Plan& P, const double* u, void* vals, int precision, hipStream_t st);
void assemble_diagonal(Plan& P, const double* u, double* diag, hipStream_t st);

// krylov.hip
struct KState {
  double gamma, pAp, alpha, beta, rr, atol2, bb, rr0;
  double dot_out;         // scratch result of dot / norm kernels
  int k, maxiter, done, stagnated;
  unsigned int counter[4];
};

void spmv(Plan& P, const void* vals, const void* x, void* y, int precision, hipStream_t st);
void jacobi_inverse(Plan& P, const void* vals, void* inv_diag, int precision, hipStream_t st);
// Jacobi-PCG; workspace taken from the plan.
apx_krylov_stats pcg(Plan& P, const void* vals, const void* inv_diag, const void* b, void* x, int precision,
                     double tol, double atol, int maxiter, hipStream_t st);
// ||r[red2full]||_2 of a full-length FP64 vector (deterministic).
double free_norm(Plan& P, const double* r_full, hipStream_t st);
void gather_rhs(Plan& P, const double* r_full, void* b, int precision, hipStream_t st);
void scatter_update(Plan& P, double* u, const void* x, double damping, int precision, hipStream_t st);
void impose_dirichlet(Plan& P, double* u, const double* values, hipStream_t st);

// timing helpers
struct ScopedTimer {
  Plan& P;
  const char* name;
  hipStream_t st;
  int64_t bytes;
  ScopedTimer(Plan& p, const char* n, hipStream_t s, int64_t b);
  ~ScopedTimer();
};

}  // namespace apx
