// Internal interfaces shared by the library's translation units.
#pragma once
#include "plan.hpp"

namespace apx {

template <class T>
void fill(T* p, int64_t n, T v, hipStream_t s);
// dst = src rounded to FP32 (nearest), element by element.
void round_to_f32(const double* src, float* dst, int64_t n, hipStream_t s);

void assemble_residual(Plan& P, const double* u, double* r, hipStream_t st);
void assemble_tangent(Plan& P, const double* u, void* vals, int precision, hipStream_t st);
// precision code of "FP64 element arithmetic, values rounded once to FP32"
constexpr int kFP64RoundedToFP32 = APX_FP64_ROUNDED;
// The FP64 tangent rounded once to FP32 into vals (FP32, SELL order): written directly by the
// owner-tile kernels, else assembled in FP64 into the plan's scratch values and rounded.
void assemble_tangent_rounded(Plan& P, const double* u, float* vals, hipStream_t st);
void assemble_diagonal(Plan& P, const double* u, double* diag, hipStream_t st);
// full tangent as BCOO with duplicates (data only; the reference's element-major order)
int64_t tangent_coo_entries(const Plan& P);
void assemble_tangent_coo(Plan& P, const double* u, void* data, int precision, hipStream_t st);
int64_t residual_bytes(const Plan& P);
int64_t tangent_bytes(const Plan& P, int precision);
// one set's residual through the colour kernels (r += ...)
void assemble_set_colour_residual(Plan& P, ElemSet& S, const double* u, double* r, hipStream_t st);

// structured.hip: owner-tile kernels for structured Q1 grids
// mode: 0 residual, 1 tangent, 2 diagonal
bool tile_supported(const Plan& P, const ElemSet& S, int mode);
bool tangent_by_tiles(const Plan& P);
void tile_residual(Plan& P, ElemSet& S, const double* u, double* r, bool overwrite, hipStream_t st);
void tile_tangent(Plan& P, ElemSet& S, const double* u, void* vals, int precision, hipStream_t st);
void tile_diagonal(Plan& P, ElemSet& S, const double* u, double* diag, bool overwrite, hipStream_t st);

// krylov.hip
struct KState {
  double gamma, gamma_prev, pAp, alpha, beta, rr, atol2, bb, rr0, rmax2, delta2;
  double dot_out;  // scratch result of dot / norm kernels
  double rsq;      // flexible CG: sum of squares of the network's input, r scaled by D^-1/2
  int k, maxiter, done, reliable;  // contiguous: copied to the host as one block
};

void ensure_krylov_scratch(Plan& P, hipStream_t st);
// reduction partials of the Krylov kernels: 4 rows of kMaxBlocks = 8192 doubles
constexpr int64_t kKrylovPartials = 4 * 8192;
void spmv(Plan& P, const void* vals, const void* x, void* y, int precision, hipStream_t st);
void sell_to_csr(Plan& P, const void* sell, void* csr, int precision, hipStream_t st);
void csr_to_sell(Plan& P, const void* csr, void* sell, int precision, hipStream_t st);
void jacobi_inverse(Plan& P, const void* vals, void* inv_diag, int precision, hipStream_t st);
// FP64 inverse diagonal of an FP32 SELL matrix
void jacobi_inverse_f32_to_f64(Plan& P, const float* vals, double* inv_diag, hipStream_t st);

// One preconditioned-CG solve on the plan's SELL matrix with values of type M
// and Krylov vectors and scalars of type T (M = float, T = double: an FP32
// matrix with FP64 vectors).  Workspace (r, z, two directions, q, x) comes
// from the plan.  run() returns when the device sets the stop flag or requests
// a reliable update (residual drop by delta since the last anchor); reanchor()
// then installs a new true residual.
template <class M, class T = M>
struct PcgRun {
  Plan& P;
  const M* vals;
  const T* inv;
  hipStream_t st;
  int64_t n = 0;
  T *r = nullptr, *z = nullptr, *pa = nullptr, *pb = nullptr, *q = nullptr, *xw = nullptr, *x = nullptr;
  KState* S = nullptr;
  double* part = nullptr;  // reduction partials
  // external = true: workspace, partials and scalars come from the caller, and the run
  // creates no events, records no timing and polls the stop flag synchronously (allocates
  // nothing)
  bool external = false;
  int maxiter = 0, iterations = 0;
  int64_t launched = 0;
  size_t counted = 0;
  int64_t spmv_bytes = 0, upd_bytes = 0;
  hipEvent_t evs[2] = {nullptr, nullptr};
  std::vector<EventPair> ev_spmv, ev_upd;

  PcgRun(Plan& P, const M* vals, const T* inv, hipStream_t st);
  // caller-provided workspace: work = 6 n_free vectors, partials = kKrylovPartials doubles,
  // state = one KState (its fields are all written by init before they are read)
  PcgRun(Plan& P, const M* vals, const T* inv, hipStream_t st, T* work, double* partials, KState* state);
  ~PcgRun();
  void init(const T* b, T* x, double tol, double atol, int maxiter, double delta);
  void launch_iteration();
  KState run();
  void reanchor(const double* r_full);
  void collect_timing();
};

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
