// Krylov kernels on the reduced system in SELL-64 layout: SpMV, Jacobi-PCG
// with fused vector updates and deterministic reductions, and the
// reliable-update hook used by mixed-precision CG.
//
// Replaces the matrix-free jvp + jax.scipy.sparse.linalg.cg of
// linear_solve_jax (solver.py:1044-1116) and the host spsolve of
// linear_solve_scipy (:1493-1539) for the 'hip' backend.
//
// One PCG iteration = two streaming kernels, each followed by a one-block
// finalize kernel:
//   k_pcg_spmv  : p = z + beta p_old fused into the gather, q = A p, block
//                 partials of p.q;  k_pcg_alpha_fin forms alpha = gamma / p.q.
//   k_pcg_update: x += alpha p, r -= alpha q, z = D^-1 r, partials of r.z and
//                 r.r;  k_pcg_beta_fin forms beta, gamma, the stop flag and
//                 the reliable-update request.
// Scalars live on the device.  Kernels exit at entry once a flag is set, so
// the host enqueues iterations in batches and reads the flags one batch late.
// The two direction buffers alternate by the device iteration counter.
// Reductions: fixed grid, per-block shuffle tree, the finalize kernel sums
// the partials in index order -> bitwise reproducible.  (A last-block-done
// reduction inside the streaming kernel would need a device-scope fence per
// block, which writes back the L2 of every XCD and cost more than the extra
// launch.)
//
// Bandwidth: every thread issues kUnroll independent element loads per pass
// (vector kernels) and one wavefront owns one 64-row SELL chunk with its
// value / column loads unrolled by 4, so enough requests are in flight to
// cover HBM latency; grids are sized to about one pass per thread.
#include <chrono>
#include <cstring>
#include <type_traits>

#include "assembly.hpp"

namespace apx {

constexpr int kBS = 256;        // vector kernels
constexpr int kSpmvBS = 512;    // SpMV kernels: 8 wavefronts = 8 chunks per block
constexpr int kUnroll = 4;
constexpr int kMaxBlocks = 8192;

ScopedTimer::ScopedTimer(Plan& p, const char* n, hipStream_t s, int64_t b) : P(p), name(n), st(s), bytes(b) {
  if (P.timing) (void)hipEventRecord(P.ev0, st);
}
ScopedTimer::~ScopedTimer() {
  if (!P.timing) return;
  (void)hipEventRecord(P.ev1, st);
  (void)hipEventSynchronize(P.ev1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, P.ev0, P.ev1);
  auto& t = P.timers[name];
  t.total_ms += ms;
  t.count += 1;
  t.bytes = bytes;
}

static inline int grid_for(int64_t n) {
  const int64_t per_block = (int64_t)kBS * kUnroll;
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, kMaxBlocks));
}
static inline int grid_chunks(int64_t n_chunks) {
  constexpr int wpb = kSpmvBS / 64;
  return (int)std::min<int64_t>(round_xcd(std::max<int64_t>(1, (n_chunks + wpb - 1) / wpb)), kMaxBlocks);
}

static_assert(kKrylovPartials == 4 * kMaxBlocks, "partials layout: 4 rows of kMaxBlocks");

void ensure_krylov_scratch(Plan& P, hipStream_t st) {
  if (P.partials.n < (size_t)(4 * kMaxBlocks)) P.partials.alloc(4 * kMaxBlocks);
  if (P.kstate.n < 64) {
    P.kstate.alloc(64);
    APX_HIP(hipMemsetAsync(P.kstate.p, 0, P.kstate.bytes(), st));
  }
}

// Cross-block reductions are split in two launches on the same stream: every
// block writes its partial sums (plain stores), then a one-block finalize
// kernel adds them in index order and updates the device scalars.  The kernel
// boundary makes the partials visible across the XCDs' separate L2 caches; an
// in-kernel last-block scheme would need an agent-scope release per block,
// which writes back the whole L2 each time.
template <int BS, int NV>
__device__ inline void block_partials(double (&v)[NV], double* partials) {
  __shared__ double lds[BS / 64];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double s = block_sum<BS>(v[j], lds);
    if (threadIdx.x == 0) partials[j * kMaxBlocks + blockIdx.x] = s;
  }
}

constexpr int kFinBS = 1024;

// Sum NV rows of nb partials (fixed order) -> valid in all threads.
template <int NV>
__device__ inline void finalize_sums(const double* partials, int nb, double (&sums)[NV]) {
  __shared__ double lds[kFinBS / 64];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double s = 0.0;
#pragma unroll 8
    for (int i = threadIdx.x; i < nb; i += kFinBS) s += partials[j * kMaxBlocks + i];
    sums[j] = block_sum<kFinBS>(s, lds);
  }
}

// Column of a stored entry: an int32 column as stored, or an int16 offset from the row (base)
__device__ __forceinline__ int32_t sell_col(int32_t c, int32_t) { return c; }
__device__ __forceinline__ int32_t sell_col(int16_t d, int32_t base) { return base + (int32_t)d; }
// The row an int16 offset is taken from: the lane's row, or the last row for lanes past n
__device__ __forceinline__ int32_t sell_base(int64_t row, int64_t n) { return (int32_t)(row < n ? row : n - 1); }

// Row gather of one SELL chunk: sum_k vals[k] * f(col[k]) for this lane's row,
// values of type M, accumulated in V (FP32 values with FP64 vectors: each
// value is widened exactly and the products and sums are FP64).
// U entries per step: all their column and value loads are issued before the gathers (U = 8
// for the wide rows of 3-DOF elasticity, 4 otherwise); the sum order is k = 0, 1, 2, ... either way.
template <class M, class V = M, int U = 4, class C, class F>
__device__ __forceinline__ V sell_row(const M* __restrict__ v, const C* __restrict__ ci, int w, int32_t base, F f) {
  V s = 0;
  int k = 0;
  for (; k + U <= w; k += U) {
    // values and columns are read once per SpMV: non-temporal, so L2 keeps the gathered vector
    C c[U];
    M vv[U];
#pragma unroll
    for (int j = 0; j < U; ++j) c[j] = __builtin_nontemporal_load(ci + (k + j) * 64);
#pragma unroll
    for (int j = 0; j < U; ++j) vv[j] = __builtin_nontemporal_load(v + (k + j) * 64);
#pragma unroll
    for (int j = 0; j < U; ++j) s += (V)vv[j] * f(sell_col(c[j], base));
  }
  for (; k < w; ++k)
    s += (V)__builtin_nontemporal_load(v + k * 64) * f(sell_col(__builtin_nontemporal_load(ci + k * 64), base));
  return s;
}

// ---- SpMV (SELL-64: one wavefront per chunk, lane = row) --------------------------
template <class T, class C>
__global__ __launch_bounds__(kSpmvBS) void k_spmv_sell(int64_t n, int64_t n_chunks, const int32_t* __restrict__ sptr,
                                                       const C* __restrict__ cols, const T* __restrict__ vals,
                                                       const T* __restrict__ x, T* __restrict__ y) {
  constexpr int wpb = kSpmvBS / 64;
  const int lane = threadIdx.x & 63;
  for (int64_t c = blockIdx.x * (int64_t)wpb + (threadIdx.x >> 6); c < n_chunks; c += (int64_t)gridDim.x * wpb) {
    const int32_t off = sptr[c], w = (sptr[c + 1] - off) >> 6;
    const int64_t row = c * 64 + lane;
    const T s = sell_row<T>(vals + off + lane, cols + off + lane, w, sell_base(row, n),
                            [&](int32_t col) { return x[col]; });
    if (row < n) y[row] = s;
  }
}

void spmv(Plan& P, const void* vals, const void* x, void* y, int precision, hipStream_t st) {
  const int g = grid_chunks(P.n_chunks);
  auto run = [&](auto cols) {
    using C = std::remove_const_t<std::remove_pointer_t<decltype(cols)>>;
    if (precision == APX_FP64)
      hipLaunchKernelGGL((k_spmv_sell<double, C>), dim3(g), dim3(kSpmvBS), 0, st, P.n_free, P.n_chunks, P.sell_ptr.p,
                         cols, (const double*)vals, (const double*)x, (double*)y);
    else
      hipLaunchKernelGGL((k_spmv_sell<float, C>), dim3(g), dim3(kSpmvBS), 0, st, P.n_free, P.n_chunks, P.sell_ptr.p,
                         cols, (const float*)vals, (const float*)x, (float*)y);
  };
  if (P.sell_dcol.p)
    run((const int16_t*)P.sell_dcol.p);
  else
    run((const int32_t*)P.sell_cols.p);
  APX_LAUNCH_CHECK();
}

// ---- layout conversion --------------------------------------------------------------
template <class T>
__global__ void k_sell_to_csr(int64_t nnz, const int32_t* csr2sell, const T* sell, T* csr) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    csr[k] = sell[csr2sell[k]];
}
template <class T>
__global__ void k_csr_to_sell(int64_t nnz, const int32_t* csr2sell, const T* csr, T* sell) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    sell[csr2sell[k]] = csr[k];
}

void sell_to_csr(Plan& P, const void* sell, void* csr, int precision, hipStream_t st) {
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((P.nnz + 255) / 256, 65536));
  if (precision == APX_FP64)
    hipLaunchKernelGGL(k_sell_to_csr<double>, dim3(g), dim3(256), 0, st, P.nnz, P.csr2sell.p, (const double*)sell,
                       (double*)csr);
  else
    hipLaunchKernelGGL(k_sell_to_csr<float>, dim3(g), dim3(256), 0, st, P.nnz, P.csr2sell.p, (const float*)sell,
                       (float*)csr);
  APX_LAUNCH_CHECK();
}

void csr_to_sell(Plan& P, const void* csr, void* sell, int precision, hipStream_t st) {
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((P.nnz + 255) / 256, 65536));
  if (precision == APX_FP64) {
    fill<double>((double*)sell, P.sell_total, 0.0, st);
    hipLaunchKernelGGL(k_csr_to_sell<double>, dim3(g), dim3(256), 0, st, P.nnz, P.csr2sell.p, (const double*)csr,
                       (double*)sell);
  } else {
    fill<float>((float*)sell, P.sell_total, 0.0f, st);
    hipLaunchKernelGGL(k_csr_to_sell<float>, dim3(g), dim3(256), 0, st, P.nnz, P.csr2sell.p, (const float*)csr,
                       (float*)sell);
  }
  APX_LAUNCH_CHECK();
}

// ---- Jacobi ------------------------------------------------------------------------
template <class M, class V = M>
__global__ void k_inv_diag(int64_t n, const int32_t* diag_slot, const M* vals, V* inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    inv[i] = V(1) / (V)vals[diag_slot[i]];
}

void jacobi_inverse(Plan& P, const void* vals, void* inv, int precision, hipStream_t st) {
  const int g = std::max(1, (int)std::min<int64_t>((P.n_free + 255) / 256, 65536));
  if (precision == APX_FP64)
    hipLaunchKernelGGL(k_inv_diag<double>, dim3(g), dim3(256), 0, st, P.n_free, P.diag_slot.p, (const double*)vals,
                       (double*)inv);
  else
    hipLaunchKernelGGL(k_inv_diag<float>, dim3(g), dim3(256), 0, st, P.n_free, P.diag_slot.p, (const float*)vals,
                       (float*)inv);
  APX_LAUNCH_CHECK();
}

void jacobi_inverse_f32_to_f64(Plan& P, const float* vals, double* inv, hipStream_t st) {
  const int g = std::max(1, (int)std::min<int64_t>((P.n_free + 255) / 256, 65536));
  hipLaunchKernelGGL((k_inv_diag<float, double>), dim3(g), dim3(256), 0, st, P.n_free, P.diag_slot.p, vals, inv);
  APX_LAUNCH_CHECK();
}

// ---- PCG kernels --------------------------------------------------------------------
__device__ __forceinline__ bool halted(const KState* S) { return S->done | S->reliable; }

// Elementwise pass with kUnroll independent elements per thread and pass.
#define APX_FOR_ELEMENTS(n, BODY_LOAD, BODY_STORE)                                         \
  for (int64_t base = blockIdx.x * (int64_t)kBS * kUnroll + threadIdx.x; base < (n);        \
       base += (int64_t)gridDim.x * kBS * kUnroll) {                                        \
    _Pragma("unroll") for (int j = 0; j < kUnroll; ++j) {                                   \
      const int64_t i = base + (int64_t)j * kBS;                                            \
      if (i < (n)) { BODY_LOAD }                                                            \
    }                                                                                       \
    _Pragma("unroll") for (int j = 0; j < kUnroll; ++j) {                                   \
      const int64_t i = base + (int64_t)j * kBS;                                            \
      if (i < (n)) { BODY_STORE }                                                           \
    }                                                                                       \
  }

template <class T>
__global__ __launch_bounds__(kBS) void k_pcg_init(int64_t n, const T* __restrict__ b, const T* __restrict__ inv,
                                                  T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                  T* __restrict__ pa, KState* S, double* partials, double tol,
                                                  double atol, int maxiter, double delta) {
  double v[2] = {0.0, 0.0};
  T bv[kUnroll], iv[kUnroll];
  APX_FOR_ELEMENTS(n, { bv[j] = b[i]; iv[j] = inv[i]; }, {
    const T zi = iv[j] * bv[j];
    x[i] = 0;
    r[i] = bv[j];
    z[i] = zi;
    pa[i] = 0;
    v[0] += (double)bv[j] * (double)zi;
    v[1] += (double)bv[j] * (double)bv[j];
  })
  block_partials<kBS, 2>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_pcg_init_fin(KState* S, const double* partials, int nb, double tol,
                                                         double atol, int maxiter, double delta) {
  double s[2];
  finalize_sums<2>(partials, nb, s);
  if (threadIdx.x == 0) {
    S->gamma = s[0];
    S->gamma_prev = s[0];
    S->rr = s[1];
    S->rr0 = s[1];
    S->bb = s[1];
    S->rmax2 = s[1];
    S->delta2 = delta * delta;
    S->atol2 = fmax(tol * tol * s[1], atol * atol);
    S->beta = 0.0;
    S->k = 0;
    S->maxiter = maxiter;
    S->reliable = 0;
    S->done = !(s[1] > S->atol2 && 0 < maxiter);
  }
}

template <class M, class T, int U, class C>
__global__ __launch_bounds__(kSpmvBS) void k_pcg_spmv(int64_t n, int64_t n_chunks, const int32_t* __restrict__ sptr,
                                                      const C* __restrict__ cols, const M* __restrict__ vals,
                                                      const T* __restrict__ z, T* __restrict__ pa, T* __restrict__ pb,
                                                      T* __restrict__ q, KState* S, double* partials) {
  if (halted(S)) return;
  constexpr int wpb = kSpmvBS / 64;
  const T beta = (T)S->beta;
  const bool odd = S->k & 1;
  const T* __restrict__ po = odd ? pb : pa;
  T* __restrict__ pn = odd ? pa : pb;
  const int lane = threadIdx.x & 63;
  double v[1] = {0.0};
  for (int64_t c = blockIdx.x * (int64_t)wpb + (threadIdx.x >> 6); c < n_chunks; c += (int64_t)gridDim.x * wpb) {
    const int32_t off = sptr[c], w = (sptr[c + 1] - off) >> 6;
    const int64_t row = c * 64 + lane;
    const T s = sell_row<M, T, U>(vals + off + lane, cols + off + lane, w, sell_base(row, n),
                                  [&](int32_t col) { return z[col] + beta * po[col]; });
    if (row < n) {
      const T pi = z[row] + beta * po[row];
      pn[row] = pi;
      q[row] = s;
      v[0] += (double)pi * (double)s;
    }
  }
  block_partials<kSpmvBS, 1>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_pcg_alpha_fin(KState* S, const double* partials, int nb) {
  if (halted(S)) return;
  double tot[1];
  finalize_sums<1>(partials, nb, tot);
  if (threadIdx.x == 0) {
    S->pAp = tot[0];
    S->alpha = S->gamma / tot[0];
  }
}

template <class T>
__global__ __launch_bounds__(kBS) void k_pcg_update(int64_t n, const T* __restrict__ inv, const T* __restrict__ pa,
                                                    const T* __restrict__ pb, const T* __restrict__ q,
                                                    T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                    KState* S, double* partials) {
  if (halted(S)) return;
  const T alpha = (T)S->alpha;
  const T* __restrict__ p = (S->k & 1) ? pa : pb;  // written by k_pcg_spmv of this iteration
  double v[2] = {0.0, 0.0};
  T pv[kUnroll], qv[kUnroll], xv[kUnroll], rv[kUnroll], iv[kUnroll];
  APX_FOR_ELEMENTS(n, {
    pv[j] = p[i];
    qv[j] = q[i];
    xv[j] = x[i];
    rv[j] = r[i];
    iv[j] = inv[i];
  }, {
    x[i] = xv[j] + alpha * pv[j];
    const T ri = rv[j] - alpha * qv[j];
    const T zi = iv[j] * ri;
    r[i] = ri;
    z[i] = zi;
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  })
  block_partials<kBS, 2>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_pcg_beta_fin(KState* S, const double* partials, int nb) {
  if (halted(S)) return;
  double s[2];
  finalize_sums<2>(partials, nb, s);
  if (threadIdx.x == 0) {
    S->gamma_prev = S->gamma;
    S->beta = s[0] / S->gamma;
    S->gamma = s[0];
    S->rr = s[1];
    S->k += 1;
    S->done = !(s[1] > S->atol2 && S->k < S->maxiter);
    if (!S->done && S->delta2 > 0.0 && s[1] < S->delta2 * S->rmax2) S->reliable = 1;
  }
}

// Replace the recursive residual by b_new = -r_full[free] (a true residual
// computed in FP64 by the caller) and zero the correction x; the direction p
// is kept (reliable update, Sleijpen & van der Vorst 1996).
template <class T>
__global__ __launch_bounds__(kBS) void k_pcg_reanchor(int64_t n, const int32_t* __restrict__ red2full,
                                                      const double* __restrict__ rfull, const T* __restrict__ inv,
                                                      T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                      KState* S, double* partials) {
  double v[2] = {0.0, 0.0};
  T rv[kUnroll], iv[kUnroll];
  APX_FOR_ELEMENTS(n, { rv[j] = (T)(-rfull[red2full[i]]); iv[j] = inv[i]; }, {
    const T zi = iv[j] * rv[j];
    x[i] = 0;
    r[i] = rv[j];
    z[i] = zi;
    v[0] += (double)rv[j] * (double)zi;
    v[1] += (double)rv[j] * (double)rv[j];
  })
  block_partials<kBS, 2>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_pcg_reanchor_fin(KState* S, const double* partials, int nb) {
  double s[2];
  finalize_sums<2>(partials, nb, s);
  if (threadIdx.x == 0) {
    S->beta = s[0] / S->gamma_prev;
    S->gamma = s[0];
    S->rr = s[1];
    S->rmax2 = s[1];
    S->reliable = 0;
    S->done = !(s[1] > S->atol2 && S->k < S->maxiter);
  }
}

// ---- PCG driver ----------------------------------------------------------------------
template <class M, class T>
PcgRun<M, T>::PcgRun(Plan& P_, const M* vals_, const T* inv_, hipStream_t st_)
    : P(P_), vals(vals_), inv(inv_), st(st_) {
  n = P.n_free;
  T* w;
  if constexpr (sizeof(T) == 8) {
    if (P.w64.n < (size_t)(6 * n)) P.w64.alloc(6 * n);
    w = (T*)P.w64.p;
  } else {
    if (P.w32.n < (size_t)(6 * n)) P.w32.alloc(6 * n);
    w = (T*)P.w32.p;
  }
  r = w; z = w + n; pa = w + 2 * n; pb = w + 3 * n; q = w + 4 * n; xw = w + 5 * n;
  ensure_krylov_scratch(P, st);
  S = (KState*)P.kstate.p;
  part = P.partials.p;
  APX_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
  APX_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
  // SELL values + columns (int16 offsets or int32), chunk offsets, z and p_old read, p_new and q written
  spmv_bytes = P.sell_total * (int64_t)(sizeof(M) + (P.sell_dcol.p ? 2 : 4)) + (P.n_chunks + 1) * 4 +
               n * 4 * (int64_t)sizeof(T);
  upd_bytes = n * 8 * (int64_t)sizeof(T);  // reads D^-1, p, q, x, r; writes x, r, z
}

template <class M, class T>
PcgRun<M, T>::PcgRun(Plan& P_, const M* vals_, const T* inv_, hipStream_t st_, T* w, double* partials, KState* state)
    : P(P_), vals(vals_), inv(inv_), st(st_), S(state), part(partials), external(true) {
  n = P.n_free;
  r = w; z = w + n; pa = w + 2 * n; pb = w + 3 * n; q = w + 4 * n; xw = w + 5 * n;
}

template <class M, class T>
PcgRun<M, T>::~PcgRun() {
  if (evs[0]) (void)hipEventDestroy(evs[0]);
  if (evs[1]) (void)hipEventDestroy(evs[1]);
}

template <class M, class T>
void PcgRun<M, T>::init(const T* b, T* x_, double tol, double atol, int maxiter_, double delta) {
  x = x_ ? x_ : xw;
  maxiter = maxiter_ <= 0 ? (int)std::min<int64_t>(10 * n, INT32_MAX / 2) : maxiter_;
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(g), dim3(kBS), 0, st, n, b, inv, x, r, z, pa, S, part, tol, atol,
                     maxiter, delta);
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(kFinBS), 0, st, S, part, g, tol, atol, maxiter, delta);
  APX_LAUNCH_CHECK();
}

template <class M, class T>
void PcgRun<M, T>::launch_iteration() {
  const int gs = grid_chunks(P.n_chunks), gu = grid_for(n);
  const bool timed = P.timing && !external;
  // timed: the pairs take the SpMV's and the update's own start and end timestamps
  const EventPair e1 = timed ? timing_pair(P) : EventPair{}, e2 = timed ? timing_pair(P) : EventPair{};
  t_launch_events = timed ? &e1 : nullptr;
  auto spmv = [&](auto cols) {
    using C = std::remove_const_t<std::remove_pointer_t<decltype(cols)>>;
    if (P.sell_total >= 32 * (int64_t)P.n_free)  // rows of 32+ stored entries: deeper load batches
      launch_kernel((k_pcg_spmv<M, T, 8, C>), dim3(gs), dim3(kSpmvBS), 0, st, n, P.n_chunks, P.sell_ptr.p, cols, vals,
                    z, pa, pb, q, S, part);
    else
      launch_kernel((k_pcg_spmv<M, T, 4, C>), dim3(gs), dim3(kSpmvBS), 0, st, n, P.n_chunks, P.sell_ptr.p, cols, vals,
                    z, pa, pb, q, S, part);
  };
  if (P.sell_dcol.p)
    spmv((const int16_t*)P.sell_dcol.p);
  else
    spmv((const int32_t*)P.sell_cols.p);
  hipLaunchKernelGGL(k_pcg_alpha_fin, dim3(1), dim3(kFinBS), 0, st, S, part, gs);
  t_launch_events = timed ? &e2 : nullptr;
  launch_kernel(k_pcg_update<T>, dim3(gu), dim3(kBS), 0, st, n, inv, pa, pb, q, x, r, z, S, part);
  hipLaunchKernelGGL(k_pcg_beta_fin, dim3(1), dim3(kFinBS), 0, st, S, part, gu);
  if (timed) {
    ev_spmv.push_back(e1);
    ev_upd.push_back(e2);
  }
}

template <class M, class T>
KState PcgRun<M, T>::run() {
  constexpr int kBatch = 16;
  volatile int* flag = (volatile int*)P.pinned;  // [slot*4 + {k, maxiter, done, reliable}]
  auto enqueue_batch = [&](int slot) {
    for (int i = 0; i < kBatch; ++i) launch_iteration();
    APX_LAUNCH_CHECK();
    APX_HIP(hipMemcpyAsync((void*)(flag + 4 * slot), &S->k, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    APX_HIP(hipEventRecord(evs[slot], st));
    launched += kBatch;
  };
  auto halted_host = [&](int slot) { return flag[4 * slot + 2] != 0 || flag[4 * slot + 3] != 0; };
  // flags after init / reanchor
  APX_HIP(hipMemcpyAsync((void*)(flag + 8), &S->k, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
  APX_HIP(hipStreamSynchronize(st));
  if (external && !halted_host(2)) {
    // no events: enqueue a batch, copy the flags, wait for the stream, test
    const int64_t limit = (int64_t)launched + maxiter + kBatch;
    for (;;) {
      for (int i = 0; i < kBatch; ++i) launch_iteration();
      APX_LAUNCH_CHECK();
      launched += kBatch;
      APX_HIP(hipMemcpyAsync((void*)flag, &S->k, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
      APX_HIP(hipStreamSynchronize(st));
      if (halted_host(0) || launched >= limit) break;
    }
  } else if (!halted_host(2)) {
    const int64_t limit = (int64_t)launched + maxiter + 2 * kBatch;
    enqueue_batch(0);
    int slot = 0;
    for (;;) {
      enqueue_batch(slot ^ 1);
      APX_HIP(hipEventSynchronize(evs[slot]));
      if (halted_host(slot) || launched >= limit) break;
      slot ^= 1;
    }
  }
  KState hs;
  if (external) {
    // through the plan's pinned block (bytes 768..1023), so no pageable staging is involved
    static_assert(sizeof(KState) <= 256, "KState must fit the pinned slot");
    char* slot = (char*)P.pinned + 768;
    APX_HIP(hipMemcpyAsync(slot, S, sizeof(KState), hipMemcpyDeviceToHost, st));
    APX_HIP(hipStreamSynchronize(st));
    std::memcpy(&hs, slot, sizeof(KState));
  } else {
    APX_HIP(hipStreamSynchronize(st));
    APX_HIP(hipMemcpy(&hs, S, sizeof(KState), hipMemcpyDeviceToHost));
  }
  iterations = hs.k;
  return hs;
}

template <class M, class T>
void PcgRun<M, T>::reanchor(const double* r_full) {
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_pcg_reanchor<T>, dim3(g), dim3(kBS), 0, st, n, P.red2full.p, r_full, inv, x, r, z, S, part);
  hipLaunchKernelGGL(k_pcg_reanchor_fin, dim3(1), dim3(kFinBS), 0, st, S, part, g);
  APX_LAUNCH_CHECK();
}

template <class M, class T>
void PcgRun<M, T>::collect_timing() {
  if (!P.timing || external) return;
  const bool f64 = sizeof(T) == 8;
  // only launches that executed an iteration are counted; the rest exited at entry
  const size_t k = (size_t)iterations - counted;
  timing_collect(P, sizeof(M) == sizeof(T) ? (f64 ? "pcg_spmv_f64" : "pcg_spmv_f32") : "pcg_spmv_f32m_f64v", ev_spmv,
                 spmv_bytes, k);
  timing_collect(P, f64 ? "pcg_update_f64" : "pcg_update_f32", ev_upd, upd_bytes, k);
  counted = iterations;
}

template struct PcgRun<double>;
template struct PcgRun<float>;
template struct PcgRun<float, double>;

template <class T>
static apx_krylov_stats pcg_t(Plan& P, const T* vals, const T* inv, const T* b, T* x, double tol, double atol,
                              int maxiter, hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  PcgRun<T> run(P, vals, inv, st);
  run.init(b, x, tol, atol, maxiter, 0.0);
  KState hs = run.run();
  run.collect_timing();
  apx_krylov_stats out{};
  out.iterations = hs.k;
  out.converged = hs.rr <= hs.atol2;
  out.initial_residual = sqrt(hs.rr0);
  out.final_residual = sqrt(hs.rr);
  out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return out;
}

apx_krylov_stats pcg(Plan& P, const void* vals, const void* inv, const void* b, void* x, int precision, double tol,
                     double atol, int maxiter, hipStream_t st) {
  if (precision == APX_FP64)
    return pcg_t<double>(P, (const double*)vals, (const double*)inv, (const double*)b, (double*)x, tol, atol,
                         maxiter, st);
  return pcg_t<float>(P, (const float*)vals, (const float*)inv, (const float*)b, (float*)x, tol, atol, maxiter, st);
}

// ---- Newton helpers -----------------------------------------------------------------
__global__ __launch_bounds__(kBS) void k_free_norm(int64_t n, const int32_t* red2full, const double* r, KState* S,
                                                   double* partials) {
  double v[1] = {0.0};
  double rv[kUnroll];
  APX_FOR_ELEMENTS(n, { rv[j] = r[red2full[i]]; }, { v[0] += rv[j] * rv[j]; })
  block_partials<kBS, 1>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_norm_fin(KState* S, const double* partials, int nb) {
  double s[1];
  finalize_sums<1>(partials, nb, s);
  if (threadIdx.x == 0) S->dot_out = sqrt(s[0]);
}

double free_norm(Plan& P, const double* r_full, hipStream_t st) {
  ensure_krylov_scratch(P, st);
  KState* S = (KState*)P.kstate.p;
  const int g = grid_for(P.n_free);
  hipLaunchKernelGGL(k_free_norm, dim3(g), dim3(kBS), 0, st, P.n_free, P.red2full.p, r_full, S, P.partials.p);
  hipLaunchKernelGGL(k_norm_fin, dim3(1), dim3(kFinBS), 0, st, S, P.partials.p, g);
  APX_LAUNCH_CHECK();
  APX_HIP(hipMemcpyAsync(P.pinned + 16, &S->dot_out, sizeof(double), hipMemcpyDeviceToHost, st));
  APX_HIP(hipStreamSynchronize(st));
  return P.pinned[16];
}

template <class T>
__global__ void k_gather_rhs(int64_t n, const int32_t* red2full, const double* r, T* b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (T)(-r[red2full[i]]);
}

void gather_rhs(Plan& P, const double* r_full, void* b, int precision, hipStream_t st) {
  const int g = std::max(1, (int)std::min<int64_t>((P.n_free + 255) / 256, 65536));
  if (precision == APX_FP64)
    hipLaunchKernelGGL(k_gather_rhs<double>, dim3(g), dim3(256), 0, st, P.n_free, P.red2full.p, r_full, (double*)b);
  else
    hipLaunchKernelGGL(k_gather_rhs<float>, dim3(g), dim3(256), 0, st, P.n_free, P.red2full.p, r_full, (float*)b);
  APX_LAUNCH_CHECK();
}

template <class T>
__global__ void k_scatter_update(int64_t n, const int32_t* red2full, double* u, const T* x, double damping) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[red2full[i]] += damping * (double)x[i];
}

void scatter_update(Plan& P, double* u, const void* x, double damping, int precision, hipStream_t st) {
  const int g = std::max(1, (int)std::min<int64_t>((P.n_free + 255) / 256, 65536));
  if (precision == APX_FP64)
    hipLaunchKernelGGL(k_scatter_update<double>, dim3(g), dim3(256), 0, st, P.n_free, P.red2full.p, u,
                       (const double*)x, damping);
  else
    hipLaunchKernelGGL(k_scatter_update<float>, dim3(g), dim3(256), 0, st, P.n_free, P.red2full.p, u,
                       (const float*)x, damping);
  APX_LAUNCH_CHECK();
}

__global__ void k_impose(int64_t n, const uint8_t* mask, double* u, const double* vals) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (mask[i]) u[i] = vals[i];
}

void impose_dirichlet(Plan& P, double* u, const double* values, hipStream_t st) {
  const int g = std::max(1, (int)std::min<int64_t>((P.n_dofs + 255) / 256, 65536));
  hipLaunchKernelGGL(k_impose, dim3(g), dim3(256), 0, st, P.n_dofs, P.dmask.p, u, values);
  APX_LAUNCH_CHECK();
}

}  // namespace apx
