// Krylov kernels on the reduced CSR system: SpMV, Jacobi-PCG with fused
// vector updates, deterministic reductions.
//
// Replaces the matrix-free jvp + jax.scipy.sparse.linalg.cg of
// linear_solve_jax (solver.py:1044-1116) and the host spsolve of
// linear_solve_scipy (:1493-1539) for the 'hip' backend.
//
// One PCG iteration = two kernels:
//   k_pcg_spmv : p_new = z + beta p_old (fused into the row gather), q = A p_new,
//                block partials of p.q; the last block forms alpha = gamma / p.q
//   k_pcg_update: x += alpha p, r -= alpha q, z = D^-1 r, partials of r.z and r.r;
//                the last block forms beta, the new gamma and the stop flag.
// Scalars live on the device; kernels exit at entry once the stop flag is set,
// so the host enqueues iterations in batches and reads the flag one batch late.
// Reductions: fixed grid, per-block shuffle tree, last block sums the
// partials in index order -> bitwise reproducible.
#include <chrono>
#include <cstring>

#include "assembly.hpp"

namespace apx {

constexpr int kBS = 256;
constexpr int kMaxBlocks = 1024;

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

static inline int grid_for(int64_t work_items, int bs = kBS) {
  int64_t b = (work_items + bs - 1) / bs;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kMaxBlocks));
}

// Last-block-done reduction of per-block partials (NV values per block).
// Returns true in every thread of the last block; sums[] then holds totals.
template <int NV>
__device__ inline bool last_block_reduce(double (&v)[NV], double* partials, unsigned int* counter,
                                         double (&sums)[NV]) {
  __shared__ double lds[kBS / 64];
  __shared__ bool is_last;
  double bs[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) bs[j] = block_sum<kBS>(v[j], lds);
  // Partials cross XCDs (separate L2s): agent-scope atomics + fences make the
  // stores of every block visible to the last one.
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      __hip_atomic_store(&partials[j * kMaxBlocks + blockIdx.x], bs[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double s = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kBS)
      s += __hip_atomic_load(&partials[j * kMaxBlocks + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sums[j] = block_sum<kBS>(s, lds);
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

// ---- SpMV ------------------------------------------------------------------------
template <class T, int G>
__global__ __launch_bounds__(kBS) void k_spmv(int64_t n, const int32_t* __restrict__ indptr,
                                              const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                              const T* __restrict__ x, T* __restrict__ y) {
  const int lane = threadIdx.x % G;
  for (int64_t row = (blockIdx.x * (int64_t)kBS + threadIdx.x) / G; row < n;
       row += (int64_t)gridDim.x * (kBS / G)) {
    const int32_t b = indptr[row], e = indptr[row + 1];
    T s = 0;
    for (int32_t k = b + lane; k < e; k += G) s += vals[k] * x[indices[k]];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (lane == 0) y[row] = s;
  }
}

static int lanes_per_row(const Plan& P) {
  const double avg = P.n_free ? (double)P.nnz / (double)P.n_free : 1.0;
  if (avg <= 12) return 4;
  if (avg <= 24) return 8;
  if (avg <= 48) return 16;
  return 32;
}

template <class T>
static void spmv_t(Plan& P, const T* vals, const T* x, T* y, hipStream_t st) {
  const int G = lanes_per_row(P);
  const int64_t rows_per_block = kBS / G;
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((P.n_free + rows_per_block - 1) / rows_per_block, 1 << 20));
  switch (G) {
    case 4: hipLaunchKernelGGL((k_spmv<T, 4>), dim3(blocks), dim3(kBS), 0, st, P.n_free, P.indptr.p, P.indices.p, vals, x, y); break;
    case 8: hipLaunchKernelGGL((k_spmv<T, 8>), dim3(blocks), dim3(kBS), 0, st, P.n_free, P.indptr.p, P.indices.p, vals, x, y); break;
    case 16: hipLaunchKernelGGL((k_spmv<T, 16>), dim3(blocks), dim3(kBS), 0, st, P.n_free, P.indptr.p, P.indices.p, vals, x, y); break;
    default: hipLaunchKernelGGL((k_spmv<T, 32>), dim3(blocks), dim3(kBS), 0, st, P.n_free, P.indptr.p, P.indices.p, vals, x, y); break;
  }
  APX_LAUNCH_CHECK();
}

void spmv(Plan& P, const void* vals, const void* x, void* y, int precision, hipStream_t st) {
  if (precision == APX_FP64)
    spmv_t<double>(P, (const double*)vals, (const double*)x, (double*)y, st);
  else
    spmv_t<float>(P, (const float*)vals, (const float*)x, (float*)y, st);
}

// ---- Jacobi ------------------------------------------------------------------------
template <class T>
__global__ void k_inv_diag(int64_t n, const int32_t* diag_slot, const T* vals, T* inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    inv[i] = T(1) / vals[diag_slot[i]];
}

void jacobi_inverse(Plan& P, const void* vals, void* inv, int precision, hipStream_t st) {
  const int g = (int)std::min<int64_t>((P.n_free + 255) / 256, 65536);
  if (precision == APX_FP64)
    hipLaunchKernelGGL(k_inv_diag<double>, dim3(std::max(g, 1)), dim3(256), 0, st, P.n_free, P.diag_slot.p,
                       (const double*)vals, (double*)inv);
  else
    hipLaunchKernelGGL(k_inv_diag<float>, dim3(std::max(g, 1)), dim3(256), 0, st, P.n_free, P.diag_slot.p,
                       (const float*)vals, (float*)inv);
  APX_LAUNCH_CHECK();
}

// ---- PCG ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBS) void k_pcg_init(int64_t n, const T* __restrict__ b, const T* __restrict__ inv,
                                                  T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                  T* __restrict__ p_old, KState* S, double* partials,
                                                  double tol, double atol, int maxiter) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const T bi = b[i];
    const T zi = inv[i] * bi;
    x[i] = 0;
    r[i] = bi;
    z[i] = zi;
    p_old[i] = 0;
    v[0] += (double)bi * (double)zi;
    v[1] += (double)bi * (double)bi;
  }
  double s[2];
  if (last_block_reduce<2>(v, partials, &S->counter[0], s) && threadIdx.x == 0) {
    S->gamma = s[0];
    S->rr = s[1];
    S->rr0 = s[1];
    S->bb = s[1];
    S->atol2 = fmax(tol * tol * s[1], atol * atol);
    S->beta = 0.0;
    S->k = 0;
    S->maxiter = maxiter;
    S->stagnated = 0;
    S->done = !(s[1] > S->atol2 && 0 < maxiter);
  }
}

template <class T, int G>
__global__ __launch_bounds__(kBS) void k_pcg_spmv(int64_t n, const int32_t* __restrict__ indptr,
                                                  const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                                  const T* __restrict__ z, const T* __restrict__ p_old,
                                                  T* __restrict__ p_new, T* __restrict__ q, KState* S,
                                                  double* partials) {
  if (S->done) return;
  const T beta = (T)S->beta;
  const int lane = threadIdx.x % G;
  double v[1] = {0.0};
  for (int64_t row = (blockIdx.x * (int64_t)kBS + threadIdx.x) / G; row < n;
       row += (int64_t)gridDim.x * (kBS / G)) {
    const int32_t b = indptr[row], e = indptr[row + 1];
    T s = 0;
    for (int32_t k = b + lane; k < e; k += G) {
      const int32_t c = indices[k];
      s += vals[k] * (z[c] + beta * p_old[c]);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (lane == 0) {
      const T pi = z[row] + beta * p_old[row];
      p_new[row] = pi;
      q[row] = s;
      v[0] += (double)pi * (double)s;
    }
  }
  double tot[1];
  if (last_block_reduce<1>(v, partials, &S->counter[1], tot) && threadIdx.x == 0) {
    S->pAp = tot[0];
    S->alpha = S->gamma / tot[0];
  }
}

template <class T>
__global__ __launch_bounds__(kBS) void k_pcg_update(int64_t n, const T* __restrict__ inv, const T* __restrict__ p,
                                                    const T* __restrict__ q, T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, KState* S, double* partials) {
  if (S->done) return;
  const T alpha = (T)S->alpha;
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    const T zi = inv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  }
  double s[2];
  if (last_block_reduce<2>(v, partials, &S->counter[2], s) && threadIdx.x == 0) {
    S->beta = s[0] / S->gamma;
    S->gamma = s[0];
    S->rr = s[1];
    S->k += 1;
    S->done = !(s[1] > S->atol2 && S->k < S->maxiter);
  }
}

template <class T>
static apx_krylov_stats pcg_t(Plan& P, const T* vals, const T* inv, const T* b, T* x, double tol, double atol,
                              int maxiter, hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  const int64_t n = P.n_free;
  // workspace: r, z, pA, pB, q
  T* w;
  if constexpr (sizeof(T) == 8) {
    if (P.w64.n < (size_t)(5 * n)) P.w64.alloc(5 * n);
    w = (T*)P.w64.p;
  } else {
    if (P.w32.n < (size_t)(5 * n)) P.w32.alloc(5 * n);
    w = (T*)P.w32.p;
  }
  T *r = w, *z = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *q = w + 4 * n;
  if (P.partials.n < (size_t)(4 * kMaxBlocks)) P.partials.alloc(4 * kMaxBlocks);
  if (P.kstate.n < 64) {
    P.kstate.alloc(64);
    APX_HIP(hipMemsetAsync(P.kstate.p, 0, P.kstate.bytes(), st));
  }
  KState* S = (KState*)P.kstate.p;
  if (maxiter <= 0) maxiter = (int)std::min<int64_t>(10 * n, INT32_MAX / 2);

  const int gv = grid_for(n);
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(gv), dim3(kBS), 0, st, n, b, inv, x, r, z, pa, S, P.partials.p, tol, atol,
                     maxiter);
  APX_LAUNCH_CHECK();

  const int G = lanes_per_row(P);
  const int gs = grid_for(n * G);
  const int64_t spmv_bytes = P.nnz * (int64_t)(sizeof(T) + 4) + (n + 1) * 4 + n * 4 * (int64_t)sizeof(T);
  const int64_t upd_bytes = n * 9 * (int64_t)sizeof(T);
  constexpr int kBatch = 16;
  int launched = 0;
  int polled_batches = 0;
  // pinned flag slots: one per batch parity
  volatile double* flag = (volatile double*)P.pinned;
  auto launch_iter = [&](T* po, T* pn) {
    switch (G) {
      case 4: hipLaunchKernelGGL((k_pcg_spmv<T, 4>), dim3(gs), dim3(kBS), 0, st, n, P.indptr.p, P.indices.p, vals, z, po, pn, q, S, P.partials.p); break;
      case 8: hipLaunchKernelGGL((k_pcg_spmv<T, 8>), dim3(gs), dim3(kBS), 0, st, n, P.indptr.p, P.indices.p, vals, z, po, pn, q, S, P.partials.p); break;
      case 16: hipLaunchKernelGGL((k_pcg_spmv<T, 16>), dim3(gs), dim3(kBS), 0, st, n, P.indptr.p, P.indices.p, vals, z, po, pn, q, S, P.partials.p); break;
      default: hipLaunchKernelGGL((k_pcg_spmv<T, 32>), dim3(gs), dim3(kBS), 0, st, n, P.indptr.p, P.indices.p, vals, z, po, pn, q, S, P.partials.p); break;
    }
    hipLaunchKernelGGL(k_pcg_update<T>, dim3(gv), dim3(kBS), 0, st, n, inv, pn, q, x, r, z, S, P.partials.p);
  };
  hipEvent_t evs[2];
  APX_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
  APX_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
  bool timing = P.timing;
  if (timing) APX_HIP(hipEventRecord(P.ev0, st));
  // Double-buffered polling: batch i+1 is queued before the flag of batch i is read.
  auto enqueue_batch = [&](int slot) {
    for (int i = 0; i < kBatch; ++i) {
      T* po = (launched & 1) ? pb : pa;
      T* pn = (launched & 1) ? pa : pb;
      launch_iter(po, pn);
      ++launched;
    }
    APX_LAUNCH_CHECK();
    APX_HIP(hipMemcpyAsync((void*)(flag + 2 * slot), &S->done, sizeof(int), hipMemcpyDeviceToHost, st));
    APX_HIP(hipEventRecord(evs[slot], st));
  };
  // done flag from init
  int done_host = 0;
  APX_HIP(hipMemcpyAsync((void*)(flag + 4), &S->done, sizeof(int), hipMemcpyDeviceToHost, st));
  APX_HIP(hipStreamSynchronize(st));
  std::memcpy(&done_host, (const void*)(flag + 4), sizeof(int));
  if (!done_host) {
    enqueue_batch(0);
    int slot = 0;
    for (;;) {
      if (launched < maxiter) enqueue_batch(slot ^ 1);
      APX_HIP(hipEventSynchronize(evs[slot]));
      ++polled_batches;
      std::memcpy(&done_host, (const void*)(flag + 2 * slot), sizeof(int));
      if (done_host || launched >= maxiter + kBatch) break;
      slot ^= 1;
    }
  }
  APX_HIP(hipStreamSynchronize(st));
  if (timing) {
    APX_HIP(hipEventRecord(P.ev1, st));
    APX_HIP(hipEventSynchronize(P.ev1));
  }
  (void)hipEventDestroy(evs[0]);
  (void)hipEventDestroy(evs[1]);
  // odd iteration count -> the final direction sits in pb, irrelevant for x.
  KState hs;
  APX_HIP(hipMemcpy(&hs, S, sizeof(KState), hipMemcpyDeviceToHost));
  apx_krylov_stats out{};
  out.iterations = hs.k;
  out.converged = hs.rr <= hs.atol2;
  out.initial_residual = sqrt(hs.rr0);
  out.final_residual = sqrt(hs.rr);
  out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (timing) {
    float ms = 0.f;
    APX_HIP(hipEventElapsedTime(&ms, P.ev0, P.ev1));
    auto& t = P.timers["pcg_iteration"];
    t.total_ms += ms;
    t.count += std::max(hs.k, 1);
    t.bytes = spmv_bytes + upd_bytes;
  }
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
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double ri = r[red2full[i]];
    v[0] += ri * ri;
  }
  double s[1];
  if (last_block_reduce<1>(v, partials, &S->counter[3], s) && threadIdx.x == 0) S->dot_out = sqrt(s[0]);
}

double free_norm(Plan& P, const double* r_full, hipStream_t st) {
  if (P.partials.n < (size_t)(4 * kMaxBlocks)) P.partials.alloc(4 * kMaxBlocks);
  if (P.kstate.n < 64) {
    P.kstate.alloc(64);
    APX_HIP(hipMemsetAsync(P.kstate.p, 0, P.kstate.bytes(), st));
  }
  KState* S = (KState*)P.kstate.p;
  hipLaunchKernelGGL(k_free_norm, dim3(grid_for(P.n_free)), dim3(kBS), 0, st, P.n_free, P.red2full.p, r_full, S,
                     P.partials.p);
  APX_LAUNCH_CHECK();
  APX_HIP(hipMemcpyAsync(P.pinned + 8, &S->dot_out, sizeof(double), hipMemcpyDeviceToHost, st));
  APX_HIP(hipStreamSynchronize(st));
  return P.pinned[8];
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
