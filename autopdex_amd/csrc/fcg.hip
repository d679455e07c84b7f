// Flexible CG with the learned preconditioner (Notay 2000, truncation 1):
//   z_k = M(r_k)                        (U-Net on the node grid)
//   beta = -(z_k . q_k-1) / (p_k-1 . q_k-1)   (0 for k = 0)
//   p_k = z_k + beta p_k-1 ; q_k = A p_k
//   alpha = (p_k . r_k) / (p_k . q_k) ; x += alpha p_k ; r -= alpha q_k
// Krylov vectors are FP64; the matrix values are FP64 or FP32 (SELL layout of
// the plan); the network runs in bf16 with FP32 accumulation.  As in the PCG
// driver, scalars stay on the device and the host polls a stop flag one batch
// late.  The host watches the residual for stagnation (no 1 % decrease over
// 10 iterations), which triggers the Jacobi-PCG fallback in the Newton driver.
#include <chrono>
#include <cmath>
#include <cstring>

#include "assembly.hpp"
#include "unet.hpp"

namespace apx {

void net_input_from_reduced_img(Plan& P, UNet& N, const void* r, int r_f64, const double* g_nodal,
                                const double* rr_src, int img, hipStream_t st);

namespace {
constexpr int kBS = 256;
constexpr int kFinBS = 1024;
constexpr int kSpmvBS = 512;
constexpr int kMaxBlocks = 8192;  // must match the partial-buffer layout of the Krylov kernels

int grid_vec(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 1023) / 1024, kMaxBlocks)); }
int grid_chunks(int64_t c) { return (int)std::min<int64_t>(round_xcd(std::max<int64_t>(1, (c + 7) / 8)), kMaxBlocks); }

template <int BS, int NV>
__device__ inline void partials_out(double (&v)[NV], double* partials) {
  __shared__ double lds[BS / 64];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double s = block_sum<BS>(v[j], lds);
    if (threadIdx.x == 0) partials[j * kMaxBlocks + blockIdx.x] = s;
  }
}

template <int NV>
__device__ inline void sums_in(const double* partials, int nb, double (&out)[NV]) {
  __shared__ double lds[kFinBS / 64];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kFinBS) s += partials[j * kMaxBlocks + i];
    out[j] = block_sum<kFinBS>(s, lds);
  }
}

// partials of r.r and of the network input's sum of squares, (D^-1/2 r).(D^-1/2 r) (dsq = D^-1/2,
// or null: no scaling), so that the network's input scale needs no pass of its own
__global__ void k_fcg_init(int64_t n, const double* b, const double* dsq, double* x, double* r, double* pa,
                           double* partials) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    pa[i] = 0.0;
    v[0] += bi * bi;
    const double si = dsq ? bi * dsq[i] : bi;
    v[1] += si * si;
  }
  partials_out<kBS, 2>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_fcg_init_fin(KState* S, const double* partials, int nb, double tol,
                                                         double atol, int maxiter) {
  double s[2];
  sums_in<2>(partials, nb, s);
  if (threadIdx.x == 0) {
    S->rr = s[0];
    S->rsq = s[1];
    S->rr0 = s[0];
    S->bb = s[0];
    S->atol2 = fmax(tol * tol * s[0], atol * atol);
    S->k = 0;
    S->maxiter = maxiter;
    S->reliable = 0;
    S->beta = 0.0;
    S->gamma_prev = 1.0;  // p.q of the previous direction
    S->done = !(s[0] > S->atol2 && 0 < maxiter);
  }
}

// z = scale * dsq * y (grid -> free DOFs) and partials of z . q_prev
__global__ void k_fcg_z(int64_t n, const int32_t* red2full, const float* y, const double* dsq, const double* r_skip,
                        const float* scale, const double* q, double* z, KState* S, double* partials) {
  if (S->done) return;
  const float s = scale[0];
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    double zi = (double)(s * y[red2full[i]]) * (dsq ? dsq[i] : 1.0);
    if (r_skip) zi += r_skip[i] * dsq[i] * dsq[i];  // Jacobi part D^-1 r
    z[i] = zi;
    v[0] += zi * q[i];
  }
  partials_out<kBS, 1>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_fcg_beta_fin(KState* S, const double* partials, int nb) {
  if (S->done) return;
  double s[1];
  sums_in<1>(partials, nb, s);
  if (threadIdx.x == 0) S->beta = S->k > 0 ? -s[0] / S->gamma_prev : 0.0;
}

// columns: int32 as stored, or int16 offsets from the lane's row (the last row for lanes past n)
__device__ __forceinline__ int32_t fcg_col(int32_t c, int32_t) { return c; }
__device__ __forceinline__ int32_t fcg_col(int16_t d, int32_t base) { return base + (int32_t)d; }

template <class T, class C>
__global__ __launch_bounds__(kSpmvBS) void k_fcg_spmv(int64_t n, int64_t n_chunks, const int32_t* __restrict__ sptr,
                                                      const C* __restrict__ cols, const T* __restrict__ vals,
                                                      const double* __restrict__ z, double* __restrict__ pa,
                                                      double* __restrict__ pb, double* __restrict__ q,
                                                      const double* __restrict__ r, KState* S, double* partials) {
  if (S->done) return;
  const double beta = S->beta;
  const bool odd = S->k & 1;
  const double* __restrict__ po = odd ? pb : pa;
  double* __restrict__ pn = odd ? pa : pb;
  const int lane = threadIdx.x & 63;
  double v[2] = {0.0, 0.0};
  for (int64_t c = blockIdx.x * (int64_t)(kSpmvBS / 64) + (threadIdx.x >> 6); c < n_chunks;
       c += (int64_t)gridDim.x * (kSpmvBS / 64)) {
    const int32_t off = sptr[c], w = (sptr[c + 1] - off) >> 6;
    const int64_t row = c * 64 + lane;
    const int32_t base = (int32_t)(row < n ? row : n - 1);
    const T* vv = vals + off + lane;
    const C* ci = cols + off + lane;
    // values and columns are read once: non-temporal; four entries' loads are issued before
    // their gathers (more loads in flight), the sum keeps the entry order
    double s = 0.0;
    int k = 0;
    for (; k + 4 <= w; k += 4) {
      const int32_t c0 = fcg_col(__builtin_nontemporal_load(ci + (k + 0) * 64), base),
                    c1 = fcg_col(__builtin_nontemporal_load(ci + (k + 1) * 64), base),
                    c2 = fcg_col(__builtin_nontemporal_load(ci + (k + 2) * 64), base),
                    c3 = fcg_col(__builtin_nontemporal_load(ci + (k + 3) * 64), base);
      const T v0 = __builtin_nontemporal_load(vv + (k + 0) * 64), v1 = __builtin_nontemporal_load(vv + (k + 1) * 64),
              v2 = __builtin_nontemporal_load(vv + (k + 2) * 64), v3 = __builtin_nontemporal_load(vv + (k + 3) * 64);
      s += (double)v0 * (z[c0] + beta * po[c0]);
      s += (double)v1 * (z[c1] + beta * po[c1]);
      s += (double)v2 * (z[c2] + beta * po[c2]);
      s += (double)v3 * (z[c3] + beta * po[c3]);
    }
    for (; k < w; ++k) {
      const int32_t col = fcg_col(__builtin_nontemporal_load(ci + k * 64), base);
      s += (double)__builtin_nontemporal_load(vv + k * 64) * (z[col] + beta * po[col]);
    }
    if (row < n) {
      const double pi = z[row] + beta * po[row];
      pn[row] = pi;
      q[row] = s;
      v[0] += pi * s;
      v[1] += pi * r[row];
    }
  }
  partials_out<kSpmvBS, 2>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_fcg_alpha_fin(KState* S, const double* partials, int nb) {
  if (S->done) return;
  double s[2];
  sums_in<2>(partials, nb, s);
  if (threadIdx.x == 0) {
    S->pAp = s[0];
    S->gamma_prev = s[0];
    S->alpha = s[1] / s[0];
    // breakdown (zero curvature along p, non-finite network output): stop
    // before the update; the last iterate stays valid.  The sign of p.Ap is
    // free: the reference's Q1 tangents are negative definite (det J < 0).
    if (!(s[0] != 0.0) || !isfinite(s[0]) || !isfinite(s[1])) {
      S->reliable = 2;
      S->done = 1;
    }
  }
}

__global__ void k_fcg_update(int64_t n, const double* pa, const double* pb, const double* q, const double* dsq,
                             double* x, double* r, KState* S, double* partials) {
  if (S->done) return;
  const double alpha = S->alpha;
  const double* p = (S->k & 1) ? pa : pb;
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] += ri * ri;
    const double si = dsq ? ri * dsq[i] : ri;
    v[1] += si * si;
  }
  partials_out<kBS, 2>(v, partials);
}

__global__ __launch_bounds__(kFinBS) void k_fcg_update_fin(KState* S, const double* partials, int nb) {
  if (S->done) return;
  double s[2];
  sums_in<2>(partials, nb, s);
  if (threadIdx.x == 0) {
    S->rr = s[0];
    S->rsq = s[1];
    S->k += 1;
    S->done = !(s[0] > S->atol2 && S->k < S->maxiter);
  }
}
}  // namespace

// Solve A x = b (x0 = 0) with the network preconditioner.  b, x: FP64 (n_free).
template <class T>
FcgResult fcg_solve(Plan& P, UNet& N, const T* vals, const double* g_nodal, const double* b, double* x, double tol,
                    double atol, int maxiter, hipStream_t st) {
  const int64_t n = P.n_free;
  if (P.w64.n < (size_t)(6 * n)) P.w64.alloc(6 * n);
  double* r = P.w64.p;
  double* z = r + n;
  double* pa = r + 2 * n;
  double* pb = r + 3 * n;
  double* q = r + 4 * n;
  ensure_krylov_scratch(P, st);
  KState* S = (KState*)P.kstate.p;
  if (maxiter <= 0) maxiter = 10000;
  const int gv = grid_vec(n), gs = grid_chunks(P.n_chunks);
  net_set_diag_scaling(P, vals, sizeof(T) == 8 ? APX_FP64 : APX_FP32, st);  // network sees D^-1/2 A D^-1/2
  APX_HIP(hipMemsetAsync(q, 0, n * sizeof(double), st));
  const double* dsq = P.net_dsq_set ? P.net_dsq.p : nullptr;  // the network's input scaling
  hipLaunchKernelGGL(k_fcg_init, dim3(gv), dim3(kBS), 0, st, n, b, dsq, x, r, pa, P.partials.p);
  hipLaunchKernelGGL(k_fcg_init_fin, dim3(1), dim3(kFinBS), 0, st, S, P.partials.p, gv, tol, atol, maxiter);
  APX_LAUNCH_CHECK();
  std::vector<EventPair> ev_net, ev_spmv;
  // for this solve only: the network times its forward passes, and iterations enqueued after
  // convergence skip it (restored on every exit)
  struct NetScope {
    UNet& N;
    ~NetScope() {
      N.timer_plan = nullptr;
      N.stop = nullptr;
    }
  } net_scope{N};
  N.timer_plan = P.timing ? &P : nullptr;
  N.stop = &S->done;
  auto iteration = [&] {
    EventPair e0 = timing_begin(P, st);
    net_input_from_reduced_img(P, N, r, 1, g_nodal, &S->rsq, 0, st);  // its scale from the last update
    N.forward(st, false);
    hipLaunchKernelGGL(k_fcg_z, dim3(gv), dim3(kBS), 0, st, n, P.red2full.p, N.y.p, P.net_dsq.p,
                       N.jacobi_skip ? r : nullptr, N.scale.p, q, z, S, P.partials.p);
    timing_end(P, e0, st);
    hipLaunchKernelGGL(k_fcg_beta_fin, dim3(1), dim3(kFinBS), 0, st, S, P.partials.p, gv);
    // timed: the pair takes the SpMV's own start and end timestamps
    const EventPair e1 = P.timing ? timing_pair(P) : EventPair{};
    t_launch_events = P.timing ? &e1 : nullptr;
    if (P.sell_dcol.p)
      launch_kernel(k_fcg_spmv<T, int16_t>, dim3(gs), dim3(kSpmvBS), 0, st, n, P.n_chunks, P.sell_ptr.p,
                    (const int16_t*)P.sell_dcol.p, vals, z, pa, pb, q, r, S, P.partials.p);
    else
      launch_kernel(k_fcg_spmv<T, int32_t>, dim3(gs), dim3(kSpmvBS), 0, st, n, P.n_chunks, P.sell_ptr.p,
                    (const int32_t*)P.sell_cols.p, vals, z, pa, pb, q, r, S, P.partials.p);
    hipLaunchKernelGGL(k_fcg_alpha_fin, dim3(1), dim3(kFinBS), 0, st, S, P.partials.p, gs);
    hipLaunchKernelGGL(k_fcg_update, dim3(gv), dim3(kBS), 0, st, n, pa, pb, q, dsq, x, r, S, P.partials.p);
    hipLaunchKernelGGL(k_fcg_update_fin, dim3(1), dim3(kFinBS), 0, st, S, P.partials.p, gv);
    APX_LAUNCH_CHECK();
    if (P.timing) {
      ev_net.push_back(e0);
      ev_spmv.push_back(e1);
    }
  };
  // host polling: one iteration per batch, its state copied to a pinned slot and read while
  // the next one runs (an iteration is ~1 ms of device work against ~0.3 ms to enqueue it).
  // An iteration queued past convergence costs ~70 early-exit launches (~0.3 ms), once per
  // inner solve, so when the last read residual, extrapolated by its last ratio, says that
  // the queued iteration converges, nothing more is queued until it is read; a wrong guess
  // costs the host's wake-up latency, since the device starts on the first relaunched kernel
  // while the host enqueues the rest
  constexpr int kBatch = 1;
  volatile char* pin = (volatile char*)P.pinned + 512;  // slot s: KState copy at byte 512 + 128 s
  static_assert(sizeof(KState) <= 128, "KState must fit a 128-byte pinned slot");
  hipEvent_t evs[2];
  APX_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
  APX_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
  auto enqueue = [&](int slot) {
    for (int i = 0; i < kBatch; ++i) iteration();
    APX_HIP(hipMemcpyAsync((void*)(pin + 128 * slot), S, sizeof(KState), hipMemcpyDeviceToHost, st));
    APX_HIP(hipEventRecord(evs[slot], st));
  };
  FcgResult res;
  KState hs;
  APX_HIP(hipMemcpyAsync(&hs, S, sizeof(KState), hipMemcpyDeviceToHost, st));
  APX_HIP(hipStreamSynchronize(st));
  res.rr0 = hs.rr0;
  std::vector<double> hist;  // rr after each polled batch
  std::vector<int> hist_k;
  if (!hs.done) {
    // at most two batches in flight, one per pinned slot: a slot is reused only after it is read
    int next_slot = 0, read_slot = 0, in_flight = 0;
    int64_t launched = 0;
    bool hold = false;
    double rr_prev = hs.rr;
    for (;;) {
      while (in_flight < 2 && !(hold && in_flight > 0)) {
        enqueue(next_slot);
        next_slot ^= 1;
        ++in_flight;
        launched += kBatch;
      }
      APX_HIP(hipEventSynchronize(evs[read_slot]));
      std::memcpy(&hs, (const void*)(pin + 128 * read_slot), sizeof(KState));
      read_slot ^= 1;
      --in_flight;
      if (hs.done) break;
      // the queued batch ends the solve: predicted to converge (rr falling by its last ratio
      // per iteration) or reaching maxiter
      const double ratio = rr_prev > 0.0 ? hs.rr / rr_prev : 1.0;
      hold = in_flight > 0 && ((ratio < 1.0 && hs.rr * std::pow(ratio, (double)kBatch) <= hs.atol2) ||
                               hs.k + kBatch >= hs.maxiter);
      rr_prev = hs.rr;
      hist.push_back(hs.rr);
      hist_k.push_back(hs.k);
      // stagnation: the residual norm has not reached a new minimum (by 1 %) for
      // kStagWindow iterations (flexible CG residuals need not decrease monotonically)
      constexpr int kStagWindow = 12;
      double best_before = INFINITY, best_since = INFINITY;
      for (size_t j = 0; j < hist.size(); ++j) {
        if (hist_k.back() - hist_k[j] >= kStagWindow)
          best_before = std::min(best_before, hist[j]);
        else
          best_since = std::min(best_since, hist[j]);
      }
      if (std::isfinite(best_before) && !(best_since < 0.98 * best_before)) res.stagnated = true;  // rr: 0.99^2
      if (hs.rr > 1e8 * hs.rr0) res.stagnated = true;  // divergence
      if (!std::isfinite(hs.rr)) res.stagnated = true;
      if (res.stagnated || launched > (int64_t)maxiter + 4 * kBatch) break;
    }
  }
  APX_HIP(hipStreamSynchronize(st));
  APX_HIP(hipMemcpy(&hs, S, sizeof(KState), hipMemcpyDeviceToHost));
  (void)hipEventDestroy(evs[0]);
  (void)hipEventDestroy(evs[1]);
  res.iterations = hs.k;
  res.rr = hs.rr;
  res.converged = hs.rr <= hs.atol2;
  if (hs.reliable == 2) res.stagnated = true;  // breakdown
  if (P.timing) {
    const size_t k = (size_t)hs.k;
    const int64_t spmv_bytes =
        P.sell_total * (int64_t)(sizeof(T) + (P.sell_dcol.p ? 2 : 4)) + (P.n_chunks + 1) * 4 + n * 6 * 8;
    timing_collect(P, "fcg_precond", ev_net, 0, k);
    timing_collect(P, sizeof(T) == 8 ? "fcg_spmv_f64" : "fcg_spmv_f32", ev_spmv, spmv_bytes, k);
    N.collect_timing(k);  // the forward passes of the executed iterations
  }
  return res;
}

template FcgResult fcg_solve<double>(Plan&, UNet&, const double*, const double*, const double*, double*, double,
                                     double, int, hipStream_t);
template FcgResult fcg_solve<float>(Plan&, UNet&, const float*, const double*, const double*, double*, double, double,
                                    int, hipStream_t);

}  // namespace apx
