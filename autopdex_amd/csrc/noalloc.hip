// Allocation-free Jacobi-PCG: apx_pcg_noalloc.
//
// The same solve as apx_pcg (the same kernels, launch shapes and stopping rule,
// so the same iterates and bits), for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Everything apx_pcg allocates -- the SELL copy of the
// matrix values, the inverse diagonal, six Krylov vectors, the reduction
// partials and the device scalars -- is carved out of one file-scope device
// array instead, and the stop flag is polled without events.
//
// Limits: the array holds kScratchBytes = 32 MiB per device, which it keeps
// until the process ends.  A solve needs
//   256 B (scalars) + 8 x 32768 B (partials) + sell_entries x T + 7 n_free x T
// rounded up to 256-byte blocks (T = 8 for FP64, 4 for FP32); larger solves are
// rejected with APX_ERR_ARG.  Every call on a device shares that device's
// array, so calls that overlap on one device race.
#include <cstring>
#include <mutex>

#include "assembly.hpp"

namespace apx {

constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
// The array's device address on each device.  In host code `scratch` is only a
// stand-in for the symbol, never a device address.
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

void* scratch_on_current_device() {
  int dev = 0;
  APX_HIP(hipGetDevice(&dev));
  APX_REQUIRE(dev >= 0 && dev < kMaxDevices, APX_ERR_INTERNAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    APX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    APX_REQUIRE(ptr != nullptr, APX_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch[dev] = ptr;
  }
  return g_scratch[dev];
}

inline size_t blocks256(size_t bytes) { return (bytes + 255) / 256 * 256; }

__global__ void k_zero_bytes(unsigned char* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

template <class T>
apx_krylov_stats pcg_noalloc(Plan& P, const T* vals_csr, const T* b, T* x, int precond, double tol, double atol,
                             int maxiter, hipStream_t st) {
  const size_t ts = sizeof(T);
  const size_t need_state = 256, need_part = blocks256(kKrylovPartials * sizeof(double));
  const size_t need_vals = blocks256((size_t)P.sell_total * ts), need_inv = blocks256((size_t)P.n_free * ts);
  const size_t need_work = blocks256((size_t)6 * P.n_free * ts);
  const size_t need = need_state + need_part + need_vals + need_inv + need_work;
  if (need > kScratchBytes)
    throw Error(APX_ERR_ARG, "apx_pcg_noalloc: the solve needs " + std::to_string(need) + " bytes of scratch (" +
                                 std::to_string(P.n_free) + " rows, " + std::to_string(P.sell_total) +
                                 " SELL entries), more than the " + std::to_string(kScratchBytes) +
                                 " bytes it holds; use apx_pcg");
  unsigned char* base = (unsigned char*)scratch_on_current_device();
  KState* state = (KState*)base;
  double* part = (double*)(base + need_state);
  T* sell = (T*)(base + need_state + need_part);
  T* inv = (T*)(base + need_state + need_part + need_vals);
  T* work = (T*)(base + need_state + need_part + need_vals + need_inv);
  // the array keeps the previous call's contents: the scalars start from zero (every field
  // is also written by the init kernels before it is read)
  hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(256), 0, st, base, need_state);
  APX_LAUNCH_CHECK();
  csr_to_sell(P, vals_csr, sell, sizeof(T) == 8 ? APX_FP64 : APX_FP32, st);  // padding filled by a kernel
  if (precond == APX_PRECOND_JACOBI)
    jacobi_inverse(P, sell, inv, sizeof(T) == 8 ? APX_FP64 : APX_FP32, st);
  else
    fill<T>(inv, P.n_free, T(1), st);
  PcgRun<T> run(P, sell, inv, st, work, part, state);
  run.init(b, x, tol, atol, maxiter, 0.0);
  KState hs = run.run();
  apx_krylov_stats out{};
  out.iterations = hs.k;
  out.converged = hs.rr <= hs.atol2;
  out.initial_residual = sqrt(hs.rr0);
  out.final_residual = sqrt(hs.rr);
  out.seconds = 0.0;
  return out;
}
}  // namespace

}  // namespace apx

using namespace apx;

struct apx_plan : public Plan {};

extern "C" int apx_pcg_noalloc(apx_plan* plan, const void* vals_dev, const void* b_dev, void* x_dev, int precision,
                               int precond, double tol, double atol, int maxiter, apx_krylov_stats* stats,
                               void* stream) {
  return guarded([&] {
    APX_REQUIRE(plan != nullptr, APX_ERR_ARG, "null plan");
    Plan& P = *plan;
    APX_REQUIRE(P.finalized, APX_ERR_STATE, "plan not finalized (call apx_plan_finalize)");
    APX_REQUIRE(vals_dev && b_dev && x_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32, APX_ERR_ARG, "precision must be FP64 or FP32");
    APX_REQUIRE(precond == APX_PRECOND_JACOBI || precond == APX_PRECOND_NONE, APX_ERR_UNSUPPORTED,
                "apx_pcg_noalloc supports Jacobi or no preconditioner");
    APX_REQUIRE(P.pinned != nullptr, APX_ERR_STATE, "plan has no pinned flag block");
    hipStream_t st = (hipStream_t)stream;
    apx_krylov_stats ks =
        precision == APX_FP64
            ? pcg_noalloc<double>(P, (const double*)vals_dev, (const double*)b_dev, (double*)x_dev, precond, tol,
                                  atol, maxiter, st)
            : pcg_noalloc<float>(P, (const float*)vals_dev, (const float*)b_dev, (float*)x_dev, precond, tol, atol,
                                 maxiter, st);
    if (stats) *stats = ks;
  });
}
