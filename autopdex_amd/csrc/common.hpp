// Shared helpers for the autopdex_amd HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/autopdex_amd.h"

namespace apx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define APX_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      throw ::apx::Error(APX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define APX_REQUIRE(cond, code, msg)                   \
  do {                                                 \
    if (!(cond)) throw ::apx::Error((code), (msg));    \
  } while (0)

#define APX_LAUNCH_CHECK() APX_HIP(hipGetLastError())

// Run `body` and translate exceptions into status codes for the C ABI.
template <class F>
int guarded(F&& body) {
  try {
    body();
    return APX_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return APX_ERR_INTERNAL;
  }
}

constexpr int kWave = 64;

// Device buffer with RAII.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    APX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T* host, size_t count) {
    alloc(count);
    if (count) APX_HIP(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---- wave / block reductions (deterministic: fixed shuffle tree) ---------------
// XCD-aware workgroup order.  MI355X deals the workgroups of a launch
// round-robin to its 8 XCDs, each with its own L2.  Renumbering block b as
// (b % 8) * (G / 8) + b / 8 gives every XCD one contiguous range of logical
// blocks, so work whose inputs overlap (neighbouring SELL chunks gathering
// the same vector entries, vertically adjacent tiles sharing halo rows) meets
// in one L2.  G must be a multiple of 8 for the renumbering (otherwise the
// identity is used).
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t G) {
  return (G % kXcds) ? b : (b % kXcds) * (G / kXcds) + b / kXcds;
}
__host__ inline int64_t round_xcd(int64_t g) { return (g + kXcds - 1) / kXcds * kXcds; }

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block sum for blockDim.x == BS (multiple of 64); result valid in all threads.
template <int BS, class T>
__device__ inline T block_sum(T v, T* lds /* BS/64 entries */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x < 64) {
    s = (threadIdx.x < BS / 64) ? lds[threadIdx.x] : T(0);
    s = wave_sum(s);
    if (threadIdx.x == 0) lds[0] = s;
  }
  __syncthreads();
  s = lds[0];
  __syncthreads();
  return s;
}

}  // namespace apx
