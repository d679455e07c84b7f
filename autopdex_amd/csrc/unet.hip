// Learned preconditioner runtime: U-Net construction, forward, backward,
// Adam, and the conversions between reduced DOF vectors and the node grid.
//
// Architecture (levels l = 0..L-1, channels C_l, grids H_l = ceil(H_{l-1}/2)):
//   a_0   = act(conv_in(x_0))
//   d_l   = act(conv_e[l](a_l))            rd_l = R(d_l)
//   a_l+1 = act(conv_down[l](rd_l))
//   u_L-1 = act(conv_b(a_L-1))
//   t_l   = conv_up[l](u_l+1)              s_l = d_l + P(t_l)
//   u_l   = act(conv_d[l](s_l))
//   y     = conv_out(u_0)                  (FP32)
// R is full weighting and P bilinear interpolation between node-centred grids
// (R = P^T, weights (1/2, 1, 1/2) per axis), so the network has the data flow
// of a multigrid V-cycle with learned smoothing / transfer convolutions.
// `act` is ReLU or the identity (a linear network, i.e. a fixed linear
// preconditioner).  There is no counterpart in the reference.
#include <cmath>
#include <cstring>

#include "assembly.hpp"
#include "unet.hpp"
#include "unet_kernels.hpp"

namespace apx {
using namespace nn;

namespace {
constexpr int kWgradBlocks = 2048;  // rows of the weight-gradient partial buffer

int kblocks(int ci) { return ci >= 16 ? 9 * (ci / 16) : (9 * ci + 15) / 16; }

// ---- conv dispatch ------------------------------------------------------------------
// Grids with fewer tiles than kSplitTiles split the output channels over
// 16-channel blocks (more blocks, 1/NT of the weights and MFMAs each).
constexpr int64_t kSplitTiles = 128;

// Blocks of one launch: at most what the device holds at once (the occupancy of
// the kernel: LDS and registers), so that no block starts after the first round
// has finished (a second partial round would run at low occupancy).
template <class K>
int64_t resident_blocks(K kernel, int threads = kConvThreads) {
  int per_cu = 0, dev = 0, cus = 0;
  APX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  APX_HIP(hipGetDevice(&dev));
  APX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return (int64_t)std::max(1, cus) * std::max(1, per_cu);
}

template <int CI, int CO, bool F32, bool XM>
void launch_conv_t(const ConvArgs& a, hipStream_t st) {
  constexpr int NT = round16(CO) / 16;
  const int64_t tiles = (int64_t)a.B * ((a.W + kTileW - 1) / kTileW) * ((a.H + kTileH - 1) / kTileH);
  if (NT > 1 && tiles < kSplitTiles) {
    static const int64_t cap = resident_blocks(k_conv3x3<CI, CO, F32, 1, 1, XM>);
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap / NT));
    launch_kernel((k_conv3x3<CI, CO, F32, 1, 1, XM>), dim3(g, NT), dim3(kConvThreads), 0, st, a);
  } else if constexpr (NT == 4) {
    // four channel tiles: two waves per tile row (WS = 2), two tiles each
    static const int64_t cap = resident_blocks(k_conv3x3<CI, CO, F32, NT, 2, XM>, 2 * kConvThreads);
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
    launch_kernel((k_conv3x3<CI, CO, F32, NT, 2, XM>), dim3(g), dim3(2 * kConvThreads), 0, st, a);
  } else {
    static const int64_t cap = resident_blocks(k_conv3x3<CI, CO, F32, NT, 1, XM>);
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
    launch_kernel((k_conv3x3<CI, CO, F32, NT, 1, XM>), dim3(g), dim3(kConvThreads), 0, st, a);
  }
}

// the ReLU-masked input (backward data pass) and the plain one are separate kernels
template <int CI, int CO, bool F32>
void launch_conv_t(const ConvArgs& a, hipStream_t st) {
  if (a.xmask)
    launch_conv_t<CI, CO, F32, true>(a, st);
  else
    launch_conv_t<CI, CO, F32, false>(a, st);
}

template <int CI, int CO, int TR, int IN = 0>
void launch_conv_ws_t(const ConvArgs& a, hipStream_t st) {
  const int64_t tiles = (int64_t)a.B * ((a.W + kTileW - 1) / kTileW) * ((a.H + TR - 1) / TR);
  APX_REQUIRE(tiles < (int64_t(1) << 30) && (int64_t)(TR + 2) * a.W * CI < (int64_t(1) << 31), APX_ERR_UNSUPPORTED,
              "convolution grid too large for 32-bit tile indices");
  APX_REQUIRE(IN == 0 || (a.xin && a.win && !a.xmask && (int64_t)a.H * a.W * IN < INT32_MAX), APX_ERR_UNSUPPORTED,
              "convolution of the input layer's output: no image, or the image is too large");
  static const int64_t cap = resident_blocks(k_conv3x3_ws<CI, CO, TR, IN>);
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
  launch_kernel((k_conv3x3_ws<CI, CO, TR, IN>), dim3(g), dim3(kConvThreads), 0, st, a);
}

// Four-row tiles: measured faster than eight-row tiles (which hold twice the accumulators and
// run at lower occupancy) on the 2049 x 2049 level-0 layers.
template <int CI, int CO>
void launch_conv_ws(const ConvArgs& a, hipStream_t st) {
  launch_conv_ws_t<CI, CO, 4>(a, st);
}

// weight gradient: the blocks the device holds at once (at most kWgradBlocks, the partial
// buffer's rows); returns the number launched, whose partials k_reduce_wgrad sums
template <int CI, int CO, bool YM>
int launch_wgrad_t(const WgradArgs& a, hipStream_t st) {
  const int64_t tiles = (int64_t)a.B * ((a.W + kTileW - 1) / kTileW) * ((a.H + kTileH - 1) / kTileH);
  static const int64_t cap = std::min<int64_t>(resident_blocks(k_conv3x3_wgrad<CI, CO, YM>), kWgradBlocks);
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
  hipLaunchKernelGGL((k_conv3x3_wgrad<CI, CO, YM>), dim3(g), dim3(kConvThreads), 0, st, a);
  return g;
}
template <int CI, int CO>
int launch_wgrad_t(const WgradArgs& a, hipStream_t st) {
  return a.ymask ? launch_wgrad_t<CI, CO, true>(a, st) : launch_wgrad_t<CI, CO, false>(a, st);
}

#define APX_CONV_CASES(X) \
  X(1, 16) X(2, 16) X(16, 16) X(16, 32) X(32, 16) X(32, 32) X(32, 64) X(64, 32) X(64, 64) X(16, 64) X(64, 16)

void launch_conv(int ci, int co, bool f32, const ConvArgs& a, hipStream_t st) {
  if (f32) {
    const int64_t tiles = (int64_t)a.B * ((a.W + 63) / 64) * ((a.H + 3) / 4);
    APX_REQUIRE(tiles < INT32_MAX, APX_ERR_UNSUPPORTED, "output layer: grid too large");
    if (ci == 16 && co == 1)
      launch_kernel(k_conv_out<1>, dim3((int)tiles), dim3(256), 0, st, a);
    else if (ci == 16 && co == 2)
      launch_kernel(k_conv_out<2>, dim3((int)tiles), dim3(256), 0, st, a);
    else
      throw Error(APX_ERR_UNSUPPORTED, "FP32-output convolution only for the 16 -> 1|2 output layer");
    return;
  }
  if (ci == 16 && co == 16) return launch_conv_ws<16, 16>(a, st);
  if (ci == 16 && co == 32) return launch_conv_ws<16, 32>(a, st);
  if (ci == 32 && co == 16) return launch_conv_ws<32, 16>(a, st);
  if (ci == 32 && co == 32) return launch_conv_ws<32, 32>(a, st);
  if ((ci == 1 || ci == 2) && co == 16 && !a.xmask) {  // the input layer
    const int64_t tiles = (int64_t)a.B * ((a.W + kTileW - 1) / kTileW) * ((a.H + kTileH - 1) / kTileH);
    APX_REQUIRE(tiles < INT32_MAX, APX_ERR_UNSUPPORTED, "input layer: grid too large");
    if (ci == 1)
      launch_kernel(k_conv_in<1>, dim3((int)tiles), dim3(kConvThreads), 0, st, a);
    else
      launch_kernel(k_conv_in<2>, dim3((int)tiles), dim3(kConvThreads), 0, st, a);
    return;
  }
#define APX_X(A, B) \
  if (ci == A && co == B) return launch_conv_t<A, B, false>(a, st);
  APX_CONV_CASES(APX_X)
#undef APX_X
  throw Error(APX_ERR_UNSUPPORTED, "no convolution kernel for " + std::to_string(ci) + " -> " + std::to_string(co) +
                                       " channels");
}

int launch_wgrad(int ci, int co, const WgradArgs& a, hipStream_t st) {
#define APX_X(A, B) \
  if (ci == A && co == B) return launch_wgrad_t<A, B>(a, st);
  APX_CONV_CASES(APX_X)
#undef APX_X
  if (ci == 16 && co == 1) return launch_wgrad_t<16, 1>(a, st);
  if (ci == 16 && co == 2) return launch_wgrad_t<16, 2>(a, st);
  if (ci == 1 && co == 16) return launch_wgrad_t<1, 16>(a, st);
  if (ci == 2 && co == 16) return launch_wgrad_t<2, 16>(a, st);
  throw Error(APX_ERR_UNSUPPORTED, "no weight-gradient kernel for " + std::to_string(ci) + " -> " +
                                       std::to_string(co) + " channels");
}

// The template name of the kernel launch_conv runs for these arguments (the names rocprofv3
// shows, without namespace and argument list), for per-launch timers.
std::string conv_kernel_name(int ci, int co, bool f32, const ConvArgs& a) {
  auto s = [](int v) { return std::to_string(v); };
  if (f32) return "k_conv_out<" + s(co) + ">";
  const std::string xm = a.xmask ? "true>" : "false>";
  if ((ci == 16 || ci == 32) && (co == 16 || co == 32)) return "k_conv3x3_ws<" + s(ci) + ", " + s(co) + ", 4, 0>";
  if ((ci == 1 || ci == 2) && co == 16 && !a.xmask) return "k_conv_in<" + s(ci) + ">";
  const int nt = round16(co) / 16;
  const int64_t tiles = (int64_t)a.B * ((a.W + kTileW - 1) / kTileW) * ((a.H + kTileH - 1) / kTileH);
  const int nb = (nt > 1 && tiles < kSplitTiles) ? 1 : nt;
  return "k_conv3x3<" + s(ci) + ", " + s(co) + ", false, " + s(nb) + (nb == 4 ? ", 2, " : ", 1, ") + xm;
}

// ---- grid transfer ---------------------------------------------------------------------
// P: fine(i, j) = sum_{I,J} w(i - 2I) w(j - 2J) coarse(I, J), w(0) = 1, w(+-1) = 1/2.
// out = skip + P(coarse) (skip may alias out).  One thread per (pixel, 8 channels):
// 16-byte loads and stores, grid (row chunks, rows, images), 32-bit index math.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bf8_unpack(const u32x4& q, float f[8]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f[2 * e] = __uint_as_float(q[e] << 16);
    f[2 * e + 1] = __uint_as_float(q[e] & 0xffff0000u);
  }
}
__device__ __forceinline__ u32x4 bf8_pack(const float f[8]) {
  u32x4 q;
#pragma unroll
  for (int e = 0; e < 4; ++e) q[e] = f2bf2(f[2 * e], f[2 * e + 1]);
  return q;
}

// One 8-channel chunk of out = skip + P(coarse) at fine node (i, j); pointers at the image,
// C channels per node.
// coarse node (I, J) at coarse + I * crs + J * cps, skip node (i, j) at skip + i * srs + j * sps
template <int C>
__device__ __forceinline__ u32x4 prolong_chunk_s(const uint16_t* __restrict__ coarse, int crs, int cps,
                                                 const uint16_t* skip, int srs, int sps, int i, int j, int c, int Hc,
                                                 int Wc) {
  const int I0 = i >> 1, J0 = j >> 1;
  const int I1 = min((i & 1) ? I0 + 1 : I0, Hc - 1), J1 = min((j & 1) ? J0 + 1 : J0, Wc - 1);
  const uint16_t* cb = coarse + c;
  float v[8], t[8];
  auto acc = [&](int I, int J, float w) {
    bf8_unpack(*(const u32x4*)(cb + I * crs + J * cps), t);
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] += w * t[e];
  };
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = 0.f;
  const bool oi = i & 1, oj = j & 1;
  const bool hi = oi && I0 + 1 < Hc, hj = oj && J0 + 1 < Wc;  // outside the coarse grid: zero
  const float wi0 = oi ? 0.5f : 1.f, wj0 = oj ? 0.5f : 1.f;
  // every load is in bounds (I1, J1 clamped) and unconditional, so all four are in flight at
  // once; a missing term adds w * 0 = +0, which leaves v unchanged (v starts at +0 and a sum
  // of this form never becomes -0), so the bits match skipping it
  auto acc_if = [&](bool on, int I, int J, float w) {
    bf8_unpack(*(const u32x4*)(cb + I * crs + J * cps), t);
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] += w * (on ? t[e] : 0.f);
  };
  acc(I0, J0, wi0 * wj0);
  acc_if(hj, I0, J1, wi0 * 0.5f);
  acc_if(hi, I1, J0, 0.5f * wj0);
  acc_if(hi && hj, I1, J1, 0.25f);
  if (skip) {
    bf8_unpack(*(const u32x4*)(skip + i * srs + j * sps + c), t);
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] += t[e];
  }
  return bf8_pack(v);
}

template <int C>
__device__ __forceinline__ u32x4 prolong_chunk(const uint16_t* __restrict__ coarse, const uint16_t* skip, int i,
                                               int j, int c, int W, int Hc, int Wc) {
  return prolong_chunk_s<C>(coarse, Wc * C, C, skip, W * C, C, i, j, c, Hc, Wc);
}

// out = skip + P(coarse), one thread per coarse node and 8-channel chunk, writing the (up to)
// four fine nodes (2I + di, 2J + dj): each coarse chunk is loaded and unpacked once.  The sums
// are those of prolong_chunk with its zero terms dropped (adding w * 0 = +0 leaves a sum that
// is never -0 unchanged), in the same order, so the bits agree with it.
template <int C8>
__global__ __launch_bounds__(256) void k_prolong_add(const uint16_t* __restrict__ coarse, const uint16_t* skip,
                                                     uint16_t* out, int H, int W, int Hc, int Wc, const int* stop) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= Wc * C8 || (stop && *stop)) return;
  const int J = idx / C8, c = (idx % C8) * 8, I = blockIdx.y, b = blockIdx.z;
  constexpr int C = 8 * C8;
  const uint16_t* cb = coarse + (size_t)Hc * Wc * C * b + c;
  const size_t fb = (size_t)H * W * C * b + c;
  const bool hi = I + 1 < Hc, hj = J + 1 < Wc;      // the next coarse row / column exists
  const bool ri = 2 * I + 1 < H, rj = 2 * J + 1 < W;  // the odd fine row / column exists
  const int I1 = hi ? I + 1 : I, J1 = hj ? J + 1 : J;
  const int i0 = 2 * I, j0 = 2 * J, i1 = ri ? i0 + 1 : i0, j1 = rj ? j0 + 1 : j0;
  // every load unconditional and in bounds, all in flight at once
  const u32x4 q00 = *(const u32x4*)(cb + (I * Wc + J) * C), q01 = *(const u32x4*)(cb + (I * Wc + J1) * C);
  const u32x4 q10 = *(const u32x4*)(cb + (I1 * Wc + J) * C), q11 = *(const u32x4*)(cb + (I1 * Wc + J1) * C);
  const size_t o00 = fb + ((size_t)i0 * W + j0) * C, o01 = fb + ((size_t)i0 * W + j1) * C;
  const size_t o10 = fb + ((size_t)i1 * W + j0) * C, o11 = fb + ((size_t)i1 * W + j1) * C;
  u32x4 s00{0u, 0u, 0u, 0u}, s01 = s00, s10 = s00, s11 = s00;
  if (skip) {
    s00 = *(const u32x4*)(skip + o00);
    s01 = *(const u32x4*)(skip + o01);
    s10 = *(const u32x4*)(skip + o10);
    s11 = *(const u32x4*)(skip + o11);
  }
  float t00[8], t01[8], t10[8], t11[8], sk[8], v[8];
  bf8_unpack(q00, t00);
  bf8_unpack(q01, t01);
  bf8_unpack(q10, t10);
  bf8_unpack(q11, t11);
  auto finish = [&](const u32x4& sq, size_t o) {
    if (skip) {
      bf8_unpack(sq, sk);
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] += sk[e];
    }
    *(u32x4*)(out + o) = bf8_pack(v);
  };
  // (2I, 2J): weight 1 on (I, J)
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = 0.f + t00[e];
  finish(s00, o00);
  if (rj) {  // (2I, 2J + 1): 1/2 on (I, J) and (I, J + 1)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = 0.f + 0.5f * t00[e];
      if (hj) v[e] += 0.5f * t01[e];
    }
    finish(s01, o01);
  }
  if (ri) {  // (2I + 1, 2J): 1/2 on (I, J) and (I + 1, J)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = 0.f + 0.5f * t00[e];
      if (hi) v[e] += 0.5f * t10[e];
    }
    finish(s10, o10);
  }
  if (ri && rj) {  // (2I + 1, 2J + 1): 1/4 on the four
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = 0.f + 0.25f * t00[e];
      if (hj) v[e] += 0.25f * t01[e];
      if (hi) v[e] += 0.25f * t10[e];
      if (hi && hj) v[e] += 0.25f * t11[e];
    }
    finish(s11, o11);
  }
}

// R = P^T: coarse(I, J) = sum_{di,dj in -1..1} w(di) w(dj) fine(2I+di, 2J+dj); one 8-channel
// chunk at coarse node (I, J), pointer at the image.
// Node (i, j) of a tensor at base + i * rs + j * ps (elements): rs = W C, ps = C for the HBM
// tensors; the coarse-level launch's padded LDS buffers have rs = (W + 2) kTailS, ps = kTailS,
// base past their one-node border.
template <int C>
__device__ __forceinline__ u32x4 restrict_chunk_s(const uint16_t* __restrict__ fine, int I, int J, int c, int H, int W,
                                                  int rs, int ps) {
  const uint16_t* fb = fine + c;
  float v[8], t[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = 0.f;
  // nodes outside the fine grid: the load is clamped into the grid and its term is w * 0 = +0,
  // which leaves v unchanged (v starts at +0 and never becomes -0), so the bits match skipping
  // it; all nine loads are unconditional and in flight at once
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj) {
      const int i = 2 * I + di, j = 2 * J + dj;
      const int ic = min(max(i, 0), H - 1), jc = min(max(j, 0), W - 1);
      const bool in = i == ic && j == jc;
      const float w = (di ? 0.5f : 1.f) * (dj ? 0.5f : 1.f);
      bf8_unpack(*(const u32x4*)(fb + ic * rs + jc * ps), t);
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] += w * (in ? t[e] : 0.f);
    }
  return bf8_pack(v);
}

template <int C>
__device__ __forceinline__ u32x4 restrict_chunk(const uint16_t* __restrict__ fine, int I, int J, int c, int H, int W) {
  return restrict_chunk_s<C>(fine, I, J, c, H, W, W * C, C);
}

template <int C8>
__global__ __launch_bounds__(256) void k_restrict(const uint16_t* __restrict__ fine, uint16_t* __restrict__ out, int H,
                                                  int W, int Hc, int Wc, const int* stop) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= Wc * C8 || (stop && *stop)) return;
  const int J = idx / C8, c = (idx % C8) * 8, I = blockIdx.y, b = blockIdx.z;
  constexpr int C = 8 * C8;
  const u32x4 v = restrict_chunk<C>(fine + (size_t)H * W * C * b, I, J, c, H, W);
  *(u32x4*)(out + (((size_t)b * Hc + I) * Wc + J) * C + c) = v;
}

// ---- coarse-level tail of the forward pass ----------------------------------------------
// One workgroup per image runs the steps in order, with a block barrier after each.  The
// activations of the coarse levels stay in LDS for the whole launch: each level has up to three
// buffers in a padded layout ((H + 2) x (W + 2) nodes, kTailS elements per node, zero border),
// so a convolution reads its 3 x 3 windows straight from its input buffer and writes its output
// into the interior of another.  Only the first step (the restriction of level c, from HBM) and
// the last (the prolongation into level c, to HBM) touch global memory.
// Convolution steps form the same products in the same accumulation order as
// k_conv3x3<64, 64, false, NB> (A = 16 pixels x 32 K, B = 32 K x 16 output channels, K
// blocks tap-major, 32 channels each, on the 16x16x32 MFMA), with the same epilogue, and the
// transfers use the arithmetic of k_restrict / the prolongation, so the results are
// bit-identical to the per-layer launches.  Pixels are grouped 16 at a time in row-major order
// over the whole grid; the weights are staged as [co][k] rows of stride kTailSW.  Wavefront w
// computes output channels 32 (w & 1) .. 32 (w & 1) + 31 of the pixel groups w / 2,
// w / 2 + 4, ...
// Weights arrive with 16-byte loads spread over the whole workgroup (a wavefront gathering
// its own B fragments from global memory touches 16 cache lines per load, which on one
// compute unit costs far more than the matrix work).  The next convolution's weights are
// loaded into registers while the current step computes: they are issued after the step's
// bias loads (loads complete in issue order as far as waits go), unconditionally (the last
// convolution reloads its own), and the barrier between the weight stores and the matrix work
// waits for LDS only.
constexpr int kTailThreads = 512;
constexpr int kTailWChunks = 64 * 576 / 8 / kTailThreads;  // weight chunks per thread
static_assert(64 * 576 / 8 % kTailThreads == 0, "weight chunks must tile the workgroup");

template <int NM>
__device__ __forceinline__ void tail_conv_groups(const uint16_t* sx, const uint16_t* sw, int m0, int H, int W, int cp,
                                                 float bias0, float bias1, int relu, uint16_t* Y, uint16_t* Yg, int lr,
                                                 int lg) {
  const int WP = W + 2, HW = H * W;
  int lp[NM];
#pragma unroll
  for (int mm = 0; mm < NM; ++mm) {
    const int p = min((m0 + 4 * mm) * 16 + lr, HW - 1);  // lanes past the grid repeat its last node
    lp[mm] = ((p / W) * WP + p % W) * kTailS + 8 * lg;   // top-left node of the 3 x 3 window
  }
  const uint16_t* wb0 = sw + (32 * cp + lr) * kTailSW + 8 * lg;
  const uint16_t* wb1 = wb0 + 16 * kTailSW;
  f32x4 acc[NM][2];
#pragma unroll
  for (int mm = 0; mm < NM; ++mm) acc[mm][0] = acc[mm][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 18; ++kb) {
    const int t = kb >> 1;  // tap 3 dy + dx, then 32-channel block kb & 1
    const int off = ((t / 3) * WP + t % 3) * kTailS + (kb & 1) * 32;
    const bf16x8 b0 = *(const bf16x8*)(wb0 + kb * 32), b1 = *(const bf16x8*)(wb1 + kb * 32);
#pragma unroll
    for (int mm = 0; mm < NM; ++mm) {
      const bf16x8 af = *(const bf16x8*)(sx + lp[mm] + off);
      acc[mm][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, b0, acc[mm][0], 0, 0, 0);
      acc[mm][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, b1, acc[mm][1], 0, 0, 0);
    }
  }
  // output: node p of the grid -> interior node (p / W + 1, p % W + 1) of the output buffer, and
  // (training, Yg set) node p of the HBM tensor, which the backward pass reads
#pragma unroll
  for (int n = 0; n < 2; ++n) {
    const int co = 32 * cp + 16 * n + lr;
    const float bb = n ? bias1 : bias0;
#pragma unroll
    for (int mm = 0; mm < NM; ++mm)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = (m0 + 4 * mm) * 16 + 4 * lg + j;
        float v = acc[mm][n][j] + bb;
        if (relu) v = fmaxf(v, 0.f);
        if (p < HW) {
          const uint16_t h = f2bf(v);
          Y[((p / W + 1) * WP + p % W + 1) * kTailS + co] = h;
          if (Yg) Yg[p * 64 + co] = h;
        }
      }
  }
}

// store: also write every step's output to its HBM tensor (training: the backward pass reads them)
__global__ __launch_bounds__(kTailThreads) void k_unet_tail(TailArgs args, const int* stop, int store) {
  if (stop && *stop) return;  // uniform over the grid: set before the launch
  __shared__ __attribute__((aligned(16))) uint16_t act[kTailPool];
  __shared__ __attribute__((aligned(16))) uint16_t sw[64 * kTailSW];
  const int img = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int cp = wave & 1, ph = wave >> 1;
  u32x4 qw[kTailWChunks];
  auto fetch_weights = [&](const uint16_t* w) {
#pragma unroll
    for (int r = 0; r < kTailWChunks; ++r) {
      const int i = threadIdx.x + r * kTailThreads;
      qw[r] = *(const u32x4*)(w + (i / 72) * 576 + (i % 72) * 8);
    }
  };
  fetch_weights(args.w_first);
  // zero pool: the buffers' borders stay zero (steps write interiors only)
  for (int i = 8 * (int)threadIdx.x; i < args.pool; i += 8 * kTailThreads) *(u32x4*)(act + i) = u32x4{0u, 0u, 0u, 0u};
  __syncthreads();
  for (int si = 0; si < args.n; ++si) {
    const TailStep& st = args.step[si];
    if (st.op == kTailConv) {
      const int H = st.H, W = st.W;
      // the prefetched weights -> LDS (every reader of the previous convolution passed its barrier)
#pragma unroll
      for (int r = 0; r < kTailWChunks; ++r) {
        const int i = threadIdx.x + r * kTailThreads;
        *(u32x4*)(sw + (i / 72) * kTailSW + (i % 72) * 8) = qw[r];
      }
      const float bias0 = st.bias[32 * cp + lr], bias1 = st.bias[32 * cp + 16 + lr];
      fetch_weights(st.w_next);  // in flight through this step's matrix work
      lds_barrier();
      const uint16_t* X = act + st.xo;
      uint16_t* Y = act + st.yo;
      uint16_t* Yg = store ? st.y + (int64_t)img * H * W * 64 : nullptr;
      const int NG = (H * W + 15) / 16;
      for (int m0 = ph; m0 < NG; m0 += 16) {
        switch (min(4, (NG - m0 + 3) / 4)) {
          case 4: tail_conv_groups<4>(X, sw, m0, H, W, cp, bias0, bias1, st.relu, Y, Yg, lr, lg); break;
          case 3: tail_conv_groups<3>(X, sw, m0, H, W, cp, bias0, bias1, st.relu, Y, Yg, lr, lg); break;
          case 2: tail_conv_groups<2>(X, sw, m0, H, W, cp, bias0, bias1, st.relu, Y, Yg, lr, lg); break;
          default: tail_conv_groups<1>(X, sw, m0, H, W, cp, bias0, bias1, st.relu, Y, Yg, lr, lg); break;
        }
      }
    } else if (st.op == kTailRestrict) {
      // fine grid H x W (LDS buffer, or level c in HBM) -> coarse Hc x Wc (LDS buffer interior)
      const int n8 = st.Hc * st.Wc * 8;
      const int crs = (st.Wc + 2) * kTailS;
      uint16_t* Yc = act + st.yo + crs + kTailS;
      uint16_t* Yg = store ? st.y + (int64_t)img * st.Hc * st.Wc * 64 : nullptr;
      if (st.xo >= 0) {
        const int frs = (st.W + 2) * kTailS;
        const uint16_t* F = act + st.xo + frs + kTailS;
        for (int i = threadIdx.x; i < n8; i += kTailThreads) {
          const int node = i >> 3, I = node / st.Wc, J = node % st.Wc;
          const u32x4 r = restrict_chunk_s<64>(F, I, J, (i & 7) * 8, st.H, st.W, frs, kTailS);
          *(u32x4*)(Yc + I * crs + J * kTailS + (i & 7) * 8) = r;
          if (Yg) *(u32x4*)(Yg + node * 64 + (i & 7) * 8) = r;
        }
      } else {
        const uint16_t* F = st.x + (int64_t)img * st.H * st.W * 64;
        for (int i0 = threadIdx.x; i0 < n8; i0 += 4 * kTailThreads) {
          u32x4 r[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kTailThreads, node = i >> 3;
            r[u] = i < n8 ? restrict_chunk<64>(F, node / st.Wc, node % st.Wc, (i & 7) * 8, st.H, st.W)
                          : u32x4{0u, 0u, 0u, 0u};
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kTailThreads, node = i >> 3;
            if (i < n8) {
              *(u32x4*)(Yc + (node / st.Wc) * crs + (node % st.Wc) * kTailS + (i & 7) * 8) = r[u];
              if (Yg) *(u32x4*)(Yg + node * 64 + (i & 7) * 8) = r[u];
            }
          }
        }
      }
    } else {
      // coarse grid Hc x Wc (LDS buffer) + skip -> fine grid H x W (LDS buffer interior, or level c
      // in HBM with its skip from HBM)
      const int n8 = st.H * st.W * 8;
      const int crs = (st.Wc + 2) * kTailS;
      const uint16_t* Cg = act + st.xo + crs + kTailS;
      if (st.yo >= 0) {
        const int frs = (st.W + 2) * kTailS;
        const uint16_t* Sk = act + st.so + frs + kTailS;
        uint16_t* Yf = act + st.yo + frs + kTailS;
        uint16_t* Yg = store ? st.y + (int64_t)img * st.H * st.W * 64 : nullptr;
        for (int i = threadIdx.x; i < n8; i += kTailThreads) {
          const int node = i >> 3, fi = node / st.W, fj = node % st.W;
          const u32x4 r = prolong_chunk_s<64>(Cg, crs, kTailS, Sk, frs, kTailS, fi, fj, (i & 7) * 8, st.Hc, st.Wc);
          *(u32x4*)(Yf + fi * frs + fj * kTailS + (i & 7) * 8) = r;
          if (Yg) *(u32x4*)(Yg + node * 64 + (i & 7) * 8) = r;
        }
      } else {
        const int64_t io = (int64_t)img * st.H * st.W * 64;
        const uint16_t* Sk = st.skip + io;
        uint16_t* Yg = st.y + io;
        for (int i0 = threadIdx.x; i0 < n8; i0 += 4 * kTailThreads) {
          u32x4 r[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kTailThreads, node = i >> 3;
            r[u] = i < n8 ? prolong_chunk_s<64>(Cg, crs, kTailS, Sk, st.W * 64, 64, node / st.W, node % st.W,
                                                (i & 7) * 8, st.Hc, st.Wc)
                          : u32x4{0u, 0u, 0u, 0u};
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kTailThreads;
            if (i < n8) *(u32x4*)(Yg + (i >> 3) * 64 + (i & 7) * 8) = r[u];
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---- parameters ---------------------------------------------------------------------------
__global__ void k_cast_weights(const float* theta, uint16_t* wf, int64_t w_off, int64_t wf_off, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    wf[wf_off + i] = f2bf(theta[w_off + i]);
}

// dgrad weights: wd[ci][t'*CO + co] = W[co][(8 - t')*CI + ci]
__global__ void k_flip_weights(const float* theta, uint16_t* wd, int64_t w_off, int64_t wd_off, int CI, int CO,
                               int wk, int d_cop, int d_wk) {
  const int64_t n = (int64_t)d_cop * d_wk;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int ci = (int)(i / d_wk), k = (int)(i % d_wk);
    float v = 0.f;
    if (ci < CI && k < 9 * CO) {
      const int tp = k / CO, co = k % CO;
      v = theta[w_off + (int64_t)co * wk + (8 - tp) * CI + ci];
    }
    wd[wd_off + i] = f2bf(v);
  }
}

// the weight-gradient partials summed over the nblk blocks (row b at b * stride: COP*WK weight
// entries, then the biases): 64 outputs per block, each of its 16 waves summing a contiguous
// range of the rows in order, the 16 range sums added in order (deterministic)
constexpr int kReduceWaves = 16;
__global__ __launch_bounds__(64 * kReduceWaves) void k_reduce_wgrad(const float* partial, int nblk, int64_t stride,
                                                                    int cop, int wk, int co_real, float* grad,
                                                                    int64_t w_off, int64_t b_off) {
  __shared__ float red[kReduceWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t n = (int64_t)cop * wk + co_real;
  const int64_t i = blockIdx.x * 64LL + lane;
  const int per = (nblk + kReduceWaves - 1) / kReduceWaves, b0 = w * per, b1 = min(nblk, b0 + per);
  float s = 0.f;
  if (i < n)
    for (int b = b0; b < b1; ++b) s += partial[(int64_t)b * stride + i];
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && i < n) {
    float t = 0.f;
    for (int k = 0; k < kReduceWaves; ++k) t += red[k][lane];
    if (i < (int64_t)cop * wk)
      grad[w_off + i] = t;
    else
      grad[b_off + (i - (int64_t)cop * wk)] = t;
  }
}

__global__ void k_adam(float* theta, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// ---- loss ---------------------------------------------------------------------------------------
constexpr int kLossBlocks = 256;
__global__ void k_loss_partials(const float* y, const float* t, int64_t per_img, int B, double* part) {
  // part[(b * kLossBlocks + blk) * 2 + {0: sum (y-t)^2, 1: sum t^2}]
  __shared__ double lds[4];
  for (int b = 0; b < B; ++b) {
    double e2 = 0.0, t2 = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < per_img; i += (int64_t)gridDim.x * blockDim.x) {
      const double ti = t[(int64_t)b * per_img + i], di = (double)y[(int64_t)b * per_img + i] - ti;
      e2 += di * di;
      t2 += ti * ti;
    }
    e2 = block_sum<256>(e2, lds);
    t2 = block_sum<256>(t2, lds);
    if (threadIdx.x == 0) {
      part[((int64_t)b * kLossBlocks + blockIdx.x) * 2 + 0] = e2;
      part[((int64_t)b * kLossBlocks + blockIdx.x) * 2 + 1] = t2;
    }
  }
}

__global__ void k_loss_fin(double* part, int B, int nblk) {
  // totals in part[B*kLossBlocks*2 + b*2 + {0,1}], loss in part[B*kLossBlocks*2 + 2B]
  if (threadIdx.x != 0) return;
  double loss = 0.0;
  for (int b = 0; b < B; ++b) {
    double e2 = 0.0, t2 = 0.0;
    for (int k = 0; k < nblk; ++k) {
      e2 += part[((int64_t)b * kLossBlocks + k) * 2 + 0];
      t2 += part[((int64_t)b * kLossBlocks + k) * 2 + 1];
    }
    part[(int64_t)B * kLossBlocks * 2 + b * 2 + 0] = e2;
    part[(int64_t)B * kLossBlocks * 2 + b * 2 + 1] = t2;
    loss += e2 / fmax(t2, 1e-300);
  }
  part[(int64_t)B * kLossBlocks * 2 + 2 * B] = loss / B;
}

__global__ void k_loss_grad(const float* y, const float* t, int64_t per_img, int B, const double* part,
                            uint16_t* gy) {
  const int64_t n = per_img * B;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(i / per_img);
    const double t2 = part[(int64_t)B * kLossBlocks * 2 + b * 2 + 1];
    const double g = 2.0 * ((double)y[i] - (double)t[i]) / (fmax(t2, 1e-300) * B);
    gy[i] = f2bf((float)g);
  }
}

void grid1d(int64_t n, int& g) { g = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }

}  // namespace

// ---- UNet -------------------------------------------------------------------------------------------
void UNet::build(int L, const int* chans, int cin, int cout, int act, int B, int H0, int W0) {
  APX_REQUIRE(L >= 1 && L <= kMaxLevels, APX_ERR_ARG, "levels must be 1..12");
  APX_REQUIRE(cin == 1 || cin == 2, APX_ERR_UNSUPPORTED, "input channels must be 1 or 2");
  APX_REQUIRE(cout == 1 || cout == 2, APX_ERR_UNSUPPORTED, "output channels must be 1 or 2");
  APX_REQUIRE(B >= 1 && H0 >= 3 && W0 >= 3, APX_ERR_ARG, "invalid grid");
  levels = L;
  c_in = cin;
  c_out = cout;
  relu = act ? 1 : 0;
  batch = B;
  for (int l = 0; l < L; ++l) {
    APX_REQUIRE(chans[l] == 16 || chans[l] == 32 || chans[l] == 64, APX_ERR_UNSUPPORTED,
                "channels per level must be 16, 32 or 64");
    ch[l] = chans[l];
    H[l] = l ? (H[l - 1] + 1) / 2 : H0;
    W[l] = l ? (W[l - 1] + 1) / 2 : W0;
    APX_REQUIRE(H[l] >= 2 && W[l] >= 2, APX_ERR_ARG, "too many levels for the grid (coarsest level below 2 x 2)");
  }
  APX_REQUIRE(ch[0] == 16, APX_ERR_UNSUPPORTED, "the finest level must have 16 channels");
  layers.clear();
  auto add = [&](int ci, int co, int r, int f32) {
    NetLayer ly;
    ly.ci = ci;
    ly.co = co;
    ly.cop = round16(co);
    ly.wk = kblocks(ci) * 16;
    ly.relu = r;
    ly.out_f32 = f32;
    ly.d_cop = round16(ci);
    ly.d_wk = kblocks(co) * 16;
    layers.push_back(ly);
    return (int)layers.size() - 1;
  };
  L_in = add(c_in, ch[0], relu, 0);
  for (int l = 0; l + 1 < L; ++l) {
    L_e[l] = add(ch[l], ch[l], relu, 0);
    L_down[l] = add(ch[l], ch[l + 1], relu, 0);
  }
  L_b = add(ch[L - 1], ch[L - 1], relu, 0);
  for (int l = L - 2; l >= 0; --l) {
    L_up[l] = add(ch[l + 1], ch[l], 0, 0);
    L_d[l] = add(ch[l], ch[l], relu, 0);
  }
  L_out = add(ch[0], c_out, 0, 1);
  n_params = n_wf = n_wd = 0;
  for (auto& ly : layers) {
    ly.w_off = n_params;
    n_params += (int64_t)ly.cop * ly.wk;
    ly.b_off = n_params;
    n_params += ly.co;
    ly.wf_off = n_wf;
    n_wf += (int64_t)ly.cop * ly.wk;
    ly.wd_off = n_wd;
    n_wd += (int64_t)ly.d_cop * ly.d_wk;
  }
  theta.alloc(n_params);
  grad.alloc(n_params);
  adam_m.alloc(n_params);
  adam_v.alloc(n_params);
  APX_HIP(hipMemset(theta.p, 0, theta.bytes()));
  APX_HIP(hipMemset(grad.p, 0, grad.bytes()));
  APX_HIP(hipMemset(adam_m.p, 0, adam_m.bytes()));
  APX_HIP(hipMemset(adam_v.p, 0, adam_v.bytes()));
  wf.alloc(n_wf);
  wd.alloc(n_wd);
  int64_t maxpart = 0;
  for (auto& ly : layers) maxpart = std::max<int64_t>(maxpart, (int64_t)ly.cop * ly.wk + ly.cop);
  wpart.alloc(maxpart * kWgradBlocks);
  x0.alloc(pixels(0) * c_in);
  for (int l = 0; l < L; ++l) {
    a[l].alloc(pixels(l) * ch[l]);
    u[l].alloc(pixels(l) * ch[l]);
    gA[l].alloc(pixels(l) * ch[l]);
    gU[l].alloc(pixels(l) * ch[l]);
    if (l + 1 < L) {
      d[l].alloc(pixels(l) * ch[l]);
      s[l].alloc(pixels(l) * ch[l]);
      gD[l].alloc(pixels(l) * ch[l]);
      gS[l].alloc(pixels(l) * ch[l]);
      rd[l].alloc(pixels(l + 1) * ch[l]);
      t[l].alloc(pixels(l + 1) * ch[l]);
      gR[l].alloc(pixels(l + 1) * ch[l]);
      gT[l].alloc(pixels(l + 1) * ch[l]);
    }
  }
  y.alloc(pixels(0) * c_out);
  gy.alloc(pixels(0) * c_out);
  loss_part.alloc((int64_t)batch * kLossBlocks * 2 + 2 * batch + 1);
  scale.alloc(batch);
  weights_ready = false;
  build_tail();
}

// The tail starts at the finest level c whose restriction begins a run of 64-channel levels
// with convolution grids of at most kTailMaxGrid nodes a side (and a step list that fits).
void UNet::build_tail() {
  const int L = levels;
  tail = TailArgs{};
  tail_level = -1;
  for (int c = 1; c + 1 < L && tail_level < 0; ++c) {
    bool ok = H[c + 1] <= kTailMaxGrid && W[c + 1] <= kTailMaxGrid && 6 * (L - 1 - c) - 1 <= kTailMaxSteps;
    for (int l = c; l < L; ++l) ok = ok && ch[l] == 64;
    if (ok) tail_level = c;
  }
  if (tail_level < 0) return;
  const int c = tail_level;
  // LDS buffers: level l in (c, L) has X_l, Y_l and (l < L - 1) the skip buffer D_l, each
  // (H + 2) x (W + 2) nodes of kTailS elements; an offset of -1 stands for the HBM tensor
  int X[kMaxLevels], Y[kMaxLevels], D[kMaxLevels];
  int pool = 0;
  for (int l = c + 1; l < L; ++l) {
    const int sz = (H[l] + 2) * (W[l] + 2) * kTailS;  // a multiple of 8
    X[l] = pool;
    Y[l] = pool + sz;
    D[l] = l < L - 1 ? pool + 2 * sz : -1;
    pool += (l < L - 1 ? 3 : 2) * sz;
  }
  if (pool > kTailPool) {  // the coarse levels do not fit in LDS: per-layer launches
    tail_level = -1;
    tail = TailArgs{};
    return;
  }
  tail.pool = pool;
  auto conv = [&](int li, int level, int xo, uint16_t* y, int yo) {
    const NetLayer& ly = layers[li];
    TailStep& ts = tail.step[tail.n++];
    ts.op = kTailConv;
    ts.y = y;
    ts.xo = xo;
    ts.yo = yo;
    ts.so = -1;
    ts.w = wf.p + ly.wf_off;
    ts.bias = theta.p + ly.b_off;
    ts.relu = ly.relu;
    ts.H = H[level];
    ts.W = W[level];
  };
  auto transfer = [&](int op, int l, const uint16_t* x, int xo, const uint16_t* skip, int so, uint16_t* y, int yo) {
    TailStep& ts = tail.step[tail.n++];
    ts.op = op;
    ts.x = x;
    ts.xo = xo;
    ts.skip = skip;
    ts.so = so;
    ts.y = y;
    ts.yo = yo;
    ts.H = H[l];
    ts.W = W[l];
    ts.Hc = H[l + 1];
    ts.Wc = W[l + 1];
  };
  // (each step's HBM tensor is also named: training writes it for the backward pass)
  transfer(kTailRestrict, c, d[c].p, -1, nullptr, -1, rd[c].p, X[c + 1]);      // rd_c -> X_{c+1}
  conv(L_down[c], c + 1, X[c + 1], a[c + 1].p, Y[c + 1]);                     // a_{c+1} -> Y_{c+1}
  for (int l = c + 1; l + 1 < L; ++l) {
    conv(L_e[l], l, Y[l], d[l].p, D[l]);                                      // d_l -> D_l
    transfer(kTailRestrict, l, nullptr, D[l], nullptr, -1, rd[l].p, X[l + 1]);  // rd_l -> X_{l+1}
    conv(L_down[l], l + 1, X[l + 1], a[l + 1].p, Y[l + 1]);                   // a_{l+1} -> Y_{l+1}
  }
  conv(L_b, L - 1, Y[L - 1], u[L - 1].p, X[L - 1]);                           // u_{L-1} -> X_{L-1}
  for (int l = L - 2; l > c; --l) {
    conv(L_up[l], l + 1, X[l + 1], t[l].p, Y[l + 1]);                         // t_l -> Y_{l+1}
    transfer(kTailProlong, l, nullptr, Y[l + 1], nullptr, D[l], s[l].p, Y[l]);  // s_l -> Y_l
    conv(L_d[l], l, Y[l], u[l].p, X[l]);                                      // u_l -> X_l
  }
  conv(L_up[c], c + 1, X[c + 1], t[c].p, Y[c + 1]);                           // t_c -> Y_{c+1}
  transfer(kTailProlong, c, nullptr, Y[c + 1], d[c].p, -1, s[c].p, -1);       // s_c (HBM)
  APX_REQUIRE(tail.n == 6 * (L - 1 - c) - 1, APX_ERR_STATE, "coarse-level tail: step count");
  // each convolution prefetches the weights of the next one (the last one its own)
  const uint16_t* next = nullptr;
  for (int i = tail.n - 1; i >= 0; --i)
    if (tail.step[i].op == kTailConv) {
      tail.step[i].w_next = next ? next : tail.step[i].w;
      next = tail.step[i].w;
    }
  tail.w_first = next;
}

void UNet::prepare_weights(hipStream_t st) {
  for (auto& ly : layers) {
    int g;
    grid1d((int64_t)ly.cop * ly.wk, g);
    hipLaunchKernelGGL(k_cast_weights, dim3(g), dim3(256), 0, st, theta.p, wf.p, ly.w_off, ly.wf_off,
                       (int64_t)ly.cop * ly.wk);
    grid1d((int64_t)ly.d_cop * ly.d_wk, g);
    hipLaunchKernelGGL(k_flip_weights, dim3(g), dim3(256), 0, st, theta.p, wd.p, ly.w_off, ly.wd_off, ly.ci, ly.co,
                       ly.wk, ly.d_cop, ly.d_wk);
  }
  APX_LAUNCH_CHECK();
  weights_ready = true;
}

namespace {
void conv_fwd(UNet& N, int li, int level, const uint16_t* x, void* y, hipStream_t st) {
  const NetLayer& ly = N.layers[li];
  ConvArgs a{};
  a.x = x;
  a.xmask = nullptr;
  a.w = N.wf.p + ly.wf_off;
  a.bias = N.theta.p + ly.b_off;
  a.y = y;
  a.yscale = nullptr;
  a.B = N.batch;
  a.H = N.H[level];
  a.W = N.W[level];
  a.relu = ly.relu;
  a.stop = N.stop;
  launch_conv(ly.ci, ly.co, ly.out_f32 != 0, a, st);
}

// dX = conv^T(dY masked by relu(Y) > 0)
void conv_dgrad(UNet& N, int li, int level, const uint16_t* dy, const uint16_t* ymask, uint16_t* dx, hipStream_t st) {
  const NetLayer& ly = N.layers[li];
  ConvArgs a{};
  a.x = dy;
  a.xmask = ly.relu ? ymask : nullptr;
  a.w = N.wd.p + ly.wd_off;
  a.bias = nullptr;
  a.y = dx;
  a.B = N.batch;
  a.H = N.H[level];
  a.W = N.W[level];
  a.relu = 0;
  launch_conv(ly.co, ly.ci, false, a, st);
}

void conv_wgrad(UNet& N, int li, int level, const uint16_t* x, const uint16_t* dy, const uint16_t* ymask,
                hipStream_t st) {
  const NetLayer& ly = N.layers[li];
  WgradArgs a{};
  a.x = x;
  a.dy = dy;
  a.ymask = ly.relu ? ymask : nullptr;
  a.partial = N.wpart.p;
  a.B = N.batch;
  a.H = N.H[level];
  a.W = N.W[level];
  const int nblk = launch_wgrad(ly.ci, ly.co, a, st);
  const int64_t n_out = (int64_t)ly.cop * ly.wk + ly.co;
  hipLaunchKernelGGL(k_reduce_wgrad, dim3((unsigned)((n_out + 63) / 64)), dim3(64 * kReduceWaves), 0, st, N.wpart.p,
                     nblk, (int64_t)ly.cop * ly.wk + ly.cop, ly.cop, ly.wk, ly.co, N.grad.p, ly.w_off, ly.b_off);
}

// d_0 = conv_e[0](conv_in(x0)) without storing conv_in's output: the input layer is formed while
// staging the level-0 tiles (bit-identical to the two launches)
void conv_fwd_in_e0(UNet& N, hipStream_t st) {
  const NetLayer& li = N.layers[N.L_in];
  const NetLayer& le = N.layers[N.L_e[0]];
  ConvArgs a{};
  a.x = nullptr;
  a.xmask = nullptr;
  a.xin = N.x0.p;
  a.win = N.wf.p + li.wf_off;
  a.bin = N.theta.p + li.b_off;
  a.relu_in = li.relu;
  a.w = N.wf.p + le.wf_off;
  a.bias = N.theta.p + le.b_off;
  a.y = N.d[0].p;
  a.yscale = nullptr;
  a.B = N.batch;
  a.H = N.H[0];
  a.W = N.W[0];
  a.relu = le.relu;
  a.stop = N.stop;
  APX_REQUIRE(li.co == 16 && le.ci == 16 && le.co == 16, APX_ERR_UNSUPPORTED, "fused input layer: 16 channels");
  if (N.c_in == 1)
    launch_conv_ws_t<16, 16, 4, 1>(a, st);
  else if (N.c_in == 2)
    launch_conv_ws_t<16, 16, 4, 2>(a, st);
  else
    throw Error(APX_ERR_UNSUPPORTED, "fused input layer: 1 or 2 input channels");
}

// y = out(conv_d[0](s_0)) without storing conv_d's output (bit-identical to the two launches)
void conv_fwd_d_out(UNet& N, hipStream_t st) {
  const NetLayer& ld = N.layers[N.L_d[0]];
  const NetLayer& lo = N.layers[N.L_out];
  APX_REQUIRE(ld.ci == 16 && ld.co == 16 && lo.ci == 16 && (lo.co == 1 || lo.co == 2), APX_ERR_UNSUPPORTED,
              "fused output layer: 16 -> 16 -> 1 or 2 channels");
  ConvArgs a{};
  a.x = N.s[0].p;
  a.w = N.wf.p + ld.wf_off;
  a.bias = N.theta.p + ld.b_off;
  a.relu = ld.relu;
  a.w2 = N.wf.p + lo.wf_off;
  a.bias2 = N.theta.p + lo.b_off;
  a.relu2 = lo.relu;
  a.y = N.y.p;
  a.yscale = nullptr;
  a.B = N.batch;
  a.H = N.H[0];
  a.W = N.W[0];
  a.stop = N.stop;
  const int64_t tiles = (int64_t)a.B * ((a.W + kTileW - 3) / (kTileW - 2)) * ((a.H + 3) / 4);
  APX_REQUIRE(tiles < (int64_t(1) << 30) && (int64_t)a.H * a.W * 16 < INT32_MAX, APX_ERR_UNSUPPORTED,
              "fused output layer: grid too large for 32-bit indices");
  auto go = [&](auto kernel) {
    static const int64_t cap = resident_blocks(kernel);
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
    launch_kernel(kernel, dim3(g), dim3(kConvThreads), 0, st, a);
  };
  if (lo.co == 1)
    go(k_conv_d_out<1>);
  else
    go(k_conv_d_out<2>);
}

void prolong_add(UNet& N, int level_fine, int C, const uint16_t* coarse, const uint16_t* skip, uint16_t* out,
                 hipStream_t st) {
  const int H = N.H[level_fine], W = N.W[level_fine];
  const int Hc = N.H[level_fine + 1], Wc = N.W[level_fine + 1];
  APX_REQUIRE(C % 8 == 0 && Hc < 65536 && N.batch < 65536 && (int64_t)H * W * C < INT32_MAX, APX_ERR_UNSUPPORTED,
              "grid transfer shape");
  const dim3 grid((Wc * (C / 8) + 255) / 256, Hc, N.batch);  // one thread per coarse node and chunk
  if (C == 16)
    launch_kernel(k_prolong_add<2>, grid, dim3(256), 0, st, coarse, skip, out, H, W, Hc, Wc, N.stop);
  else if (C == 32)
    launch_kernel(k_prolong_add<4>, grid, dim3(256), 0, st, coarse, skip, out, H, W, Hc, Wc, N.stop);
  else if (C == 64)
    launch_kernel(k_prolong_add<8>, grid, dim3(256), 0, st, coarse, skip, out, H, W, Hc, Wc, N.stop);
  else
    throw Error(APX_ERR_UNSUPPORTED, "grid transfer: channels must be 16, 32 or 64");
}

void restrict_(UNet& N, int level_fine, int C, const uint16_t* fine, uint16_t* out, hipStream_t st) {
  const int H = N.H[level_fine], W = N.W[level_fine];
  const int Hc = N.H[level_fine + 1], Wc = N.W[level_fine + 1];
  APX_REQUIRE(C % 8 == 0 && Hc < 65536 && N.batch < 65536 && (int64_t)H * W * C < INT32_MAX, APX_ERR_UNSUPPORTED,
              "grid transfer shape");
  const dim3 grid((Wc * (C / 8) + 255) / 256, Hc, N.batch);
  if (C == 16)
    launch_kernel(k_restrict<2>, grid, dim3(256), 0, st, fine, out, H, W, Hc, Wc, N.stop);
  else if (C == 32)
    launch_kernel(k_restrict<4>, grid, dim3(256), 0, st, fine, out, H, W, Hc, Wc, N.stop);
  else if (C == 64)
    launch_kernel(k_restrict<8>, grid, dim3(256), 0, st, fine, out, H, W, Hc, Wc, N.stop);
  else
    throw Error(APX_ERR_UNSUPPORTED, "grid transfer: channels must be 16, 32 or 64");
}
}  // namespace

int64_t UNet::conv_bytes(int li, int level) const {
  const NetLayer& ly = layers[li];
  return pixels(level) * (ly.ci * 2 + ly.co * (ly.out_f32 ? 4 : 2)) + (int64_t)ly.cop * ly.wk * 2;
}

int64_t UNet::forward_bytes() const {
  const int L = levels;
  int64_t b = conv_bytes(L_in, 0) + conv_bytes(L_b, L - 1) + conv_bytes(L_out, 0);
  for (int l = 0; l + 1 < L; ++l) {
    b += conv_bytes(L_e[l], l) + conv_bytes(L_down[l], l + 1) + conv_bytes(L_up[l], l + 1) + conv_bytes(L_d[l], l);
    b += (pixels(l) + pixels(l + 1)) * ch[l] * 2;          // restriction
    b += (2 * pixels(l) + pixels(l + 1)) * ch[l] * 2;      // prolongation + skip
  }
  return b;
}

void UNet::collect_timing(size_t count) {
  if (!timer_plan) return;
  timing_collect(*timer_plan, "unet_conv_d0", ev_conv0, levels > 1 ? conv_bytes(L_d[0], 0) : 0, count);
  timing_collect(*timer_plan, "unet_forward", ev_fwd, forward_bytes(), count);
  // per-launch kernel records of the first `count` forward passes
  Plan& P = *timer_plan;
  for (auto& kv : ev_launch) {
    auto& t = P.timers[kv.first];
    double ms_sum = t.total_ms;
    int64_t n0 = t.count, bytes_sum = t.bytes * t.count, flops_sum = t.flops * t.count;
    for (auto& rec : kv.second) {
      APX_HIP(hipEventSynchronize(rec.e.b));
      if (rec.fwd < count) {
        float ms = 0.f;
        APX_HIP(hipEventElapsedTime(&ms, rec.e.a, rec.e.b));
        ms_sum += ms;
        bytes_sum += rec.bytes;
        flops_sum += rec.flops;
        ++n0;
      }
      P.evpool.push_back(rec.e.a);
      P.evpool.push_back(rec.e.b);
    }
    t.total_ms = ms_sum;
    t.count = n0;
    t.bytes = n0 ? bytes_sum / n0 : 0;
    t.flops = n0 ? flops_sum / n0 : 0;
    kv.second.clear();
  }
  n_fwd = 0;
}

namespace {
// filter: kernel names separated by ';' (the names themselves contain commas)
bool filter_has(const std::string& filter, const std::string& name) {
  size_t pos = 0;
  while (pos <= filter.size()) {
    const size_t end = std::min(filter.find(';', pos), filter.size());
    if (filter.compare(pos, end - pos, name) == 0) return true;
    pos = end + 1;
  }
  return false;
}
}  // namespace

void UNet::forward(hipStream_t st, bool keep_activations) {
  APX_REQUIRE(weights_ready, APX_ERR_STATE, "network weights not prepared");
  const int L = levels;
  const bool timed = timer_plan && timer_plan->timing;
  EventPair ef = timed ? timing_begin(*timer_plan, st) : EventPair{};
  // per-launch records: every kernel at timing level 2, the filtered names at level 1
  const bool rec_all = timed && timer_plan->timing >= 2;
  const bool rec_some = timed && !timer_plan->timing_filter.empty();
  const size_t fwd_index = n_fwd;
  auto record = [&](const std::string& name, int64_t bytes, int64_t flops, auto&& launch) {
    if (!(rec_all || (rec_some && filter_has(timer_plan->timing_filter, name)))) return launch();
    // the pair takes the kernel's own start and end timestamps (launch_kernel)
    EventPair e = timing_pair(*timer_plan);
    struct Pending {
      ~Pending() { t_launch_events = nullptr; }
    } pending;
    t_launch_events = &e;
    launch();
    APX_REQUIRE(t_launch_events == nullptr, APX_ERR_STATE, "timed network launch: the kernel did not take its events");
    ev_launch[name].push_back(LaunchRec{e, fwd_index, bytes, flops});
  };
  auto conv = [&](int li, int level, const uint16_t* x, void* y) {
    if (!(rec_all || rec_some)) return conv_fwd(*this, li, level, x, y, st);
    const NetLayer& ly = layers[li];
    ConvArgs probe{};
    probe.B = batch;
    probe.H = H[level];
    probe.W = W[level];
    record(conv_kernel_name(ly.ci, ly.co, ly.out_f32 != 0, probe), conv_bytes(li, level),
           (int64_t)18 * ly.ci * ly.co * pixels(level), [&] { conv_fwd(*this, li, level, x, y, st); });
  };
  auto transfer = [&](bool restriction, int l) {
    const int C8 = ch[l] / 8;
    const int64_t bytes = (restriction ? pixels(l) : 2 * pixels(l)) * ch[l] * 2 + pixels(l + 1) * ch[l] * 2;
    const std::string name = std::string(restriction ? "k_restrict<" : "k_prolong_add<") + std::to_string(C8) + ">";
    if (restriction)
      record(name, bytes, 0, [&] { restrict_(*this, l, ch[l], d[l].p, rd[l].p, st); });
    else
      record(name, bytes, 0, [&] { prolong_add(*this, l, ch[l], t[l].p, d[l].p, s[l].p, st); });
  };
  // levels >= c from the restriction of level c to its prolongation: one launch (the tail)
  const bool fused = fuse_tail && tail_level >= 0;
  const int c = fused ? tail_level : L;
  // inference: the input layer is formed inside the level-0 conv_e launch (its output a_0 is
  // read by nothing else; training keeps it for the backward pass)
  const bool fused_in = fuse_tail && !keep_activations && L > 1 && ch[0] == 16;
  if (fused_in) {
    const NetLayer& le = layers[L_e[0]];
    record("k_conv3x3_ws<16, 16, 4, " + std::to_string(c_in) + ">",
           pixels(0) * (c_in * 2 + le.co * 2) + (int64_t)(layers[L_in].cop * layers[L_in].wk + le.cop * le.wk) * 2,
           (int64_t)18 * (c_in * 16 + 16 * le.co) * pixels(0), [&] { conv_fwd_in_e0(*this, st); });
  } else {
    conv(L_in, 0, x0.p, a[0].p);
  }
  for (int l = 0; l + 1 < L && l <= c; ++l) {
    if (l > 0 || !fused_in) conv(L_e[l], l, a[l].p, d[l].p);
    if (l == c) break;
    transfer(true, l);
    conv(L_down[l], l + 1, rd[l].p, a[l + 1].p);
  }
  if (fused)
    record("k_unet_tail", 0, 0,
           [&] {
             launch_kernel(k_unet_tail, dim3(batch), dim3(kTailThreads), 0, st, tail, stop, keep_activations ? 1 : 0);
           });
  else
    conv(L_b, L - 1, a[L - 1].p, u[L - 1].p);
  // inference: the level-0 conv_d and the output layer in one launch (conv_d's output u_0 is read
  // by nothing else; training keeps it for the backward pass)
  const NetLayer& lo = layers[L_out];
  const bool fused_out = fuse_tail && !keep_activations && L > 1 && ch[0] == 16 && layers[L_d[0]].ci == 16 &&
                         layers[L_d[0]].co == 16 && lo.ci == 16 && (lo.co == 1 || lo.co == 2);
  for (int l = std::min(L - 2, c); l >= 0; --l) {
    if (l != c) {
      conv(L_up[l], l + 1, u[l + 1].p, t[l].p);
      transfer(false, l);
    }
    if (l == 0 && fused_out) break;
    EventPair ec{};
    if (timed && l == 0) ec = timing_begin(*timer_plan, st);
    conv(L_d[l], l, s[l].p, u[l].p);
    if (timed && l == 0) {
      timing_end(*timer_plan, ec, st);
      ev_conv0.push_back(ec);
    }
  }
  if (fused_out)
    record("k_conv_d_out<" + std::to_string(lo.co) + ">",
           pixels(0) * (16 * 2 + lo.co * 4) + (int64_t)(layers[L_d[0]].cop * layers[L_d[0]].wk + lo.cop * lo.wk) * 2,
           (int64_t)18 * 16 * (16 + lo.co) * pixels(0), [&] { conv_fwd_d_out(*this, st); });
  else
    conv(L_out, 0, u[0].p, y.p);
  APX_LAUNCH_CHECK();
  if (timed) {
    timing_end(*timer_plan, ef, st);
    ev_fwd.push_back(ef);
    ++n_fwd;
  }
}

double UNet::loss_backward(const float* target, hipStream_t st) {
  const int L = levels;
  const int64_t per_img = (int64_t)H[0] * W[0] * c_out;
  hipLaunchKernelGGL(k_loss_partials, dim3(kLossBlocks), dim3(256), 0, st, y.p, target, per_img, batch, loss_part.p);
  hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, st, loss_part.p, batch, kLossBlocks);
  int g;
  grid1d(per_img * batch, g);
  hipLaunchKernelGGL(k_loss_grad, dim3(g), dim3(256), 0, st, y.p, target, per_img, batch, loss_part.p, gy.p);
  backward(st);
  double loss = 0.0;
  APX_HIP(hipMemcpyAsync(&loss, loss_part.p + (int64_t)batch * kLossBlocks * 2 + 2 * batch, sizeof(double),
                         hipMemcpyDeviceToHost, st));
  APX_HIP(hipStreamSynchronize(st));
  return loss;
}

void UNet::backward(hipStream_t st) {
  const int L = levels;
  // backward through the decoder
  conv_wgrad(*this, L_out, 0, u[0].p, gy.p, nullptr, st);
  conv_dgrad(*this, L_out, 0, gy.p, nullptr, gU[0].p, st);
  for (int l = 0; l + 1 < L; ++l) {
    conv_wgrad(*this, L_d[l], l, s[l].p, gU[l].p, u[l].p, st);
    conv_dgrad(*this, L_d[l], l, gU[l].p, u[l].p, gS[l].p, st);
    restrict_(*this, l, ch[l], gS[l].p, gT[l].p, st);  // P^T = R
    conv_wgrad(*this, L_up[l], l + 1, u[l + 1].p, gT[l].p, nullptr, st);
    conv_dgrad(*this, L_up[l], l + 1, gT[l].p, nullptr, gU[l + 1].p, st);
  }
  conv_wgrad(*this, L_b, L - 1, a[L - 1].p, gU[L - 1].p, u[L - 1].p, st);
  conv_dgrad(*this, L_b, L - 1, gU[L - 1].p, u[L - 1].p, gA[L - 1].p, st);
  for (int l = L - 2; l >= 0; --l) {
    conv_wgrad(*this, L_down[l], l + 1, rd[l].p, gA[l + 1].p, a[l + 1].p, st);
    conv_dgrad(*this, L_down[l], l + 1, gA[l + 1].p, a[l + 1].p, gR[l].p, st);
    prolong_add(*this, l, ch[l], gR[l].p, gS[l].p, gD[l].p, st);  // R^T = P, plus the skip
    conv_wgrad(*this, L_e[l], l, a[l].p, gD[l].p, d[l].p, st);
    conv_dgrad(*this, L_e[l], l, gD[l].p, d[l].p, gA[l].p, st);
  }
  conv_wgrad(*this, L_in, 0, x0.p, gA[0].p, a[0].p, st);
  if (!relu) {
    // linear network: no biases, so the preconditioner is linear (and odd)
    for (auto& ly : layers) APX_HIP(hipMemsetAsync(grad.p + ly.b_off, 0, ly.co * sizeof(float), st));
  }
  APX_LAUNCH_CHECK();
}

void UNet::adam(float lr, float b1, float b2, float eps, int step, hipStream_t st) {
  int g;
  grid1d(n_params, g);
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(g), dim3(256), 0, st, theta.p, grad.p, adam_m.p, adam_v.p, n_params, lr, b1, b2, eps,
                     c1, c2);
  APX_LAUNCH_CHECK();
  prepare_weights(st);
}

// ---- reduced vector <-> node grid ----------------------------------------------------------------------
namespace {
template <class T>
__global__ void k_sq_partials(int64_t n, const T* r, const double* dsq, double* part, const int* stop) {
  __shared__ double lds[4];
  if (stop && *stop) return;
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = dsq ? (double)r[i] * dsq[i] : (double)r[i];
    s += v * v;
  }
  s = block_sum<256>(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_scale_fin(const double* part, int nb, const double* rr_src, int64_t n, float* scale, int img,
                            const int* stop) {
  __shared__ double lds[4];
  if (stop && *stop) return;
  double s = 0.0;
  if (rr_src) {
    s = *rr_src;
  } else {
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
    s = block_sum<256>(s, lds);
  }
  if (threadIdx.x == 0) {
    const double rms = sqrt(s / (double)(n > 0 ? n : 1));
    scale[img] = rms > 0.0 ? (float)rms : 1.f;
  }
}

template <class T>
__global__ void k_to_grid(int64_t n_nodes, int d, int cin, const uint8_t* dmask, const int32_t* full2red, const T* r,
                          const double* dsq, const double* g, const float* scale, int img, uint16_t* x,
                          const int* stop) {
  const float inv = 1.f / scale[img];  // issued together with the stop flag's load
  if (stop && *stop) return;
  for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < n_nodes;
       node += (int64_t)gridDim.x * blockDim.x) {
    uint16_t* xo = x + ((int64_t)img * n_nodes + node) * cin;
    for (int c = 0; c < d; ++c) {
      // loads unconditional (a Dirichlet DOF reads reduced entry 0 and stores zero), so that
      // they are not serialised behind a per-lane branch
      const int64_t gd = node * d + c;
      const bool fixed = dmask[gd] != 0;
      const int32_t i0 = full2red[gd];
      const int32_t i = fixed ? 0 : i0;
      const float v = (float)((double)r[i] * (dsq ? dsq[i] : 1.0)) * inv;
      xo[c] = fixed ? (uint16_t)0 : f2bf(v);
    }
    if (cin > d) xo[d] = f2bf(g ? (float)g[node] : 0.f);
  }
}

template <class T>
__global__ void k_from_grid(int64_t n_free, int d, const int32_t* red2full, const float* y, const double* dsq,
                            const double* r_skip, const float* scale, int img, int64_t n_nodes, T* z) {
  const float s = scale[img];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_free; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t gd = red2full[i];
    const double v = (double)(s * y[(int64_t)img * n_nodes * d + gd]);
    double zi = dsq ? v * dsq[i] : v;
    if (r_skip) zi += r_skip[i] * dsq[i] * dsq[i];  // Jacobi part D^-1 r
    z[i] = (T)zi;
  }
}

template <class T>
__global__ void k_dsq(int64_t n, const int32_t* diag_slot, const T* vals, double* dsq) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = fabs((double)vals[diag_slot[i]]);
    dsq[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
  }
}
}  // namespace

void net_set_diag_scaling(Plan& P, const void* vals_sell, int precision, hipStream_t st) {
  if (P.net_dsq.n < (size_t)P.n_free) P.net_dsq.alloc(P.n_free);
  int g;
  grid1d(P.n_free, g);
  if (precision == APX_FP64)
    hipLaunchKernelGGL(k_dsq<double>, dim3(g), dim3(256), 0, st, P.n_free, P.diag_slot.p, (const double*)vals_sell,
                       P.net_dsq.p);
  else
    hipLaunchKernelGGL(k_dsq<float>, dim3(g), dim3(256), 0, st, P.n_free, P.diag_slot.p, (const float*)vals_sell,
                       P.net_dsq.p);
  APX_LAUNCH_CHECK();
  P.net_dsq_set = true;
}

void net_input_from_reduced_img(Plan& P, UNet& N, const void* r, int r_f64, const double* g_nodal,
                                const double* rr_src, int img, hipStream_t st) {
  APX_REQUIRE((int64_t)N.H[0] * N.W[0] == P.n_nodes, APX_ERR_ARG, "network grid does not match the plan's nodes");
  APX_REQUIRE(N.c_in == P.d || N.c_in == P.d + 1, APX_ERR_ARG, "network input channels do not match the DOFs");
  APX_REQUIRE(img >= 0 && img < N.batch, APX_ERR_ARG, "image index out of range");
  APX_REQUIRE(P.n_free > 0, APX_ERR_ARG, "network input: the plan has no free DOFs");
  ensure_krylov_scratch(P, st);
  // the scale is the RMS of the network's input (r scaled by D^-1/2 when the plan has the
  // scaling); rr_src, when given, is that input's sum of squares on the device
  const double* dsq = P.net_dsq_set ? P.net_dsq.p : nullptr;
  int nb = 1;
  if (!rr_src) {
    grid1d(P.n_free, nb);
    nb = std::min(nb, 1024);
    if (r_f64)
      hipLaunchKernelGGL(k_sq_partials<double>, dim3(nb), dim3(256), 0, st, P.n_free, (const double*)r, dsq,
                         P.partials.p, N.stop);
    else
      hipLaunchKernelGGL(k_sq_partials<float>, dim3(nb), dim3(256), 0, st, P.n_free, (const float*)r, dsq,
                         P.partials.p, N.stop);
  }
  hipLaunchKernelGGL(k_scale_fin, dim3(1), dim3(256), 0, st, P.partials.p, nb, rr_src, P.n_free, N.scale.p, img,
                     N.stop);
  int g;
  grid1d(P.n_nodes, g);
  if (r_f64)
    hipLaunchKernelGGL(k_to_grid<double>, dim3(g), dim3(256), 0, st, P.n_nodes, P.d, N.c_in, P.dmask.p, P.full2red.p,
                       (const double*)r, dsq, g_nodal, N.scale.p, img, N.x0.p, N.stop);
  else
    hipLaunchKernelGGL(k_to_grid<float>, dim3(g), dim3(256), 0, st, P.n_nodes, P.d, N.c_in, P.dmask.p, P.full2red.p,
                       (const float*)r, dsq, g_nodal, N.scale.p, img, N.x0.p, N.stop);
  APX_LAUNCH_CHECK();
}

namespace {
__global__ void k_target_grid(int64_t n_dofs, const uint8_t* dmask, const int32_t* full2red, const double* e,
                              const double* dsq, const float* scale, int img, float* t) {
  const float inv = 1.f / scale[img];
  for (int64_t gd = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; gd < n_dofs; gd += (int64_t)gridDim.x * blockDim.x) {
    if (dmask[gd]) {
      t[(int64_t)img * n_dofs + gd] = 0.f;
    } else {
      const int32_t i = full2red[gd];
      t[(int64_t)img * n_dofs + gd] = (float)(e[i] / (dsq ? dsq[i] : 1.0)) * inv;
    }
  }
}
}  // namespace

void net_target_from_reduced_img(Plan& P, UNet& N, const double* e, float* target, int img, hipStream_t st) {
  APX_REQUIRE(N.c_out == P.d, APX_ERR_ARG, "network output channels do not match the DOFs");
  APX_REQUIRE(img >= 0 && img < N.batch, APX_ERR_ARG, "image index out of range");
  int g;
  grid1d(P.n_dofs, g);
  hipLaunchKernelGGL(k_target_grid, dim3(g), dim3(256), 0, st, P.n_dofs, P.dmask.p, P.full2red.p, e,
                     P.net_dsq_set ? P.net_dsq.p : nullptr, N.scale.p, img, target);
  APX_LAUNCH_CHECK();
}

void net_output_to_reduced_img(Plan& P, UNet& N, const double* r, void* z, int z_f64, int img, hipStream_t st) {
  APX_REQUIRE(N.c_out == P.d, APX_ERR_ARG, "network output channels do not match the DOFs");
  int g;
  grid1d(P.n_free, g);
  const double* dsq = P.net_dsq_set ? P.net_dsq.p : nullptr;
  const double* r_skip = nullptr;
  if (N.jacobi_skip) {
    APX_REQUIRE(dsq && r, APX_ERR_STATE, "the Jacobi part of the network needs the matrix diagonal and the residual");
    r_skip = r;
  }
  if (z_f64)
    hipLaunchKernelGGL(k_from_grid<double>, dim3(g), dim3(256), 0, st, P.n_free, P.d, P.red2full.p, N.y.p, dsq,
                       r_skip, N.scale.p, img, P.n_nodes, (double*)z);
  else
    hipLaunchKernelGGL(k_from_grid<float>, dim3(g), dim3(256), 0, st, P.n_free, P.d, P.red2full.p, N.y.p, dsq,
                       r_skip, N.scale.p, img, P.n_nodes, (float*)z);
  APX_LAUNCH_CHECK();
}

// ---- residual-loss training step -------------------------------------------------------------
// For image b with residual r_b (reduced, FP64) and optional error e_b = A^-1 r_b:
//   z_b = s_b y_b on the free DOFs,  L_b = ||r_b - A z_b||^2 / ||r_b||^2 + w_e ||z_b - e_b||^2 / ||e_b||^2
// (loss = mean_b L_b).  The residual term weights every eigenmode of A alike
// (|lambda mu - 1|^2 for network gain mu), which is what a preconditioner
// needs; the error term alone would fit only the smooth modes.
namespace {
constexpr int kRedBlocks = 512;

__global__ void k_res_terms(int64_t n, const double* r, const double* w, const double* z, const double* e,
                            double* res, double* part) {
  __shared__ double lds[4];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = r[i], di = ri - w[i];
    res[i] = di;
    s[0] += di * di;
    s[1] += ri * ri;
    if (e) {
      const double de = z[i] - e[i];
      s[2] += de * de;
      s[3] += e[i] * e[i];
    }
  }
  for (int j = 0; j < 4; ++j) {
    const double t = block_sum<256>(s[j], lds);
    if (threadIdx.x == 0) part[j * kRedBlocks + blockIdx.x] = t;
  }
}

// part[j * kRedBlocks + block] = sum a_j . b_j (j = 0..2; a null a_j gives 0)
__global__ void k_dots3(int64_t n, const double* a0, const double* b0, const double* a1, const double* b1,
                        const double* a2, const double* b2, double* part) {
  __shared__ double lds[4];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    s0 += a0[i] * b0[i];
    if (a1) s1 += a1[i] * b1[i];
    if (a2) s2 += a2[i] * b2[i];
  }
  s0 = block_sum<256>(s0, lds);
  s1 = block_sum<256>(s1, lds);
  s2 = block_sum<256>(s2, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kRedBlocks + blockIdx.x] = s1;
    part[2 * kRedBlocks + blockIdx.x] = s2;
  }
}

// dL/dy at the network output: z = s * dsq * y, so dL/dy = s * dsq * dL/dz
__global__ void k_grad_out(int64_t n, const int32_t* red2full, const double* Ares, const double* z, const double* e,
                           double cr, double ce, const double* dsq, const float* scale, int img, float inv_batch,
                           int64_t n_dofs, uint16_t* gy) {
  const double s = scale[img];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double g = -cr * Ares[i];
    if (e) g += ce * (z[i] - e[i]);
    if (dsq) g *= dsq[i];
    gy[(int64_t)img * n_dofs + red2full[i]] = f2bf((float)(s * g * inv_batch));
  }
}
}  // namespace

double unet_residual_loss_backward(Plan& P, UNet& N, const double* vals_sell, const double* const* r,
                                   const double* const* e, double w_e, int mode, hipStream_t st) {
  const int64_t n = P.n_free;
  if (P.tw.n < (size_t)(4 * n)) P.tw.alloc(4 * n);
  double *z = P.tw.p, *w = z + n, *res = z + 2 * n, *Ares = z + 3 * n;
  if (P.tpart.n < (size_t)(4 * kRedBlocks)) P.tpart.alloc(4 * kRedBlocks);
  std::vector<double> hp(4 * kRedBlocks);
  int g;
  grid1d(n, g);
  const int gr = std::min(g, kRedBlocks);
  APX_HIP(hipMemsetAsync(N.gy.p, 0, N.gy.bytes(), st));
  net_set_diag_scaling(P, vals_sell, APX_FP64, st);  // the network sees D^-1/2 A D^-1/2
  // sign of the (definite) matrix from its first diagonal entry
  double d0 = 1.0;
  {
    int32_t slot0 = 0;
    APX_HIP(hipMemcpy(&slot0, P.diag_slot.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    APX_HIP(hipMemcpy(&d0, vals_sell + slot0, sizeof(double), hipMemcpyDeviceToHost));
  }
  const double sA = d0 < 0.0 ? -1.0 : 1.0;
  std::vector<float> hscale(N.batch);
  APX_HIP(hipMemcpy(hscale.data(), N.scale.p, N.batch * sizeof(float), hipMemcpyDeviceToHost));
  double loss = 0.0;
  for (int b = 0; b < N.batch; ++b) {
    net_output_to_reduced_img(P, N, r[b], z, 1, b, st);
    spmv(P, vals_sell, z, w, APX_FP64, st);
    hipLaunchKernelGGL(k_res_terms, dim3(gr), dim3(256), 0, st, n, r[b], w, z, e ? e[b] : nullptr, res,
                       P.tpart.p);
    APX_HIP(hipMemcpyAsync(hp.data(), P.tpart.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    APX_HIP(hipStreamSynchronize(st));
    double t[4] = {0, 0, 0, 0};
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < gr; ++k) t[j] += hp[j * kRedBlocks + k];
    const double rr = std::max(t[1], 1e-300), ee = std::max(t[3], 1e-300);
    const double* eb = (e && e[b]) ? e[b] : nullptr;
    if (mode == 1 || mode == 2) {
      double d[3] = {0.0, 0.0, 0.0};
      auto dots = [&](const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                      const double* b2) {
        hipLaunchKernelGGL(k_dots3, dim3(gr), dim3(256), 0, st, n, a0, b0, a1, b1, a2, b2, P.tpart.p);
        APX_HIP(hipMemcpyAsync(hp.data(), P.tpart.p, 3 * kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
        APX_HIP(hipStreamSynchronize(st));
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < gr; ++k) d[j] += hp[j * kRedBlocks + k];
      };
      double cr;
      if (mode == 1) {
        // energy functional J(z) = sA (z.Az / 2 - z.r) / (s^2 n), gradient sA (A z - r) / (s^2 n)
        const double sb = hscale[b], norm = sb * sb * (double)n;
        dots(z, w, z, r[b], nullptr, nullptr);
        loss += sA * (0.5 * d[0] - d[1]) / norm + (eb ? w_e * t[2] / ee : 0.0);
        cr = sA / norm;
      } else {
        // relative A-norm error (z - e).Ahat(z - e) / e.Ahat e, Ahat = sA A SPD, A e = r:
        // Ahat (z - e) = -sA res, so the value is -sA (z - e).res / E and the gradient -2 sA res / E
        APX_REQUIRE(eb != nullptr, APX_ERR_ARG, "the A-norm loss needs the exact errors e");
        dots(z, res, eb, res, eb, r[b]);
        const double E = std::max(sA * d[2], 1e-300);
        loss += -sA * (d[0] - d[1]) / E + w_e * t[2] / ee;
        cr = 2.0 * sA / E;
      }
      // g = -cr res (+ the error term) through the Ares slot of k_grad_out
      hipLaunchKernelGGL(k_grad_out, dim3(g), dim3(256), 0, st, n, P.red2full.p, res, z, eb, cr, 2.0 * w_e / ee,
                         P.net_dsq.p, N.scale.p, b, 1.f / N.batch, P.n_dofs, N.gy.p);
    } else {
      loss += t[0] / rr + (eb ? w_e * t[2] / ee : 0.0);
      spmv(P, vals_sell, res, Ares, APX_FP64, st);
      // d/dz: -2 A (r - A z) / ||r||^2 + 2 w_e (z - e) / ||e||^2 ; d/dy = s d/dz
      hipLaunchKernelGGL(k_grad_out, dim3(g), dim3(256), 0, st, n, P.red2full.p, Ares, z, eb, 2.0 / rr,
                         2.0 * w_e / ee, P.net_dsq.p, N.scale.p, b, 1.f / N.batch, P.n_dofs, N.gy.p);
    }
    APX_LAUNCH_CHECK();
  }
  N.backward(st);
  APX_HIP(hipStreamSynchronize(st));
  return loss / N.batch;
}

}  // namespace apx
