// Convolution, grid-transfer and training kernels of the learned
// preconditioner (multigrid-structured U-Net on the structured node grid).
//
// Tensors are NHWC (channel fastest), batch of B images of H x W nodes,
// activations and activation gradients stored as bf16, accumulation in FP32
// on the matrix cores (v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][4(l/16)+j],
// B[4(l/16)+j][l%16] and D[4(l/16)+j][l%16]; layers with 32 or 64 input channels use
// v_mfma_f32_16x16x32_bf16, whose lanes hold 8 K elements of A and B and D as above).
//
// 3x3 convolution (stride 1, zero padding) as implicit GEMM:
//   Y[p, co] = act( sum_{t<9} sum_ci X[p + o_t, ci] W[t, co, ci] + b[co] )
// M = pixels (16-pixel tiles along a row), N = output channels (16-wide
// tiles), K = 9 taps x input channels.  Inputs with fewer than 16 channels
// (the residual layer) pack (tap, channel) into K.  The backward data pass is
// the same kernel on dY with weights flipped in the taps and transposed in the
// channels; the weight gradient is a second MFMA kernel with K = pixels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "common.hpp"

namespace apx {
namespace nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));  // operands of the 16x16x32 MFMA
typedef short short8 __attribute__((ext_vector_type(8)));
// two 4-element fragments as one 8-element operand (lo: elements 0..3)
__device__ __forceinline__ bf16x8 pair8(bf16x4 lo, bf16x4 hi) {
  return __builtin_bit_cast(bf16x8, (short8)__builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
// FP32 -> bf16 with round to nearest even (NaN stays a quiet NaN): the gfx950 conversion
// instruction (v_cvt_pk_bf16_f32), one value or a packed pair (low half = a).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2_t));
}

// Block barrier that orders LDS accesses only.  __syncthreads() also waits for every
// outstanding global load and store of the wave (vmcnt 0), which puts the previous tile's
// output stores and any prefetch in series with the barrier.  The waves of k_conv3x3's tile
// loop and of the coarse-level launch share only LDS, so their barriers need only this one.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kTileW = 64;  // output pixels per tile row (4 MFMA row tiles)
constexpr int kTileH = 4;   // tile rows = wavefronts per block
constexpr int kConvThreads = 64 * kTileH;

__host__ __device__ constexpr int round16(int c) { return (c + 15) / 16 * 16; }
// LDS pixel stride (bf16 elements) of a staged tile: padded by 4 to spread banks.
__host__ __device__ constexpr int lds_stride(int c) { return c >= 16 ? c + 4 : 16 + 4; }

struct ConvArgs {
  const uint16_t* x;      // B x H x W x CI (bf16)
  const uint16_t* xmask;  // optional: X is multiplied by (xmask > 0) elementwise (ReLU backward)
  const uint16_t* w;      // 9 x CO_pad x CI_pad (bf16), CI_pad = round16(CI) or packed K
  const float* bias;      // CO or null
  void* y;                // B x H x W x CO (bf16, or float when OUT_F32)
  const float* yscale;    // optional: per-image output scale (float y only)
  int B, H, W;
  int relu;
  const int* stop;        // optional device flag: nonzero -> the launch does nothing
  // weight-stationary kernel with IN > 0: the input is act(W_in P + b_in) of the IN-channel image
  // xin (the input layer, formed while staging; x is not read): W_in [16][16 KB_in], b_in [16]
  const uint16_t* xin;
  const uint16_t* win;
  const float* bin;
  int relu_in;
  // k_conv_d_out: the output layer applied to this convolution's output (weights as k_conv_out
  // reads them: [co][9 taps x 16 channels], bias, ReLU); y is then the output layer's FP32 image
  const uint16_t* w2;
  const float* bias2;
  int relu2;
};

// ---- forward / backward-data convolution ------------------------------------
// CI, CO: stored channel counts.  K packing: CI >= 16 -> per tap CI/16 blocks of
// 16 channels; CI < 16 -> k = t*CI + c over ceil(9 CI / 16) blocks.
//
// Memory path: the input tile (one-pixel halo) and the weights are staged in
// LDS with 16-byte global loads and 16-byte LDS stores.  bf16 outputs leave the
// accumulators directly: with the weights as the MFMA's A operand a lane holds four
// consecutive channels of one pixel, one 8-byte store (four lanes cover 32 contiguous
// bytes of a pixel), so a tile needs no output pass through LDS and no barriers
// around one.
//
// LDS strides.  A 16-byte LDS read (the fragments of the 16x16x32 MFMA) is served
// for 32 lanes at a time over 128 four-byte banks (measured on gfx950 with
// rocprofv3's SQ_LDS_BANK_CONFLICT on lane patterns).  Lane (lr, lg), lr = 0..15,
// lg = 0..3, reads row lr at byte offset 16 lg; lanes 0..31 (lg = 0, 1) are
// conflict-free exactly when the row stride is 32 mod 64 bytes (it is then an odd
// multiple of 32 bytes, so 16 lr + ... covers every 16-byte bank group once).  So
// the pixel stride of 32- and 64-channel tiles is CI + 16 elements and the weight
// rows of the 16x16x32 path are WK + 16 (WK a multiple of 32).  The CI + 8 stride
// they replace (144 bytes at 64 channels) made every fragment read 2-way.  Blocks
// with one 16-channel output tile (NB = 1) keep CI + 8 for the input tile: the wider
// tile would take a 64-channel block past 80 KB of LDS and to one block per compute
// unit, and on the small grids those blocks run at their latency floor.
// Rows stay 16-byte aligned either way.  16-channel tiles (8-byte fragment reads
// of the 16x16x16 MFMA) keep CI + 8.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int CI, int NB>
__host__ __device__ constexpr int conv_lds_stride() {
  return (CI >= 32 && NB > 1) ? CI + 16 : CI >= 16 ? CI + 8 : 8;
}

// NB: 16-wide output-channel tiles per block; blockIdx.y selects the channel
// group (gridDim.y = NT / NB).  Small grids use NB = 1 so that more blocks
// share the work and each stages only its own weight rows.  A block walks its
// tiles with the next tile's input loads in flight (registers) during the
// current tile's matrix work.  WS: wavefronts per tile row (block of 256 WS
// threads); with WS = 2 each row's NB channel tiles are split between two waves,
// so that two waves share each SIMD and one issues MFMAs while the other waits
// on its LDS fragments (the block's LDS, which allows one block per compute unit,
// is the same).  XM: the input is multiplied by (xmask > 0) (the ReLU backward of the data
// pass); a template parameter, so that each kernel has one load path (loads on two paths of a
// runtime branch made the compiler wait for all of them where the paths join, before the MFMAs
// they were meant to overlap).
template <int CI, int CO, bool OUT_F32, int NB, int WS = 1, bool XM = false>
__global__ __launch_bounds__(kConvThreads * WS) void k_conv3x3(ConvArgs a) {
  constexpr int BS = kConvThreads * WS;
  constexpr int NBW = NB / WS;  // channel tiles per wavefront
  static_assert(NB % WS == 0, "the waves of a row split its channel tiles evenly");
  constexpr int COP = round16(CO);
  constexpr int S = conv_lds_stride<CI, NB>();
  constexpr int KB = CI >= 16 ? 9 * (CI / 16) : (9 * CI + 15) / 16;  // K blocks of 16
  constexpr int WK = KB * 16;                                        // packed K length (per co)
  constexpr int SWK = CI % 32 == 0 ? WK + 16 : WK + 8;               // LDS weight row stride (above)
  constexpr int NT = COP / 16;
  static_assert(NT % NB == 0, "channel groups must tile the output channels");
  constexpr int CB = (NB == NT) ? CO : 16 * NB;  // output channels of this block
  constexpr int SX = (kTileH + 2) * (kTileW + 2) * S;
  // bf16 output: the MFMAs form D^T = W X^T (the weights as the A operand), so that a lane holds
  // output channels 4 lg .. 4 lg + 3 of pixel lr and stores them with one 8-byte write, straight
  // from the accumulators.  FP32 output: D = X W^T (a lane holds one channel of pixels
  // 4 lg .. 4 lg + 3), written through LDS rows.  The products and their order are the same.
  constexpr bool kDirect = !OUT_F32;
  static_assert(!kDirect || CO % 4 == 0, "direct stores write four channels at a time");
  // output staging (FP32 output only): pixel stride CB
  constexpr int SO = kDirect ? 0 : kTileH * kTileW * CB * 2;  // uint16 units
  constexpr int SBUF = SX > SO ? SX : SO;
  constexpr int CH8 = (CI + 7) / 8;  // 16-byte input chunks per pixel
  constexpr int NP = (kTileH + 2) * (kTileW + 2);
  constexpr int NC = NP * CH8, RC = (NC + BS - 1) / BS;
  static_assert(OUT_F32 || CB % 8 == 0, "bf16 outputs are stored in 8-channel chunks");
  __shared__ __attribute__((aligned(16))) uint16_t sbuf[SBUF];
  __shared__ __attribute__((aligned(16))) uint16_t sw[16 * NB * SWK];
  if (a.stop && *a.stop) return;  // uniform over the grid: set before the launch
  uint16_t* sx = sbuf;
  const int co0 = blockIdx.y * 16 * NB;  // first output channel of this block
  // weights -> LDS once per block: global [co][WK] -> LDS [co - co0][SWK], 16-byte chunks; the
  // weight loads and the first tile's loads are all issued before the weights are stored, so
  // that the two round trips overlap (it matters on small grids, where a launch has one tile)
  constexpr int NW = 16 * NB * (WK / 8), RW = (NW + BS - 1) / BS;
  // every load below is unconditional, at an address clamped into its array, so that all of a
  // thread's loads are in flight at once (a load under a per-lane condition makes the compiler
  // wait for it at the branch join); what lies outside is dropped or zeroed at the LDS store
  // the biases of this lane's output channels, once per block and ahead of the weights, so that
  // the wait for the weights covers them (a load used in the epilogue would otherwise be waited
  // for with every tile, behind the next tile's prefetch)
  const int wave = (threadIdx.x >> 6) % kTileH, part = (threadIdx.x >> 6) / kTileH, lane = threadIdx.x & 63;
  const int lr = lane & 15, lg = lane >> 4;
  const int n0 = part * NBW;  // this wave's first channel tile
  constexpr int NBIAS = kDirect ? 4 : 1;  // output channels per lane and channel tile
  float bias[NBW][NBIAS];
#pragma unroll
  for (int n = 0; n < NBW; ++n)
#pragma unroll
    for (int j = 0; j < NBIAS; ++j) {
      const int cl = (n0 + n) * 16 + (kDirect ? 4 * lg + j : lr);
      bias[n][j] = (a.bias && cl < CB && co0 + cl < CO) ? a.bias[co0 + cl] : 0.f;
    }
  u32x4 qw[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int i = min((int)threadIdx.x + r * BS, NW - 1);
    qw[r] = *(const u32x4*)(a.w + (int64_t)(co0 + i / (WK / 8)) * WK + (i % (WK / 8)) * 8);
  }
  const int tiles_x = (a.W + kTileW - 1) / kTileW, tiles_y = (a.H + kTileH - 1) / kTileH;
  const int64_t ntiles = (int64_t)a.B * tiles_x * tiles_y;
  // input tile loads with a one-pixel halo into registers; bit r of qin: chunk r lies inside
  // the image (otherwise it is stored as zeros)
  static_assert(RC <= 32, "one validity bit per staged chunk");
  u32x4 qs[RC];
  uint32_t qin = 0;
  auto load_tile = [&](int64_t tile, auto masked) __attribute__((always_inline)) {
    constexpr bool M = decltype(masked)::value;  // ReLU backward: X *= (xmask > 0)
    const int img = (int)(tile / ((int64_t)tiles_x * tiles_y));
    const int trem = (int)(tile % ((int64_t)tiles_x * tiles_y));
    const int ty0 = (trem / tiles_x) * kTileH, tx0 = (trem % tiles_x) * kTileW;
    const int64_t img_off = (int64_t)img * a.H * a.W * CI;
    const uint16_t* xb = a.x + img_off;
    qin = 0;
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const int i = threadIdx.x + r * BS;
      const int pix = i / CH8, ch = (i % CH8) * 8;
      const int yy = ty0 - 1 + pix / (kTileW + 2), xx = tx0 - 1 + pix % (kTileW + 2);
      qin |= (uint32_t)(i < NC && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) << r;
      const int yc = min(max(yy, 0), a.H - 1), xc = min(max(xx, 0), a.W - 1);
      const int64_t o = ((int64_t)yc * a.W + xc) * CI + ch;
      u32x4 q;
      if constexpr (CI % 8 == 0) {
        q = *(const u32x4*)(xb + o);
        if constexpr (M) {
          const u32x4 m = *(const u32x4*)(a.xmask + img_off + o);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            uint32_t keep = 0u;
            if (bf2f((uint16_t)(m[e] & 0xffffu)) > 0.f) keep |= 0xffffu;
            if (bf2f((uint16_t)(m[e] >> 16)) > 0.f) keep |= 0xffff0000u;
            q[e] &= keep;
          }
        }
      } else {
        uint16_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (ch + e < CI) {
            v[e] = xb[o + e];
            if constexpr (M)
              if (!(bf2f(a.xmask[img_off + o + e]) > 0.f)) v[e] = 0;
          }
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = (uint32_t)v[2 * e] | ((uint32_t)v[2 * e + 1] << 16);
      }
      qs[r] = q;
    }
  };
  auto load_tile_any = [&](int64_t tile) __attribute__((always_inline)) {
    load_tile(tile, std::integral_constant<bool, XM>{});
  };
  const int64_t tile0 = xcd_block(blockIdx.x, gridDim.x);  // vertically adjacent tiles share an L2
  if (tile0 < ntiles) load_tile_any(tile0);
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int i = threadIdx.x + r * BS;
    if (i < NW) *(u32x4*)(sw + (i / (WK / 8)) * SWK + (i % (WK / 8)) * 8) = qw[r];
  }
  for (int64_t tile = tile0; tile < ntiles; tile += gridDim.x) {
    const int img = (int)(tile / ((int64_t)tiles_x * tiles_y));
    const int trem = (int)(tile % ((int64_t)tiles_x * tiles_y));
    const int ty0 = (trem / tiles_x) * kTileH, tx0 = (trem % tiles_x) * kTileW;
    // every wave is done with the previous tile (also orders the weight stores); the direct
    // output stores of that tile stay in flight
    if constexpr (kDirect)
      lds_barrier();
    else
      __syncthreads();
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const int i = threadIdx.x + r * BS;
      if (i < NC) *(u32x4*)(sx + (i / CH8) * S + (i % CH8) * 8) = ((qin >> r) & 1u) ? qs[r] : u32x4{0u, 0u, 0u, 0u};
    }
    // every global access issued so far has completed (the staging needed the loads; the previous
    // tile's stores are older still): an unconditional wait here, because the compiler's wait for
    // the last chunk sits under its per-lane branch, and otherwise the first fragment reads below
    // (into registers that the previous tile's stores read from) wait for the prefetch too
    __builtin_amdgcn_s_waitcnt(0);
    if constexpr (kDirect)
      lds_barrier();
    else
      __syncthreads();
    if (tile + gridDim.x < ntiles) load_tile_any(tile + gridDim.x);  // in flight during the MFMAs
    // each wavefront: one tile row, 4 row tiles of 16 pixels, NBW column tiles
    f32x4 acc[4][NBW];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < NBW; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (CI % 32 == 0) {
      // 32-wide K blocks (tap-major, 32 channels each) on the double-K MFMA: lane (lr, lg) holds
      // K elements 8 lg .. 8 lg + 7 of both operands, 16-byte LDS reads
#pragma unroll
      for (int kb = 0; kb < 9 * CI / 32; ++kb) {
        bf16x8 bfrag[NBW];
#pragma unroll
        for (int n = 0; n < NBW; ++n) bfrag[n] = *(const bf16x8*)(sw + ((n0 + n) * 16 + lr) * SWK + kb * 32 + 8 * lg);
        const int t = kb / (CI / 32), cb = (kb % (CI / 32)) * 32;
        const int dy = t / 3, dx = t % 3;  // halo offsets 0..2
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const bf16x8 af =
              *(const bf16x8*)(sx + ((wave + dy) * (kTileW + 2) + m * 16 + lr + dx) * S + cb + 8 * lg);
#pragma unroll
          for (int n = 0; n < NBW; ++n)
            acc[m][n] = kDirect ? __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfrag[n], af, acc[m][n], 0, 0, 0)
                                : __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfrag[n], acc[m][n], 0, 0, 0);
        }
      }
    } else
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      bf16x4 bfrag[NBW];
#pragma unroll
      for (int n = 0; n < NBW; ++n) bfrag[n] = *(const bf16x4*)(sw + ((n0 + n) * 16 + lr) * SWK + kb * 16 + 4 * lg);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        bf16x4 af;
        const int px = m * 16 + lr;  // pixel column inside the tile
        if constexpr (CI >= 16) {
          const int t = kb / (CI / 16), cb = (kb % (CI / 16)) * 16;
          const int dy = t / 3, dx = t % 3;  // halo offsets 0..2
          af = *(const bf16x4*)(sx + ((wave + dy) * (kTileW + 2) + px + dx) * S + cb + 4 * lg);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int k = kb * 16 + 4 * lg + j;
            short v = 0;
            if (k < 9 * CI) {
              const int t = k / CI, c = k % CI;
              v = (short)sx[((wave + t / 3) * (kTileW + 2) + px + t % 3) * S + c];
            }
            af[j] = v;
          }
        }
#pragma unroll
        for (int n = 0; n < NBW; ++n)
          acc[m][n] = kDirect ? __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bfrag[n], af, acc[m][n], 0, 0, 0)
                              : __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfrag[n], acc[m][n], 0, 0, 0);
      }
    }
    const int yy = ty0 + wave;
    if constexpr (kDirect) {
      // bias, activation, bf16, one 8-byte store per (row tile, channel tile): lane (lr, lg) holds
      // channels 4 lg .. 4 lg + 3 of pixel m * 16 + lr
      if (yy < a.H) {
        uint16_t* yrow = (uint16_t*)a.y + (((int64_t)img * a.H + yy) * a.W + tx0) * CO + co0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int px = m * 16 + lr;
          if (tx0 + px < a.W) {
#pragma unroll
            for (int n = 0; n < NBW; ++n) {
              const int cl = (n0 + n) * 16 + 4 * lg;  // first of this lane's four channels
              if (cl < CB && co0 + cl < CO) {
                float v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                  v[j] = acc[m][n][j] + bias[n][j];
                  if (a.relu) v[j] = fmaxf(v[j], 0.f);
                }
                *(uint2*)(yrow + (int64_t)px * CO + cl) = uint2{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])};
              }
            }
          }
        }
      }
    } else {
      __syncthreads();  // every wavefront is done with the input tile: reuse it for the output rows
      // epilogue: bias, activation, scale -> LDS row of this wavefront (lane holds co = co0+n*16+lr,
      // pixels 4*lg+j)
      const float scale = a.yscale ? a.yscale[img] : 1.f;
#pragma unroll
      for (int n = 0; n < NBW; ++n) {
        const int cl = (n0 + n) * 16 + lr;  // channel inside the block's group
        if (cl < CB && co0 + cl < CO) {
#pragma unroll
          for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              float v = acc[m][n][j] + bias[n][0];
              if (a.relu) v = fmaxf(v, 0.f);
              ((float*)sbuf)[(wave * kTileW + m * 16 + 4 * lg + j) * CB + cl] = v * scale;
            }
        }
      }
      __syncthreads();
      // copy out: each wavefront's row holds 64 pixels x CB channels (contiguous in NHWC when CB = CO)
      const int npx = min(kTileW, a.W - tx0);
      if (yy < a.H) {
        const int64_t row0 = (((int64_t)img * a.H + yy) * a.W + tx0) * CO + co0;
        const float* src = (const float*)sbuf + wave * kTileW * CB;
        float* dst = (float*)a.y + row0;
        for (int i = lane + 64 * part; i < npx * CB; i += 64 * WS) dst[(i / CB) * CO + i % CB] = src[i];
      }
    }
  }
}

// ---- weight-stationary convolution for 16- and 32-channel layers ---------------
// The same operator as k_conv3x3 (bf16 output) for CI, CO in {16, 32}, arranged
// for less on-chip traffic per MFMA:
//  - the weights are the A operand (M = output channels) and stay in registers
//    for the whole launch: lane (l%16, l/16) holds W[16n + l%16][t CI + 16c + 4(l/16) + j];
//  - pixels are the N dimension: a B fragment is 16 consecutive pixels of one
//    staged row, and the D fragment (pixel l%16, channels 16n + 4(l/16) + j) is
//    stored straight to global memory, 16 pixels x 32 B contiguous per store;
//  - each wavefront owns a 16-pixel column of the TR x 64 tile and all TR output
//    rows, so one fragment of staged row yi feeds output rows yi-2 .. yi;
//  - the staged tile is planar in channel quads: plane q holds channels 4q..4q+3
//    of every staged pixel (8 B each), so the 16 lanes of one quad read 128
//    contiguous bytes, and consecutive planes start 128 B apart modulo 256 B.
// Two barriers per tile; the next tile's global loads are in flight in registers
// during the current tile's MFMAs.
template <int TR>
__host__ __device__ constexpr int ws_plane() {
  // (TR + 2) x (kTileW + 2) staged pixels, rounded up to 16 (mod 32)
  return (((TR + 2) * (kTileW + 2) + 15) / 32) * 32 + 16;
}

// IN > 0 (CI = 16): the staged input tile is the input layer's output, computed from an LDS
// tile of the IN-channel image with the operands and order of k_conv_in (bit-identical to it),
// zero outside the image; the input layer's output never goes to HBM.
template <int CI, int CO, int TR, int IN = 0>
__global__ __launch_bounds__(kConvThreads) void k_conv3x3_ws(ConvArgs a) {
  static_assert((CI == 16 || CI == 32) && (CO == 16 || CO == 32), "16- or 32-channel layers");
  static_assert(IN == 0 || CI == 16, "the input layer has 16 output channels");
  constexpr int NC = CI / 16, NO = CO / 16;  // K blocks per tap, output channel blocks
  constexpr int TWP = kTileW + 2;            // staged row length (pixels)
  constexpr int NP = (TR + 2) * TWP;         // staged pixels
  constexpr int PP = ws_plane<TR>();         // pixels per plane
  static_assert(PP >= NP, "plane holds the staged tile");
  constexpr int CH8 = CI / 8;                // 16-byte chunks per pixel
  constexpr int NL = (NP + 15) / 16 * 16 * CH8;
  constexpr int RC = (NL + kConvThreads - 1) / kConvThreads;
  constexpr int WK = 9 * CI;                 // weight row length
  __shared__ __attribute__((aligned(16))) uint16_t sx[(CI / 4) * PP * 4];
  if (a.stop && *a.stop) return;  // uniform over the grid: set before the launch
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;
  bf16x4 wa[9][NC][NO];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int n = 0; n < NO; ++n)
        wa[t][c][n] = *(const bf16x4*)(a.w + (int64_t)(16 * n + lr) * WK + t * CI + 16 * c + 4 * lq);
  float bb[NO][4];
#pragma unroll
  for (int n = 0; n < NO; ++n)
#pragma unroll
    for (int j = 0; j < 4; ++j) bb[n][j] = a.bias ? a.bias[16 * n + 4 * lq + j] : 0.f;

  const int tiles_x = (a.W + kTileW - 1) / kTileW, tiles_y = (a.H + TR - 1) / TR;
  const int tiles_img = tiles_x * tiles_y;
  const int ntiles = a.B * tiles_img;  // the launcher checks that tile indices fit in 32 bits
  // load slot i: 16-lane group g = i / 16 takes chunk g % CH8 of pixels 16 (g / CH8) + i % 16.
  // A slot's place in the staged tile does not change from tile to tile: its row and column,
  // chunk and LDS offset are computed once (row -> a large negative value for unused slots).
  // s_off: the slot's element offset from the tile's staged origin (row ty0 - 1, column tx0 - 1)
  int s_row[RC], s_col[RC], s_off[RC], s_lds[RC];
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int i = threadIdx.x + r * kConvThreads;
    const int g = i >> 4, h = g % CH8, pix = (g / CH8) * 16 + (i & 15);
    const bool used = i < NL && pix < NP;
    s_row[r] = used ? pix / TWP : -(1 << 20);
    s_col[r] = pix % TWP;
    s_off[r] = ((pix / TWP) * a.W + pix % TWP) * CI + 8 * h;
    s_lds[r] = used ? ((2 * h) * PP + pix) * 4 : -1;
  }
  u32x4 qs[RC];
  // IN > 0: the image tile (rows ty0 - 2 .. ty0 + TR + 1, columns tx0 - 2 .. tx0 + 65) in LDS,
  // element ((row * XW) + column) * IN + channel; bit r of qxin: element r is inside the image
  constexpr int INC = IN > 0 ? IN : 1, XW = kTileW + 4, XT = (TR + 4) * XW * INC;
  constexpr int RX = (XT + kConvThreads - 1) / kConvThreads, KBI = (9 * INC + 15) / 16;
  static_assert(RX <= 32, "one validity bit per staged image element");
  __shared__ uint16_t xs[IN > 0 ? RX * kConvThreads : 1];
  uint16_t qx[IN > 0 ? RX : 1];
  uint32_t qxin = 0;
  bf16x4 wi[KBI];
  float bi[4];
  if constexpr (IN > 0) {
#pragma unroll
    for (int kb = 0; kb < KBI; ++kb) wi[kb] = *(const bf16x4*)(a.win + lr * (KBI * 16) + kb * 16 + 4 * lq);
#pragma unroll
    for (int j = 0; j < 4; ++j) bi[j] = a.bin ? a.bin[4 * lq + j] : 0.f;
  }
  auto load_xtile = [&](int tile) {
    const int img = tile / tiles_img, trem = tile - img * tiles_img;
    const int tyi = trem / tiles_x;
    const int ty0 = tyi * TR, tx0 = (trem - tyi * tiles_x) * kTileW;
    const uint16_t* xb = a.xin + (int64_t)img * a.H * a.W * INC;
    qxin = 0;
#pragma unroll
    for (int r = 0; r < RX; ++r) {
      const int i = threadIdx.x + r * kConvThreads;
      const int c = i % INC, px = (i / INC) % XW, row = i / (INC * XW);
      const int yy = ty0 - 2 + row, xx = tx0 - 2 + px;
      qxin |= (uint32_t)(i < XT && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) << r;
      const int yc = min(max(yy, 0), a.H - 1), xc = min(max(xx, 0), a.W - 1);
      qx[r] = xb[(yc * a.W + xc) * INC + c];
    }
  };
  auto load_tile = [&](int tile) {
    const int img = tile / tiles_img, trem = tile - img * tiles_img;
    const int tyi = trem / tiles_x;
    const int ty0 = tyi * TR, tx0 = (trem - tyi * tiles_x) * kTileW;
    // element offset of the staged origin (may be negative; every loaded slot lands inside the image)
    const int64_t origin = ((int64_t)img * a.H * a.W + (int64_t)(ty0 - 1) * a.W + (tx0 - 1)) * CI;
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const int yy = ty0 - 1 + s_row[r], xx = tx0 - 1 + s_col[r];
      u32x4 q = u32x4{0u, 0u, 0u, 0u};
      if ((unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W) {
        const int64_t o = origin + s_off[r];
        q = *(const u32x4*)(a.x + o);
        if (a.xmask) {
          const u32x4 m = *(const u32x4*)(a.xmask + o);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            uint32_t keep = 0u;
            if (bf2f((uint16_t)(m[e] & 0xffffu)) > 0.f) keep |= 0xffffu;
            if (bf2f((uint16_t)(m[e] >> 16)) > 0.f) keep |= 0xffff0000u;
            q[e] &= keep;
          }
        }
      }
      qs[r] = q;
    }
  };
  const int tile0 = (int)xcd_block(blockIdx.x, gridDim.x);  // vertically adjacent tiles share an L2
  if (tile0 < ntiles) {
    if constexpr (IN > 0)
      load_xtile(tile0);
    else
      load_tile(tile0);
  }
  for (int tile = tile0; tile < ntiles; tile += gridDim.x) {
    const int img = tile / tiles_img, trem = tile - img * tiles_img;
    const int tyi = trem / tiles_x;
    const int ty0 = tyi * TR, tx0 = (trem - tyi * tiles_x) * kTileW;
    __syncthreads();  // every wavefront is done reading the previous tile
    if constexpr (IN > 0) {
#pragma unroll
      for (int r = 0; r < RX; ++r) xs[threadIdx.x + r * kConvThreads] = ((qxin >> r) & 1u) ? qx[r] : (uint16_t)0;
      __syncthreads();
      // the staged tile = the input layer at the staged pixels, 16 pixels per MFMA (D: channels
      // 4 lq .. 4 lq + 3 of pixel lr = plane lq), zero outside the image
      constexpr int NG = (NP + 15) / 16;  // 16-pixel groups; wavefront w takes w, w + 4, ...
#pragma unroll
      for (int k = 0; k < (NG + 3) / 4; ++k) {
        const int gi = wave + 4 * k;
        const int pix = 16 * gi + lr, pc = min(pix, NP - 1);  // past the tile: recomputes its last pixel
        const int sr = pc / TWP, sc = pc - sr * TWP;
        f32x4 d = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBI; ++kb) {
          bf16x4 b;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int k = 16 * kb + 4 * lq + j, kc = min(k, 9 * INC - 1);
            const int t = kc / INC, c = kc % INC;
            const uint16_t v = xs[((sr + t / 3) * XW + sc + t % 3) * INC + c];
            b[j] = k < 9 * INC ? (short)v : (short)0;
          }
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wi[kb], b, d, 0, 0, 0);
        }
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v[j] = d[j] + bi[j];
          if (a.relu_in) v[j] = fmaxf(v[j], 0.f);
        }
        const int yy = ty0 - 1 + sr, xx = tx0 - 1 + sc;
        const bool inside = (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W;
        if (pix < NP)
          *(uint2*)(sx + (lq * PP + pix) * 4) = inside ? uint2{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])} : uint2{0u, 0u};
      }
    } else {
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        if (s_lds[r] >= 0) {
          *(uint2*)(sx + s_lds[r]) = uint2{qs[r][0], qs[r][1]};
          *(uint2*)(sx + s_lds[r] + PP * 4) = uint2{qs[r][2], qs[r][3]};
        }
      }
    }
    __syncthreads();
    if (tile + (int)gridDim.x < ntiles) {  // in flight during the MFMAs
      if constexpr (IN > 0)
        load_xtile(tile + (int)gridDim.x);
      else
        load_tile(tile + (int)gridDim.x);
    }
    f32x4 acc[TR][NO];
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int n = 0; n < NO; ++n) acc[r][n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int yi = 0; yi < TR + 2; ++yi)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        if constexpr (CI == 32) {
          // one 16x16x32 MFMA per tap: lane group lq holds channels 4 lq .. 4 lq + 3 (plane lq)
          // and 16 + 4 lq .. 16 + 4 lq + 3 (plane lq + 4) of both operands
          const int o = (yi * TWP + 16 * wave + lr + dx) * 4;
          const bf16x8 f = pair8(*(const bf16x4*)(sx + lq * PP * 4 + o), *(const bf16x4*)(sx + (lq + 4) * PP * 4 + o));
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) {
            const int r = yi - dy;  // output row fed by staged row yi through tap row dy
            if (r >= 0 && r < TR) {
#pragma unroll
              for (int n = 0; n < NO; ++n)
                acc[r][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pair8(wa[dy * 3 + dx][0][n], wa[dy * 3 + dx][1][n]),
                                                                    f, acc[r][n], 0, 0, 0);
            }
          }
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const bf16x4 f = *(const bf16x4*)(sx + ((4 * c + lq) * PP + yi * TWP + 16 * wave + lr + dx) * 4);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
              const int r = yi - dy;  // output row fed by staged row yi through tap row dy
              if (r >= 0 && r < TR) {
#pragma unroll
                for (int n = 0; n < NO; ++n)
                  acc[r][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[dy * 3 + dx][c][n], f, acc[r][n], 0, 0, 0);
              }
            }
          }
        }
      }
    const int xx = tx0 + 16 * wave + lr;
    if (xx < a.W) {
      uint16_t* dst0 = (uint16_t*)a.y + (((int64_t)img * a.H + ty0) * a.W + xx) * CO + 4 * lq;
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int yy = ty0 + r;
        if (yy < a.H) {
          uint16_t* dst = dst0 + (int64_t)r * a.W * CO;
#pragma unroll
          for (int n = 0; n < NO; ++n) {
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              v[j] = acc[r][n][j] + bb[n][j];
              if (a.relu) v[j] = fmaxf(v[j], 0.f);
            }
            *(uint2*)(dst + 16 * n) = uint2{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])};
          }
        }
      }
    }
  }
}

// ---- output layer ------------------------------------------------------------------
// 16 -> CO channels (CO = 1 or 2), FP32 output scaled per image:
//   y[p, co] = s_img act(b[co] + sum_t Z[p + o_t, 9 co + t]),  Z[q, 9 co + t] = sum_c X[q, c] W[co][16 t + c].
// The per-tap partials Z of every node of a TR x 64 tile plus its one-node halo come from one
// MFMA per 16 nodes and 16 partial columns (A = 16 nodes x 16 channels, loaded straight from
// HBM: 8 B per lane, 512 B contiguous per 16 nodes; B = the 16 x 9 CO weights); each output
// then sums its nine neighbours' partials from LDS.  The implicit GEMM of k_conv3x3 would use
// CO of its 16 output columns: this form does 9x fewer MFMAs and stages no input in LDS.
template <int CO>
__global__ __launch_bounds__(256) void k_conv_out(ConvArgs a) {
  constexpr int NN = 9 * CO;                // partial columns
  constexpr int NT = (NN + 15) / 16;        // MFMA N-tiles
  constexpr int ZS = NN + 1;                // LDS floats per staged node (odd: no bank conflicts)
  constexpr int TR = 4, SR = TR + 2;        // output rows, staged rows
  constexpr int SG = 5, SC = 16 * SG;       // 16-node groups per staged row, staged columns (>= 66)
  constexpr int NG = SR * SG, GPW = (NG + 3) / 4;  // groups, groups per wavefront
  __shared__ float sz[SR * SC * ZS];
  if (a.stop && *a.stop) return;  // uniform over the grid: set before the launch
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  const int tiles_x = (a.W + 63) / 64, tiles_img = tiles_x * ((a.H + TR - 1) / TR);
  const int img = blockIdx.x / tiles_img, trem = blockIdx.x - img * tiles_img;
  const int ty0 = (trem / tiles_x) * TR, tx0 = (trem % tiles_x) * 64;
  // B fragments: lane (lr, lq) holds W[co][16 t + 4 lq + j] of column n = 16 nt + lr = 9 co + t
  bf16x4 wb[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int n = min(16 * nt + lr, NN - 1);
    wb[nt] = *(const bf16x4*)(a.w + (n / 9) * 144 + (n % 9) * 16 + 4 * lq);
    if (16 * nt + lr >= NN) wb[nt] = bf16x4{0, 0, 0, 0};
  }
  // A fragments of this wavefront's groups g = wave + 4 k: node lr of the group, channels 4 lq..
  // (unconditional loads, clamped into the image; nodes outside contribute zero partials)
  bf16x4 xa[GPW];
  uint32_t in = 0;
#pragma unroll
  for (int k = 0; k < GPW; ++k) {
    const int g = min(wave + 4 * k, NG - 1);
    const int yy = ty0 - 1 + g / SG, xx = tx0 - 1 + 16 * (g % SG) + lr;
    in |= (uint32_t)(wave + 4 * k < NG && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) << k;
    const int yc = min(max(yy, 0), a.H - 1), xc = min(max(xx, 0), a.W - 1);
    xa[k] = *(const bf16x4*)(a.x + (((int64_t)img * a.H + yc) * a.W + xc) * 16 + 4 * lq);
  }
  float bias[CO];
#pragma unroll
  for (int co = 0; co < CO; ++co) bias[co] = a.bias ? a.bias[co] : 0.f;
  const float scale = a.yscale ? a.yscale[img] : 1.f;
#pragma unroll
  for (int k = 0; k < GPW; ++k) {
    const int g = wave + 4 * k;
    if (g < NG) {
      const bf16x4 af = ((in >> k) & 1u) ? xa[k] : bf16x4{0, 0, 0, 0};
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, wb[nt], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        const int n = 16 * nt + lr;  // D: node 4 lq + j of the group, column n
        if (n < NN) {
#pragma unroll
          for (int j = 0; j < 4; ++j) sz[((g / SG) * SC + 16 * (g % SG) + 4 * lq + j) * ZS + n] = d[j];
        }
      }
    }
  }
  __syncthreads();
  const int r = threadIdx.x >> 6, x = threadIdx.x & 63;  // output node (ty0 + r, tx0 + x)
  const int yy = ty0 + r, xx = tx0 + x;
  if (yy < a.H && xx < a.W) {
    float* dst = (float*)a.y + (((int64_t)img * a.H + yy) * a.W + xx) * CO;
#pragma unroll
    for (int co = 0; co < CO; ++co) {
      float v = 0.f;
#pragma unroll
      for (int t = 0; t < 9; ++t) v += sz[((r + t / 3) * SC + x + t % 3) * ZS + 9 * co + t];
      v += bias[co];
      if (a.relu) v = fmaxf(v, 0.f);
      dst[co] = v * scale;
    }
  }
}

// ---- level-0 conv_d fused with the output layer (inference) ---------------------------
// y = conv_out(conv_d(s)) without storing conv_d's output: each block computes conv_d (16 -> 16,
// the arithmetic of k_conv3x3_ws<16, 16, TR + 2, 0>) on the DR = TR + 2 rows x 64 columns around
// its TR x 62 output tile, keeps that bf16 tile in LDS (zero outside the image, which is what
// k_conv_out reads there), then forms k_conv_out's per-tap partials from it by the same MFMA and
// sums them in the same order.  The result is bit-identical to the two launches; the 134 MB
// intermediate image is neither written nor read back.  Tiles overlap by two columns.
template <int CO>
__global__ __launch_bounds__(kConvThreads, 4) void k_conv_d_out(ConvArgs a) {
  constexpr int TR = 4, DR = TR + 2;          // output rows, conv_d rows
  constexpr int OW = kTileW - 2;              // output columns per tile
  constexpr int TWP = kTileW + 2;             // staged row length (pixels)
  constexpr int NP = (DR + 2) * TWP;          // staged pixels
  constexpr int PP = ws_plane<DR>();          // pixels per plane
  static_assert(PP >= NP, "plane holds the staged tile");
  constexpr int NL = (NP + 15) / 16 * 16 * 2;  // 16-byte load slots (two per pixel)
  constexpr int RC = (NL + kConvThreads - 1) / kConvThreads;
  constexpr int NN = 9 * CO, ZS = NN + 1;     // output-layer partial columns; LDS floats per node
  constexpr int NT = (NN + 15) / 16;          // their MFMA N-tiles
  constexpr int SXE = 4 * PP * 4;             // staged tile (bf16 elements)
  constexpr int SZE = DR * kTileW * ZS * 2;   // partials (bf16-element units: 2 per float)
  // the partials reuse the staged tile's space (it is dead once conv_d's MFMAs are done)
  __shared__ __attribute__((aligned(16))) uint16_t sx[SXE > SZE ? SXE : SZE];
  __shared__ __attribute__((aligned(16))) uint16_t sd[DR * kTileW * 16];  // conv_d output tile
  float* sz = (float*)sx;
  if (a.stop && *a.stop) return;  // uniform over the grid: set before the launch
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;
  bf16x4 wa[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) wa[t] = *(const bf16x4*)(a.w + (int64_t)lr * 144 + t * 16 + 4 * lq);
  float bb[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bb[j] = a.bias ? a.bias[4 * lq + j] : 0.f;
  // output layer: B fragments as in k_conv_out (column n = 9 co + t), bias
  bf16x4 wb[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int n = min(16 * nt + lr, NN - 1);
    wb[nt] = *(const bf16x4*)(a.w2 + (n / 9) * 144 + (n % 9) * 16 + 4 * lq);
    if (16 * nt + lr >= NN) wb[nt] = bf16x4{0, 0, 0, 0};
  }
  float b2[CO];
#pragma unroll
  for (int co = 0; co < CO; ++co) b2[co] = a.bias2 ? a.bias2[co] : 0.f;

  const int tiles_x = (a.W + OW - 1) / OW, tiles_y = (a.H + TR - 1) / TR;
  const int tiles_img = tiles_x * tiles_y;
  const int ntiles = a.B * tiles_img;  // the launcher checks that tile indices fit in 32 bits
  // load slots as in k_conv3x3_ws (two 16-byte chunks per pixel, planar in channel quads): slot
  // i = threadIdx.x + r kConvThreads takes chunk h of staged pixel pix; computed where used, not
  // kept in registers (that keeps the kernel within 128 registers: 4 waves per SIMD)
  auto slot = [&](int r, int& h, int& pix, bool& used) __attribute__((always_inline)) {
    const int i = (int)threadIdx.x + r * kConvThreads;
    h = (i >> 4) & 1;
    pix = ((i >> 5) << 4) + (i & 15);
    used = i < NL && pix < NP;
  };
  u32x4 qs[RC];
  // staged origin: row ty0 - 2, column tx0 - 2 (conv_d rows ty0 - 1 .., columns tx0 - 1 ..)
  auto load_tile = [&](int tile) {
    const int img = tile / tiles_img, trem = tile - img * tiles_img;
    const int tyi = trem / tiles_x;
    const int ty0 = tyi * TR, tx0 = (trem - tyi * tiles_x) * OW;
    const int64_t origin = ((int64_t)img * a.H * a.W + (int64_t)(ty0 - 2) * a.W + (tx0 - 2)) * 16;
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      int h, pix;
      bool used;
      slot(r, h, pix, used);
      const int row = pix / TWP, col = pix - row * TWP;
      const int yy = ty0 - 2 + row, xx = tx0 - 2 + col;
      u32x4 q = u32x4{0u, 0u, 0u, 0u};
      if (used && (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W)
        q = *(const u32x4*)(a.x + origin + ((int64_t)row * a.W + col) * 16 + 8 * h);
      qs[r] = q;
    }
  };
  const int tile0 = (int)xcd_block(blockIdx.x, gridDim.x);  // vertically adjacent tiles share an L2
  if (tile0 < ntiles) load_tile(tile0);
  for (int tile = tile0; tile < ntiles; tile += gridDim.x) {
    const int img = tile / tiles_img, trem = tile - img * tiles_img;
    const int tyi = trem / tiles_x;
    const int ty0 = tyi * TR, tx0 = (trem - tyi * tiles_x) * OW;
    __syncthreads();  // every wavefront is done with the previous tile's partials (sx)
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      int h, pix;
      bool used;
      slot(r, h, pix, used);
      if (used) {
        const int o = ((2 * h) * PP + pix) * 4;
        *(uint2*)(sx + o) = uint2{qs[r][0], qs[r][1]};
        *(uint2*)(sx + o + PP * 4) = uint2{qs[r][2], qs[r][3]};
      }
    }
    __syncthreads();
    if (tile + (int)gridDim.x < ntiles) load_tile(tile + (int)gridDim.x);  // in flight during the MFMAs
    // conv_d: wavefront w owns conv_d columns 16 w .. 16 w + 15 of all DR rows
    f32x4 acc[DR];
#pragma unroll
    for (int r = 0; r < DR; ++r) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int yi = 0; yi < DR + 2; ++yi)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const bf16x4 f = *(const bf16x4*)(sx + (lq * PP + yi * TWP + 16 * wave + lr + dx) * 4);
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          const int r = yi - dy;  // conv_d row fed by staged row yi through tap row dy
          if (r >= 0 && r < DR) acc[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[dy * 3 + dx], f, acc[r], 0, 0, 0);
        }
      }
    // conv_d's output tile -> LDS (lane: channels 4 lq .. 4 lq + 3 of column 16 w + lr), zero outside
    {
      const int xx = tx0 - 1 + 16 * wave + lr;
#pragma unroll
      for (int r = 0; r < DR; ++r) {
        const int yy = ty0 - 1 + r;
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v[j] = acc[r][j] + bb[j];
          if (a.relu) v[j] = fmaxf(v[j], 0.f);
        }
        const bool inside = (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W;
        *(uint2*)(sd + ((r * kTileW + 16 * wave + lr) * 16 + 4 * lq)) =
            inside ? uint2{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])} : uint2{0u, 0u};
      }
    }
    __syncthreads();  // the conv_d tile is complete and sx is free
    // output-layer partials of the DR x 64 nodes: group g = (row g / 4, columns 16 (g % 4) ..)
#pragma unroll
    for (int k = 0; k < DR; ++k) {
      const int g = wave + 4 * k, gr = g / 4, gc = 16 * (g % 4);
      const bf16x4 af = *(const bf16x4*)(sd + ((gr * kTileW + gc + lr) * 16 + 4 * lq));
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, wb[nt], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        const int n = 16 * nt + lr;  // D: node 4 lq + j of the group, column n
        if (n < NN) {
#pragma unroll
          for (int j = 0; j < 4; ++j) sz[(gr * kTileW + gc + 4 * lq + j) * ZS + n] = d[j];
        }
      }
    }
    __syncthreads();
    // y: thread (r, x) sums the nine partials around conv_d node (r + 1, x + 1) in tap order
    const int r = threadIdx.x >> 6, x = threadIdx.x & 63;
    const int yy = ty0 + r, xx = tx0 + x;
    if (x < OW && yy < a.H && xx < a.W) {
      const float sc = a.yscale ? a.yscale[img] : 1.f;
      float* dst = (float*)a.y + (((int64_t)img * a.H + yy) * a.W + xx) * CO;
#pragma unroll
      for (int co = 0; co < CO; ++co) {
        float v = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) v += sz[((r + t / 3) * kTileW + x + t % 3) * ZS + 9 * co + t];
        v += b2[co];
        if (a.relu2) v = fmaxf(v, 0.f);
        dst[co] = v * sc;
      }
    }
  }
}

// ---- input layer -------------------------------------------------------------------
// CI = 1 or 2 -> 16 channels (bf16 output), weight-stationary like k_conv3x3_ws: A = W (16
// output channels x 16 K, packed k = t CI + c, K blocks of 16 as in k_conv3x3), B = 16 pixels'
// 3 x 3 patches read from the staged input tile in LDS; lane (lr, lq) ends with channels
// 4 lq .. 4 lq + 3 of pixel lr, stored straight to HBM (8 B per lane, 512 B contiguous per 16
// pixels; the layer writes 16x what it reads).  One 4 x 64 tile per block, wavefront w owns
// pixel columns 16 w .. 16 w + 15.
template <int CI>
__global__ __launch_bounds__(kConvThreads) void k_conv_in(ConvArgs a) {
  constexpr int KB = (9 * CI + 15) / 16;
  constexpr int TR = kTileH, SR = TR + 2, SW = kTileW + 2;  // output rows, staged rows and columns
  constexpr int NS = SR * SW * CI, RS = (NS + kConvThreads - 1) / kConvThreads;
  __shared__ uint16_t sx[RS * kConvThreads];
  if (a.stop && *a.stop) return;  // uniform over the grid: set before the launch
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  const int tiles_x = (a.W + kTileW - 1) / kTileW, tiles_img = tiles_x * ((a.H + TR - 1) / TR);
  const int tile = (int)xcd_block(blockIdx.x, gridDim.x);  // vertically adjacent tiles share an L2
  const int img = tile / tiles_img, trem = tile - img * tiles_img;
  const int ty0 = (trem / tiles_x) * TR, tx0 = (trem % tiles_x) * kTileW;
  const uint16_t* xb = a.x + (int64_t)img * a.H * a.W * CI;
  // staged element i = (row, column, channel), row-major; unconditional clamped loads, zero outside
  uint16_t q[RS];
  uint32_t qin = 0;
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int i = threadIdx.x + r * kConvThreads;
    const int c = i % CI, px = (i / CI) % SW, row = i / (CI * SW);
    const int yy = ty0 - 1 + row, xx = tx0 - 1 + px;
    qin |= (uint32_t)(i < NS && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) << r;
    const int yc = min(max(yy, 0), a.H - 1), xc = min(max(xx, 0), a.W - 1);
    q[r] = xb[((int64_t)yc * a.W + xc) * CI + c];
  }
  bf16x4 wa[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) wa[kb] = *(const bf16x4*)(a.w + lr * (KB * 16) + kb * 16 + 4 * lq);
  float bias[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bias[j] = a.bias ? a.bias[4 * lq + j] : 0.f;
#pragma unroll
  for (int r = 0; r < RS; ++r) sx[threadIdx.x + r * kConvThreads] = ((qin >> r) & 1u) ? q[r] : (uint16_t)0;
  __syncthreads();
  const int xo = 16 * wave + lr;  // this lane's pixel column in the tile
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      bf16x4 b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 16 * kb + 4 * lq + j, kc = min(k, 9 * CI - 1);
        const int t = kc / CI, c = kc % CI;
        const uint16_t v = sx[((r + t / 3) * SW + xo + t % 3) * CI + c];
        b[j] = k < 9 * CI ? (short)v : (short)0;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[kb], b, acc, 0, 0, 0);
    }
    const int yy = ty0 + r, xx = tx0 + xo;
    if (yy < a.H && xx < a.W) {
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = acc[j] + bias[j];
        if (a.relu) v[j] = fmaxf(v[j], 0.f);
      }
      *(uint2*)((uint16_t*)a.y + (((int64_t)img * a.H + yy) * a.W + xx) * 16 + 4 * lq) =
          uint2{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])};
    }
  }
}

// ---- weight gradient ---------------------------------------------------------
// dW[t, co, ci] = sum_p dY[p, co] X[p + o_t, ci]   (dY masked by ymask > 0 when YM)
// db[co] = sum_p dY[p, co].  Each block accumulates over its tiles and writes
// one partial [KB*16 x COP] + [COP] in FP32; k_reduce_wgrad sums the blocks
// in a fixed order.  MFMA: A = dY^T (16 co x 16 px), B = X_t (16 px x 16 k).
// A tile's loads (the X tile with its one-pixel halo and the dY tile; 8 channels, or one
// pixel of at most 2 channels, per load) are all in flight at once: every address is clamped
// into the image and what lies outside is stored as zeros, so that no load sits under a
// per-lane condition (the compiler would wait for each at its branch join).
struct WgradArgs {
  const uint16_t* x;      // B x H x W x CI
  const uint16_t* dy;     // B x H x W x CO
  const uint16_t* ymask;  // ReLU output of the layer (YM only)
  float* partial;         // gridDim x (COP*WK + COP)
  int B, H, W;
};

// zero the 16-bit halves of q whose mask half is not > 0 (ReLU backward)
__device__ __forceinline__ uint32_t relu_keep(uint32_t q, uint32_t m) {
  uint32_t keep = 0u;
  if (bf2f((uint16_t)(m & 0xffffu)) > 0.f) keep |= 0xffffu;
  if (bf2f((uint16_t)(m >> 16)) > 0.f) keep |= 0xffff0000u;
  return q & keep;
}

template <int CI, int CO, bool YM>
__global__ __launch_bounds__(kConvThreads) void k_conv3x3_wgrad(WgradArgs a) {
  constexpr int S = lds_stride(CI);
  constexpr int COP = round16(CO);
  constexpr int KB = CI >= 16 ? 9 * (CI / 16) : (9 * CI + 15) / 16;
  constexpr int WK = KB * 16;
  constexpr int NT = COP / 16;
  constexpr int SD = COP + 4;  // LDS stride of the dY tile
  // fragments (co tile n, k block kb) are distributed over the 4 wavefronts
  constexpr int NF = NT * KB;
  constexpr int FPW = (NF + kTileH - 1) / kTileH;
  static_assert(CI % 8 == 0 || CI <= 2, "X loads: 8 channels, or one pixel of at most 2");
  static_assert(CO % 16 == 0 || CO <= 2, "dY loads: 8 channels, or one pixel of at most 2");
  constexpr int XL = CI % 8 == 0 ? CI / 8 : 1, DL = CO % 8 == 0 ? CO / 8 : 1;  // loads per pixel
  constexpr int NPX = (kTileH + 2) * (kTileW + 2), NPD = kTileH * kTileW;       // staged pixels
  constexpr int NXL = NPX * XL, NDL = NPD * DL;
  constexpr int RX = (NXL + kConvThreads - 1) / kConvThreads, RD = (NDL + kConvThreads - 1) / kConvThreads;
  static_assert(RX <= 32 && RD <= 32, "one validity bit per load");
  // LDS stores of 8 channels are two 8-byte writes: pixel strides are multiples of 4 elements
  static_assert(S % 4 == 0 && SD % 4 == 0, "8-byte aligned LDS chunks");
  __shared__ __attribute__((aligned(16))) uint16_t sx[NPX * S];
  __shared__ __attribute__((aligned(16))) uint16_t sd[NPD * SD];
  __shared__ float sbias[kTileH][COP];
  const int tiles_x = (a.W + kTileW - 1) / kTileW, tiles_y = (a.H + kTileH - 1) / kTileH;
  const int64_t ntiles = (int64_t)a.B * tiles_x * tiles_y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lg = lane >> 4;
  // dY channels CO .. COP - 1 (CO <= 2) are read as zeros: the staging writes channels < CO only
  if constexpr (CO < COP)
    for (int i = threadIdx.x; i < NPD * (COP - CO); i += kConvThreads)
      sd[(i / (COP - CO)) * SD + CO + i % (COP - CO)] = 0;
  f32x4 acc[FPW];
#pragma unroll
  for (int f = 0; f < FPW; ++f) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.f;  // bias partial: thread -> (row = wave, co = lane) when lane < COP
  u32x4 qx[RX], qd[RD];
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int img = (int)(tile / ((int64_t)tiles_x * tiles_y));
    const int trem = (int)(tile % ((int64_t)tiles_x * tiles_y));
    const int ty0 = (trem / tiles_x) * kTileH, tx0 = (trem % tiles_x) * kTileW;
    const int64_t pix0 = (int64_t)img * a.H * a.W;
    // loads into registers; bit r of inx / ind: load r lies inside the tile's arrays and image
    uint32_t inx = 0, ind = 0;
#pragma unroll
    for (int r = 0; r < RX; ++r) {
      const int i0 = (int)threadIdx.x + r * kConvThreads, i = min(i0, NXL - 1);
      const int pix = i / XL, h = i % XL;
      const int yy = ty0 - 1 + pix / (kTileW + 2), xx = tx0 - 1 + pix % (kTileW + 2);
      inx |= (uint32_t)(i0 < NXL && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) << r;
      const int yc = min(max(yy, 0), a.H - 1), xc = min(max(xx, 0), a.W - 1);
      const uint16_t* src = a.x + (pix0 + (int64_t)yc * a.W + xc) * CI + h * 8;
      if constexpr (CI % 8 == 0)
        qx[r] = *(const u32x4*)src;
      else if constexpr (CI == 2)
        qx[r] = u32x4{*(const uint32_t*)src, 0u, 0u, 0u};
      else
        qx[r] = u32x4{(uint32_t)*src, 0u, 0u, 0u};
    }
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int i0 = (int)threadIdx.x + r * kConvThreads, i = min(i0, NDL - 1);
      const int pix = i / DL, h = i % DL;
      const int yy = ty0 + pix / kTileW, xx = tx0 + pix % kTileW;
      ind |= (uint32_t)(i0 < NDL && yy < a.H && xx < a.W) << r;
      const int yc = min(yy, a.H - 1), xc = min(xx, a.W - 1);
      const int64_t o = (pix0 + (int64_t)yc * a.W + xc) * CO + h * 8;
      u32x4 q;
      if constexpr (CO % 8 == 0) {
        q = *(const u32x4*)(a.dy + o);
        if constexpr (YM) {
          const u32x4 m = *(const u32x4*)(a.ymask + o);
#pragma unroll
          for (int e = 0; e < 4; ++e) q[e] = relu_keep(q[e], m[e]);
        }
      } else if constexpr (CO == 2) {
        q = u32x4{*(const uint32_t*)(a.dy + o), 0u, 0u, 0u};
        if constexpr (YM) q[0] = relu_keep(q[0], *(const uint32_t*)(a.ymask + o));
      } else {
        q = u32x4{(uint32_t)a.dy[o], 0u, 0u, 0u};
        if constexpr (YM) q[0] = relu_keep(q[0], (uint32_t)a.ymask[o]);
      }
      qd[r] = q;
    }
    lds_barrier();  // every wave is done with the previous tile's LDS
#pragma unroll
    for (int r = 0; r < RX; ++r) {
      const int i = (int)threadIdx.x + r * kConvThreads;
      if (i < NXL) {
        const u32x4 q = ((inx >> r) & 1u) ? qx[r] : u32x4{0u, 0u, 0u, 0u};
        uint16_t* d = sx + (i / XL) * S + (i % XL) * 8;
        if constexpr (CI % 8 == 0) {
          *(uint2*)d = uint2{q[0], q[1]};
          *(uint2*)(d + 4) = uint2{q[2], q[3]};
        } else if constexpr (CI == 2) {
          *(uint32_t*)d = q[0];
        } else {
          *d = (uint16_t)q[0];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int i = (int)threadIdx.x + r * kConvThreads;
      if (i < NDL) {
        const u32x4 q = ((ind >> r) & 1u) ? qd[r] : u32x4{0u, 0u, 0u, 0u};
        uint16_t* d = sd + (i / DL) * SD + (i % DL) * 8;
        if constexpr (CO % 8 == 0) {
          *(uint2*)d = uint2{q[0], q[1]};
          *(uint2*)(d + 4) = uint2{q[2], q[3]};
        } else if constexpr (CO == 2) {
          *(uint32_t*)d = q[0];
        } else {
          *d = (uint16_t)q[0];
        }
      }
    }
    lds_barrier();
    if (lane < COP) {
      float s = 0.f;
      for (int px = 0; px < kTileW; ++px) s += bf2f(sd[(wave * kTileW + px) * SD + lane]);
      bacc += s;
    }
    // K = pixels of the tile: 4 rows x 64 px = 16 blocks of 16 px.  A wave's fragments cover at
    // most two co tiles, n0 and n0 + 1 (FPW <= KB), so their A fragments are read once per pixel
    // block and selected per fragment (the compiler cannot see that fid / KB is constant over f)
    constexpr int kAN = 2;
    static_assert(FPW <= KB, "a wave's fragments span at most two co tiles");
    const int n0 = min(wave * FPW / KB, NT - 1);
    for (int pb = 0; pb < kTileH * (kTileW / 16); ++pb) {
      const int row = pb / (kTileW / 16), px0 = (pb % (kTileW / 16)) * 16;
      // A fragment: A[co = n*16 + lr][px = px0 + 4 lg + j] = dY[px][co]
      bf16x4 an[kAN];
#pragma unroll
      for (int q = 0; q < kAN; ++q) {
        const int nq = min(n0 + q, NT - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) an[q][j] = (short)sd[(row * kTileW + px0 + 4 * lg + j) * SD + nq * 16 + lr];
      }
#pragma unroll
      for (int f = 0; f < FPW; ++f) {
        const int fid = wave * FPW + f;
        if (fid < NF) {
        const int n = fid / KB, kb = fid % KB;
        bf16x4 bf;
        const bf16x4 af = n == n0 ? an[0] : an[kAN - 1];
        // B fragment: B[px = px0 + 4 lg + j][k = kb*16 + lr] = X[px + o_t][ci]
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = kb * 16 + lr;
          const int px = px0 + 4 * lg + j;
          short v = 0;
          if constexpr (CI >= 16) {
            const int t = k / CI, c = k % CI;
            v = (short)sx[((row + t / 3) * (kTileW + 2) + px + t % 3) * S + c];
          } else if (k < 9 * CI) {
            const int t = k / CI, c = k % CI;
            v = (short)sx[((row + t / 3) * (kTileW + 2) + px + t % 3) * S + c];
          }
          bf[j] = v;
        }
        acc[f] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf, acc[f], 0, 0, 0);
        }
      }
    }
  }
  // write this block's partial: D[co][k] (lane: k = kb*16 + lr, co = n*16 + 4 lg + j)
  float* out = a.partial + (int64_t)blockIdx.x * (COP * WK + COP);
#pragma unroll
  for (int f = 0; f < FPW; ++f) {
    const int fid = wave * FPW + f;
    if (fid < NF) {
      const int n = fid / KB, kb = fid % KB;
#pragma unroll
      for (int j = 0; j < 4; ++j) out[(n * 16 + 4 * lg + j) * WK + kb * 16 + lr] = acc[f][j];
    }
  }
  if (lane < COP) sbias[wave][lane] = bacc;
  __syncthreads();
  if (threadIdx.x < COP) {
    float s = 0.f;
    for (int w = 0; w < kTileH; ++w) s += sbias[w][threadIdx.x];
    out[COP * WK + threadIdx.x] = s;
  }
}

}  // namespace nn
}  // namespace apx
