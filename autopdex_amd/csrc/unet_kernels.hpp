// Convolution, grid-transfer and training kernels of the learned
// preconditioner (multigrid-structured U-Net on the structured node grid).
//
// Tensors are NHWC (channel fastest), batch of B images of H x W nodes,
// activations and activation gradients stored as bf16, accumulation in FP32
// on the matrix cores (v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][4(l/16)+j],
// B[4(l/16)+j][l%16] and D[4(l/16)+j][l%16]).
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

namespace apx {
namespace nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even; NaN kept quiet
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
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
};

// ---- forward / backward-data convolution ------------------------------------
// CI, CO: stored channel counts.  K packing: CI >= 16 -> per tap CI/16 blocks of
// 16 channels; CI < 16 -> k = t*CI + c over ceil(9 CI / 16) blocks.
template <int CI, int CO, bool OUT_F32>
__global__ __launch_bounds__(kConvThreads) void k_conv3x3(ConvArgs a) {
  constexpr int CIP = CI >= 16 ? CI : 16;
  constexpr int COP = round16(CO);
  constexpr int S = lds_stride(CI);  // LDS pixel stride
  constexpr int KB = CI >= 16 ? 9 * (CI / 16) : (9 * CI + 15) / 16;  // K blocks of 16
  constexpr int WK = KB * 16;                                        // packed K length (per co)
  constexpr int NT = COP / 16;
  __shared__ uint16_t sx[(kTileH + 2) * (kTileW + 2) * S];
  __shared__ uint16_t sw[COP * (WK + 4)];
  // weights -> LDS once per block (layout [co][k], k packed as in the kernel's K order)
  for (int i = threadIdx.x; i < COP * WK; i += kConvThreads) {
    const int co = i / WK, k = i % WK;
    sw[co * (WK + 4) + k] = a.w[i];
  }
  const int tiles_x = (a.W + kTileW - 1) / kTileW, tiles_y = (a.H + kTileH - 1) / kTileH;
  const int64_t ntiles = (int64_t)a.B * tiles_x * tiles_y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lg = lane >> 4;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int img = (int)(tile / ((int64_t)tiles_x * tiles_y));
    const int trem = (int)(tile % ((int64_t)tiles_x * tiles_y));
    const int ty0 = (trem / tiles_x) * kTileH, tx0 = (trem % tiles_x) * kTileW;
    __syncthreads();  // previous tile's LDS reads done (also orders the weight load)
    // stage the input tile with a one-pixel halo; zero outside the image
    {
      constexpr int CH8 = (CI + 7) / 8;  // 16-byte chunks per pixel
      constexpr int NP = (kTileH + 2) * (kTileW + 2);
      const uint16_t* xb = a.x + (int64_t)img * a.H * a.W * CI;
      for (int i = threadIdx.x; i < NP * CH8; i += kConvThreads) {
        const int pix = i / CH8, ch = (i % CH8) * 8;
        const int yy = ty0 - 1 + pix / (kTileW + 2), xx = tx0 - 1 + pix % (kTileW + 2);
        uint16_t v[8];
        const bool in = yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0;
        if (in) {
          const int64_t o = ((int64_t)yy * a.W + xx) * CI + ch;
          if constexpr (CI % 8 == 0) {
            const uint4 q = *(const uint4*)(xb + o);
            v[0] = q.x & 0xffff; v[1] = q.x >> 16; v[2] = q.y & 0xffff; v[3] = q.y >> 16;
            v[4] = q.z & 0xffff; v[5] = q.z >> 16; v[6] = q.w & 0xffff; v[7] = q.w >> 16;
            if (a.xmask) {
              const uint4 m = *(const uint4*)(a.xmask + (int64_t)img * a.H * a.W * CI + o);
              const uint32_t mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const uint16_t me = (uint16_t)(mm[e / 2] >> (16 * (e & 1)));
                if (!(bf2f(me) > 0.f)) v[e] = 0;
              }
            }
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (ch + e < CI) {
                v[e] = xb[o + e];
                if (a.xmask && !(bf2f(a.xmask[(int64_t)img * a.H * a.W * CI + o + e]) > 0.f)) v[e] = 0;
              }
          }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (ch + e < S) sx[pix * S + ch + e] = v[e];
      }
    }
    __syncthreads();
    // each wavefront: one tile row, 4 row tiles of 16 pixels, NT column tiles
    f32x4 acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      bf16x4 bfrag[NT];
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const uint16_t* wp = sw + (n * 16 + lr) * (WK + 4) + kb * 16 + 4 * lg;
        bfrag[n] = bf16x4{(short)wp[0], (short)wp[1], (short)wp[2], (short)wp[3]};
      }
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        bf16x4 af;
        const int px = m * 16 + lr;  // pixel column inside the tile
        if constexpr (CI >= 16) {
          const int t = kb / (CI / 16), cb = (kb % (CI / 16)) * 16;
          const int dy = t / 3, dx = t % 3;  // halo offsets 0..2
          const uint16_t* xp = sx + ((wave + dy) * (kTileW + 2) + px + dx) * S + cb + 4 * lg;
          af = bf16x4{(short)xp[0], (short)xp[1], (short)xp[2], (short)xp[3]};
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
        for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfrag[n], acc[m][n], 0, 0, 0);
      }
    }
    // epilogue: bias, activation, store (lane holds column co = n*16+lr, pixels 4*lg+j)
    const int yy = ty0 + wave;
    if (yy < a.H) {
      const float scale = (OUT_F32 && a.yscale) ? a.yscale[img] : 1.f;
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int co = n * 16 + lr;
        if (co >= CO) continue;
        const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int xx = tx0 + m * 16 + 4 * lg + j;
            if (xx >= a.W) continue;
            float v = acc[m][n][j] + bb;
            if (a.relu) v = fmaxf(v, 0.f);
            const int64_t o = (((int64_t)img * a.H + yy) * a.W + xx) * CO + co;
            if constexpr (OUT_F32)
              ((float*)a.y)[o] = v * scale;
            else
              ((uint16_t*)a.y)[o] = f2bf(v);
          }
      }
    }
  }
}

// ---- weight gradient ---------------------------------------------------------
// dW[t, co, ci] = sum_p dY[p, co] X[p + o_t, ci]   (dY masked by ymask > 0 when given)
// db[co] = sum_p dY[p, co].  Each block accumulates over its tiles and writes
// one partial [KB*16 x COP] + [COP] in FP32; k_reduce_partials sums the blocks
// in index order.  MFMA: A = dY^T (16 co x 16 px), B = X_t (16 px x 16 k).
struct WgradArgs {
  const uint16_t* x;      // B x H x W x CI
  const uint16_t* dy;     // B x H x W x CO
  const uint16_t* ymask;  // optional (ReLU output of the layer)
  float* partial;         // gridDim x (COP*WK + COP)
  int B, H, W;
};

template <int CI, int CO>
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
  __shared__ uint16_t sx[(kTileH + 2) * (kTileW + 2) * S];
  __shared__ uint16_t sd[kTileH * kTileW * SD];
  __shared__ float sbias[kTileH][COP];
  const int tiles_x = (a.W + kTileW - 1) / kTileW, tiles_y = (a.H + kTileH - 1) / kTileH;
  const int64_t ntiles = (int64_t)a.B * tiles_x * tiles_y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lg = lane >> 4;
  f32x4 acc[FPW];
#pragma unroll
  for (int f = 0; f < FPW; ++f) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.f;  // bias partial: thread -> (row = wave, co = lane) when lane < COP
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int img = (int)(tile / ((int64_t)tiles_x * tiles_y));
    const int trem = (int)(tile % ((int64_t)tiles_x * tiles_y));
    const int ty0 = (trem / tiles_x) * kTileH, tx0 = (trem % tiles_x) * kTileW;
    __syncthreads();
    {
      constexpr int NP = (kTileH + 2) * (kTileW + 2);
      const uint16_t* xb = a.x + (int64_t)img * a.H * a.W * CI;
      for (int i = threadIdx.x; i < NP * S; i += kConvThreads) {
        const int pix = i / S, c = i % S;
        const int yy = ty0 - 1 + pix / (kTileW + 2), xx = tx0 - 1 + pix % (kTileW + 2);
        uint16_t v = 0;
        if (c < CI && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = xb[((int64_t)yy * a.W + xx) * CI + c];
        sx[i] = v;
      }
      const uint16_t* db = a.dy + (int64_t)img * a.H * a.W * CO;
      const uint16_t* mb = a.ymask ? a.ymask + (int64_t)img * a.H * a.W * CO : nullptr;
      for (int i = threadIdx.x; i < kTileH * kTileW * SD; i += kConvThreads) {
        const int pix = i / SD, c = i % SD;
        const int yy = ty0 + pix / kTileW, xx = tx0 + pix % kTileW;
        uint16_t v = 0;
        if (c < CO && yy < a.H && xx < a.W) {
          const int64_t o = ((int64_t)yy * a.W + xx) * CO + c;
          v = db[o];
          if (mb && !(bf2f(mb[o]) > 0.f)) v = 0;
        }
        sd[i] = v;
      }
    }
    __syncthreads();
    if (lane < COP) {
      float s = 0.f;
      for (int px = 0; px < kTileW; ++px) s += bf2f(sd[(wave * kTileW + px) * SD + lane]);
      bacc += s;
    }
    // K = pixels of the tile: 4 rows x 64 px = 16 blocks of 16 px
    for (int pb = 0; pb < kTileH * (kTileW / 16); ++pb) {
      const int row = pb / (kTileW / 16), px0 = (pb % (kTileW / 16)) * 16;
      // A fragment: A[co = n*16 + lr][px = px0 + 4 lg + j] = dY[px][co]
#pragma unroll
      for (int f = 0; f < FPW; ++f) {
        const int fid = wave * FPW + f;
        if (fid < NF) {
        const int n = fid / KB, kb = fid % KB;
        bf16x4 af, bf;
#pragma unroll
        for (int j = 0; j < 4; ++j) af[j] = (short)sd[(row * kTileW + px0 + 4 * lg + j) * SD + n * 16 + lr];
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
