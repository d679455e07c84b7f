// Owner-tile assembly kernels for structured Q1 grids (the connectivity of
// mesher.structured_mesh, autopdex/mesher.py:88-113).
//
// A workgroup of 512 threads owns a 7 x 63 block of nodes.  It stages the
// 9 x 65 nodes around it (coordinates, DOFs, nodal fields, Dirichlet mask) in
// LDS; each thread then evaluates one of the 8 x 64 elements that touch the
// block -- all its Gauss points, accumulated in registers through the same
// element math as the colour kernels -- and leaves the element's nodal
// residuals (or its element matrix) in LDS.  Every thread of the owner phase
// sums the four element contributions of one node in a fixed order and
// writes them once.  No colours, no atomics, no slot map: every assembly is
// bitwise reproducible and every output is written exactly once.  Elements on
// the tile border are evaluated by both neighbouring tiles (512 element
// evaluations per 441 owned nodes).
//
// Tangent rows (one DOF per node) go straight to the SELL layout: the columns
// of a reduced row are the free nodes of the 3 x 3 neighbourhood in node order
// (the reduced index is monotone in the node index), so a row's k-th stored
// entry is the k-th free in-grid neighbour.  The kernel also zeroes the
// chunk padding of its rows, so the values need no separate fill.
//
// Replaces the per-element jacrev / jacfwd(jacrev) + scatter of
// user_potential_assemble_residual / _tangent / _tangent_diagonal
// (autopdex/assembler.py:1064-1149) for such grids.
#include "assembly.hpp"
#include "elements.hpp"

namespace apx {

namespace {
constexpr int kTI = 7, kTJ = 63;             // owned nodes per tile
constexpr int kEJ = kTJ + 1;                 // element columns per tile (64)
constexpr int kNE = (kTI + 1) * kEJ;         // elements per tile (512 = threads)
constexpr int kNI = kTI + 2, kNJ = kTJ + 2;  // staged nodes (9 x 65)
constexpr int kNN = kNI * kNJ;
constexpr int kTBS = kNE;

enum { TILE_RES = 0, TILE_TAN = 1, TILE_DIAG = 2 };

struct TileArgs {
  const double* coords;  // SoA, 2 x n_nodes
  int64_t n_nodes;
  int64_t nx, ny;        // elements per direction
  const double* u;       // n_dofs
  ModelParams mp;
  const double* load;    // optional: u-independent residual part (n_dofs), added to r
  const uint8_t* dmask;  // n_dofs
  const int32_t* full2red;
  const int32_t* sell_ptr;
  double* r;             // residual or diagonal (n_dofs)
  void* vals;            // SELL values (tangent)
  int overwrite;         // 1: r = sum, 0: r += sum
  int want_source;       // residual: evaluate the u-independent part in the kernel
  int64_t tiles, tiles_j;  // tiles in the grid, tiles per tile row
};

template <class M, class T, int MODE>
__global__ __launch_bounds__(kTBS) void k_tile(TileArgs A) {
  constexpr int D = M::D, NPE = 4;
  static_assert(M::NPE == 4 && M::DIM == 2, "owner tiles are for Q1 quads");
  static_assert(MODE == TILE_RES || D == 1, "tangent / diagonal tiles: one DOF per node");
  __shared__ double sx[kNN][2];
  __shared__ double su[kNN][D];
  __shared__ double sn[M::NODAL ? kNN : 1][2];
  __shared__ uint8_t sm[MODE == TILE_TAN ? kNN : 1];
  constexpr int NR = (MODE == TILE_RES) ? kNE * NPE * D : kNE * NPE * NPE;
  __shared__ T er[NR];  // residual rows er[el][a][c]; element matrices er[el][a][b]

  // XCD-aware tile order (1D grid rounded up to a multiple of 8): consecutive logical tiles,
  // row-major over the tile grid, share an L2
  const int64_t lin = xcd_block(blockIdx.x, gridDim.x);
  if (lin >= A.tiles) return;  // whole workgroup: no barrier is left waiting
  const int64_t i0 = (lin / A.tiles_j) * kTI, j0 = (lin % A.tiles_j) * kTJ;
  const int64_t nx = A.nx, ny = A.ny;
  const int tid = threadIdx.x;
  // stage the 9 x 65 nodes around the tile (rows of 65 consecutive nodes)
  for (int idx = tid; idx < kNN; idx += kTBS) {
    const int li = idx / kNJ, lj = idx % kNJ;
    const int64_t I = i0 - 1 + li, J = j0 - 1 + lj;
    const bool in = I >= 0 && I <= nx && J >= 0 && J <= ny;
    const int64_t node = in ? I * (ny + 1) + J : 0;
    sx[idx][0] = in ? A.coords[node] : 0.0;
    sx[idx][1] = in ? A.coords[A.n_nodes + node] : 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) su[idx][c] = in ? A.u[node * D + c] : 0.0;
    if constexpr (M::NODAL) {
      sn[idx][0] = in ? A.mp.nodal[0][node] : 0.0;
      sn[idx][1] = in ? A.mp.nodal[1][node] : 0.0;
    }
    if constexpr (MODE == TILE_TAN) sm[idx] = in ? (A.dmask[node] ? 1 : 0) : 2;  // 2: outside the grid
  }
  __syncthreads();
  // element phase: thread -> element (i0 - 1 + lei, j0 - 1 + lej); local node a of it sits at
  // staged node (lei + kDi[a], lej + kDj[a])
  constexpr int kDi[4] = {0, 0, 1, 1}, kDj[4] = {0, 1, 1, 0};
  {
    const int el = tid;
    const int lei = el / kEJ, lej = el % kEJ;
    const int64_t ei = i0 - 1 + lei, ej = j0 - 1 + lej;
    if (ei >= 0 && ei < nx && ej >= 0 && ej < ny) {
      double xl[NPE][2], ul[NPE][D], nl[NPE][2];
#pragma unroll
      for (int a = 0; a < NPE; ++a) {
        const int s = (lei + kDi[a]) * kNJ + lej + kDj[a];
        xl[a][0] = sx[s][0];
        xl[a][1] = sx[s][1];
#pragma unroll
        for (int c = 0; c < D; ++c) ul[a][c] = su[s][c];
        if constexpr (M::NODAL) {
          nl[a][0] = sn[s][0];
          nl[a][1] = sn[s][1];
        } else {
          nl[a][0] = nl[a][1] = 0.0;
        }
      }
      if constexpr (MODE == TILE_RES) {
        double out[NPE][D];
#pragma unroll
        for (int a = 0; a < NPE; ++a)
#pragma unroll
          for (int c = 0; c < D; ++c) out[a][c] = 0.0;
#pragma unroll
        for (int q = 0; q < M::NQ; ++q) {
          typename M::template QP<double> Q;
          M::template qp<double>(q, xl, ul, nl, A.mp, A.want_source != 0, Q);
          M::res_gp(q, Q, A.mp, out);
        }
#pragma unroll
        for (int a = 0; a < NPE; ++a)
#pragma unroll
          for (int c = 0; c < D; ++c) er[(el * NPE + a) * D + c] = out[a][c];
      } else {
        T K[NPE][NPE];
#pragma unroll
        for (int a = 0; a < NPE; ++a)
#pragma unroll
          for (int b = 0; b < NPE; ++b) K[a][b] = 0;
#pragma unroll
        for (int q = 0; q < M::NQ; ++q) {
          typename M::template QP<T> Q;
          M::template qp<T>(q, xl, ul, nl, A.mp, false, Q);
          M::template tan_gp<T>(Q, A.mp, K);
        }
#pragma unroll
        for (int a = 0; a < NPE; ++a)
#pragma unroll
          for (int b = 0; b < NPE; ++b) er[(el * NPE + a) * NPE + b] = K[a][b];
      }
    }
  }
  __syncthreads();
  // owner phase: thread -> node (i0 + li, j0 + lj)
  if (tid >= kTI * kTJ) return;
  const int li = tid / kTJ, lj = tid % kTJ;
  const int64_t I = i0 + li, J = j0 + lj;
  if (I > nx || J > ny) return;
  const int64_t node = I * (ny + 1) + J;
  // the four elements around the node, in element order, with the node's local index:
  // (I-1, J-1): 2, (I-1, J): 3, (I, J-1): 1, (I, J): 0
  const bool ok[4] = {I >= 1 && J >= 1, I >= 1 && J < ny, I < nx && J >= 1, I < nx && J < ny};
  const int elx[4] = {li * kEJ + lj, li * kEJ + lj + 1, (li + 1) * kEJ + lj, (li + 1) * kEJ + lj + 1};
  constexpr int kLoc[4] = {2, 3, 1, 0};
  if constexpr (MODE == TILE_RES) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (ok[k]) s += (double)er[(elx[k] * NPE + kLoc[k]) * D + c];
      if (A.load) s += A.load[node * D + c];
      double* ro = A.r + node * D + c;
      *ro = A.overwrite ? s : *ro + s;
    }
  } else {
    // v[dI + 1][dJ + 1]: coupling of the node with neighbour (I + dI, J + dJ)
    T v[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) v[a][b] = 0;
    constexpr int kEi[4] = {-1, -1, 0, 0}, kEjo[4] = {-1, 0, -1, 0};  // element offsets from the node
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!ok[k]) continue;
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        const int dI = kEi[k] + kDi[b], dJ = kEjo[k] + kDj[b];
        v[dI + 1][dJ + 1] += er[(elx[k] * NPE + kLoc[k]) * NPE + b];
      }
    }
    if constexpr (MODE == TILE_DIAG) {
      double* ro = A.r + node;
      *ro = A.overwrite ? (double)v[1][1] : *ro + (double)v[1][1];
    } else {
      if (sm[(li + 1) * kNJ + lj + 1]) return;  // Dirichlet node: no reduced row
      const int32_t row = A.full2red[node];
      const int32_t base = A.sell_ptr[row >> 6], width = (A.sell_ptr[(row >> 6) + 1] - base) >> 6;
      T* vals = (T*)A.vals + base + (row & 63);
      int k = 0;
#pragma unroll
      for (int dI = -1; dI <= 1; ++dI)
#pragma unroll
        for (int dJ = -1; dJ <= 1; ++dJ) {
          if (sm[(li + 1 + dI) * kNJ + lj + 1 + dJ] == 0) {  // free node inside the grid
            vals[(int64_t)k * 64] = v[dI + 1][dJ + 1];
            ++k;
          }
        }
      for (; k < width; ++k) vals[(int64_t)k * 64] = 0;  // chunk padding
    }
  }
}

template <class M, class T, int MODE>
void run_tile(Plan& P, ElemSet& S, const double* u, double* r, void* vals, const double* load, bool overwrite,
              hipStream_t st) {
  TileArgs A;
  A.coords = P.coords.p;
  A.n_nodes = P.n_nodes;
  A.nx = S.gnx;
  A.ny = S.gny;
  A.u = u;
  A.mp = S.params;
  A.load = load;
  A.dmask = P.dmask.p;
  A.full2red = P.full2red.p;
  A.sell_ptr = P.sell_ptr.p;
  A.r = r;
  A.vals = vals;
  A.overwrite = overwrite ? 1 : 0;
  A.want_source = load ? 0 : 1;
  if constexpr (M::NODAL) {
    APX_REQUIRE(S.params.nodal[0] && S.params.nodal[1], APX_ERR_STATE,
                "variable-coefficient Poisson set needs nodal fields g and f bound");
  }
  A.tiles_j = (S.gny + 1 + kTJ - 1) / kTJ;
  A.tiles = A.tiles_j * ((S.gnx + 1 + kTI - 1) / kTI);
  const int64_t g = round_xcd(A.tiles);
  APX_REQUIRE(g < ((int64_t)1 << 31), APX_ERR_UNSUPPORTED, "structured grid too large for the owner-tile launch");
  hipLaunchKernelGGL((k_tile<M, T, MODE>), dim3((unsigned)g), dim3(kTBS), 0, st, A);
  APX_LAUNCH_CHECK();
}
}  // namespace

// The u-independent part of a precomputed-source model's residual (its value
// at u = 0), assembled once per parameter set by the colour kernels.
static const double* source_load(Plan& P, ElemSet& S, hipStream_t st) {
  if (S.model != APX_MODEL_POISSON_SOURCE) return nullptr;
  if (!S.load_valid) {
    if (S.load.n < (size_t)P.n_dofs) S.load.alloc(P.n_dofs);
    DevBuf<double> zero;
    zero.alloc(P.n_dofs);
    fill<double>(zero.p, P.n_dofs, 0.0, st);
    fill<double>(S.load.p, P.n_dofs, 0.0, st);
    assemble_set_colour_residual(P, S, zero.p, S.load.p, st);
    APX_HIP(hipStreamSynchronize(st));  // `zero` is freed on return
    S.load_valid = true;
  }
  return S.load.p;
}

bool tile_supported(const Plan& P, const ElemSet& S, int mode) {
  if (!S.structured) return false;
  const bool scalar = S.model == APX_MODEL_POISSON_SOURCE || S.model == APX_MODEL_POISSON_VARK;
  if (mode == 0) return scalar || S.model == APX_MODEL_NEO_HOOKE;
  return scalar && P.d == 1;
}

void tile_residual(Plan& P, ElemSet& S, const double* u, double* r, bool overwrite, hipStream_t st) {
  switch (S.model) {
    case APX_MODEL_POISSON_SOURCE:
      return run_tile<PoissonSource<4, 2>, double, TILE_RES>(P, S, u, r, nullptr, source_load(P, S, st), overwrite,
                                                             st);
    case APX_MODEL_POISSON_VARK:
      return run_tile<PoissonVarK<4, 2>, double, TILE_RES>(P, S, u, r, nullptr, nullptr, overwrite, st);
    case APX_MODEL_NEO_HOOKE:
      return run_tile<NeoHooke<4, 2>, double, TILE_RES>(P, S, u, r, nullptr, nullptr, overwrite, st);
    default:
      throw Error(APX_ERR_INTERNAL, "owner-tile residual: unsupported model");
  }
}

void tile_tangent(Plan& P, ElemSet& S, const double* u, void* vals, int precision, hipStream_t st) {
  const bool f64 = precision == APX_FP64;
  switch (S.model) {
    case APX_MODEL_POISSON_SOURCE:
      if (f64) return run_tile<PoissonSource<4, 2>, double, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
      return run_tile<PoissonSource<4, 2>, float, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
    case APX_MODEL_POISSON_VARK:
      if (f64) return run_tile<PoissonVarK<4, 2>, double, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
      return run_tile<PoissonVarK<4, 2>, float, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
    default:
      throw Error(APX_ERR_INTERNAL, "owner-tile tangent: unsupported model");
  }
}

void tile_diagonal(Plan& P, ElemSet& S, const double* u, double* diag, bool overwrite, hipStream_t st) {
  switch (S.model) {
    case APX_MODEL_POISSON_SOURCE:
      return run_tile<PoissonSource<4, 2>, double, TILE_DIAG>(P, S, u, diag, nullptr, nullptr, overwrite, st);
    case APX_MODEL_POISSON_VARK:
      return run_tile<PoissonVarK<4, 2>, double, TILE_DIAG>(P, S, u, diag, nullptr, nullptr, overwrite, st);
    default:
      throw Error(APX_ERR_INTERNAL, "owner-tile diagonal: unsupported model");
  }
}

}  // namespace apx
