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

// ---- parallelogram fast path for the Poisson models ---------------------------------
// A Q1 element whose nodes form a parallelogram (x0 - x1 + x2 - x3 = 0) has a constant
// Jacobian, so every Gauss point shares det J and B = J^-1 J^-T, and the 2 x 2-point
// stiffness is K_ab = det J sum_q k_q sum_jl dN_q[a][j] B_jl dN_q[b][l] with the reference
// gradients dN_q of Shape<4, 2> (constants).  The mesher's grids of rectangles and
// parallelograms take this path: one 2 x 2 inverse per element instead of four, and no
// per-point gradient.  Elements that deviate by more than 1e-13 of their size use the
// general per-point evaluation; within that bound J differs between the points by at most
// 1e-13 relative, below the FP64 agreement the element tests ask for.
__host__ __device__ constexpr double q1_gx(int i) { return i ? kGaussB : kGaussA; }
// dN_a/dxi_i at Gauss point q, the arithmetic of Shape<4, 2>::eval
__host__ __device__ constexpr double q1_dN(int q, int a, int i) {
  const double x0 = 0.5 * q1_gx(q & 1), x2 = 0.5 * q1_gx(q >> 1);
  const double x1 = 0.5 - x0, x3 = 0.5 - x2, x4 = x0 + 0.5, x5 = x2 + 0.5;
  switch (a * 2 + i) {
    case 0: return -0.5 * x3;
    case 1: return -0.5 * x1;
    case 2: return 0.5 * x3;
    case 3: return -0.5 * x4;
    case 4: return 0.5 * x5;
    case 5: return 0.5 * x4;
    case 6: return -0.5 * x5;
    default: return 0.5 * x1;
  }
}
__host__ __device__ constexpr double q1_N(int q, int a) {
  const double x0 = 0.5 * q1_gx(q & 1), x2 = 0.5 * q1_gx(q >> 1);
  const double x1 = 0.5 - x0, x3 = 0.5 - x2, x4 = x0 + 0.5, x5 = x2 + 0.5;
  return a == 0 ? x1 * x3 : a == 1 ? x3 * x4 : a == 2 ? x4 * x5 : x1 * x5;
}
// sum over the Gauss points (weights 1) of dN_q[a][j] dN_q[b][l] + (j != l: the transposed term)
__host__ __device__ constexpr double q1_S(int a, int b, int j, int l) {
  double s = 0.0;
  for (int q = 0; q < 4; ++q) {
    s += q1_dN(q, a, j) * q1_dN(q, b, l);
    if (j != l) s += q1_dN(q, a, l) * q1_dN(q, b, j);
  }
  return s;
}

// det J and B = J^-1 J^-T (B01 = B10) of a parallelogram element; false otherwise
__device__ inline bool q1_parallelogram(const double (*x)[2], double& det, double& B00, double& B01, double& B11) {
  const double h0 = x[0][0] - x[1][0] + x[2][0] - x[3][0], h1 = x[0][1] - x[1][1] + x[2][1] - x[3][1];
  // J_ij = sum_a x_ai dN_aj at the centre, dN(0) = 0.25 (-1,-1), (1,-1), (1,1), (-1,1)
  const double J00 = 0.25 * (-x[0][0] + x[1][0] + x[2][0] - x[3][0]);
  const double J01 = 0.25 * (-x[0][0] - x[1][0] + x[2][0] + x[3][0]);
  const double J10 = 0.25 * (-x[0][1] + x[1][1] + x[2][1] - x[3][1]);
  const double J11 = 0.25 * (-x[0][1] - x[1][1] + x[2][1] + x[3][1]);
  const double size = fabs(J00) + fabs(J01) + fabs(J10) + fabs(J11);
  if (!(fabs(h0) + fabs(h1) <= 1e-13 * size)) return false;
  det = J00 * J11 - J01 * J10;
  const double id = 1.0 / det;
  const double i00 = J11 * id, i01 = -J01 * id, i10 = -J10 * id, i11 = J00 * id;
  B00 = i00 * i00 + i01 * i01;
  B01 = i00 * i10 + i01 * i11;
  B11 = i10 * i10 + i11 * i11;
  return true;
}

// K_ab (a <= b) of a parallelogram with a constant coefficient, in T, from the summed tables
template <class T>
__device__ __forceinline__ T q1_affine_entry(int a, int b, T d, T b00, T b01, T b11) {
  return d * (b00 * (T)q1_S(a, b, 0, 0) + b01 * (T)q1_S(a, b, 0, 1) + b11 * (T)q1_S(a, b, 1, 1));
}

// Element stiffness of a parallelogram in T: constant coefficient (KQ = false) from the
// summed tables, or k_q per Gauss point
template <class T, bool KQ>
__device__ inline void q1_affine_stiffness(double det, double B00, double B01, double B11, const double* kq,
                                           T K[4][4]) {
  const T d = (T)det, b00 = (T)B00, b01 = (T)B01, b11 = (T)B11;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b) {
      T v;
      if constexpr (!KQ) {
        v = q1_affine_entry<T>(a, b, d, b00, b01, b11);
      } else {
        v = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const T c00 = (T)(q1_dN(q, a, 0) * q1_dN(q, b, 0)), c11 = (T)(q1_dN(q, a, 1) * q1_dN(q, b, 1));
          const T c01 = (T)(q1_dN(q, a, 0) * q1_dN(q, b, 1) + q1_dN(q, a, 1) * q1_dN(q, b, 0));
          v += (T)kq[q] * (b00 * c00 + b01 * c01 + b11 * c11);
        }
        v *= d;
      }
      K[a][b] = v;
      K[b][a] = v;
    }
}

template <class M>
struct AffineModel {  // which models take the parallelogram path, and how
  static constexpr bool ok = false, nodal = false;
};
template <>
struct AffineModel<PoissonSource<4, 2>> {
  static constexpr bool ok = true, nodal = false;
};
template <>
struct AffineModel<PoissonVarK<4, 2>> {
  static constexpr bool ok = true, nodal = true;
};

// Node, DOF and tile indices are 32-bit: tile_supported admits grids whose DOF count fits.
struct TileArgs {
  const double* cx;      // node coordinates x (SoA)
  const double* cy;      // node coordinates y
  int nx, ny;            // elements per direction
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
  int tiles, tiles_j;    // tiles in the grid, tiles per tile row
};

// index of K_ab (a <= b, or swapped) in the upper triangle of a symmetric 4 x 4 matrix
__host__ __device__ constexpr int sym_entry(int a, int b) {
  return a <= b ? a * 4 - a * (a - 1) / 2 + (b - a) : sym_entry(b, a);
}

// an element matrix into the LDS planes: the upper triangle (tangent) or the diagonal
template <int MODE, class T>
__device__ inline void store_matrix(T* er, int el, const T (&K)[4][4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if constexpr (MODE == TILE_DIAG) {
      er[a * kNE + el] = K[a][a];
    } else {
#pragma unroll
      for (int b = a; b < 4; ++b) er[sym_entry(a, b) * kNE + el] = K[a][b];
    }
  }
}

// AFF: every element of the grid is a parallelogram (checked once per plan by
// k_parallelogram_check with the same predicate), so only the parallelogram path is compiled;
// the general per-point path holds registers that cost the kernel occupancy.
// TO: type of the stored tangent values (FP32 for the FP64 tangent rounded once, TO = float
// with T = double: the same bits as rounding the FP64 values afterwards)
template <class M, class T, int MODE, bool AFF, class TO = T>
__global__ __launch_bounds__(kTBS) void k_tile(TileArgs A) {
  constexpr int D = M::D, NPE = 4;
  static_assert(M::NPE == 4 && M::DIM == 2, "owner tiles are for Q1 quads");
  static_assert(MODE == TILE_RES || D == 1, "tangent / diagonal tiles: one DOF per node");
  static_assert(!AFF || AffineModel<M>::ok, "the parallelogram-only kernel is for the Poisson models");
  __shared__ double sx[kNN][2];
  // the tile tangent models (Poisson, constant or nodal k) are linear in u: only the
  // residual stages it
  __shared__ double su[MODE == TILE_RES ? kNN : 1][D];
  __shared__ double sn[M::NODAL ? kNN : 1][2];
  __shared__ uint8_t sm[MODE == TILE_TAN ? kNN : 1];
  // Element matrices of the tile models are symmetric, bit for bit (K_ab and K_ba are the same
  // products summed in the same order), so the tangent keeps a <= b only and the diagonal
  // keeps K_aa only.
  // GEOM: a constant-coefficient parallelogram tangent keeps only (det J, B00, B01, B11) per
  // element, and each owner forms the entries of its row from them (the same expression, so the
  // same bits): 4 planes instead of 10
  constexpr bool GEOM = AFF && MODE == TILE_TAN && !AffineModel<M>::nodal;
  constexpr int NE_ = MODE == TILE_RES ? NPE * D : GEOM ? 4 : MODE == TILE_TAN ? NPE * (NPE + 1) / 2 : NPE;
  // one plane of kNE words per element entry: entry e of element el sits at er[e * kNE + el]
  // (e = a * D + c for residual rows, sym_entry(a, b) for matrices, a for diagonals), so the
  // lanes of a wave touch consecutive words in both phases and the LDS banks do not conflict
  __shared__ T er[kNE * NE_];

  // XCD-aware tile order (1D grid rounded up to a multiple of 8): consecutive logical tiles,
  // row-major over the tile grid, share an L2
  const int lin = (int)xcd_block(blockIdx.x, gridDim.x);
  if (lin >= A.tiles) return;  // whole workgroup: no barrier is left waiting
  const int i0 = (lin / A.tiles_j) * kTI, j0 = (lin % A.tiles_j) * kTJ;
  const int nx = A.nx, ny = A.ny;
  const int tid = threadIdx.x;
  // owner-phase inputs first, so that their latency overlaps the staging loads and the
  // element phase: thread tid < 441 owns node (i0 + tid / 63, j0 + tid % 63)
  const int oli = tid / kTJ, olj = tid % kTJ;
  const int oI = i0 + oli, oJ = j0 + olj;
  const bool owner = tid < kTI * kTJ && oI <= nx && oJ <= ny;
  const int onode = owner ? oI * (ny + 1) + oJ : 0;
  double pl[D], pr[D];  // the node's load and previous r (residual), previous value (diagonal)
#pragma unroll
  for (int c = 0; c < D; ++c) pl[c] = pr[c] = 0.0;
  int32_t orow = -1;  // reduced row of the node, -1 for a Dirichlet node
  if (owner) {
    if constexpr (MODE == TILE_TAN) {
      orow = A.full2red[onode];
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        if (MODE == TILE_RES && A.load) pl[c] = A.load[onode * D + c];
        if (!A.overwrite) pr[c] = A.r[onode * D + c];
      }
    }
  }
  // stage the 9 x 65 nodes around the tile (rows of 65 consecutive nodes): all loads of a
  // thread are issued before its first LDS store
  constexpr int kSR = (kNN + kTBS - 1) / kTBS;
  double gx[kSR][2], gu[kSR][D], gn[kSR][2];
  uint8_t gm[kSR];
#pragma unroll
  for (int q = 0; q < kSR; ++q) {
    const int idx = tid + q * kTBS;
    const int li = idx / kNJ, lj = idx % kNJ;
    const int I = i0 - 1 + li, J = j0 - 1 + lj;
    const bool in = idx < kNN && I >= 0 && I <= nx && J >= 0 && J <= ny;
    const int node = in ? I * (ny + 1) + J : 0;
    gx[q][0] = in ? A.cx[node] : 0.0;
    gx[q][1] = in ? A.cy[node] : 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) gu[q][c] = (MODE == TILE_RES && in) ? A.u[node * D + c] : 0.0;
    gn[q][0] = (M::NODAL && in) ? A.mp.nodal[0][node] : 0.0;
    gn[q][1] = (M::NODAL && in) ? A.mp.nodal[1][node] : 0.0;
    gm[q] = (MODE == TILE_TAN && in) ? (A.dmask[node] ? 1 : 0) : 2;  // 2: outside the grid
  }
#pragma unroll
  for (int q = 0; q < kSR; ++q) {
    const int idx = tid + q * kTBS;
    if (idx >= kNN) continue;
    sx[idx][0] = gx[q][0];
    sx[idx][1] = gx[q][1];
    if constexpr (MODE == TILE_RES) {
#pragma unroll
      for (int c = 0; c < D; ++c) su[idx][c] = gu[q][c];
    }
    if constexpr (M::NODAL) {
      sn[idx][0] = gn[q][0];
      sn[idx][1] = gn[q][1];
    }
    if constexpr (MODE == TILE_TAN) sm[idx] = gm[q];
  }
  // the SELL chunk of the owned row, read while the elements are evaluated
  int32_t obase = 0, owidth = 0;
  if constexpr (MODE == TILE_TAN) {
    if (orow >= 0) {
      obase = A.sell_ptr[orow >> 6];
      owidth = (A.sell_ptr[(orow >> 6) + 1] - obase) >> 6;
    }
  }
  __syncthreads();
  // element phase: thread -> element (i0 - 1 + lei, j0 - 1 + lej); local node a of it sits at
  // staged node (lei + kDi[a], lej + kDj[a])
  constexpr int kDi[4] = {0, 0, 1, 1}, kDj[4] = {0, 1, 1, 0};
  {
    const int el = tid;
    const int lei = el / kEJ, lej = el % kEJ;
    const int ei = i0 - 1 + lei, ej = j0 - 1 + lej;
    if (ei >= 0 && ei < nx && ej >= 0 && ej < ny) {
      double xl[NPE][2], ul[NPE][D], nl[NPE][2];
#pragma unroll
      for (int a = 0; a < NPE; ++a) {
        const int s = (lei + kDi[a]) * kNJ + lej + kDj[a];
        xl[a][0] = sx[s][0];
        xl[a][1] = sx[s][1];
#pragma unroll
        for (int c = 0; c < D; ++c) ul[a][c] = MODE == TILE_RES ? su[s][c] : 0.0;
        if constexpr (M::NODAL) {
          nl[a][0] = sn[s][0];
          nl[a][1] = sn[s][1];
        } else {
          nl[a][0] = nl[a][1] = 0.0;
        }
      }
      bool affine = false;
      if constexpr (AffineModel<M>::ok) {
        double det, B00, B01, B11;
        // the residual of the source model with the source evaluated here needs the points
        const bool par = q1_parallelogram(xl, det, B00, B01, B11);
        if (AFF || ((MODE != TILE_RES || AffineModel<M>::nodal || A.want_source == 0) && par)) {
          affine = true;
          double kq[4] = {1.0, 1.0, 1.0, 1.0};
          if constexpr (AffineModel<M>::nodal) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              double g = 0.0;
#pragma unroll
              for (int a = 0; a < NPE; ++a) g += q1_N(q, a) * nl[a][0];
              kq[q] = exp(g);
            }
          }
          if constexpr (MODE == TILE_RES) {
            double K[4][4];
            q1_affine_stiffness<double, AffineModel<M>::nodal>(det, B00, B01, B11, kq, K);
#pragma unroll
            for (int a = 0; a < NPE; ++a) {
              double r = 0.0;
#pragma unroll
              for (int b = 0; b < NPE; ++b) r += K[a][b] * ul[b][0];
              if constexpr (AffineModel<M>::nodal) {
                // - det sum_q f_q N_q[a], f_q = N_q . f
                double src = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                  double f = 0.0;
#pragma unroll
                  for (int c = 0; c < NPE; ++c) f += q1_N(q, c) * nl[c][1];
                  src += f * q1_N(q, a);
                }
                r -= det * src;
              }
              er[a * D * kNE + el] = r;
            }
          } else if constexpr (GEOM) {
            er[el] = (T)det;
            er[kNE + el] = (T)B00;
            er[2 * kNE + el] = (T)B01;
            er[3 * kNE + el] = (T)B11;
          } else {
            T K[4][4];
            q1_affine_stiffness<T, AffineModel<M>::nodal>(det, B00, B01, B11, kq, K);
            store_matrix<MODE>(er, el, K);
          }
        }
      }
      if constexpr (AFF) {
        // every element takes the parallelogram path
      } else if (affine) {
        // element done by the parallelogram path
      } else if constexpr (MODE == TILE_RES) {
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
          for (int c = 0; c < D; ++c) er[(a * D + c) * kNE + el] = out[a][c];
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
        store_matrix<MODE>(er, el, K);
      }
    }
  }
  __syncthreads();
  // owner phase: thread -> node (i0 + li, j0 + lj)
  if (!owner) return;
  const int li = oli, lj = olj;
  const int I = oI, J = oJ;
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
        if (ok[k]) s += (double)er[(kLoc[k] * D + c) * kNE + elx[k]];
      if (A.load) s += pl[c];
      A.r[onode * D + c] = A.overwrite ? s : pr[c] + s;
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
      if constexpr (GEOM) {
        const T d = er[elx[k]], b00 = er[kNE + elx[k]], b01 = er[2 * kNE + elx[k]], b11 = er[3 * kNE + elx[k]];
#pragma unroll
        for (int b = 0; b < NPE; ++b) {
          const int dI = kEi[k] + kDi[b], dJ = kEjo[k] + kDj[b];
          const int lo = kLoc[k] < b ? kLoc[k] : b, hi = kLoc[k] < b ? b : kLoc[k];
          v[dI + 1][dJ + 1] += q1_affine_entry<T>(lo, hi, d, b00, b01, b11);
        }
        continue;
      }
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        const int dI = kEi[k] + kDi[b], dJ = kEjo[k] + kDj[b];
        if (MODE == TILE_DIAG && b != kLoc[k]) continue;
        v[dI + 1][dJ + 1] += er[(MODE == TILE_DIAG ? kLoc[k] : sym_entry(kLoc[k], b)) * kNE + elx[k]];
      }
    }
    if constexpr (MODE == TILE_DIAG) {
      A.r[onode] = A.overwrite ? (double)v[1][1] : pr[0] + (double)v[1][1];
    } else {
      if (orow < 0) return;  // Dirichlet node: no reduced row
      TO* vals = (TO*)A.vals + obase + (orow & 63);
      int k = 0;
#pragma unroll
      for (int dI = -1; dI <= 1; ++dI)
#pragma unroll
        for (int dJ = -1; dJ <= 1; ++dJ) {
          if (sm[(li + 1 + dI) * kNJ + lj + 1 + dJ] == 0) {  // free node inside the grid
            vals[k * 64] = (TO)v[dI + 1][dJ + 1];
            ++k;
          }
        }
      for (; k < owidth; ++k) vals[k * 64] = 0;  // chunk padding
    }
  }
}

// one thread per element of the structured grid: flag any element that is not a parallelogram
__global__ void k_parallelogram_check(const double* coords, int64_t n_nodes, int64_t nx, int64_t ny, int* irregular) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nx * ny) return;
  const int64_t b = (e / ny) * (ny + 1) + e % ny;
  const int64_t nd[4] = {b, b + 1, b + ny + 2, b + ny + 1};  // local nodes 0..3 as in k_tile
  double x[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    x[a][0] = coords[nd[a]];
    x[a][1] = coords[n_nodes + nd[a]];
  }
  double det, B00, B01, B11;
  if (!q1_parallelogram(x, det, B00, B01, B11)) *irregular = 1;  // every writer stores the same value
}

// Whether every element of the set is a parallelogram; decided on the device once per set (the
// coordinates are fixed for the plan's lifetime).
bool all_parallelograms(Plan& P, ElemSet& S, hipStream_t st) {
  if (S.all_parallelograms < 0) {
    DevBuf<int> flag;
    flag.alloc(1);
    fill<int>(flag.p, 1, 0, st);
    const int64_t ne = S.gnx * S.gny;
    if (ne > 0) {
      hipLaunchKernelGGL(k_parallelogram_check, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, P.coords.p,
                         P.n_nodes, S.gnx, S.gny, flag.p);
      APX_LAUNCH_CHECK();
    }
    int h = 1;
    APX_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    APX_HIP(hipStreamSynchronize(st));
    S.all_parallelograms = h ? 0 : 1;
  }
  return S.all_parallelograms == 1;
}

template <class M, class T, int MODE, class TO = T>
void run_tile(Plan& P, ElemSet& S, const double* u, double* r, void* vals, const double* load, bool overwrite,
              hipStream_t st) {
  TileArgs A;
  A.cx = P.coords.p;
  A.cy = P.coords.p + P.n_nodes;
  A.nx = (int)S.gnx;
  A.ny = (int)S.gny;
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
  const int64_t tiles_j = (S.gny + 1 + kTJ - 1) / kTJ, tiles = tiles_j * ((S.gnx + 1 + kTI - 1) / kTI);
  const int64_t g = round_xcd(tiles);
  APX_REQUIRE(g < ((int64_t)1 << 31), APX_ERR_UNSUPPORTED, "structured grid too large for the owner-tile launch");
  A.tiles_j = (int)tiles_j;
  A.tiles = (int)tiles;
  bool aff = false;
  if constexpr (AffineModel<M>::ok) {
    // the source model's residual takes the parallelogram path only with its source part given
    aff = (MODE != TILE_RES || AffineModel<M>::nodal || A.want_source == 0) && all_parallelograms(P, S, st);
  }
  if (aff) {
    if constexpr (AffineModel<M>::ok)
      hipLaunchKernelGGL((k_tile<M, T, MODE, true, TO>), dim3((unsigned)g), dim3(kTBS), 0, st, A);
  } else {
    hipLaunchKernelGGL((k_tile<M, T, MODE, false, TO>), dim3((unsigned)g), dim3(kTBS), 0, st, A);
  }
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
  if (P.n_dofs >= ((int64_t)1 << 31) - 64) return false;  // the owner-tile kernels index in 32 bits
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
  const bool f64 = precision == APX_FP64, rounded = precision == kFP64RoundedToFP32;
  switch (S.model) {
    case APX_MODEL_POISSON_SOURCE:
      if (f64) return run_tile<PoissonSource<4, 2>, double, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
      if (rounded)
        return run_tile<PoissonSource<4, 2>, double, TILE_TAN, float>(P, S, u, nullptr, vals, nullptr, true, st);
      return run_tile<PoissonSource<4, 2>, float, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
    case APX_MODEL_POISSON_VARK:
      if (f64) return run_tile<PoissonVarK<4, 2>, double, TILE_TAN>(P, S, u, nullptr, vals, nullptr, true, st);
      if (rounded)
        return run_tile<PoissonVarK<4, 2>, double, TILE_TAN, float>(P, S, u, nullptr, vals, nullptr, true, st);
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
