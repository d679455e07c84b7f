// Element-loop assembly kernels: residual, reduced-CSR tangent values and
// tangent diagonal for the 'user potential' mode.
//
// Replaces jacrev / jacfwd(jacrev) of the element potential under lax.map and
// the .at[].add scatter (assembler.py:1064-1149, :219-436).
//
// Determinism: elements are stored colour-major (no two elements of a colour
// share a node), one launch per colour in fixed order, and every output entry
// is updated by a plain read-add-write from exactly one lane per launch.  No
// floating-point atomics, so every assembly is bitwise reproducible.
#include "assembly.hpp"
#include "elements.hpp"

namespace apx {

enum { MODE_RES = 0, MODE_TAN = 1, MODE_DIAG = 2, MODE_COO = 3 };

struct AsmArgs {
  const double* coords;  // SoA, dim x n_nodes
  int64_t n_nodes;
  const int32_t* conn;   // colour-major, n_elems x npe
  int64_t e0, e1;        // element range of one colour
  const double* u;       // n_dofs
  ModelParams mp;
  const int32_t* slots;  // colour-major, n_elems x (npe d)^2
  double* r;             // residual or diagonal output (n_dofs)
  void* vals;            // reduced CSR values
};

template <class M, class T, int MODE, int BS>
__global__ __launch_bounds__(BS) void k_team(AsmArgs A) {
  constexpr int NPE = M::NPE, DIM = M::DIM, D = M::D, ND = NPE * D;
  constexpr int EPB = BS / NPE;
  using QPT = typename M::template QP<T>;
  __shared__ double sx[EPB][NPE][DIM];
  __shared__ double su[EPB][NPE][D];
  __shared__ double sn[EPB][NPE][2];
  __shared__ QPT sq[EPB][NPE];

  const int le = threadIdx.x / NPE, t = threadIdx.x % NPE;
  const int64_t e = A.e0 + (int64_t)blockIdx.x * EPB + le;
  const bool active = e < A.e1;
  int32_t node = 0;
  if (active) {
    node = A.conn[e * NPE + t];
#pragma unroll
    for (int k = 0; k < DIM; ++k) sx[le][t][k] = A.coords[k * A.n_nodes + node];
#pragma unroll
    for (int i = 0; i < D; ++i) su[le][t][i] = A.u[(int64_t)node * D + i];
    if constexpr (M::NODAL) {
      sn[le][t][0] = A.mp.nodal[0][node];
      sn[le][t][1] = A.mp.nodal[1][node];
    }
  }
  __syncthreads();
  if (active) M::template qp<T>(t, sx[le], su[le], sn[le], A.mp, MODE == MODE_RES, sq[le][t]);
  __syncthreads();
  if (!active) return;

  if constexpr (MODE == MODE_RES) {
    double out[D];
    M::res_row(t, (const typename M::template QP<double>*)sq[le], A.mp, out);
#pragma unroll
    for (int i = 0; i < D; ++i) A.r[(int64_t)node * D + i] += out[i];
  } else {
    T out[D][ND];
    M::template tan_rows<T>(t, sq[le], A.mp, out);
    if constexpr (MODE == MODE_DIAG) {
#pragma unroll
      for (int i = 0; i < D; ++i) A.r[(int64_t)node * D + i] += (double)out[i][t * D + i];
    } else if constexpr (MODE == MODE_COO) {
      // element-major, row-major block (the reference's BCOO order); every entry written once
      T* data = (T*)A.vals + ((int64_t)e * ND + (int64_t)t * D) * ND;
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int B = 0; B < ND; ++B) data[i * ND + B] = out[i][B];
    } else {
      T* vals = (T*)A.vals;
      const int32_t* sl = A.slots + (e * ND + (int64_t)t * D) * ND;
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int B = 0; B < ND; ++B) {
          const int32_t s = sl[i * ND + B];
          if (s >= 0) vals[s] += out[i][B];
        }
    }
  }
}

template <class T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <class T>
void fill(T* p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, n, v);
  APX_LAUNCH_CHECK();
}
template void fill<double>(double*, int64_t, double, hipStream_t);
template void fill<float>(float*, int64_t, float, hipStream_t);
template void fill<int>(int*, int64_t, int, hipStream_t);

__global__ void k_round_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];  // round to nearest
}

void round_to_f32(const double* src, float* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_round_to_f32, dim3(blocks), dim3(256), 0, s, src, dst, n);
  APX_LAUNCH_CHECK();
}

template <class M, class T, int MODE>
static void run_set(Plan& P, ElemSet& S, const double* u, double* r, void* vals, hipStream_t st) {
  // whole elements per block: every thread's element slot le = threadIdx.x / NPE is < EPB
  constexpr int EPB = (M::NPE >= 8 ? 64 : 128) / M::NPE;
  constexpr int BS = EPB * M::NPE;
  AsmArgs A;
  A.coords = P.coords.p;
  A.n_nodes = P.n_nodes;
  A.conn = S.conn.p;
  A.u = u;
  A.mp = S.params;
  A.slots = S.slots.p;
  A.r = r;
  A.vals = vals;
  if constexpr (M::NODAL) {
    APX_REQUIRE(S.params.nodal[0] && S.params.nodal[1], APX_ERR_STATE,
                "variable-coefficient Poisson set needs nodal fields g and f bound");
  }
  if constexpr (MODE == MODE_COO) {
    // element blocks go to disjoint slots: one launch over the original element order
    if (S.conn_orig.n < (size_t)(S.n_elems * S.npe)) {
      std::vector<int32_t> c32(S.conn_host.begin(), S.conn_host.end());
      S.conn_orig.upload(c32.data(), c32.size());
    }
    A.conn = S.conn_orig.p;
    A.e0 = 0;
    A.e1 = S.n_elems;
    hipLaunchKernelGGL((k_team<M, T, MODE, BS>), dim3((S.n_elems + EPB - 1) / EPB), dim3(BS), 0, st, A);
    APX_LAUNCH_CHECK();
    return;
  }
  for (size_t c = 0; c + 1 < S.colour_offsets.size(); ++c) {
    A.e0 = S.colour_offsets[c];
    A.e1 = S.colour_offsets[c + 1];
    const int64_t ne = A.e1 - A.e0;
    if (ne <= 0) continue;
    hipLaunchKernelGGL((k_team<M, T, MODE, BS>), dim3((ne + EPB - 1) / EPB), dim3(BS), 0, st, A);
    APX_LAUNCH_CHECK();
  }
}

template <class T, int MODE>
static void dispatch(Plan& P, ElemSet& S, const double* u, double* r, void* vals, hipStream_t st) {
  const int key = S.model * 100 + S.npe * 10 + P.dim;
  switch (key) {
    case APX_MODEL_POISSON_SOURCE * 100 + 42:
      APX_REQUIRE(P.d == 1, APX_ERR_ARG, "Poisson model needs one DOF per node");
      return run_set<PoissonSource<4, 2>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_POISSON_SOURCE * 100 + 83:
      APX_REQUIRE(P.d == 1, APX_ERR_ARG, "Poisson model needs one DOF per node");
      return run_set<PoissonSource<8, 3>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_POISSON_VARK * 100 + 42:
      APX_REQUIRE(P.d == 1, APX_ERR_ARG, "Poisson model needs one DOF per node");
      return run_set<PoissonVarK<4, 2>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_POISSON_VARK * 100 + 83:
      APX_REQUIRE(P.d == 1, APX_ERR_ARG, "Poisson model needs one DOF per node");
      return run_set<PoissonVarK<8, 3>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_NEO_HOOKE * 100 + 42:
      APX_REQUIRE(P.d == 2, APX_ERR_ARG, "plane-strain Neo-Hooke needs 2 DOFs per node");
      return run_set<NeoHooke<4, 2>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_NEO_HOOKE * 100 + 83:
      APX_REQUIRE(P.d == 3, APX_ERR_ARG, "3D Neo-Hooke needs 3 DOFs per node");
      return run_set<NeoHooke<8, 3>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_NEO_HOOKE * 100 + 92:
      APX_REQUIRE(P.d == 2, APX_ERR_ARG, "plane-strain Neo-Hooke needs 2 DOFs per node");
      return run_set<NeoHooke<9, 2>, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_LINE_TRACTION * 100 + 22:
      APX_REQUIRE(P.d == 2, APX_ERR_ARG, "line traction needs 2 DOFs per node");
      return run_set<LineTraction, T, MODE>(P, S, u, r, vals, st);
    case APX_MODEL_LINE_TRACTION * 100 + 32:
      APX_REQUIRE(P.d == 2, APX_ERR_ARG, "line traction needs 2 DOFs per node");
      return run_set<LineTraction3, T, MODE>(P, S, u, r, vals, st);
    default:
      throw Error(APX_ERR_UNSUPPORTED, "no kernel for model " + std::to_string(S.model) + " with " +
                                           std::to_string(S.npe) + "-node elements in " +
                                           std::to_string(P.dim) + "D");
  }
}

void assemble_set_colour_residual(Plan& P, ElemSet& S, const double* u, double* r, hipStream_t st) {
  dispatch<double, MODE_RES>(P, S, u, r, nullptr, st);
}

// Algorithmic (compulsory) bytes of one assembly: what the kernels that run
// must move at least once -- coordinates and DOFs read, outputs written; the
// colour path also reads the connectivity (and, for the tangent, the slot
// map is not counted: it is an implementation choice), the owner-tile path
// derives the connectivity from the grid and reads the precomputed source
// part instead.
int64_t residual_bytes(const Plan& P) {
  int64_t b = P.n_nodes * P.dim * 8 + P.n_dofs * 8 * 2;
  for (auto& s : P.sets) {
    if (tile_supported(P, s, 0))
      b += s.model == APX_MODEL_POISSON_SOURCE ? P.n_dofs * 8 : 0;
    else
      b += s.n_elems * s.npe * 4;
    if (s.model == APX_MODEL_POISSON_VARK) b += P.n_nodes * 16;
  }
  return b;
}

int64_t tangent_bytes(const Plan& P, int precision) {
  const int vb = precision == APX_FP64 ? 8 : 4;
  int64_t b = P.n_nodes * P.dim * 8;
  if (tangent_by_tiles(P))  // values incl. padding, maps, mask; the tile models' tangent does not read u
    b += P.sell_total * vb + P.n_free * 4 + (P.n_chunks + 1) * 4 + P.n_dofs;
  else
    b += P.n_dofs * 8 + P.nnz * vb;
  for (auto& s : P.sets) {
    if (!tangent_by_tiles(P)) b += s.n_elems * s.npe * 4;
    if (s.model == APX_MODEL_POISSON_VARK) b += P.n_nodes * 16;
  }
  return b;
}

// One set carries the whole tangent, on a structured grid whose reduced pattern
// is the 3 x 3 node stencil: the owner-tile kernel writes every stored entry.
bool tangent_by_tiles(const Plan& P) {
  if (!P.stencil_pattern) return false;
  const ElemSet* only = nullptr;
  for (auto& s : P.sets) {
    if (s.model == APX_MODEL_LINE_TRACTION) continue;  // zero tangent
    if (only) return false;
    only = &s;
  }
  return only && tile_supported(P, *only, 1);
}

void assemble_residual(Plan& P, const double* u, double* r, hipStream_t st) {
  bool have = false;  // r holds this assembly's partial sums
  for (auto& S : P.sets) {
    if (tile_supported(P, S, 0)) {
      tile_residual(P, S, u, r, !have, st);
    } else {
      if (!have) fill<double>(r, P.n_dofs, 0.0, st);
      dispatch<double, MODE_RES>(P, S, u, r, nullptr, st);
    }
    have = true;
  }
  if (!have) fill<double>(r, P.n_dofs, 0.0, st);
}

// Tangent values in the plan's SELL layout (P.sell_total entries, padding 0).
void assemble_tangent(Plan& P, const double* u, void* vals, int precision, hipStream_t st) {
  if (tangent_by_tiles(P)) {
    for (auto& S : P.sets)
      if (S.model != APX_MODEL_LINE_TRACTION) tile_tangent(P, S, u, vals, precision, st);
    return;
  }
  if (precision == APX_FP64) {
    fill<double>((double*)vals, P.sell_total, 0.0, st);
    for (auto& S : P.sets)
      if (S.model != APX_MODEL_LINE_TRACTION) dispatch<double, MODE_TAN>(P, S, u, nullptr, vals, st);
  } else {
    fill<float>((float*)vals, P.sell_total, 0.0f, st);
    for (auto& S : P.sets)
      if (S.model != APX_MODEL_LINE_TRACTION) dispatch<float, MODE_TAN>(P, S, u, nullptr, vals, st);
  }
}

void assemble_tangent_rounded(Plan& P, const double* u, float* vals, hipStream_t st) {
  if (tangent_by_tiles(P)) {
    for (auto& S : P.sets)
      if (S.model != APX_MODEL_LINE_TRACTION) tile_tangent(P, S, u, vals, kFP64RoundedToFP32, st);
    return;
  }
  if (P.vals64.n < (size_t)P.sell_total) P.vals64.alloc(P.sell_total);
  assemble_tangent(P, u, P.vals64.p, APX_FP64, st);
  round_to_f32(P.vals64.p, vals, P.sell_total, st);
}

int64_t tangent_coo_entries(const Plan& P) {
  int64_t n = 0;
  for (auto& S : P.sets) n += S.n_elems * ndofs_elem(P, S) * ndofs_elem(P, S);
  return n;
}

// The full tangent as the reference's BCOO with duplicates: per set, element-major blocks
// (rows repeat(gdofs), columns tile(gdofs)); sets concatenated in order.
void assemble_tangent_coo(Plan& P, const double* u, void* data, int precision, hipStream_t st) {
  int64_t off = 0;
  for (auto& S : P.sets) {
    const int64_t nd = ndofs_elem(P, S);
    if (precision == APX_FP64)
      dispatch<double, MODE_COO>(P, S, u, nullptr, (double*)data + off, st);
    else
      dispatch<float, MODE_COO>(P, S, u, nullptr, (float*)data + off, st);
    off += S.n_elems * nd * nd;
  }
}

void assemble_diagonal(Plan& P, const double* u, double* diag, hipStream_t st) {
  bool have = false;
  for (auto& S : P.sets) {
    if (S.model == APX_MODEL_LINE_TRACTION) continue;  // zero tangent
    if (tile_supported(P, S, 2)) {
      tile_diagonal(P, S, u, diag, !have, st);
    } else {
      if (!have) fill<double>(diag, P.n_dofs, 0.0, st);
      dispatch<double, MODE_DIAG>(P, S, u, diag, nullptr, st);
    }
    have = true;
  }
  if (!have) fill<double>(diag, P.n_dofs, 0.0, st);
}

}  // namespace apx
