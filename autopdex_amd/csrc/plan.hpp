// Plan: mesh, DOF layout, reduced CSR pattern and device workspace.
#pragma once
#include <map>
#include <memory>
#include <string>

#include <hip/hip_ext.h>

#include "common.hpp"

namespace apx {

constexpr int kMaxParams = 8;

struct ModelParams {
  double p[kMaxParams];
  const double* nodal[2];  // device nodal fields (var-k: g, f)
};

struct ElemSet {
  int npe = 0;
  int model = 0;
  int64_t n_elems = 0;
  ModelParams params{};
  std::vector<int64_t> conn_host;  // original order, n_elems * npe
  // colour-permuted device data
  DevBuf<int32_t> conn;            // n_elems * npe, colour-major order
  DevBuf<int32_t> slots;           // n_elems * (npe*d)^2 reduced CSR slots, -1 = Dirichlet
  DevBuf<int32_t> conn_orig;       // original element order (made on first BCOO export)
  std::vector<int64_t> colour_offsets;  // n_colours + 1
  // Structured Q1 grid (the connectivity of mesher.structured_mesh): element
  // e = i * gny + j has nodes b, b + 1, b + gny + 2, b + gny + 1 with
  // b = i * (gny + 1) + j.  Such sets use the owner-tile kernels.
  int structured = 0;
  int64_t gnx = 0, gny = 0;
  DevBuf<double> load;     // u-independent residual part (precomputed-source models), n_dofs
  bool load_valid = false;
  int all_parallelograms = -1;  // structured grid of parallelograms only: -1 not yet checked, 0 no, 1 yes
};

// Timing of a kernel family: HIP events on the caller's stream.
struct KernelTimer {
  double total_ms = 0.0;
  int64_t count = 0;
  int64_t bytes = 0;  // algorithmic bytes per launch (the average over the launches timed)
  int64_t flops = 0;  // algorithmic flops per launch (matrix-core work; 0 for streaming kernels)
};

struct Plan {
  int dim = 0, d = 0;
  int64_t n_nodes = 0, n_dofs = 0, n_free = 0, nnz = 0;
  bool finalized = false;
  double build_seconds = 0.0;

  std::vector<double> coords_host;   // n_nodes * dim row-major
  std::vector<uint8_t> dmask_host;   // n_dofs
  std::vector<ElemSet> sets;

  // host copies of the reduced pattern
  std::vector<int32_t> indptr_h, indices_h, red2full_h, full2red_h;

  // device data
  DevBuf<double> coords;       // SoA: dim arrays of n_nodes
  DevBuf<uint8_t> dmask;       // n_dofs
  DevBuf<int32_t> full2red;    // n_dofs, -1 for Dirichlet
  DevBuf<int32_t> red2full;    // n_free
  DevBuf<int32_t> indptr;      // n_free + 1
  DevBuf<int32_t> indices;     // nnz
  // Internal matrix layout: SELL-64 (chunks of 64 rows = one wavefront; within
  // a chunk entries are stored column-major, lane = row, padded to the chunk's
  // longest row).  Element slot maps point into this layout; the C ABI
  // exchanges values in canonical CSR order through csr2sell.
  int64_t sell_total = 0, n_chunks = 0;
  bool stencil_pattern = false;  // reduced rows = free nodes of the 3 x 3 stencil of a structured grid
  std::vector<int32_t> sell_ptr_h, sell_cols_h, csr2sell_h;
  DevBuf<int32_t> sell_ptr;    // n_chunks + 1 (entry offsets)
  DevBuf<int32_t> sell_cols;   // sell_total (padding: column 0, value 0); empty when sell_dcol is used
  // The same columns as int16 offsets from the entry's row (rows past n_free: from row
  // n_free - 1; padding: offset 0), built when every offset fits: 2 bytes per stored entry
  // instead of 4 in the SpMV's stream.  Matrices of bandwidth under 2^15 (every structured grid
  // here, 3-DOF grids up to about 100^3 nodes) qualify.
  DevBuf<int16_t> sell_dcol;
  DevBuf<int32_t> csr2sell;    // nnz
  DevBuf<int32_t> diag_slot;   // n_free, position of the diagonal in SELL

  // workspace (lazily sized)
  DevBuf<double> vals64, w64;  // nnz, vectors
  DevBuf<float> vals32, w32;
  DevBuf<double> r_full, diag_full;  // n_dofs
  DevBuf<double> partials;     // reduction partials
  DevBuf<double> tw, tpart;    // network training workspace (4 n_free) and its reduction partials
  DevBuf<double> kstate;       // device Krylov scalars
  double* pinned = nullptr;    // host pinned scalars
  struct UNet* net = nullptr;  // attached learned preconditioner (not owned)
  const double* net_g = nullptr;  // nodal coefficient channel of the network input (device, optional)
  // Jacobi scaling of the network (1 / sqrt|A_ii| per free DOF): the network maps
  // D^-1/2 r to D^1/2 z, i.e. it approximates the inverse of D^-1/2 A D^-1/2.
  // Set from the matrix of the solve or training step that uses the network.
  DevBuf<double> net_dsq;
  bool net_dsq_set = false;
  // 0 off; 1 kernel-family timers of the solvers, plus one timer per launched network kernel
  // (name = its template, e.g. "k_conv3x3<64, 64, false, 4, 2>") for the names in timing_filter;
  // 2 every network kernel
  int timing = 0;
  std::string timing_filter;  // ';'-separated kernel names timed per launch at level 1
  std::map<std::string, KernelTimer> timers;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> evpool;  // recycled timing events

  ~Plan();
  int64_t device_bytes() const;
};

// Non-blocking launch timing: event pairs recorded on the launch stream and
// resolved after the work completes.
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
};
EventPair timing_begin(Plan& P, hipStream_t st);
void timing_end(Plan& P, EventPair& e, hipStream_t st);

// Single-kernel timing.  timing_pair() takes two events from the pool without recording them;
// set in t_launch_events, the pair is taken by the next launch_kernel on this thread, which
// hands it to hipExtLaunchKernel: the events then carry that dispatch's own start and end
// timestamps, so their elapsed time is the kernel's execution alone (a pair recorded on the
// stream around a launch also counts the gap between the previous kernel's end and this one's
// start).
EventPair timing_pair(Plan& P);
inline thread_local const EventPair* t_launch_events = nullptr;

template <typename... Args, typename F = void (*)(Args...)>
inline void launch_kernel(F kernel, const dim3& grid, const dim3& block, uint32_t shmem, hipStream_t st,
                          Args... args) {
  const EventPair* e = t_launch_events;
  t_launch_events = nullptr;
  if (e)
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, e->a, e->b, 0u, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
}

// Wait for the pairs, add their durations to timer `name`, recycle events.
void timing_collect(Plan& P, const char* name, std::vector<EventPair>& pairs, int64_t bytes, size_t count_limit);

inline int64_t ndofs_elem(const Plan& P, const ElemSet& s) { return (int64_t)s.npe * P.d; }

constexpr int kSellC = 64;

// plan.hip
void plan_finalize(Plan& P);
void plan_build_host(Plan& P, std::vector<std::vector<int32_t>>& conn_perm,
                     std::vector<std::vector<int32_t>>& slots_perm);

}  // namespace apx
