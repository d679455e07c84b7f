// Host-side plan construction: reduced CSR pattern, element colouring and
// element -> CSR slot maps.  Replaces, once per mesh and Dirichlet mask, the
// per-call index construction of _get_indices (assembler.py:47-141) and the
// scipy COO -> CSR -> [:, free][free] reduction of scipy_assembling
// (solver.py:1180-1222).  The reduced pattern equals scipy's canonical CSR:
// rows and columns are the free DOFs in increasing order, one entry per
// coupled (free, free) DOF pair of any element of any set, columns sorted.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "plan.hpp"

namespace apx {

Plan::~Plan() {
  if (pinned) (void)hipHostFree(pinned);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  for (auto e : evpool) (void)hipEventDestroy(e);
}

static hipEvent_t pool_event(Plan& P) {
  if (!P.evpool.empty()) {
    hipEvent_t e = P.evpool.back();
    P.evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  APX_HIP(hipEventCreate(&e));
  return e;
}

EventPair timing_begin(Plan& P, hipStream_t st) {
  EventPair e;
  if (!P.timing) return e;
  e.a = pool_event(P);
  e.b = pool_event(P);
  APX_HIP(hipEventRecord(e.a, st));
  return e;
}

EventPair timing_pair(Plan& P) {
  EventPair e;
  e.a = pool_event(P);
  e.b = pool_event(P);
  return e;
}

void timing_end(Plan& P, EventPair& e, hipStream_t st) {
  if (e.b) APX_HIP(hipEventRecord(e.b, st));
}

void timing_collect(Plan& P, const char* name, std::vector<EventPair>& pairs, int64_t bytes, size_t count_limit) {
  if (pairs.empty()) return;
  auto& t = P.timers[name];
  for (size_t i = 0; i < pairs.size(); ++i) {
    auto& e = pairs[i];
    if (!e.a) continue;
    if (i < count_limit) {
      APX_HIP(hipEventSynchronize(e.b));
      float ms = 0.f;
      APX_HIP(hipEventElapsedTime(&ms, e.a, e.b));
      t.total_ms += ms;
      t.count += 1;
      t.bytes = bytes;
    } else {
      APX_HIP(hipEventSynchronize(e.b));
    }
    P.evpool.push_back(e.a);
    P.evpool.push_back(e.b);
  }
  pairs.clear();
}

int64_t Plan::device_bytes() const {
  int64_t b = coords.bytes() + dmask.bytes() + full2red.bytes() + red2full.bytes() + indptr.bytes() +
              indices.bytes() + diag_slot.bytes() + sell_ptr.bytes() + sell_cols.bytes() + sell_dcol.bytes() +
              csr2sell.bytes() + vals64.bytes() + vals32.bytes() + w64.bytes() + w32.bytes() + r_full.bytes() +
              diag_full.bytes() + partials.bytes();
  for (auto& s : sets) b += s.conn.bytes() + s.slots.bytes();
  return b;
}

// Build the host-side reduced pattern and colour/slot data (no device calls).
void plan_build_host(Plan& P, std::vector<std::vector<int32_t>>& conn_perm,
                     std::vector<std::vector<int32_t>>& slots_perm) {
  const int d = P.d;
  const int64_t N = P.n_nodes, ndof = P.n_dofs;
  APX_REQUIRE(!P.sets.empty(), APX_ERR_STATE, "plan has no element sets");

  // 1. free-DOF numbering (monotone, so reduced columns stay sorted)
  P.full2red_h.assign(ndof, -1);
  P.red2full_h.clear();
  for (int64_t g = 0; g < ndof; ++g)
    if (!P.dmask_host[g]) {
      P.full2red_h[g] = (int32_t)P.red2full_h.size();
      P.red2full_h.push_back((int32_t)g);
    }
  P.n_free = (int64_t)P.red2full_h.size();

  // 2. node -> element incidence over all sets
  std::vector<int64_t> inc_ptr(N + 1, 0);
  for (auto& s : P.sets)
    for (int64_t i = 0; i < s.n_elems * s.npe; ++i) {
      int64_t n = s.conn_host[i];
      APX_REQUIRE(n >= 0 && n < N, APX_ERR_ARG, "connectivity references a node outside [0, n_nodes)");
      inc_ptr[n + 1]++;
    }
  for (int64_t n = 0; n < N; ++n) inc_ptr[n + 1] += inc_ptr[n];
  std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
  std::vector<std::pair<int32_t, int64_t>> inc(inc_ptr[N]);  // (set, element)
  for (int si = 0; si < (int)P.sets.size(); ++si) {
    auto& s = P.sets[si];
    for (int64_t e = 0; e < s.n_elems; ++e)
      for (int a = 0; a < s.npe; ++a) inc[fill[s.conn_host[e * s.npe + a]]++] = {si, e};
  }

  // 3. node adjacency (sorted unique) -> reduced DOF rows
  std::vector<int64_t> adj_ptr(N + 1, 0);
  std::vector<int32_t> adj;
  adj.reserve((size_t)N * (P.dim == 3 ? 27 : 9));
  std::vector<int32_t> tmp;
  for (int64_t n = 0; n < N; ++n) {
    tmp.clear();
    for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
      auto& s = P.sets[inc[k].first];
      const int64_t* c = &s.conn_host[inc[k].second * s.npe];
      for (int a = 0; a < s.npe; ++a) tmp.push_back((int32_t)c[a]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    adj.insert(adj.end(), tmp.begin(), tmp.end());
    adj_ptr[n + 1] = (int64_t)adj.size();
  }
  P.indptr_h.assign(P.n_free + 1, 0);
  int64_t nnz = 0;
  for (int64_t row = 0; row < P.n_free; ++row) {
    int64_t g = P.red2full_h[row], n = g / d;
    for (int64_t k = adj_ptr[n]; k < adj_ptr[n + 1]; ++k)
      for (int c = 0; c < d; ++c) nnz += P.dmask_host[(int64_t)adj[k] * d + c] ? 0 : 1;
    APX_REQUIRE(nnz < (int64_t)INT32_MAX, APX_ERR_UNSUPPORTED, "reduced CSR exceeds int32 indexing");
    P.indptr_h[row + 1] = (int32_t)nnz;
  }
  P.nnz = nnz;
  P.indices_h.resize(nnz);
  for (int64_t row = 0; row < P.n_free; ++row) {
    int64_t g = P.red2full_h[row], n = g / d;
    int64_t o = P.indptr_h[row];
    for (int64_t k = adj_ptr[n]; k < adj_ptr[n + 1]; ++k)
      for (int c = 0; c < d; ++c) {
        int32_t col = P.full2red_h[(int64_t)adj[k] * d + c];
        if (col >= 0) P.indices_h[o++] = col;
      }
  }

  // 4. SELL-64 layout of the reduced matrix
  P.n_chunks = (P.n_free + kSellC - 1) / kSellC;
  P.sell_ptr_h.assign(P.n_chunks + 1, 0);
  for (int64_t c = 0; c < P.n_chunks; ++c) {
    int32_t w = 0;
    for (int64_t r = c * kSellC; r < std::min<int64_t>(P.n_free, (c + 1) * kSellC); ++r)
      w = std::max(w, P.indptr_h[r + 1] - P.indptr_h[r]);
    APX_REQUIRE((int64_t)P.sell_ptr_h[c] + (int64_t)w * kSellC < (int64_t)INT32_MAX, APX_ERR_UNSUPPORTED,
                "SELL layout exceeds int32 indexing");
    P.sell_ptr_h[c + 1] = P.sell_ptr_h[c] + w * kSellC;
  }
  P.sell_total = P.sell_ptr_h[P.n_chunks];
  P.sell_cols_h.assign(P.sell_total, 0);
  P.csr2sell_h.assign(P.nnz, 0);
  for (int64_t r = 0; r < P.n_free; ++r) {
    const int64_t c = r / kSellC, lane = r % kSellC;
    for (int32_t k = P.indptr_h[r]; k < P.indptr_h[r + 1]; ++k) {
      const int64_t s = P.sell_ptr_h[c] + (int64_t)(k - P.indptr_h[r]) * kSellC + lane;
      P.sell_cols_h[s] = P.indices_h[k];
      P.csr2sell_h[k] = (int32_t)s;
    }
  }

  // 4b. structured Q1 grids (owner-tile kernels)
  for (auto& s : P.sets) {
    s.structured = 0;
    if (P.dim != 2 || s.npe != 4 || s.n_elems < 1 || s.conn_host[0] != 0) continue;
    const int64_t ny = s.conn_host[3] - 1;
    if (ny < 1 || s.n_elems % ny) continue;
    const int64_t nx = s.n_elems / ny;
    if ((nx + 1) * (ny + 1) != P.n_nodes) continue;
    bool ok = true;
    for (int64_t e = 0; e < s.n_elems && ok; ++e) {
      const int64_t i = e / ny, j = e % ny, b = i * (ny + 1) + j;
      const int64_t* c = &s.conn_host[e * 4];
      ok = c[0] == b && c[1] == b + 1 && c[2] == b + ny + 2 && c[3] == b + ny + 1;
    }
    if (ok) {
      s.structured = 1;
      s.gnx = nx;
      s.gny = ny;
    }
  }
  // 4c. is the reduced pattern the free-node 3 x 3 stencil of that grid (one DOF per node)?
  P.stencil_pattern = false;
  for (auto& s : P.sets) {
    if (!s.structured || P.d != 1) continue;
    const int64_t nx = s.gnx, ny = s.gny;
    bool ok = true;
    for (int64_t I = 0; I <= nx && ok; ++I)
      for (int64_t J = 0; J <= ny && ok; ++J) {
        const int32_t row = P.full2red_h[I * (ny + 1) + J];
        if (row < 0) continue;
        int32_t k = P.indptr_h[row];
        for (int dI = -1; dI <= 1 && ok; ++dI)
          for (int dJ = -1; dJ <= 1 && ok; ++dJ) {
            const int64_t I2 = I + dI, J2 = J + dJ;
            if (I2 < 0 || I2 > nx || J2 < 0 || J2 > ny) continue;
            const int32_t col = P.full2red_h[I2 * (ny + 1) + J2];
            if (col < 0) continue;
            ok = k < P.indptr_h[row + 1] && P.indices_h[k] == col;
            ++k;
          }
        ok = ok && k == P.indptr_h[row + 1];
      }
    P.stencil_pattern = ok;
    break;
  }

  // 5. greedy element colouring per set (node-disjoint colours) and slot maps
  //    (slots index the SELL layout)
  conn_perm.assign(P.sets.size(), {});
  slots_perm.assign(P.sets.size(), {});
  std::vector<uint64_t> used(N);
  for (int si = 0; si < (int)P.sets.size(); ++si) {
    auto& s = P.sets[si];
    std::fill(used.begin(), used.end(), 0ull);
    std::vector<uint8_t> colour(s.n_elems);
    int ncol = 0;
    for (int64_t e = 0; e < s.n_elems; ++e) {
      uint64_t m = 0;
      const int64_t* c = &s.conn_host[e * s.npe];
      for (int a = 0; a < s.npe; ++a) m |= used[c[a]];
      APX_REQUIRE(~m != 0ull, APX_ERR_UNSUPPORTED, "element colouring needs more than 64 colours");
      int col = __builtin_ctzll(~m);
      colour[e] = (uint8_t)col;
      ncol = std::max(ncol, col + 1);
      for (int a = 0; a < s.npe; ++a) used[c[a]] |= (1ull << col);
    }
    s.colour_offsets.assign(ncol + 1, 0);
    for (int64_t e = 0; e < s.n_elems; ++e) s.colour_offsets[colour[e] + 1]++;
    for (int c = 0; c < ncol; ++c) s.colour_offsets[c + 1] += s.colour_offsets[c];
    std::vector<int64_t> pos(s.colour_offsets.begin(), s.colour_offsets.end() - 1);
    std::vector<int64_t> perm(s.n_elems);
    for (int64_t e = 0; e < s.n_elems; ++e) perm[pos[colour[e]]++] = e;  // stable within a colour

    const int nd = s.npe * d;
    auto& cp = conn_perm[si];
    auto& sl = slots_perm[si];
    cp.resize(s.n_elems * s.npe);
    sl.resize((size_t)s.n_elems * nd * nd);
    for (int64_t i = 0; i < s.n_elems; ++i) {
      const int64_t* c = &s.conn_host[perm[i] * s.npe];
      for (int a = 0; a < s.npe; ++a) cp[i * s.npe + a] = (int32_t)c[a];
      for (int A = 0; A < nd; ++A) {
        int32_t row = P.full2red_h[c[A / d] * d + A % d];
        for (int B = 0; B < nd; ++B) {
          int32_t col = P.full2red_h[c[B / d] * d + B % d];
          int32_t slot = -1;
          if (row >= 0 && col >= 0) {
            const int32_t* b = P.indices_h.data() + P.indptr_h[row];
            const int32_t* e = P.indices_h.data() + P.indptr_h[row + 1];
            const int32_t* f = std::lower_bound(b, e, col);
            APX_REQUIRE(f != e && *f == col, APX_ERR_INTERNAL, "slot lookup failed");
            slot = P.csr2sell_h[f - P.indices_h.data()];
          }
          sl[((size_t)i * nd + A) * nd + B] = slot;
        }
      }
    }
  }
}

void plan_finalize(Plan& P) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::vector<int32_t>> conn_perm, slots_perm;
  plan_build_host(P, conn_perm, slots_perm);

  // SoA coordinates
  std::vector<double> soa(P.n_nodes * P.dim);
  for (int64_t n = 0; n < P.n_nodes; ++n)
    for (int k = 0; k < P.dim; ++k) soa[k * P.n_nodes + n] = P.coords_host[n * P.dim + k];
  P.coords.upload(soa.data(), soa.size());
  P.dmask.upload(P.dmask_host.data(), P.dmask_host.size());
  P.full2red.upload(P.full2red_h.data(), P.full2red_h.size());
  P.red2full.upload(P.red2full_h.data(), P.red2full_h.size());
  P.indptr.upload(P.indptr_h.data(), P.indptr_h.size());
  P.indices.upload(P.indices_h.data(), P.indices_h.size());
  std::vector<int32_t> ds(P.n_free);
  for (int64_t r = 0; r < P.n_free; ++r) {
    const int32_t* b = P.indices_h.data() + P.indptr_h[r];
    const int32_t* e = P.indices_h.data() + P.indptr_h[r + 1];
    const int32_t* f = std::lower_bound(b, e, (int32_t)r);
    APX_REQUIRE(f != e && *f == r, APX_ERR_INTERNAL, "missing diagonal entry in reduced CSR");
    ds[r] = P.csr2sell_h[f - P.indices_h.data()];
  }
  P.diag_slot.upload(ds.data(), ds.size());
  P.sell_ptr.upload(P.sell_ptr_h.data(), P.sell_ptr_h.size());
  {
    // int16 column offsets when every stored entry's offset from its row fits
    std::vector<int16_t> dcol(P.sell_total, 0);
    bool fits = true;
    for (int64_t r = 0; r < P.n_free && fits; ++r) {
      const int64_t c = r / kSellC, lane = r % kSellC;
      for (int32_t k = P.indptr_h[r]; k < P.indptr_h[r + 1]; ++k) {
        const int64_t s = P.sell_ptr_h[c] + (int64_t)(k - P.indptr_h[r]) * kSellC + lane;
        const int64_t d = (int64_t)P.sell_cols_h[s] - r;
        if (d < INT16_MIN || d > INT16_MAX) {
          fits = false;
          break;
        }
        dcol[s] = (int16_t)d;
      }
    }
    if (fits) {
      P.sell_dcol.upload(dcol.data(), dcol.size());
      P.sell_cols = DevBuf<int32_t>();
    } else {
      P.sell_dcol = DevBuf<int16_t>();
      P.sell_cols.upload(P.sell_cols_h.data(), P.sell_cols_h.size());
    }
  }
  P.csr2sell.upload(P.csr2sell_h.data(), P.csr2sell_h.size());
  std::vector<int32_t>().swap(P.sell_cols_h);
  std::vector<int32_t>().swap(P.csr2sell_h);
  for (size_t si = 0; si < P.sets.size(); ++si) {
    P.sets[si].conn.upload(conn_perm[si].data(), conn_perm[si].size());
    P.sets[si].slots.upload(slots_perm[si].data(), slots_perm[si].size());
  }
  // pinned host scalars: [0, 512) B PCG flags and norms, [512, 768) B FCG state copies,
  // [768, 1024) B the final Krylov state of an allocation-free PCG run
  if (!P.pinned) APX_HIP(hipHostMalloc((void**)&P.pinned, 128 * sizeof(double), hipHostMallocDefault));
  if (!P.ev0) APX_HIP(hipEventCreate(&P.ev0));
  if (!P.ev1) APX_HIP(hipEventCreate(&P.ev1));
  P.finalized = true;
  P.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace apx
