// extern "C" entry points of libautopdex_amd (declared in include/autopdex_amd.h).
#include <cstring>
#include <string>

#include "assembly.hpp"

namespace apx {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
void newton(Plan& P, double* u, const double* dvals, const apx_newton_options& o, apx_newton_stats& s,
            hipStream_t st);
void plan_build_host(Plan& P, std::vector<std::vector<int32_t>>& conn_perm,
                     std::vector<std::vector<int32_t>>& slots_perm);
}  // namespace apx

using namespace apx;

struct apx_plan : public Plan {};

static inline hipStream_t S(void* s) { return (hipStream_t)s; }

// The plan's internal matrix buffer (SELL layout) for a precision.
static void* sell_buffer(Plan& P, int precision) {
  if (precision == APX_FP64) {
    if (P.vals64.n < (size_t)P.sell_total) P.vals64.alloc(P.sell_total);
    return P.vals64.p;
  }
  if (P.vals32.n < (size_t)P.sell_total) P.vals32.alloc(P.sell_total);
  return P.vals32.p;
}
static inline Plan& ready(apx_plan* p) {
  APX_REQUIRE(p != nullptr, APX_ERR_ARG, "null plan");
  APX_REQUIRE(p->finalized, APX_ERR_STATE, "plan not finalized (call apx_plan_finalize)");
  return *p;
}

extern "C" {

const char* apx_last_error(void) { return g_last_error.c_str(); }
int apx_version(void) { return 100; }

int apx_plan_create(apx_plan** out, int dim, int64_t n_nodes, const double* coords_host, int dofs_per_node,
                    const uint8_t* dirichlet_mask_host) {
  return guarded([&] {
    APX_REQUIRE(out && coords_host && dirichlet_mask_host, APX_ERR_ARG, "null argument");
    APX_REQUIRE(dim == 2 || dim == 3, APX_ERR_ARG, "dim must be 2 or 3");
    APX_REQUIRE(n_nodes > 0 && n_nodes < INT32_MAX, APX_ERR_ARG, "n_nodes out of range");
    APX_REQUIRE(dofs_per_node >= 1 && dofs_per_node <= 3, APX_ERR_ARG, "dofs_per_node must be 1..3");
    auto* p = new apx_plan();
    p->dim = dim;
    p->d = dofs_per_node;
    p->n_nodes = n_nodes;
    p->n_dofs = n_nodes * dofs_per_node;
    APX_REQUIRE(p->n_dofs < INT32_MAX, APX_ERR_ARG, "n_dofs exceeds int32 indexing");
    p->coords_host.assign(coords_host, coords_host + n_nodes * dim);
    p->dmask_host.assign(dirichlet_mask_host, dirichlet_mask_host + p->n_dofs);
    *out = p;
  });
}

int apx_plan_add_set(apx_plan* plan, const int64_t* conn_host, int64_t n_elems, int npe, int model,
                     const double* params, int n_params, int* set_out) {
  return guarded([&] {
    APX_REQUIRE(plan && conn_host, APX_ERR_ARG, "null argument");
    APX_REQUIRE(!plan->finalized, APX_ERR_STATE, "plan already finalized");
    APX_REQUIRE(n_elems > 0, APX_ERR_ARG, "empty element set");
    APX_REQUIRE(npe == 2 || npe == 4 || npe == 8, APX_ERR_UNSUPPORTED, "nodes_per_elem must be 2, 4 or 8");
    APX_REQUIRE(n_params >= 0 && n_params <= kMaxParams, APX_ERR_ARG, "too many parameters");
    APX_REQUIRE(model >= APX_MODEL_POISSON_SOURCE && model <= APX_MODEL_LINE_TRACTION, APX_ERR_UNSUPPORTED,
                "unknown model id");
    ElemSet s;
    s.npe = npe;
    s.model = model;
    s.n_elems = n_elems;
    s.conn_host.assign(conn_host, conn_host + n_elems * npe);
    for (int i = 0; i < n_params; ++i) s.params.p[i] = params[i];
    plan->sets.push_back(std::move(s));
    if (set_out) *set_out = (int)plan->sets.size() - 1;
  });
}

int apx_plan_set_params(apx_plan* plan, int set, const double* params, int n_params) {
  return guarded([&] {
    APX_REQUIRE(plan && params, APX_ERR_ARG, "null argument");
    APX_REQUIRE(set >= 0 && set < (int)plan->sets.size(), APX_ERR_ARG, "set index out of range");
    APX_REQUIRE(n_params >= 0 && n_params <= kMaxParams, APX_ERR_ARG, "too many parameters");
    for (int i = 0; i < n_params; ++i) plan->sets[set].params.p[i] = params[i];
  });
}

int apx_plan_bind_nodal_field(apx_plan* plan, int set, int which, const double* field_dev) {
  return guarded([&] {
    APX_REQUIRE(plan, APX_ERR_ARG, "null plan");
    APX_REQUIRE(set >= 0 && set < (int)plan->sets.size(), APX_ERR_ARG, "set index out of range");
    APX_REQUIRE(which == 0 || which == 1, APX_ERR_ARG, "nodal field index must be 0 or 1");
    plan->sets[set].params.nodal[which] = field_dev;
  });
}

int apx_plan_build_pattern(apx_plan* plan) {
  return guarded([&] {
    APX_REQUIRE(plan, APX_ERR_ARG, "null plan");
    std::vector<std::vector<int32_t>> cp, sl;
    plan_build_host(*plan, cp, sl);
  });
}

int apx_plan_finalize(apx_plan* plan) {
  return guarded([&] {
    APX_REQUIRE(plan, APX_ERR_ARG, "null plan");
    APX_REQUIRE(!plan->finalized, APX_ERR_STATE, "plan already finalized");
    plan_finalize(*plan);
  });
}

int apx_plan_destroy(apx_plan* plan) {
  return guarded([&] { delete plan; });
}

int apx_plan_get_info(const apx_plan* plan, apx_plan_info* info) {
  return guarded([&] {
    APX_REQUIRE(plan && info, APX_ERR_ARG, "null argument");
    *info = apx_plan_info{};
    info->n_nodes = plan->n_nodes;
    info->n_dofs = plan->n_dofs;
    info->n_free = plan->n_free;
    info->nnz = plan->nnz;
    info->dim = plan->dim;
    info->dofs_per_node = plan->d;
    info->n_sets = (int)plan->sets.size();
    info->structured = 0;
    for (size_t i = 0; i < plan->sets.size() && i < 8; ++i)
      info->n_colours[i] = plan->sets[i].colour_offsets.empty() ? 0 : (int)plan->sets[i].colour_offsets.size() - 1;
    info->build_seconds = plan->build_seconds;
    info->device_bytes = plan->device_bytes();
  });
}

int apx_plan_csr_pattern(const apx_plan* plan, int32_t* indptr_host, int32_t* indices_host) {
  return guarded([&] {
    APX_REQUIRE(plan && indptr_host && indices_host, APX_ERR_ARG, "null argument");
    APX_REQUIRE(!plan->indptr_h.empty(), APX_ERR_STATE, "pattern not built");
    std::memcpy(indptr_host, plan->indptr_h.data(), plan->indptr_h.size() * 4);
    std::memcpy(indices_host, plan->indices_h.data(), plan->indices_h.size() * 4);
  });
}

int apx_plan_free_dofs(const apx_plan* plan, int32_t* red2full_host) {
  return guarded([&] {
    APX_REQUIRE(plan && red2full_host, APX_ERR_ARG, "null argument");
    APX_REQUIRE(!plan->indptr_h.empty(), APX_ERR_STATE, "pattern not built");
    std::memcpy(red2full_host, plan->red2full_h.data(), plan->red2full_h.size() * 4);
  });
}

int apx_assemble_residual(apx_plan* plan, const double* u_dev, double* r_dev, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(u_dev && r_dev, APX_ERR_ARG, "null device pointer");
    ScopedTimer t(P, "residual_call", S(stream), residual_bytes(P));
    assemble_residual(P, u_dev, r_dev, S(stream));
  });
}

int apx_assemble_tangent(apx_plan* plan, const double* u_dev, void* vals_dev, int precision, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(u_dev && vals_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32, APX_ERR_ARG, "precision must be FP64 or FP32");
    ScopedTimer t(P, precision == APX_FP64 ? "tangent_call_f64" : "tangent_call_f32", S(stream),
                  tangent_bytes(P, precision));
    void* sell = sell_buffer(P, precision);
    assemble_tangent(P, u_dev, sell, precision, S(stream));
    sell_to_csr(P, sell, vals_dev, precision, S(stream));
  });
}

int apx_assemble_diagonal(apx_plan* plan, const double* u_dev, double* diag_dev, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(u_dev && diag_dev, APX_ERR_ARG, "null device pointer");
    assemble_diagonal(P, u_dev, diag_dev, S(stream));
  });
}

int apx_spmv(apx_plan* plan, const void* vals_dev, const void* x_dev, void* y_dev, int precision, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(vals_dev && x_dev && y_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32, APX_ERR_ARG, "precision must be FP64 or FP32");
    const int64_t ts = precision == APX_FP64 ? 8 : 4;
    void* sell = sell_buffer(P, precision);
    csr_to_sell(P, vals_dev, sell, precision, S(stream));
    ScopedTimer t(P, "spmv", S(stream), P.sell_total * (ts + 4) + (P.n_chunks + 1) * 4 + P.n_free * 2 * ts);
    spmv(P, sell, x_dev, y_dev, precision, S(stream));
  });
}

int apx_pcg(apx_plan* plan, const void* vals_dev, const void* b_dev, void* x_dev, int precision, int precond,
            double tol, double atol, int maxiter, apx_krylov_stats* stats, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(vals_dev && b_dev && x_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32, APX_ERR_ARG, "precision must be FP64 or FP32");
    APX_REQUIRE(precond == APX_PRECOND_JACOBI || precond == APX_PRECOND_NONE, APX_ERR_UNSUPPORTED,
                "apx_pcg supports Jacobi or no preconditioner");
    const size_t ts = precision == APX_FP64 ? 8 : 4;
    void* sell = sell_buffer(P, precision);
    csr_to_sell(P, vals_dev, sell, precision, S(stream));
    DevBuf<char> inv;
    inv.alloc(P.n_free * ts);
    if (precond == APX_PRECOND_JACOBI)
      jacobi_inverse(P, sell, inv.p, precision, S(stream));
    else if (precision == APX_FP64)
      fill<double>((double*)inv.p, P.n_free, 1.0, S(stream));
    else
      fill<float>((float*)inv.p, P.n_free, 1.0f, S(stream));
    apx_krylov_stats ks = pcg(P, sell, inv.p, b_dev, x_dev, precision, tol, atol, maxiter, S(stream));
    APX_HIP(hipStreamSynchronize(S(stream)));
    if (stats) *stats = ks;
  });
}

int apx_newton(apx_plan* plan, double* u_dev, const double* dirichlet_values_dev, const apx_newton_options* opts,
               apx_newton_stats* stats, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(u_dev && dirichlet_values_dev && opts, APX_ERR_ARG, "null argument");
    APX_REQUIRE(opts->precision == APX_FP64 || opts->precision == APX_FP32, APX_ERR_ARG,
                "precision must be FP64 or FP32");
    APX_REQUIRE(opts->maxiter >= 0 && opts->damping > 0.0, APX_ERR_ARG, "invalid Newton options");
    apx_newton_stats s{};
    newton(P, u_dev, dirichlet_values_dev, *opts, s, S(stream));
    if (stats) *stats = s;
  });
}

int apx_set_timing(apx_plan* plan, int enabled) {
  return guarded([&] {
    APX_REQUIRE(plan, APX_ERR_ARG, "null plan");
    plan->timing = enabled != 0;
    if (!enabled) plan->timers.clear();
  });
}

int apx_get_timing(apx_plan* plan, const char* name, double* avg_ms, int64_t* count, int64_t* bytes) {
  return guarded([&] {
    APX_REQUIRE(plan && name, APX_ERR_ARG, "null argument");
    auto it = plan->timers.find(name);
    const KernelTimer t = it == plan->timers.end() ? KernelTimer{} : it->second;
    if (avg_ms) *avg_ms = t.count ? t.total_ms / (double)t.count : 0.0;
    if (count) *count = t.count;
    if (bytes) *bytes = t.bytes;
  });
}

}  // extern "C"
