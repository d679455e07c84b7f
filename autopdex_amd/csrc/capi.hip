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
    APX_REQUIRE(npe == 2 || npe == 3 || npe == 4 || npe == 8 || npe == 9, APX_ERR_UNSUPPORTED,
                "nodes_per_elem must be 2, 3, 4, 8 or 9");
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
    plan->sets[set].load_valid = false;  // the precomputed source part depends on the parameters
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
    for (auto& s : plan->sets)
      if (s.structured) info->structured |= 1;
    if (plan->finalized && tangent_by_tiles(*plan)) info->structured |= 2;
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
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32 || precision == APX_FP64_ROUNDED, APX_ERR_ARG,
                "precision must be FP64, FP32 or FP64_ROUNDED");
    const int out = precision == APX_FP64 ? APX_FP64 : APX_FP32;
    ScopedTimer t(P, precision == APX_FP64 ? "tangent_call_f64" : precision == APX_FP32 ? "tangent_call_f32"
                                                                                        : "tangent_call_f64_to_f32",
                  S(stream), tangent_bytes(P, out));
    void* sell = sell_buffer(P, out);
    if (precision == APX_FP64_ROUNDED)
      assemble_tangent_rounded(P, u_dev, (float*)sell, S(stream));
    else
      assemble_tangent(P, u_dev, sell, precision, S(stream));
    sell_to_csr(P, sell, vals_dev, out, S(stream));
  });
}

int apx_tangent_coo_entries(const apx_plan* plan, int64_t* count) {
  return guarded([&] {
    APX_REQUIRE(plan && count, APX_ERR_ARG, "null argument");
    *count = tangent_coo_entries(*plan);
  });
}

int apx_assemble_tangent_coo(apx_plan* plan, const double* u_dev, void* data_dev, int precision, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(u_dev && data_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32, APX_ERR_ARG, "precision must be FP64 or FP32");
    assemble_tangent_coo(P, u_dev, data_dev, precision, S(stream));
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
    APX_REQUIRE(enabled >= 0 && enabled <= 2, APX_ERR_ARG, "timing level must be 0, 1 or 2");
    plan->timing = enabled;
    if (!enabled) plan->timers.clear();
  });
}

int apx_set_timing_filter(apx_plan* plan, const char* names) {
  return guarded([&] {
    APX_REQUIRE(plan, APX_ERR_ARG, "null plan");
    plan->timing_filter = names ? names : "";
  });
}

int apx_get_timer(const apx_plan* plan, int index, char* name, int name_len, double* avg_ms, int64_t* count,
                  int64_t* bytes, int64_t* flops) {
  return guarded([&] {
    APX_REQUIRE(plan, APX_ERR_ARG, "null plan");
    APX_REQUIRE(index >= 0 && (size_t)index < plan->timers.size(), APX_ERR_ARG, "timer index out of range");
    auto it = plan->timers.begin();
    std::advance(it, index);
    if (name && name_len > 0) {
      std::strncpy(name, it->first.c_str(), (size_t)name_len - 1);
      name[name_len - 1] = 0;
    }
    const KernelTimer& t = it->second;
    if (avg_ms) *avg_ms = t.count ? t.total_ms / (double)t.count : 0.0;
    if (count) *count = t.count;
    if (bytes) *bytes = t.bytes;
    if (flops) *flops = t.flops;
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

// ---- learned preconditioner -------------------------------------------------------------
#include "unet.hpp"

struct apx_unet : public apx::UNet {};

static inline UNet& netref(apx_unet* n) {
  APX_REQUIRE(n != nullptr, APX_ERR_ARG, "null network");
  return *n;
}

extern "C" {

int apx_unet_create(apx_unet** out, int levels, const int* channels, int c_in, int c_out, int relu, int batch,
                    int height, int width) {
  return guarded([&] {
    APX_REQUIRE(out && channels, APX_ERR_ARG, "null argument");
    auto* n = new apx_unet();
    try {
      n->build(levels, channels, c_in, c_out, relu, batch, height, width);
    } catch (...) {
      delete n;
      throw;
    }
    *out = n;
  });
}

int apx_unet_destroy(apx_unet* net) {
  return guarded([&] { delete net; });
}

int apx_unet_num_params(const apx_unet* net, int64_t* n_params, int* n_layers) {
  return guarded([&] {
    APX_REQUIRE(net, APX_ERR_ARG, "null network");
    if (n_params) *n_params = net->n_params;
    if (n_layers) *n_layers = (int)net->layers.size();
  });
}

int apx_unet_layer(const apx_unet* net, int layer, int64_t* w_off, int64_t* b_off, int* ci, int* co, int* co_pad,
                   int* k_pad, int* relu, int* level) {
  return guarded([&] {
    APX_REQUIRE(net && layer >= 0 && layer < (int)net->layers.size(), APX_ERR_ARG, "layer index out of range");
    const NetLayer& ly = net->layers[layer];
    if (w_off) *w_off = ly.w_off;
    if (b_off) *b_off = ly.b_off;
    if (ci) *ci = ly.ci;
    if (co) *co = ly.co;
    if (co_pad) *co_pad = ly.cop;
    if (k_pad) *k_pad = ly.wk;
    if (relu) *relu = ly.relu;
    if (level) {
      int lv = 0;
      for (int l = 0; l + 1 < net->levels; ++l)
        if (layer == net->L_down[l] || layer == net->L_up[l]) lv = l + 1;
        else if (layer == net->L_e[l] || layer == net->L_d[l]) lv = l;
      if (layer == net->L_b) lv = net->levels - 1;
      *level = lv;
    }
  });
}

int apx_unet_set_params(apx_unet* net, const float* params_host) {
  return guarded([&] {
    UNet& N = netref(net);
    APX_REQUIRE(params_host, APX_ERR_ARG, "null parameters");
    APX_HIP(hipMemcpy(N.theta.p, params_host, N.n_params * sizeof(float), hipMemcpyHostToDevice));
    APX_HIP(hipMemset(N.adam_m.p, 0, N.adam_m.bytes()));
    APX_HIP(hipMemset(N.adam_v.p, 0, N.adam_v.bytes()));
    N.prepare_weights(nullptr);
    APX_HIP(hipDeviceSynchronize());
  });
}

int apx_unet_get_params(const apx_unet* net, float* params_host) {
  return guarded([&] {
    APX_REQUIRE(net && params_host, APX_ERR_ARG, "null argument");
    APX_HIP(hipDeviceSynchronize());
    APX_HIP(hipMemcpy(params_host, net->theta.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int apx_unet_get_grads(const apx_unet* net, float* grads_host) {
  return guarded([&] {
    APX_REQUIRE(net && grads_host, APX_ERR_ARG, "null argument");
    APX_HIP(hipDeviceSynchronize());
    APX_HIP(hipMemcpy(grads_host, net->grad.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int apx_unet_buffers(apx_unet* net, void** x, void** y, void** scale) {
  return guarded([&] {
    UNet& N = netref(net);
    if (x) *x = N.x0.p;
    if (y) *y = N.y.p;
    if (scale) *scale = N.scale.p;
  });
}

int apx_unet_forward(apx_unet* net, void* stream) {
  return guarded([&] { netref(net).forward(S(stream)); });
}

int apx_unet_loss_backward(apx_unet* net, const float* target_dev, double* loss, void* stream) {
  return guarded([&] {
    APX_REQUIRE(target_dev, APX_ERR_ARG, "null target");
    const double l = netref(net).loss_backward(target_dev, S(stream));
    if (loss) *loss = l;
  });
}

int apx_unet_adam(apx_unet* net, float lr, float beta1, float beta2, float eps, int step, void* stream) {
  return guarded([&] {
    APX_REQUIRE(step >= 1, APX_ERR_ARG, "Adam step counts from 1");
    netref(net).adam(lr, beta1, beta2, eps, step, S(stream));
  });
}

int apx_unet_load_pair(apx_unet* net, apx_plan* plan, const double* r_dev, const double* g_nodal_dev,
                       const double* e_dev, float* target_dev, int image, void* stream) {
  return guarded([&] {
    UNet& N = netref(net);
    Plan& P = ready(plan);
    APX_REQUIRE(r_dev, APX_ERR_ARG, "null residual");
    net_input_from_reduced_img(P, N, r_dev, 1, g_nodal_dev, nullptr, image, S(stream));
    if (e_dev) {
      APX_REQUIRE(target_dev, APX_ERR_ARG, "null target buffer");
      net_target_from_reduced_img(P, N, e_dev, target_dev, image, S(stream));
    }
  });
}

int apx_unet_apply(apx_unet* net, apx_plan* plan, const double* r_dev, const double* g_nodal_dev, double* z_dev,
                   void* stream) {
  return guarded([&] {
    UNet& N = netref(net);
    Plan& P = ready(plan);
    APX_REQUIRE(r_dev && z_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(N.batch == 1, APX_ERR_ARG, "apply needs a batch-1 network");
    net_input_from_reduced_img(P, N, r_dev, 1, g_nodal_dev, nullptr, 0, S(stream));
    N.forward(S(stream), false);
    net_output_to_reduced_img(P, N, r_dev, z_dev, 1, 0, S(stream));
  });
}

int apx_unet_set_jacobi_skip(apx_unet* net, int enable) {
  return guarded([&] {
    APX_REQUIRE(net, APX_ERR_ARG, "null network");
    net->jacobi_skip = enable ? 1 : 0;
  });
}

int apx_unet_set_fused_tail(apx_unet* net, int enable, int* tail_level) {
  return guarded([&] {
    APX_REQUIRE(net, APX_ERR_ARG, "null network");
    net->fuse_tail = enable ? 1 : 0;
    if (tail_level) *tail_level = net->tail_level;
  });
}

int apx_plan_attach_unet(apx_plan* plan, apx_unet* net, const double* g_nodal_dev) {
  return guarded([&] {
    Plan& P = ready(plan);
    if (!net) {
      P.net = nullptr;
      P.net_g = nullptr;
      return;
    }
    UNet& N = *net;
    APX_REQUIRE(N.batch == 1, APX_ERR_ARG, "the preconditioner network must have batch 1");
    APX_REQUIRE((int64_t)N.H[0] * N.W[0] == P.n_nodes, APX_ERR_ARG, "network grid does not match the plan's nodes");
    APX_REQUIRE(N.c_out == P.d, APX_ERR_ARG, "network output channels must equal the DOFs per node");
    APX_REQUIRE(N.c_in == P.d || (N.c_in == P.d + 1 && g_nodal_dev), APX_ERR_ARG,
                "network input channels: DOFs per node, plus one when a nodal field is given");
    // the grid convention: node (i, j) = i * width + j, Q1 connectivity of mesher.structured_mesh
    APX_REQUIRE(P.dim == 2 && !P.sets.empty() && P.sets[0].npe == 4, APX_ERR_UNSUPPORTED,
                "the learned preconditioner needs a structured 2D quad mesh");
    const ElemSet& s0 = P.sets[0];
    const int64_t nx = N.H[0] - 1, ny = N.W[0] - 1;
    APX_REQUIRE(s0.n_elems == nx * ny, APX_ERR_ARG, "element count does not match the network grid");
    for (int64_t e = 0; e < s0.n_elems; ++e) {
      const int64_t i = e / ny, j = e % ny, b = i * (ny + 1) + j;
      const int64_t* c = &s0.conn_host[e * 4];
      APX_REQUIRE(c[0] == b && c[1] == b + 1 && c[2] == b + ny + 2 && c[3] == b + ny + 1, APX_ERR_ARG,
                  "connectivity is not the structured grid of the network");
    }
    P.net = &N;
    P.net_g = N.c_in > P.d ? g_nodal_dev : nullptr;
  });
}

int apx_fcg(apx_plan* plan, const void* vals_dev, const double* b_dev, double* x_dev, int precision, double tol,
            double atol, int maxiter, apx_krylov_stats* stats, void* stream) {
  return guarded([&] {
    Plan& P = ready(plan);
    APX_REQUIRE(P.net, APX_ERR_STATE, "no network attached (apx_plan_attach_unet)");
    APX_REQUIRE(vals_dev && b_dev && x_dev, APX_ERR_ARG, "null device pointer");
    APX_REQUIRE(precision == APX_FP64 || precision == APX_FP32, APX_ERR_ARG, "precision must be APX_FP64 or APX_FP32");
    void* sell = sell_buffer(P, precision);
    csr_to_sell(P, vals_dev, sell, precision, S(stream));
    FcgResult r = precision == APX_FP64
                      ? fcg_solve<double>(P, *P.net, (const double*)sell, P.net_g, b_dev, x_dev, tol, atol, maxiter,
                                          S(stream))
                      : fcg_solve<float>(P, *P.net, (const float*)sell, P.net_g, b_dev, x_dev, tol, atol, maxiter,
                                         S(stream));
    if (stats) {
      stats->iterations = r.iterations;
      stats->converged = r.converged ? 1 : 0;
      stats->initial_residual = sqrt(r.rr0);
      stats->final_residual = sqrt(r.rr);
      stats->seconds = 0.0;
    }
  });
}

}  // extern "C"

extern "C" int apx_unet_train_step(apx_unet* net, apx_plan* plan, const void* vals_dev, const double* const* r_list,
                                   const double* const* g_list, const double* const* e_list, int loss_kind,
                                   double w_e, double* loss, void* stream) {
  return guarded([&] {
    UNet& N = netref(net);
    Plan& P = ready(plan);
    APX_REQUIRE(vals_dev && r_list, APX_ERR_ARG, "null argument");
    void* sell = sell_buffer(P, APX_FP64);
    csr_to_sell(P, vals_dev, sell, APX_FP64, S(stream));
    // the input scaling D^-1/2 r / s and the output map z = D^-1/2 s y use this matrix's diagonal
    net_set_diag_scaling(P, sell, APX_FP64, S(stream));
    for (int b = 0; b < N.batch; ++b) {
      APX_REQUIRE(r_list[b], APX_ERR_ARG, "null residual in batch");
      net_input_from_reduced_img(P, N, r_list[b], 1, g_list ? g_list[b] : nullptr, nullptr, b, S(stream));
    }
    N.forward(S(stream));
    APX_REQUIRE(loss_kind >= 0 && loss_kind <= 2, APX_ERR_ARG, "loss kind: 0 residual, 1 energy, 2 A-norm error");
    APX_REQUIRE(loss_kind != 2 || e_list != nullptr, APX_ERR_ARG, "the A-norm loss needs e_list");
    const double l =
        unet_residual_loss_backward(P, N, (const double*)sell, r_list, e_list, w_e, loss_kind, S(stream));
    if (loss) *loss = l;
  });
}
