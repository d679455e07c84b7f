// Learned preconditioner: multigrid-structured U-Net on the structured node grid.
#pragma once
#include "plan.hpp"

namespace apx {

constexpr int kMaxLevels = 12;

struct NetLayer {
  int ci = 0, co = 0;      // stored channel counts
  int cop = 0, wk = 0;     // padded output channels, packed K length
  int relu = 0;
  int out_f32 = 0;
  int64_t w_off = 0, b_off = 0;  // offsets into the FP32 parameter vector
  int64_t wf_off = 0;            // offset into bf16 forward weights [cop][wk]
  int64_t wd_off = 0;            // offset into bf16 backward-data weights
  int d_cop = 0, d_wk = 0;       // dgrad conv: ci' = co, co' = ci
};

// The coarse levels of the forward pass run as one launch, one workgroup per image: a list of
// 64 -> 64 channel convolutions and grid transfers (device pointers at image 0), each step
// reading what the previous ones wrote.  Convolution grids stay within 9 x 9 nodes: a step's
// padded input and weights fit in LDS, and the matrix work of one step is small enough for
// one compute unit (at 17 x 17 nodes the per-layer launches are faster).
enum TailOp : int { kTailConv = 0, kTailRestrict = 1, kTailProlong = 2 };
struct TailStep {
  const uint16_t* x;     // convolution input; restriction: fine grid; prolongation: coarse grid
  const uint16_t* skip;  // prolongation: the skip tensor added to P(x)
  uint16_t* y;           // output
  const uint16_t* w;     // convolution weights [64][576] (bf16, forward layout)
  const uint16_t* w_next;  // convolution: weights of the next convolution step (prefetched; the last: w)
  const float* bias;     // convolution bias (64)
  int op;                // TailOp
  int relu;
  int H, W;              // convolution grid; transfers: the fine grid
  int Hc, Wc;            // transfers: the coarse grid
  // LDS element offsets of the input, skip and output buffers in the launch's activation pool
  // (padded layout: (H + 2) x (W + 2) nodes, kTailS elements per node, zero border); -1: the
  // tensor is in HBM at x / skip / y (the first step's input, the last step's skip and output)
  int xo, so, yo;
};
constexpr int kTailMaxSteps = 28;  // 2 KB of kernel arguments
constexpr int kTailMaxGrid = 9;    // largest convolution grid side in the tail
// LDS of the coarse-level launch: the weights of one 64 -> 64 convolution (64 rows of 584
// bf16) and a pool of kTailPool bf16 elements for the activations, 160 KB in all
constexpr int kTailS = 72;        // LDS elements per node of an activation buffer (64 + 8)
constexpr int kTailSW = 576 + 8;  // LDS elements per weight row
constexpr int kTailPool = (160 * 1024 - 64 * kTailSW * 2) / 2;
struct TailArgs {
  TailStep step[kTailMaxSteps];
  const uint16_t* w_first;  // weights of the first convolution step
  int n;
  int pool;  // activation pool elements used (a multiple of 8)
};

struct UNet {
  int levels = 0, c_in = 0, c_out = 0, relu = 0, batch = 1;
  int ch[kMaxLevels] = {};
  int H[kMaxLevels] = {}, W[kMaxLevels] = {};
  std::vector<NetLayer> layers;
  // layer indices
  int L_in = -1, L_b = -1, L_out = -1;
  int L_e[kMaxLevels], L_down[kMaxLevels], L_up[kMaxLevels], L_d[kMaxLevels];
  int64_t n_params = 0, n_wf = 0, n_wd = 0;
  DevBuf<float> theta, grad, adam_m, adam_v, wpart;
  DevBuf<uint16_t> wf, wd;
  // activations (bf16): x0; per level a, d, u, s, t, rd
  DevBuf<uint16_t> x0;
  DevBuf<uint16_t> a[kMaxLevels], d[kMaxLevels], u[kMaxLevels], s[kMaxLevels], t[kMaxLevels], rd[kMaxLevels];
  DevBuf<float> y;                 // output (FP32), B x H x W x c_out
  // gradient buffers (bf16)
  DevBuf<uint16_t> gy, gA[kMaxLevels], gD[kMaxLevels], gU[kMaxLevels], gS[kMaxLevels], gT[kMaxLevels],
      gR[kMaxLevels];
  DevBuf<double> loss_part;
  DevBuf<float> scale;             // per-image input scale
  bool weights_ready = false;
  // launch timing (HIP events on the launch stream) when the attached plan times its kernels
  struct Plan* timer_plan = nullptr;
  std::vector<EventPair> ev_conv0, ev_fwd;  // level-0 conv_d launches, forward passes
  // per-launch records of individual kernels (timing level 2, or the plan's filter): the
  // forward pass they belong to, algorithmic bytes and flops
  struct LaunchRec {
    EventPair e;
    size_t fwd;
    int64_t bytes, flops;
  };
  std::map<std::string, std::vector<LaunchRec>> ev_launch;
  size_t n_fwd = 0;  // forward passes recorded since the last collect
  // optional device stop flag (a Krylov solver's done flag): forward-pass kernels launched
  // after it is set return at once
  const int* stop = nullptr;
  // 1: the network is a correction to Jacobi, z = D^-1 r + D^-1/2 s net(D^-1/2 r / s)
  int jacobi_skip = 0;
  // coarse-level tail: levels >= tail_level (from its restriction to its prolongation) run as
  // one launch when fuse_tail is set; tail_level = -1: no such levels
  int tail_level = -1;
  int fuse_tail = 1;
  TailArgs tail{};

  void build(int levels, const int* channels, int c_in, int c_out, int relu, int batch, int H0, int W0);
  void build_tail();                      // tail_level and the tail's step list (after build's allocations)
  void prepare_weights(hipStream_t st);   // FP32 master -> bf16 forward / backward-data layouts
  // x0 -> y; keep_activations = false (inference): the input layer's output, read only by the
  // backward pass otherwise, is formed inside the level-0 conv_e launch and not stored
  void forward(hipStream_t st, bool keep_activations = true);
  // loss = mean_b ||y_b - t_b||^2 / ||t_b||^2 ; fills grad (FP32) via backward
  double loss_backward(const float* target, hipStream_t st);
  // backward pass from the output gradient gy (bf16) into grad
  void backward(hipStream_t st);
  void adam(float lr, float b1, float b2, float eps, int step, hipStream_t st);
  int64_t pixels(int l) const { return (int64_t)batch * H[l] * W[l]; }
  // algorithmic bytes: one convolution launch (input and output once, weights once),
  // and one forward pass (every convolution and grid transfer)
  int64_t conv_bytes(int layer, int level) const;
  int64_t forward_bytes() const;
  // move the first `count` recorded forward timings into the plan's timers (the rest are dropped)
  void collect_timing(size_t count);
};

// Network scaling from the diagonal of a SELL matrix (values in `precision`).
void net_set_diag_scaling(Plan& P, const void* vals_sell, int precision, hipStream_t st);

// Preconditioner use: z = s * net(r / s) on the plan's free DOFs (grid = nodes),
// with r and z scaled by D^-1/2 when the plan's diagonal scaling is set.
// s = sqrt(sum / n_free), the sum of squares of the network's input (r, scaled by D^-1/2 when
// the scaling is set): read from rr_src (device) when given, else computed here.
void net_input_from_reduced_img(Plan& P, UNet& N, const void* r, int r_f64, const double* g_nodal,
                                const double* rr_src, int img, hipStream_t st);
// r: the residual the input came from (FP64, reduced; used by the Jacobi skip, may be null
// when the network has none)
void net_output_to_reduced_img(Plan& P, UNet& N, const double* r, void* z, int z_f64, int img, hipStream_t st);
// target t = e / s of image img (grid layout, FP32) for training
void net_target_from_reduced_img(Plan& P, UNet& N, const double* e, float* target, int img, hipStream_t st);

// Residual-loss backward for the images already run forward (inputs from
// residuals r[b]); vals_sell: FP64 matrix in the plan's SELL layout.
// mode 0: residual loss ||r - A z||^2 / ||r||^2; mode 1: energy functional
// sA (z.Az/2 - z.r) / (s^2 n) with sA the sign of A (minimizer z = A^-1 r);
// mode 2: relative A-norm error (z-e).sA A (z-e) / e.sA A e (needs e = A^-1 r).
double unet_residual_loss_backward(Plan& P, UNet& N, const double* vals_sell, const double* const* r,
                                   const double* const* e, double w_e, int mode, hipStream_t st);

struct FcgResult {
  int iterations = 0;
  bool converged = false, stagnated = false;
  double rr0 = 0.0, rr = 0.0;
};
// Flexible CG with the network preconditioner on the plan's SELL matrix (values T).
template <class T>
FcgResult fcg_solve(Plan& P, UNet& N, const T* vals, const double* g_nodal, const double* b, double* x, double tol,
                    double atol, int maxiter, hipStream_t st);

}  // namespace apx
