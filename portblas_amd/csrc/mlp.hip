// The learned tile-config selector: a small MLP trained on the GPU by these
// kernels on timings the autotuner collects, and evaluated to pick a GEMM
// configuration for shapes it has not seen.  (portBLAS picks configs with the
// hand-written arith_ratio buckets of src/interface/blas3/backend/
// amd_gpu.hpp:52-288; its exhaustive tuner is tools/auto_tuner.)
//
// Network: layer l maps width d_l -> d_{l+1}, Z = X W^T + b, ReLU on hidden
// layers, linear output.  Loss: masked mean squared error.  Row-major
// [batch x width] activations, W_l row-major [d_{l+1} x d_l].
//
// Kernels (all deterministic: every sum runs in a fixed order, no atomics):
//   linear_fwd   one thread per output element, the layer's weights in LDS;
//   loss_grad    masked MSE value and dL/dY of the output layer;
//   grad_w       dW = dY^T X and db = sum dY, one thread per weight, looping
//                over the batch;
//   grad_x       dX = (dY W) * relu'(X) for the layer below;
//   adam         one Adam step per parameter.
// BF16 inference rounds weights and each layer's input to BF16 and
// accumulates in FP32.  The per-call dispatch in the library evaluates the
// same BF16 network on the host with the same arithmetic and summation order
// (a device round trip per GEMM call would cost more than it saves).
#include "lowp.hpp"
#include "pb_internal.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kMaxWidth = 128;

__global__ __launch_bounds__(256) void linear_fwd(const float* __restrict__ X,
                                                  const float* __restrict__ W,
                                                  const float* __restrict__ b, float* Y,
                                                  int64_t batch, int din, int dout, int relu) {
  __shared__ float Ws[kMaxWidth * kMaxWidth / 2];  // up to 8192 weights
  const int nw = din * dout;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) Ws[i] = W[i];
  __syncthreads();
  const int64_t total = batch * dout;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = e / dout;
    const int o = int(e % dout);
    const float* x = X + r * din;
    const float* w = Ws + o * din;
    float s = b[o];
    for (int k = 0; k < din; ++k) s = fmaf(x[k], w[k], s);
    Y[e] = (relu && s < 0.f) ? 0.f : s;
  }
}

// BF16 inference layer: W given as BF16 bits, the input rounded to BF16.
__global__ __launch_bounds__(256) void linear_fwd_bf16(const float* __restrict__ X,
                                                       const uint16_t* __restrict__ W,
                                                       const float* __restrict__ b, float* Y,
                                                       int64_t batch, int din, int dout,
                                                       int relu) {
  __shared__ float Ws[kMaxWidth * kMaxWidth / 2];
  const int nw = din * dout;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) Ws[i] = pblowp::bf16_to_f32(W[i]);
  __syncthreads();
  const int64_t total = batch * dout;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = e / dout;
    const int o = int(e % dout);
    const float* x = X + r * din;
    const float* w = Ws + o * din;
    float s = b[o];
    for (int k = 0; k < din; ++k) s = fmaf(pblowp::bf16_to_f32(pblowp::f32_to_bf16(x[k])), w[k], s);
    Y[e] = (relu && s < 0.f) ? 0.f : s;
  }
}

// Masked MSE: loss = sum m (y - t)^2 / max(1, sum m); G = dL/dY.
// One workgroup; fixed-order partial sums.
__global__ __launch_bounds__(256) void loss_grad(const float* __restrict__ Y,
                                                 const float* __restrict__ T,
                                                 const float* __restrict__ M, float* G,
                                                 float* loss, int64_t n) {
  __shared__ double sl[256], sm[256];
  double l = 0.0, m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double w = M ? M[i] : 1.0;
    const double d = double(Y[i]) - double(T[i]);
    l += w * d * d;
    m += w;
  }
  sl[threadIdx.x] = l;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      sl[threadIdx.x] += sl[threadIdx.x + s];
      sm[threadIdx.x] += sm[threadIdx.x + s];
    }
    __syncthreads();
  }
  const double denom = sm[0] > 1.0 ? sm[0] : 1.0;
  if (threadIdx.x == 0) loss[0] = float(sl[0] / denom);
  const float scale = float(2.0 / denom);
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const float w = M ? M[i] : 1.f;
    G[i] = scale * w * (Y[i] - T[i]);
  }
}

// dW[o][k] = sum_r G[r][o] X[r][k]; db[o] = sum_r G[r][o].  One thread per
// weight (plus one per bias), batch walked in order.
__global__ __launch_bounds__(256) void grad_w(const float* __restrict__ G,
                                              const float* __restrict__ X, float* dW, float* db,
                                              int64_t batch, int din, int dout) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < din * dout) {
    const int o = e / din, k = e % din;
    float s = 0.f;
    for (int64_t r = 0; r < batch; ++r) s = fmaf(G[r * dout + o], X[r * din + k], s);
    dW[e] = s;
  } else if (e < din * dout + dout) {
    const int o = e - din * dout;
    float s = 0.f;
    for (int64_t r = 0; r < batch; ++r) s += G[r * dout + o];
    db[o] = s;
  }
}

// dX[r][k] = (sum_o G[r][o] W[o][k]) * (X[r][k] > 0): the gradient through the
// ReLU that produced X.
__global__ __launch_bounds__(256) void grad_x(const float* __restrict__ G,
                                              const float* __restrict__ W,
                                              const float* __restrict__ X, float* dX,
                                              int64_t batch, int din, int dout) {
  const int64_t total = batch * din;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = e / din;
    const int k = int(e % din);
    float s = 0.f;
    for (int o = 0; o < dout; ++o) s = fmaf(G[r * dout + o], W[o * din + k], s);
    dX[e] = X[e] > 0.f ? s : 0.f;
  }
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

unsigned grid_for(pb_handle_t h, int64_t n) {
  int64_t g = pb::ceil_div(n, 256);
  const int64_t cap = 8 * int64_t(h->num_cus);
  if (g > cap) g = cap;
  return unsigned(g < 1 ? 1 : g);
}

pb_status_t check_sizes(int n_layers, const int* sizes) {
  PB_REQUIRE(n_layers >= 1 && n_layers <= 16 && sizes != nullptr, "bad layer count");
  for (int l = 0; l <= n_layers; ++l)
    PB_REQUIRE(sizes[l] >= 1 && sizes[l] <= kMaxWidth, "layer width %d out of [1, %d]", sizes[l],
               kMaxWidth);
  for (int l = 0; l < n_layers; ++l)
    PB_REQUIRE(sizes[l] * sizes[l + 1] <= kMaxWidth * kMaxWidth / 2,
               "layer %d has more than %d weights", l, kMaxWidth * kMaxWidth / 2);
  return PB_STATUS_SUCCESS;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-side evaluation of the installed selector (BF16 arithmetic, as above).
// ---------------------------------------------------------------------------
namespace pbgemm {

int default_config(pb_handle_t h, int64_t M, int64_t N, int64_t K, int64_t batch);

// Features of one GEMM shape, in the order the selector was trained on.
void selector_features(char ta, char tb, int64_t M, int64_t N, int64_t K, int64_t batch,
                       float f[8]) {
  auto lg = [](int64_t v) { return float(std::log2(double(v > 0 ? v : 1))); };
  f[0] = lg(M);
  f[1] = lg(N);
  f[2] = lg(K);
  f[3] = (ta == 't' || ta == 'T' || ta == 'c' || ta == 'C') ? 1.f : 0.f;
  f[4] = (tb == 't' || tb == 'T' || tb == 'c' || tb == 'C') ? 1.f : 0.f;
  f[5] = 0.f;  // element type code: BF16
  f[6] = lg(batch);
  f[7] = 0.f;  // leading-dimension padding (log2 of lda / rows); packed = 0
}

// The configuration an installed network picks for features f (its
// argmin of predicted relative runtime), or -1 when none is installed.
int selector_eval(const pb_selector_state& s, const float f[8]) {
  if (s.n_layers == 0) return -1;
  uint32_t hsh = 2166136261u;  // FNV-1a over the features' bytes
  const unsigned char* fb = reinterpret_cast<const unsigned char*>(f);
  for (int i = 0; i < 32; ++i) hsh = (hsh ^ fb[i]) * 16777619u;
  pb_selector_state::Decision& slot = s.cache[hsh % pb_selector_state::kCacheSlots];
  if (slot.valid && std::memcmp(slot.f, f, sizeof(slot.f)) == 0) return slot.choice;
  std::vector<float> x(size_t(s.sizes[0]), 0.f), y;
  for (int i = 0; i < s.sizes[0] && i < 8; ++i) x[i] = (f[i] - s.mean[i]) / s.stdv[i];
  for (int l = 0; l < s.n_layers; ++l) {
    const int din = s.sizes[l], dout = s.sizes[l + 1];
    // The layer's input rounded to BF16 once; the weights are stored rounded.
    // A product of two BF16 values has at most 16 significant bits, so it is
    // exact in FP32 and `acc + x * w` rounds exactly as a fused multiply-add
    // would (the GPU inference's arithmetic): the host needs no fma, which
    // without hardware FMA in its baseline instruction set is a library call
    // (about 30 us per evaluation).
    for (float& v : x) v = pblowp::bf16_to_f32(pblowp::f32_to_bf16(v));
    y.assign(size_t(dout), 0.f);
    for (int o = 0; o < dout; ++o) {
      const float* wr = s.w[l].data() + size_t(o) * din;
      float acc = s.b[l][o];
      for (int k = 0; k < din; ++k) acc += x[k] * wr[k];
      y[o] = (l + 1 < s.n_layers && acc < 0.f) ? 0.f : acc;
    }
    x.swap(y);
  }
  int best = 0;
  for (int i = 1; i < int(x.size()); ++i)
    if (x[i] < x[best]) best = i;
  slot.valid = true;
  std::memcpy(slot.f, f, sizeof(slot.f));
  slot.choice = best;
  return best;
}

int selector_choose(pb_handle_t h, char ta, char tb, int64_t M, int64_t N, int64_t K,
                    int64_t batch) {
  float f[8];
  selector_features(ta, tb, M, N, K, batch, f);
  return selector_eval(h->selector, f);
}

// Validates and converts a network for a handle (weights rounded to BF16).
pb_status_t selector_state(int n_layers, const int* sizes, const float* const* weights,
                           const float* const* biases, const float* feat_mean, const float* feat_std,
                           int outputs, pb_selector_state& s);

}  // namespace pbgemm

extern "C" {

pb_status_t pb_tune_set_selector(pb_handle_t h, int n_layers, const int* sizes,
                                 const float* const* weights, const float* const* biases,
                                 const float* feat_mean, const float* feat_std) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  pb_selector_state s;
  PB_TRY(pbgemm::selector_state(n_layers, sizes, weights, biases, feat_mean, feat_std, 0, s));
  h->selector = std::move(s);
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_tune_set_plain_selector(pb_handle_t h, int n_layers, const int* sizes,
                                       const float* const* weights, const float* const* biases,
                                       const float* feat_mean, const float* feat_std) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  pb_selector_state s;
  PB_TRY(pbgemm::selector_state(n_layers, sizes, weights, biases, feat_mean, feat_std,
                                pb_tune_num_plain_configs(), s));
  h->plain_selector = std::move(s);
  return PB_STATUS_SUCCESS;
}

}  // extern "C"

namespace pbgemm {

pb_status_t selector_state(int n_layers, const int* sizes, const float* const* weights,
                           const float* const* biases, const float* feat_mean, const float* feat_std,
                           int outputs, pb_selector_state& s) {
  s = pb_selector_state();
  if (n_layers == 0) return PB_STATUS_SUCCESS;
  PB_TRY(check_sizes(n_layers, sizes));
  PB_REQUIRE(sizes[0] == 8, "the selector takes 8 features (got %d)", sizes[0]);
  PB_REQUIRE(outputs == 0 || sizes[n_layers] == outputs, "the selector must have %d outputs (got %d)", outputs,
             sizes[n_layers]);
  PB_REQUIRE(weights && biases && feat_mean && feat_std, "null selector array");
  s.n_layers = n_layers;
  s.sizes.assign(sizes, sizes + n_layers + 1);
  for (int l = 0; l < n_layers; ++l) {
    const size_t nw = size_t(sizes[l]) * sizes[l + 1];
    std::vector<float> w(nw);
    for (size_t i = 0; i < nw; ++i) w[i] = pblowp::bf16_to_f32(pblowp::f32_to_bf16(weights[l][i]));
    s.w.push_back(std::move(w));
    s.b.emplace_back(biases[l], biases[l] + sizes[l + 1]);
  }
  s.mean.assign(feat_mean, feat_mean + 8);
  s.stdv.assign(feat_std, feat_std + 8);
  for (float& v : s.stdv)
    if (!(v > 0.f)) v = 1.f;
  return PB_STATUS_SUCCESS;
}

}  // namespace pbgemm

extern "C" {

int pb_tune_select(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                   int64_t batch) {
  if (!h) return -1;
  const int id = pbgemm::selector_choose(h, ta, tb, m, n, k, batch);
  if (id >= 0) return id;
  return pbgemm::default_config(h, m, n, k, batch);
}

pb_status_t pb_mlp_forward(pb_handle_t h, int n_layers, const int* sizes, const float* const* w,
                           const float* const* b, float* const* acts, int64_t batch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_TRY(check_sizes(n_layers, sizes));
  PB_REQUIRE(batch > 0, "empty batch");
  for (int l = 0; l < n_layers; ++l) {
    hipLaunchKernelGGL(linear_fwd, dim3(grid_for(h, batch * sizes[l + 1])), dim3(256), 0,
                       h->stream, acts[l], w[l], b[l], acts[l + 1], batch, sizes[l], sizes[l + 1],
                       int(l + 1 < n_layers));
    PB_CHECK_LAUNCH();
  }
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_mlp_forward_bf16(pb_handle_t h, int n_layers, const int* sizes,
                                const uint16_t* const* w, const float* const* b, const float* x,
                                float* y, int64_t batch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_TRY(check_sizes(n_layers, sizes));
  PB_REQUIRE(batch > 0, "empty batch");
  int maxw = 0;
  for (int l = 1; l < n_layers; ++l) maxw = sizes[l] > maxw ? sizes[l] : maxw;
  float* bufs[2] = {nullptr, nullptr};
  if (n_layers > 1) {
    void* ws = nullptr;
    PB_TRY(pb::workspace(h, 2 * sizeof(float) * size_t(batch) * maxw + 512, &ws));
    bufs[0] = static_cast<float*>(ws);
    bufs[1] = bufs[0] + pb::align_up(size_t(batch) * maxw, 64);
  }
  const float* in = x;
  for (int l = 0; l < n_layers; ++l) {
    float* out = (l + 1 == n_layers) ? y : bufs[l & 1];
    hipLaunchKernelGGL(linear_fwd_bf16, dim3(grid_for(h, batch * sizes[l + 1])), dim3(256), 0,
                       h->stream, in, w[l], b[l], out, batch, sizes[l], sizes[l + 1],
                       int(l + 1 < n_layers));
    PB_CHECK_LAUNCH();
    in = out;
  }
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_mlp_backward(pb_handle_t h, int n_layers, const int* sizes, const float* const* w,
                            const float* const* acts, const float* target, const float* mask,
                            float* const* dw, float* const* db, float* loss, float* scratch,
                            int64_t batch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_TRY(check_sizes(n_layers, sizes));
  PB_REQUIRE(batch > 0 && scratch != nullptr && loss != nullptr, "bad backward arguments");
  int maxw = 0;
  for (int l = 0; l <= n_layers; ++l) maxw = sizes[l] > maxw ? sizes[l] : maxw;
  float* g = scratch;                       // gradient w.r.t. the current layer output
  float* gn = scratch + size_t(batch) * maxw;  // gradient for the layer below
  const int dl = sizes[n_layers];
  hipLaunchKernelGGL(loss_grad, dim3(1), dim3(256), 0, h->stream, acts[n_layers], target, mask, g,
                     loss, batch * dl);
  PB_CHECK_LAUNCH();
  for (int l = n_layers - 1; l >= 0; --l) {
    const int din = sizes[l], dout = sizes[l + 1];
    hipLaunchKernelGGL(grad_w, dim3(unsigned(pb::ceil_div(din * dout + dout, 256))), dim3(256), 0,
                       h->stream, g, acts[l], dw[l], db[l], batch, din, dout);
    if (l > 0)
      hipLaunchKernelGGL(grad_x, dim3(grid_for(h, batch * din)), dim3(256), 0, h->stream, g, w[l],
                         acts[l], gn, batch, din, dout);
    PB_CHECK_LAUNCH();
    float* t = g;
    g = gn;
    gn = t;
  }
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_mlp_adam(pb_handle_t h, float* param, const float* grad, float* m, float* v,
                        int64_t count, float lr, float beta1, float beta2, float eps, int step) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(count >= 0 && step >= 1, "bad Adam arguments");
  if (count == 0) return PB_STATUS_SUCCESS;
  const float c1 = 1.f - std::pow(beta1, float(step));
  const float c2 = 1.f - std::pow(beta2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(h, count)), dim3(256), 0, h->stream, param, grad,
                     m, v, count, lr, beta1, beta2, eps, c1, c2);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // extern "C"
