// Learned scheduler: the 12 -> 32 -> 32 -> 9 MLP that predicts, for each
// batched-step action, the log of the time to finish the wave. Forward,
// backward, gradient reduction and Adam are hand-written kernels; the batch
// kernel evaluates the trained network (mlp.hpp) at every step decision.
//
// There is no counterpart in the reference, whose visits always expand both
// frontiers top-down (bernaschi/.../make_stcon.c:960-1200); the network only
// reorders that work, so verdicts and distances are unchanged.
//
// Parameter layout (FC_SCHED_PARAMS floats):
//   W1[32][12] b1[32] W2[32][32] b2[32] W3[9][32] b3[9].
// Loss over n rows with row weights w_r (all 1 unless given):
//   L = 0.5 * sum_r w_r * sum_a (y[r][a] - target[r][a])^2 / sum_r w_r.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "mlp.hpp"

namespace fcd {

constexpr int kP = kMlpParams;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kFeatures;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kActions * kHidden;
static_assert(kOffB3 + kActions == kP, "parameter layout");

// One thread per row: forward pass and, when T is given, the backward pass
// down to the per-row layer gradients. Stores the activations and the
// gradients the parameter reduction needs. Wr (nullable: all 1) holds the
// row weights and inv_wsum is 1 / sum_r w_r.
__global__ __launch_bounds__(256) void mlp_rows_kernel(const float* __restrict__ w, const float* __restrict__ X,
                                                       const float* __restrict__ T, const float* __restrict__ Wr,
                                                       float inv_wsum, int n, float* __restrict__ Y,
                                                       float* __restrict__ H1, float* __restrict__ H2,
                                                       float* __restrict__ D1, float* __restrict__ D2,
                                                       float* __restrict__ DY, float* __restrict__ err) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= n) return;
  const float* W1 = w + kOffW1;
  const float* b1 = w + kOffB1;
  const float* W2 = w + kOffW2;
  const float* b2 = w + kOffB2;
  const float* W3 = w + kOffW3;
  const float* b3 = w + kOffB3;
  float x[kFeatures], h1[kHidden], h2[kHidden], y[kActions];
  for (int i = 0; i < kFeatures; ++i) x[i] = X[(size_t)r * kFeatures + i];
  for (int j = 0; j < kHidden; ++j) {
    float s = b1[j];
    for (int i = 0; i < kFeatures; ++i) s = __fmaf_rn(W1[j * kFeatures + i], x[i], s);
    h1[j] = s > 0.f ? s : 0.f;
  }
  for (int j = 0; j < kHidden; ++j) {
    float s = b2[j];
    for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W2[j * kHidden + i], h1[i], s);
    h2[j] = s > 0.f ? s : 0.f;
  }
  for (int a = 0; a < kActions; ++a) {
    float s = b3[a];
    for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W3[a * kHidden + i], h2[i], s);
    y[a] = s;
    Y[(size_t)r * kActions + a] = s;
  }
  if (!T) return;
  const float wr = Wr ? Wr[r] : 1.f;
  const float scale = wr * inv_wsum;
  float dy[kActions], e = 0.f;
  for (int a = 0; a < kActions; ++a) {
    const float d = y[a] - T[(size_t)r * kActions + a];
    e = __fmaf_rn(d, d, e);
    dy[a] = d * scale;
    DY[(size_t)r * kActions + a] = dy[a];
  }
  err[r] = 0.5f * wr * e;
  float d2[kHidden];
  for (int j = 0; j < kHidden; ++j) {
    float s = 0.f;
    for (int a = 0; a < kActions; ++a) s = __fmaf_rn(W3[a * kHidden + j], dy[a], s);
    d2[j] = h2[j] > 0.f ? s : 0.f;
    D2[(size_t)r * kHidden + j] = d2[j];
    H2[(size_t)r * kHidden + j] = h2[j];
  }
  for (int j = 0; j < kHidden; ++j) {
    float s = 0.f;
    for (int k = 0; k < kHidden; ++k) s = __fmaf_rn(W2[k * kHidden + j], d2[k], s);
    D1[(size_t)r * kHidden + j] = h1[j] > 0.f ? s : 0.f;
    H1[(size_t)r * kHidden + j] = h1[j];
  }
}

// One thread per parameter: its gradient summed over the rows in row order
// (deterministic).
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float* __restrict__ X, const float* __restrict__ H1,
                                                       const float* __restrict__ H2, const float* __restrict__ D1,
                                                       const float* __restrict__ D2, const float* __restrict__ DY,
                                                       int n, float* __restrict__ grad) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= kP) return;
  const float* A;  // upstream gradient column
  const float* B;  // input activation column (null: bias)
  int sa, sb, ca, cb = 0;
  if (p < kOffB1) {
    ca = p / kFeatures;
    cb = p % kFeatures;
    A = D1, sa = kHidden, B = X, sb = kFeatures;
  } else if (p < kOffW2) {
    ca = p - kOffB1;
    A = D1, sa = kHidden, B = nullptr, sb = 0;
  } else if (p < kOffB2) {
    ca = (p - kOffW2) / kHidden;
    cb = (p - kOffW2) % kHidden;
    A = D2, sa = kHidden, B = H1, sb = kHidden;
  } else if (p < kOffW3) {
    ca = p - kOffB2;
    A = D2, sa = kHidden, B = nullptr, sb = 0;
  } else if (p < kOffB3) {
    ca = (p - kOffW3) / kHidden;
    cb = (p - kOffW3) % kHidden;
    A = DY, sa = kActions, B = H2, sb = kHidden;
  } else {
    ca = p - kOffB3;
    A = DY, sa = kActions, B = nullptr, sb = 0;
  }
  float s = 0.f;
  if (B) {
    for (int r = 0; r < n; ++r) s = __fmaf_rn(A[(size_t)r * sa + ca], B[(size_t)r * sb + cb], s);
  } else {
    for (int r = 0; r < n; ++r) s += A[(size_t)r * sa + ca];
  }
  grad[p] = s;
}

// Deterministic sum of the per-row (weighted) errors times inv_wsum (one block).
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* __restrict__ err, int n, float inv_wsum,
                                                       float* __restrict__ out) {
  __shared__ float part[256];
  float s = 0.f;
  for (int r = (int)threadIdx.x; r < n; r += 256) s += err[r];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] * inv_wsum;
}

// Adam (Kingma & Ba 2015) with bias correction; t is the 1-based step.
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                       float* __restrict__ m, float* __restrict__ v, float lr,
                                                       float beta1, float beta2, float eps, float c1, float c2) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= kP) return;
  const float gi = g[p];
  const float mi = beta1 * m[p] + (1.f - beta1) * gi;
  const float vi = beta2 * v[p] + (1.f - beta2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  const float mh = mi / c1, vh = vi / c2;
  w[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Device buffers of one training / gradient call.
struct MlpBuffers {
  float *X = nullptr, *T = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr, *D1 = nullptr, *D2 = nullptr,
        *DY = nullptr, *err = nullptr, *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr,
        *wr = nullptr;
  float inv_wsum = 1.f;  // 1 / sum of the row weights
  ~MlpBuffers() {
    for (float* p : {X, T, Y, H1, H2, D1, D2, DY, err, w, g, m, v, loss, wr}) (void)hipFree(p);
  }
  // Uploads row weights (null: every row weighs 1) and sets inv_wsum.
  fc_status set_weights(const float* weight, int n, hipStream_t s) {
    double sum = 0.0;
    if (weight) {
      for (int r = 0; r < n; ++r) {
        if (!(weight[r] >= 0.f) || !std::isfinite(weight[r]))
          return fc::fail(FC_ERR_INVALID, "scheduler row weight %d is negative or not finite", r);
        sum += weight[r];
      }
      if (!(sum > 0.0)) return fc::fail(FC_ERR_INVALID, "scheduler row weights sum to zero");
      if (hipMalloc((void**)&wr, sizeof(float) * (size_t)n) != hipSuccess)
        return fc::fail(FC_ERR_NOMEM, "scheduler buffers: device allocation failed");
      FC_HIP(hipMemcpyAsync(wr, weight, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, s));
    } else {
      sum = (double)n;
    }
    inv_wsum = (float)(1.0 / sum);
    return FC_OK;
  }
  fc_status alloc(int n, bool train) {
    auto A = [](float** p, size_t cnt) { return hipMalloc((void**)p, sizeof(float) * (cnt ? cnt : 1)); };
    const size_t N = (size_t)n;
    if (A(&X, N * kFeatures) != hipSuccess || A(&T, N * kActions) != hipSuccess || A(&Y, N * kActions) != hipSuccess ||
        A(&H1, N * kHidden) != hipSuccess || A(&H2, N * kHidden) != hipSuccess || A(&D1, N * kHidden) != hipSuccess ||
        A(&D2, N * kHidden) != hipSuccess || A(&DY, N * kActions) != hipSuccess || A(&err, N) != hipSuccess ||
        A(&w, kP) != hipSuccess || A(&g, kP) != hipSuccess || A(&loss, 1) != hipSuccess)
      return fc::fail(FC_ERR_NOMEM, "scheduler buffers: device allocation failed");
    if (train && (A(&m, kP) != hipSuccess || A(&v, kP) != hipSuccess))
      return fc::fail(FC_ERR_NOMEM, "scheduler buffers: device allocation failed");
    return FC_OK;
  }
};

static hipError_t launch_rows(const MlpBuffers& b, int n, bool backward, hipStream_t s) {
  const int blocks = (n + 255) / 256;
  hipLaunchKernelGGL(mlp_rows_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, s, b.w, b.X, backward ? b.T : nullptr,
                     b.wr, b.inv_wsum, n, b.Y, b.H1, b.H2, b.D1, b.D2, b.DY, b.err);
  return hipGetLastError();
}

static hipError_t launch_grad(const MlpBuffers& b, int n, hipStream_t s) {
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((kP + 255) / 256), dim3(256), 0, s, b.X, b.H1, b.H2, b.D1, b.D2, b.DY, n,
                     b.g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, s, b.err, n, b.inv_wsum, b.loss);
  return hipGetLastError();
}

// Seeded Xavier-uniform initialisation (weights), zero biases.
static void init_params(uint64_t seed, std::vector<float>& w) {
  w.assign(kP, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double lim = std::sqrt(6.0 / (double)(fan_in + fan_out));
    for (int i = 0; i < fan_out * fan_in; ++i) w[off + i] = (float)((2.0 * uni() - 1.0) * lim);
  };
  fill(kOffW1, kHidden, kFeatures);
  fill(kOffW2, kHidden, kHidden);
  fill(kOffW3, kActions, kHidden);
}

static fc_status set_params_device(fc_graph* g, const std::vector<float>& w) {
  g->h_mlp = w;
  // the batch kernel reads a second copy in its wave-friendly layout
  std::vector<float> wt((size_t)kP);
  for (int i = 0; i < kP; ++i) wt[(size_t)mlp_t_index(i)] = w[(size_t)i];
  FC_HIP(hipMemcpyAsync(g->d_mlp, g->h_mlp.data(), sizeof(float) * (size_t)kP, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(g->d_mlp + kP, wt.data(), sizeof(float) * (size_t)kP, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  g->mlp_set = true;
  return FC_OK;
}

}  // namespace fcd

using namespace fcd;

// Parameters live on the device (read by the batch kernel's learned policy)
// with a host copy. Passing params == NULL and n == 0 removes them, which
// makes the learned policy unavailable again.
extern "C" fc_status fc_sched_set_params(fc_graph* g, const float* params, int32_t n) {
  fc::clear_error();
  if (!g) return fc::fail(FC_ERR_INVALID, "fc_sched_set_params: null graph");
  if (!params && n == 0) {
    g->mlp_set = false;
    return FC_OK;
  }
  if (!params || n != kP) return fc::fail(FC_ERR_INVALID, "fc_sched_set_params: expected %d parameters", kP);
  for (int32_t i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return fc::fail(FC_ERR_INVALID, "fc_sched_set_params: parameter %d is not finite", i);
  return set_params_device(g, std::vector<float>(params, params + n));
}

extern "C" fc_status fc_sched_get_params(fc_graph* g, float* params, int32_t n) {
  fc::clear_error();
  if (!g || !params) return fc::fail(FC_ERR_INVALID, "fc_sched_get_params: null argument");
  if (n != kP) return fc::fail(FC_ERR_INVALID, "fc_sched_get_params: expected %d parameters", kP);
  if (!g->mlp_set) return fc::fail(FC_ERR_INVALID, "fc_sched_get_params: no scheduler parameters are set");
  std::copy(g->h_mlp.begin(), g->h_mlp.end(), params);
  return FC_OK;
}

static fc_status check_rows(fc_graph* g, const float* x, int32_t n, const char* fn) {
  if (!g || (n > 0 && !x)) return fc::fail(FC_ERR_INVALID, "%s: null argument", fn);
  if (n <= 0) return fc::fail(FC_ERR_INVALID, "%s: need at least one row", fn);
  if (!g->mlp_set) return fc::fail(FC_ERR_INVALID, "%s: no scheduler parameters are set", fn);
  return FC_OK;
}

extern "C" fc_status fc_sched_forward(fc_graph* g, const float* x, int32_t n, float* y_out) {
  fc::clear_error();
  fc_status st = check_rows(g, x, n, "fc_sched_forward");
  if (st != FC_OK) return st;
  if (!y_out) return fc::fail(FC_ERR_INVALID, "fc_sched_forward: null output");
  MlpBuffers b;
  if ((st = b.alloc(n, false)) != FC_OK) return st;
  FC_HIP(hipMemcpyAsync(b.X, x, sizeof(float) * (size_t)n * kFeatures, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b.w, g->d_mlp, sizeof(float) * kP, hipMemcpyDeviceToDevice, g->stream));
  FC_HIP(launch_rows(b, n, false, g->stream));
  FC_HIP(hipMemcpyAsync(y_out, b.Y, sizeof(float) * (size_t)n * kActions, hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  return FC_OK;
}

extern "C" fc_status fc_sched_gradient(fc_graph* g, const float* x, const float* target, const float* weight,
                                       int32_t n, float* loss_out, float* grad_out) {
  fc::clear_error();
  fc_status st = check_rows(g, x, n, "fc_sched_gradient");
  if (st != FC_OK) return st;
  if (!target || !grad_out) return fc::fail(FC_ERR_INVALID, "fc_sched_gradient: null argument");
  MlpBuffers b;
  if ((st = b.alloc(n, false)) != FC_OK) return st;
  if ((st = b.set_weights(weight, n, g->stream)) != FC_OK) return st;
  FC_HIP(hipMemcpyAsync(b.X, x, sizeof(float) * (size_t)n * kFeatures, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b.T, target, sizeof(float) * (size_t)n * kActions, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b.w, g->d_mlp, sizeof(float) * kP, hipMemcpyDeviceToDevice, g->stream));
  FC_HIP(launch_rows(b, n, true, g->stream));
  FC_HIP(launch_grad(b, n, g->stream));
  float loss = 0.f;
  FC_HIP(hipMemcpyAsync(grad_out, b.g, sizeof(float) * kP, hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipMemcpyAsync(&loss, b.loss, sizeof(float), hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  if (loss_out) *loss_out = loss;
  return FC_OK;
}

// Training targets of a trace row: log(cost) minus its mean over the actions.
static bool row_targets(const fc_trace_row& r, float* tgt) {
  double lc[kActions], mean = 0.0;
  for (int a = 0; a < kActions; ++a) {
    if (!(r.cost_us[a] > 0.f) || !std::isfinite(r.cost_us[a])) return false;
    lc[a] = std::log((double)r.cost_us[a]);
    mean += lc[a];
  }
  mean /= kActions;
  for (int a = 0; a < kActions; ++a) tgt[a] = (float)(lc[a] - mean);
  return true;
}

extern "C" fc_status fc_sched_train(fc_graph* g, const fc_trace_row* rows, int32_t n_rows, int32_t epochs, float lr,
                                    uint64_t seed, int32_t weighting, float* final_loss) {
  fc::clear_error();
  if (!g || (n_rows > 0 && !rows)) return fc::fail(FC_ERR_INVALID, "fc_sched_train: null argument");
  if (n_rows <= 0 || epochs <= 0 || !(lr > 0.f))
    return fc::fail(FC_ERR_INVALID, "fc_sched_train: need rows, epochs > 0 and lr > 0");
  if (weighting != FC_SCHED_WEIGHT_UNIFORM && weighting != FC_SCHED_WEIGHT_TIME)
    return fc::fail(FC_ERR_INVALID, "fc_sched_train: unknown weighting %d", weighting);
  std::vector<float> X, T, Wt;
  X.reserve((size_t)n_rows * kFeatures);
  T.reserve((size_t)n_rows * kActions);
  int32_t n = 0;
  for (int32_t i = 0; i < n_rows; ++i) {
    float tg[kActions];
    if (!row_targets(rows[i], tg)) continue;
    for (int k = 0; k < kFeatures; ++k) {
      if (!std::isfinite(rows[i].features[k])) return fc::fail(FC_ERR_INVALID, "fc_sched_train: row %d feature %d", i, k);
      X.push_back(rows[i].features[k]);
    }
    T.insert(T.end(), tg, tg + kActions);
    Wt.push_back(*std::min_element(rows[i].cost_us, rows[i].cost_us + kActions));
    ++n;
  }
  if (n == 0) return fc::fail(FC_ERR_INVALID, "fc_sched_train: no row has positive finite costs");
  if (weighting == FC_SCHED_WEIGHT_TIME) {  // each row's best remaining time over the mean of those
    double mean = 0.0;
    for (float v : Wt) mean += v;
    mean /= (double)n;
    for (float& v : Wt) v = (float)((double)v / mean);
  }
  std::vector<float> w;
  init_params(seed, w);
  MlpBuffers b;
  fc_status st = b.alloc(n, true);
  if (st != FC_OK) return st;
  if ((st = b.set_weights(weighting == FC_SCHED_WEIGHT_TIME ? Wt.data() : nullptr, n, g->stream)) != FC_OK) return st;
  FC_HIP(hipMemcpyAsync(b.X, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b.T, T.data(), sizeof(float) * T.size(), hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b.w, w.data(), sizeof(float) * kP, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemsetAsync(b.m, 0, sizeof(float) * kP, g->stream));
  FC_HIP(hipMemsetAsync(b.v, 0, sizeof(float) * kP, g->stream));
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  double c1 = 1.0, c2 = 1.0;
  for (int32_t e = 0; e < epochs; ++e) {
    FC_HIP(launch_rows(b, n, true, g->stream));
    FC_HIP(launch_grad(b, n, g->stream));
    c1 *= beta1;
    c2 *= beta2;
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kP + 255) / 256), dim3(256), 0, g->stream, b.w, b.g, b.m, b.v, lr, beta1,
                       beta2, eps, (float)(1.0 - c1), (float)(1.0 - c2));
    FC_HIP(hipGetLastError());
  }
  // loss of the final parameters
  FC_HIP(launch_rows(b, n, true, g->stream));
  FC_HIP(launch_grad(b, n, g->stream));
  float loss = 0.f;
  FC_HIP(hipMemcpyAsync(w.data(), b.w, sizeof(float) * kP, hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipMemcpyAsync(&loss, b.loss, sizeof(float), hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  for (float v : w)
    if (!std::isfinite(v)) return fc::fail(FC_ERR_INVALID, "fc_sched_train: training diverged (lower lr)");
  if (final_loss) *final_loss = loss;
  return set_params_device(g, w);
}
