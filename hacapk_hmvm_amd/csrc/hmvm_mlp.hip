// hmvm_mlp.hip -- the learned cost model: a nin -> nh -> nh -> 1 ReLU MLP whose
// forward pass, backward pass and Adam steps run as HIP kernels.
//
// It predicts the duration of a tile (one wavefront's share of a launch) from
// the tile's features; the schedule search uses the predictions to choose
// tile widths, launch variants and the dispatch order without timing every
// candidate.  FP32 throughout.  The network is tiny (nh <= 64), so the
// kernels are plain per-sample loops with the parameters staged in LDS;
// weight gradients are reduced over samples in fixed-size chunks and the
// chunks are summed in a fixed order, so a gradient is bit-reproducible.
//
// Parameter vector: W1 (nh x nin), b1 (nh), W2 (nh x nh), b2 (nh), w3 (nh), b3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "hmvm_hip.h"

namespace {

constexpr int kMaxIn = 32;
constexpr int kMaxH = 64;
constexpr int kChunk = 256;  // samples per weight-gradient partial

struct Layout {
  int nin, nh;
  int w1, b1, w2, b2, w3, b3, total;
  __host__ __device__ Layout(int i, int h) : nin(i), nh(h) {
    w1 = 0;
    b1 = w1 + nh * nin;
    w2 = b1 + nh;
    b2 = w2 + nh * nh;
    w3 = b2 + nh;
    b3 = w3 + nh;
    total = b3 + 1;
  }
};

// Forward pass of rows idx[s] (or s when idx is NULL); stores the
// standardised input, both hidden activations and the output.
__global__ __launch_bounds__(256) void mlp_forward_kernel(int n, int nin, int nh, const float *__restrict__ X,
                                                          const int *__restrict__ idx,
                                                          const float *__restrict__ params,
                                                          const float *__restrict__ mean,
                                                          const float *__restrict__ inv_std,
                                                          float *__restrict__ XS, float *__restrict__ H1,
                                                          float *__restrict__ H2, float *__restrict__ out) {
  extern __shared__ float sp[];
  const Layout L(nin, nh);
  for (int i = threadIdx.x; i < L.total; i += blockDim.x) sp[i] = params[i];
  __syncthreads();
  const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= n) return;
  const int row = idx ? idx[s] : s;
  float xs[kMaxIn];
  for (int i = 0; i < nin; ++i) {
    xs[i] = (X[(int64_t)row * nin + i] - mean[i]) * inv_std[i];
    if (XS) XS[(int64_t)s * nin + i] = xs[i];
  }
  float h1[kMaxH];
  for (int k = 0; k < nh; ++k) {
    float a = sp[L.b1 + k];
    for (int i = 0; i < nin; ++i) a += sp[L.w1 + k * nin + i] * xs[i];
    h1[k] = a > 0.f ? a : 0.f;
    if (H1) H1[(int64_t)s * nh + k] = h1[k];
  }
  float o = sp[L.b3];
  for (int j = 0; j < nh; ++j) {
    float a = sp[L.b2 + j];
    for (int k = 0; k < nh; ++k) a += sp[L.w2 + j * nh + k] * h1[k];
    a = a > 0.f ? a : 0.f;
    if (H2) H2[(int64_t)s * nh + j] = a;
    o += sp[L.w3 + j] * a;
  }
  out[s] = o;
}

// Output error and hidden deltas of every sample:
//   e = (out - y)/n, D2[j] = e w3[j] [H2>0], D1[k] = sum_j W2[j][k] D2[j] [H1>0].
__global__ __launch_bounds__(256) void mlp_delta_kernel(int n, int nin, int nh, float inv_n,
                                                        const float *__restrict__ y, const int *__restrict__ idx,
                                                        const float *__restrict__ params,
                                                        const float *__restrict__ H1, const float *__restrict__ H2,
                                                        const float *__restrict__ out, float *__restrict__ E,
                                                        float *__restrict__ D1, float *__restrict__ D2) {
  extern __shared__ float sp[];
  const Layout L(nin, nh);
  for (int i = threadIdx.x; i < L.total; i += blockDim.x) sp[i] = params[i];
  __syncthreads();
  const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= n) return;
  const int row = idx ? idx[s] : s;
  const float e = (out[s] - y[row]) * inv_n;
  E[s] = e;
  float d2[kMaxH];
  for (int j = 0; j < nh; ++j) {
    d2[j] = H2[(int64_t)s * nh + j] > 0.f ? e * sp[L.w3 + j] : 0.f;
    D2[(int64_t)s * nh + j] = d2[j];
  }
  for (int k = 0; k < nh; ++k) {
    float a = 0.f;
    if (H1[(int64_t)s * nh + k] > 0.f)
      for (int j = 0; j < nh; ++j) a += sp[L.w2 + j * nh + k] * d2[j];
    D1[(int64_t)s * nh + k] = a;
  }
}

// Partial weight gradients over one chunk of kChunk samples: thread p of the
// grid computes parameter p of chunk blockIdx.y.
//   W1[k][i] += D1[s][k] XS[s][i]   b1[k] += D1[s][k]
//   W2[j][k] += D2[s][j] H1[s][k]   b2[j] += D2[s][j]
//   w3[j]    += E[s] H2[s][j]       b3    += E[s]
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(int n, int nin, int nh, const float *__restrict__ XS,
                                                        const float *__restrict__ H1, const float *__restrict__ H2,
                                                        const float *__restrict__ E, const float *__restrict__ D1,
                                                        const float *__restrict__ D2, float *__restrict__ partial) {
  const Layout L(nin, nh);
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= L.total) return;
  const int s0 = (int)blockIdx.y * kChunk;
  const int s1 = min(n, s0 + kChunk);
  float a = 0.f;
  if (p < L.b1) {
    const int k = p / nin, i = p - k * nin;
    for (int s = s0; s < s1; ++s) a += D1[(int64_t)s * nh + k] * XS[(int64_t)s * nin + i];
  } else if (p < L.w2) {
    const int k = p - L.b1;
    for (int s = s0; s < s1; ++s) a += D1[(int64_t)s * nh + k];
  } else if (p < L.b2) {
    const int j = (p - L.w2) / nh, k = (p - L.w2) - j * nh;
    for (int s = s0; s < s1; ++s) a += D2[(int64_t)s * nh + j] * H1[(int64_t)s * nh + k];
  } else if (p < L.w3) {
    const int j = p - L.b2;
    for (int s = s0; s < s1; ++s) a += D2[(int64_t)s * nh + j];
  } else if (p < L.b3) {
    const int j = p - L.w3;
    for (int s = s0; s < s1; ++s) a += E[s] * H2[(int64_t)s * nh + j];
  } else {
    for (int s = s0; s < s1; ++s) a += E[s];
  }
  partial[(int64_t)blockIdx.y * L.total + p] = a;
}

// grad[p] = sum over chunks in order; loss = 0.5/n sum (out - y)^2 by one
// block in a fixed order.
__global__ __launch_bounds__(256) void mlp_reduce_kernel(int total, int nchunks, const float *__restrict__ partial,
                                                         float *__restrict__ grad) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= total) return;
  float a = 0.f;
  for (int c = 0; c < nchunks; ++c) a += partial[(int64_t)c * total + p];
  grad[p] = a;
}

__global__ __launch_bounds__(256) void mlp_loss_kernel(int n, const float *__restrict__ out,
                                                       const float *__restrict__ y, const int *__restrict__ idx,
                                                       float *__restrict__ loss) {
  __shared__ double red[256];
  double a = 0;
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    const double d = (double)out[s] - (double)y[idx ? idx[s] : s];
    a += d * d;
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = (float)(0.5 * red[0] / n);
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(int total, float *__restrict__ params,
                                                       const float *__restrict__ grad, float *__restrict__ m1,
                                                       float *__restrict__ m2, float lr, float c1, float c2) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= total) return;
  const float g = grad[p];
  const float a = 0.9f * m1[p] + 0.1f * g;
  const float b = 0.999f * m2[p] + 0.001f * g * g;
  m1[p] = a;
  m2[p] = b;
  params[p] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
  hipError_t alloc(size_t bytes) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc(&p, std::max<size_t>(bytes, 4));
  }
};

int fail(const char *what, hipError_t e) {
  std::fprintf(stderr, "hmvm_mlp: %s: %s\n", what, hipGetErrorString(e));
  return HMVM_ERR_HIP;
}

#define MLP_TRY(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return fail(#expr, e_);       \
  } while (0)

}  // namespace

struct hmvm_mlp {
  Layout L;
  std::vector<float> params, mean, stdev;
  hipStream_t st = nullptr;
  hmvm_mlp(int nin, int nh) : L(nin, nh) {}
  ~hmvm_mlp() {
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Device workspace for one pass over n samples.
struct Work {
  DevBuf X, y, idx, params, mean, istd, XS, H1, H2, out, E, D1, D2, partial, grad, loss;
  int n = 0, nrows = 0;
  int setup(const hmvm_mlp *m, const float *hX, const float *hy, int rows, int batch) {
    const Layout &L = m->L;
    nrows = rows;
    n = batch;
    MLP_TRY(X.alloc(sizeof(float) * (size_t)rows * L.nin));
    MLP_TRY(hipMemcpy(X.p, hX, sizeof(float) * (size_t)rows * L.nin, hipMemcpyHostToDevice));
    if (hy) {
      MLP_TRY(y.alloc(sizeof(float) * (size_t)rows));
      MLP_TRY(hipMemcpy(y.p, hy, sizeof(float) * (size_t)rows, hipMemcpyHostToDevice));
    }
    MLP_TRY(idx.alloc(sizeof(int) * (size_t)batch));
    MLP_TRY(params.alloc(sizeof(float) * L.total));
    MLP_TRY(hipMemcpy(params.p, m->params.data(), sizeof(float) * L.total, hipMemcpyHostToDevice));
    std::vector<float> istd_h(L.nin);
    for (int i = 0; i < L.nin; i++) istd_h[i] = m->stdev[i] > 0 ? 1.0f / m->stdev[i] : 1.0f;
    MLP_TRY(mean.alloc(sizeof(float) * L.nin));
    MLP_TRY(istd.alloc(sizeof(float) * L.nin));
    MLP_TRY(hipMemcpy(mean.p, m->mean.data(), sizeof(float) * L.nin, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(istd.p, istd_h.data(), sizeof(float) * L.nin, hipMemcpyHostToDevice));
    MLP_TRY(XS.alloc(sizeof(float) * (size_t)batch * L.nin));
    MLP_TRY(H1.alloc(sizeof(float) * (size_t)batch * L.nh));
    MLP_TRY(H2.alloc(sizeof(float) * (size_t)batch * L.nh));
    MLP_TRY(out.alloc(sizeof(float) * (size_t)batch));
    MLP_TRY(E.alloc(sizeof(float) * (size_t)batch));
    MLP_TRY(D1.alloc(sizeof(float) * (size_t)batch * L.nh));
    MLP_TRY(D2.alloc(sizeof(float) * (size_t)batch * L.nh));
    const int nch = (batch + kChunk - 1) / kChunk;
    MLP_TRY(partial.alloc(sizeof(float) * (size_t)nch * L.total));
    MLP_TRY(grad.alloc(sizeof(float) * L.total));
    MLP_TRY(loss.alloc(sizeof(float)));
    return HMVM_OK;
  }
  // forward (+ loss) on nb samples selected by d_idx (NULL: rows 0..nb-1)
  void forward(const hmvm_mlp *m, int nb, const int *d_idx, hipStream_t st) {
    const Layout &L = m->L;
    const size_t lds = sizeof(float) * L.total;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((nb + 255) / 256), dim3(256), lds, st, nb, L.nin, L.nh,
                       X.as<float>(), d_idx, params.as<float>(), mean.as<float>(), istd.as<float>(), XS.as<float>(),
                       H1.as<float>(), H2.as<float>(), out.as<float>());
  }
  void backward(const hmvm_mlp *m, int nb, const int *d_idx, hipStream_t st) {
    const Layout &L = m->L;
    const size_t lds = sizeof(float) * L.total;
    hipLaunchKernelGGL(mlp_delta_kernel, dim3((nb + 255) / 256), dim3(256), lds, st, nb, L.nin, L.nh,
                       1.0f / (float)nb, y.as<float>(), d_idx, params.as<float>(), H1.as<float>(), H2.as<float>(),
                       out.as<float>(), E.as<float>(), D1.as<float>(), D2.as<float>());
    const int nch = (nb + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((L.total + 255) / 256, nch), dim3(256), 0, st, nb, L.nin, L.nh,
                       XS.as<float>(), H1.as<float>(), H2.as<float>(), E.as<float>(), D1.as<float>(), D2.as<float>(),
                       partial.as<float>());
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((L.total + 255) / 256), dim3(256), 0, st, L.total, nch,
                       partial.as<float>(), grad.as<float>());
  }
};

bool shape_ok(int nin, int nh) { return nin >= 1 && nin <= kMaxIn && nh >= 1 && nh <= kMaxH; }

}  // namespace

extern "C" {

int hmvm_mlp_create(int nin, int nh, unsigned long long seed, hmvm_mlp **out) {
  if (!out || !shape_ok(nin, nh)) return HMVM_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<hmvm_mlp> m(new hmvm_mlp(nin, nh));
  const Layout &L = m->L;
  m->params.assign(L.total, 0.f);
  m->mean.assign(nin, 0.f);
  m->stdev.assign(nin, 1.f);
  // He-uniform initialisation from a splitmix64 stream
  uint64_t s = seed ^ 0x9E3779B97F4A7C15ull;
  auto uni = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  const double a1 = std::sqrt(6.0 / nin), a2 = std::sqrt(6.0 / nh);
  for (int i = 0; i < nh * nin; i++) m->params[L.w1 + i] = (float)(a1 * uni());
  for (int i = 0; i < nh * nh; i++) m->params[L.w2 + i] = (float)(a2 * uni());
  for (int i = 0; i < nh; i++) m->params[L.w3 + i] = (float)(a2 * uni() * 0.5);
  if (hipStreamCreate(&m->st) != hipSuccess) return HMVM_ERR_HIP;
  *out = m.release();
  return HMVM_OK;
}

void hmvm_mlp_destroy(hmvm_mlp *m) { delete m; }

int hmvm_mlp_num_params(const hmvm_mlp *m) { return m ? m->L.total : 0; }

int hmvm_mlp_get_params(const hmvm_mlp *m, float *params, float *mean, float *stdev) {
  if (!m) return HMVM_ERR_ARG;
  if (params) std::memcpy(params, m->params.data(), sizeof(float) * m->L.total);
  if (mean) std::memcpy(mean, m->mean.data(), sizeof(float) * m->L.nin);
  if (stdev) std::memcpy(stdev, m->stdev.data(), sizeof(float) * m->L.nin);
  return HMVM_OK;
}

int hmvm_mlp_set_params(hmvm_mlp *m, const float *params, const float *mean, const float *stdev) {
  if (!m) return HMVM_ERR_ARG;
  if (params) std::memcpy(m->params.data(), params, sizeof(float) * m->L.total);
  if (mean) std::memcpy(m->mean.data(), mean, sizeof(float) * m->L.nin);
  if (stdev) std::memcpy(m->stdev.data(), stdev, sizeof(float) * m->L.nin);
  return HMVM_OK;
}

int hmvm_mlp_fit_scaler(hmvm_mlp *m, const float *X, int n) {
  if (!m || !X || n < 1) return HMVM_ERR_ARG;
  const int nin = m->L.nin;
  for (int i = 0; i < nin; i++) {
    double s = 0, s2 = 0;
    for (int r = 0; r < n; r++) s += X[(size_t)r * nin + i];
    const double mu = s / n;
    for (int r = 0; r < n; r++) {
      const double d = X[(size_t)r * nin + i] - mu;
      s2 += d * d;
    }
    const double sd = std::sqrt(s2 / n);
    m->mean[i] = (float)mu;
    m->stdev[i] = sd > 1e-6 ? (float)sd : 1.0f;
  }
  return HMVM_OK;
}

int hmvm_mlp_predict(hmvm_mlp *m, const float *X, int n, float *y) {
  if (!m || !X || !y || n < 1) return HMVM_ERR_ARG;
  Work w;
  int rc = w.setup(m, X, nullptr, n, n);
  if (rc) return rc;
  w.forward(m, n, nullptr, m->st);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(y, w.out.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  return HMVM_OK;
}

int hmvm_mlp_loss_grad(hmvm_mlp *m, const float *X, const float *y, int n, float *loss, float *grad) {
  if (!m || !X || !y || n < 1) return HMVM_ERR_ARG;
  Work w;
  int rc = w.setup(m, X, y, n, n);
  if (rc) return rc;
  w.forward(m, n, nullptr, m->st);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, m->st, n, w.out.as<float>(), w.y.as<float>(),
                     (const int *)nullptr, w.loss.as<float>());
  w.backward(m, n, nullptr, m->st);
  MLP_TRY(hipGetLastError());
  if (loss) MLP_TRY(hipMemcpyAsync(loss, w.loss.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
  if (grad) MLP_TRY(hipMemcpyAsync(grad, w.grad.p, sizeof(float) * m->L.total, hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  return HMVM_OK;
}

int hmvm_mlp_train(hmvm_mlp *m, const float *X, const float *y, int n, int epochs, int batch, float lr,
                   unsigned long long seed, float *final_loss) {
  if (!m || !X || !y || n < 1 || epochs < 0 || batch < 1 || !(lr > 0)) return HMVM_ERR_ARG;
  batch = std::min(batch, n);
  Work w;
  int rc = w.setup(m, X, y, n, std::max(batch, n));
  if (rc) return rc;
  const Layout &L = m->L;
  DevBuf m1, m2;
  MLP_TRY(m1.alloc(sizeof(float) * L.total));
  MLP_TRY(m2.alloc(sizeof(float) * L.total));
  MLP_TRY(hipMemsetAsync(m1.p, 0, sizeof(float) * L.total, m->st));
  MLP_TRY(hipMemsetAsync(m2.p, 0, sizeof(float) * L.total, m->st));
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1;
  auto next = [&]() {
    s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
    return s * 0x2545F4914F6CDD1Dull;
  };
  int step = 0;
  const int nbatches = (n + batch - 1) / batch;
  for (int ep = 0; ep < epochs; ep++) {
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(size_t)(next() % (uint64_t)(i + 1))]);
    MLP_TRY(hipMemcpyAsync(w.idx.p, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->st));
    for (int b = 0; b < nbatches; b++) {
      const int b0 = b * batch, nb = std::min(batch, n - b0);
      const int *d_idx = w.idx.as<int>() + b0;
      w.forward(m, nb, d_idx, m->st);
      w.backward(m, nb, d_idx, m->st);
      step++;
      const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)step));
      const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)step));
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((L.total + 255) / 256), dim3(256), 0, m->st, L.total,
                         w.params.as<float>(), w.grad.as<float>(), m1.as<float>(), m2.as<float>(), lr, c1, c2);
    }
    MLP_TRY(hipGetLastError());
    // the permutation buffer is rewritten next epoch: finish this one first
    MLP_TRY(hipStreamSynchronize(m->st));
  }
  MLP_TRY(hipMemcpyAsync(m->params.data(), w.params.p, sizeof(float) * L.total, hipMemcpyDeviceToHost, m->st));
  w.forward(m, n, nullptr, m->st);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, m->st, n, w.out.as<float>(), w.y.as<float>(),
                     (const int *)nullptr, w.loss.as<float>());
  float l = 0;
  MLP_TRY(hipMemcpyAsync(&l, w.loss.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  if (final_loss) *final_loss = l;
  return HMVM_OK;
}

}  // extern "C"
