// hmvm_mlp.hip -- the learned cost model: a nin -> nh -> nh -> 1 ReLU MLP whose
// forward pass, backward pass and Adam steps run as HIP kernels.
//
// It predicts the duration of a tile (one wavefront's share of a launch) from
// the tile's features; the schedule search uses the predictions to choose
// tile widths, launch variants and the dispatch order without timing every
// candidate.  FP32 throughout.  The network is tiny (nh <= 64): one
// workgroup takes 64 samples through the forward and backward passes with
// the parameters and activations in LDS (register-tiled matrix products),
// and one more launch adds the chunks' gradients in chunk order and applies
// the Adam step, so a training step is two launches and a gradient is
// bit-reproducible.
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
constexpr int kChunk = 64;  // samples per workgroup (one gradient partial)

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

// ---- one chunk of samples per workgroup -------------------------------------
// A workgroup of 256 threads takes kChunk = 64 samples (rows idx[s], or s
// when idx is NULL) through the whole network: the parameters and every
// activation of the chunk live in LDS, and each of the five matrix products
//   Z1 = Xs W1^T (64 x nh, K = nin)     Z2 = H1 W2^T (64 x nh, K = nh)
//   D1 = D2 W2   (64 x nh, K = nh)      gW2 = D2^T H1 (nh x nh, K = 64)
//   gW1 = D1^T Xs (nh x nin, K = 64)
// is computed with 4 x 4 register tiles (thread t: rows 4(t/16) .. +3,
// columns t%16 + 16c, c < 4), so one LDS word feeds four FMAs.  Sums over
// the chunk's samples run in sample order inside a thread; the chunk's
// gradient goes to its own row of `partial`, and the chunks are added in
// chunk order afterwards, so a gradient is bit-reproducible.
// Samples past n in the last chunk carry zero error and add nothing.
constexpr int kThreads = 256;

// LDS row strides (floats): odd, so the rows a column of threads reads fall
// in different banks.
__host__ __device__ constexpr int pad_odd(int k) { return k | 1; }

struct ChunkLds {  // float offsets into the dynamic LDS array
  int w1, b1, w2, b2, w3, b3, xs, h1, h2, e, o, yt, total;
  __host__ __device__ ChunkLds(int nin, int nh) {
    const int si = pad_odd(nin), sh = pad_odd(nh);
    w1 = 0;                 // nh rows of nin (stride si)
    b1 = w1 + nh * si;
    w2 = b1 + nh;           // nh rows of nh (stride sh)
    b2 = w2 + nh * sh;
    w3 = b2 + nh;
    b3 = w3 + nh;
    xs = b3 + 1;            // kChunk rows of nin (stride si)
    h1 = xs + kChunk * si;  // kChunk rows of nh: H1, later D1
    h2 = h1 + kChunk * sh;  // kChunk rows of nh: H2, later D2
    e = h2 + kChunk * sh;   // kChunk output errors (out - y)/n
    o = e + kChunk;         // kChunk outputs
    yt = o + kChunk;        // kChunk targets
    total = yt + kChunk;
  }
};

// C[r][c] (+)= sum_k A[r][k] * B[c][k] for r < 64, c < ncols (row strides
// sa, sb): thread t holds rows 4(t/16)+i and columns t%16 + 16j.
__device__ __forceinline__ void tile_rows_cols(const float *A, int sa, const float *B, int sb, int K, int ncols,
                                               float (&acc)[4][4]) {
  const int r0 = 4 * ((int)threadIdx.x >> 4), c0 = (int)threadIdx.x & 15;
  for (int k = 0; k < K; ++k) {
    float a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = A[(r0 + i) * sa + k];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = c0 + 16 * j < ncols ? B[(c0 + 16 * j) * sb + k] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
  }
}

// C[r][c] = sum_s A[s][r] * B[s][c] over the chunk's samples s (the
// transposed products of the weight gradients): rows r < nr, columns c < nc.
__device__ __forceinline__ void tile_grad(const float *A, int sa, int nr, const float *B, int sb, int nc,
                                          float (&acc)[4][4], int r0, int c0) {
  for (int s = 0; s < kChunk; ++s) {
    float a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = r0 + i < nr ? A[s * sa + r0 + i] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = c0 + 16 * j < nc ? B[s * sb + c0 + 16 * j] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
  }
}

// kMode 0: predict (out[s] for the chunk's samples); 1: loss and gradient
// partials of the chunk (partial[chunk * total + p], loss_part[chunk]).
template <int kMode>
__global__ __launch_bounds__(kThreads) void mlp_chunk_kernel(
    int n, int nin, int nh, float inv_n, const float *__restrict__ X, const float *__restrict__ y,
    const int *__restrict__ idx, const float *__restrict__ params, const float *__restrict__ mean,
    const float *__restrict__ inv_std, float *__restrict__ out, float *__restrict__ partial,
    float *__restrict__ loss_part) {
  extern __shared__ float lds[];
  const Layout L(nin, nh);
  const ChunkLds C(nin, nh);
  const int si = pad_odd(nin), sh = pad_odd(nh);
  const int tid = (int)threadIdx.x;
  const int s0 = (int)blockIdx.x * kChunk;
  const int ns = min(kChunk, n - s0);
  // parameters (global layout) -> LDS (padded rows)
  for (int p = tid; p < L.total; p += kThreads) {
    const float v = params[p];
    if (p < L.b1) lds[C.w1 + (p / nin) * si + p % nin] = v;
    else if (p < L.w2) lds[C.b1 + p - L.b1] = v;
    else if (p < L.b2) lds[C.w2 + ((p - L.w2) / nh) * sh + (p - L.w2) % nh] = v;
    else if (p < L.w3) lds[C.b2 + p - L.b2] = v;
    else if (p < L.b3) lds[C.w3 + p - L.w3] = v;
    else lds[C.b3] = v;
  }
  // standardised inputs (zero rows past n)
  for (int q = tid; q < kChunk * nin; q += kThreads) {
    const int s = q / nin, i = q - s * nin;
    float v = 0.f;
    if (s < ns) {
      const int row = idx ? idx[s0 + s] : s0 + s;
      v = (X[(int64_t)row * nin + i] - mean[i]) * inv_std[i];
    }
    lds[C.xs + s * si + i] = v;
  }
  __syncthreads();
  const int r0 = 4 * (tid >> 4), c0 = tid & 15;
  float acc[4][4];
  // H1 = relu(Xs W1^T + b1)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  tile_rows_cols(lds + C.xs, si, lds + C.w1, si, nin, nh, acc);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (c0 + 16 * j < nh)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds[C.h1 + (r0 + i) * sh + c0 + 16 * j] = fmaxf(acc[i][j] + lds[C.b1 + c0 + 16 * j], 0.f);
  __syncthreads();
  // H2 = relu(H1 W2^T + b2)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  tile_rows_cols(lds + C.h1, sh, lds + C.w2, sh, nh, nh, acc);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (c0 + 16 * j < nh)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds[C.h2 + (r0 + i) * sh + c0 + 16 * j] = fmaxf(acc[i][j] + lds[C.b2 + c0 + 16 * j], 0.f);
  __syncthreads();
  // out = H2 w3 + b3; training: e = (out - y)/n, zero past n
  if (tid < kChunk) {
    float o = lds[C.b3];
    for (int j = 0; j < nh; ++j) o += lds[C.w3 + j] * lds[C.h2 + tid * sh + j];
    if (kMode == 0) {
      if (tid < ns) out[s0 + tid] = o;
    } else {
      float e = 0.f, yv = 0.f;
      if (tid < ns) {
        yv = y[idx ? idx[s0 + tid] : s0 + tid];
        e = (o - yv) * inv_n;
      }
      lds[C.e + tid] = e;
      lds[C.o + tid] = o;
      lds[C.yt + tid] = yv;
    }
  }
  if (kMode == 0) return;
  __syncthreads();
  float *part = partial + (int64_t)blockIdx.x * L.total;
  if (tid == 0) {  // loss: sum of squared errors of the chunk, in sample order
    double a = 0;
    for (int s = 0; s < ns; ++s) {
      const double d = (double)lds[C.o + s] - (double)lds[C.yt + s];
      a += d * d;
    }
    loss_part[blockIdx.x] = (float)a;
  }
  // gw3[j] = sum_s e[s] H2[s][j], gb3 = sum_s e[s]  (before H2 becomes D2)
  if (tid < nh) {
    float a = 0.f;
    for (int s = 0; s < kChunk; ++s) a += lds[C.e + s] * lds[C.h2 + s * sh + tid];
    part[L.w3 + tid] = a;
  } else if (tid == 64) {
    float a = 0.f;
    for (int s = 0; s < kChunk; ++s) a += lds[C.e + s];
    part[L.b3] = a;
  }
  __syncthreads();
  // D2 = e w3^T [H2 > 0], in place of H2
  for (int q = tid; q < kChunk * nh; q += kThreads) {
    const int s = q / nh, j = q - s * nh;
    float *h = lds + C.h2 + s * sh + j;
    *h = *h > 0.f ? lds[C.e + s] * lds[C.w3 + j] : 0.f;
  }
  __syncthreads();
  // gW2[j][k] = sum_s D2[s][j] H1[s][k]; gb2[j] = sum_s D2[s][j]
  {
    // thread: rows j = 4(t/16)..+3 of gW2, columns k = t%16 + 16c
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    tile_grad(lds + C.h2, sh, nh, lds + C.h1, sh, nh, acc, r0, c0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (r0 + i < nh)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (c0 + 16 * j < nh) part[L.w2 + (r0 + i) * nh + c0 + 16 * j] = acc[i][j];
    if (tid < nh) {
      float a = 0.f;
      for (int s = 0; s < kChunk; ++s) a += lds[C.h2 + s * sh + tid];
      part[L.b2 + tid] = a;
    }
  }
  // D1 = (D2 W2) [H1 > 0]: D1[s][k] = sum_j D2[s][j] W2[j][k]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  for (int jj = 0; jj < nh; ++jj) {
    float a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = lds[C.h2 + (r0 + i) * sh + jj];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = c0 + 16 * j < nh ? lds[C.w2 + jj * sh + c0 + 16 * j] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
  }
  __syncthreads();  // every read of H1 (gW2 above) is done before D1 replaces it
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (c0 + 16 * j < nh)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float *h = lds + C.h1 + (r0 + i) * sh + c0 + 16 * j;
        *h = *h > 0.f ? acc[i][j] : 0.f;
      }
  __syncthreads();
  // gW1[k][i] = sum_s D1[s][k] Xs[s][i]; gb1[k] = sum_s D1[s][k]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  // rows k (nh) = 4(t/16)..+3, columns i (nin <= 32) = t%16 + 16c, c < 2 used
  tile_grad(lds + C.h1, sh, nh, lds + C.xs, si, nin, acc, r0, c0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (r0 + i < nh)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (c0 + 16 * j < nin) part[L.w1 + (r0 + i) * nin + c0 + 16 * j] = acc[i][j];
  if (tid < nh) {
    float a = 0.f;
    for (int s = 0; s < kChunk; ++s) a += lds[C.h1 + s * sh + tid];
    part[L.b1 + tid] = a;
  }
}

// grad[p] = sum of the chunks' partials in chunk order; with kAdam the Adam
// step (beta1 0.9, beta2 0.999, bias corrections c1, c2) is applied to the
// parameters instead of storing the gradient.  Block 0 also adds the chunks'
// squared errors: loss = 0.5/n * sum.
template <bool kAdam>
__global__ __launch_bounds__(256) void mlp_reduce_kernel(int total, int nchunks, const float *__restrict__ partial,
                                                         const float *__restrict__ loss_part, float half_inv_n,
                                                         float *__restrict__ grad, float *__restrict__ params,
                                                         float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                         float c1, float c2, float *__restrict__ loss) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (loss && p == 0) {
    double a = 0;
    for (int c = 0; c < nchunks; ++c) a += (double)loss_part[c];
    *loss = (float)(a * half_inv_n);
  }
  if (p >= total) return;
  float g = 0.f;
  for (int c = 0; c < nchunks; ++c) g += partial[(int64_t)c * total + p];
  if (kAdam) {
    const float a = 0.9f * m1[p] + 0.1f * g;
    const float b = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
  } else {
    grad[p] = g;
  }
}

__global__ void mlp_zero_kernel(int n, float *__restrict__ a) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) a[i] = 0.f;
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

// Device workspace for passes over up to `batch` samples of `rows` rows.
struct Work {
  DevBuf X, y, idx, params, mean, istd, out, partial, loss_part, grad, loss;
  int n = 0, nrows = 0;
  int setup(const hmvm_mlp *m, const float *hX, const float *hy, int rows, int batch) {
    const Layout &L = m->L;
    nrows = rows;
    n = batch;
    MLP_TRY(X.alloc(sizeof(float) * (size_t)rows * L.nin));
    {
      int dev = 0, max_lds = 0;
      MLP_TRY(hipGetDevice(&dev));
      MLP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
      if (lds_bytes(m) > (size_t)max_lds) {
        std::fprintf(stderr, "hmvm_mlp: %zu bytes of LDS per workgroup exceed the device's %d\n", lds_bytes(m),
                     max_lds);
        return HMVM_ERR_UNSUPPORTED;
      }
    }
    MLP_TRY(hipMemcpy(X.p, hX, sizeof(float) * (size_t)rows * L.nin, hipMemcpyHostToDevice));
    if (hy) {
      MLP_TRY(y.alloc(sizeof(float) * (size_t)rows));
      MLP_TRY(hipMemcpy(y.p, hy, sizeof(float) * (size_t)rows, hipMemcpyHostToDevice));
    }
    MLP_TRY(idx.alloc(sizeof(int) * (size_t)rows));
    MLP_TRY(params.alloc(sizeof(float) * L.total));
    MLP_TRY(hipMemcpy(params.p, m->params.data(), sizeof(float) * L.total, hipMemcpyHostToDevice));
    std::vector<float> istd_h(L.nin);
    for (int i = 0; i < L.nin; i++) istd_h[i] = m->stdev[i] > 0 ? 1.0f / m->stdev[i] : 1.0f;
    MLP_TRY(mean.alloc(sizeof(float) * L.nin));
    MLP_TRY(istd.alloc(sizeof(float) * L.nin));
    MLP_TRY(hipMemcpy(mean.p, m->mean.data(), sizeof(float) * L.nin, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(istd.p, istd_h.data(), sizeof(float) * L.nin, hipMemcpyHostToDevice));
    MLP_TRY(out.alloc(sizeof(float) * (size_t)rows));
    const int nch = (batch + kChunk - 1) / kChunk;
    MLP_TRY(partial.alloc(sizeof(float) * (size_t)nch * L.total));
    MLP_TRY(loss_part.alloc(sizeof(float) * (size_t)nch));
    MLP_TRY(grad.alloc(sizeof(float) * L.total));
    MLP_TRY(loss.alloc(sizeof(float)));
    return HMVM_OK;
  }
  size_t lds_bytes(const hmvm_mlp *m) const { return sizeof(float) * (size_t)ChunkLds(m->L.nin, m->L.nh).total; }
  // out[s] = f(X[s]) for the first nb rows
  void predict(const hmvm_mlp *m, int nb, hipStream_t st) {
    const Layout &L = m->L;
    hipLaunchKernelGGL(mlp_chunk_kernel<0>, dim3((nb + kChunk - 1) / kChunk), dim3(kThreads), lds_bytes(m), st, nb,
                       L.nin, L.nh, 0.f, X.as<float>(), (const float *)nullptr, (const int *)nullptr,
                       params.as<float>(), mean.as<float>(), istd.as<float>(), out.as<float>(), (float *)nullptr,
                       (float *)nullptr);
  }
  // loss and gradient partials of nb samples selected by d_idx (NULL: rows 0..nb-1)
  void chunks(const hmvm_mlp *m, int nb, const int *d_idx, hipStream_t st) {
    const Layout &L = m->L;
    hipLaunchKernelGGL(mlp_chunk_kernel<1>, dim3((nb + kChunk - 1) / kChunk), dim3(kThreads), lds_bytes(m), st, nb,
                       L.nin, L.nh, 1.0f / (float)nb, X.as<float>(), y.as<float>(), d_idx, params.as<float>(),
                       mean.as<float>(), istd.as<float>(), (float *)nullptr, partial.as<float>(),
                       loss_part.as<float>());
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
  w.predict(m, n, m->st);
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
  const Layout &L = m->L;
  w.chunks(m, n, nullptr, m->st);
  const int nch = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(mlp_reduce_kernel<false>, dim3((L.total + 255) / 256), dim3(256), 0, m->st, L.total, nch,
                     w.partial.as<float>(), w.loss_part.as<float>(), 0.5f / (float)n, w.grad.as<float>(),
                     (float *)nullptr, (float *)nullptr, (float *)nullptr, 0.f, 0.f, 0.f, w.loss.as<float>());
  MLP_TRY(hipGetLastError());
  if (loss) MLP_TRY(hipMemcpyAsync(loss, w.loss.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
  if (grad) MLP_TRY(hipMemcpyAsync(grad, w.grad.p, sizeof(float) * L.total, hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  return HMVM_OK;
}

// Minibatch Adam: one launch of the chunk kernel (forward and backward of
// the minibatch, one workgroup per 64 samples) and one reduce-and-update
// launch per step; the host only reshuffles the sample order once an epoch.
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
  hipLaunchKernelGGL(mlp_zero_kernel, dim3((L.total + 255) / 256), dim3(256), 0, m->st, L.total, m1.as<float>());
  hipLaunchKernelGGL(mlp_zero_kernel, dim3((L.total + 255) / 256), dim3(256), 0, m->st, L.total, m2.as<float>());
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
      w.chunks(m, nb, w.idx.as<int>() + b0, m->st);
      step++;
      const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)step));
      const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)step));
      hipLaunchKernelGGL(mlp_reduce_kernel<true>, dim3((L.total + 255) / 256), dim3(256), 0, m->st, L.total,
                         (nb + kChunk - 1) / kChunk, w.partial.as<float>(), w.loss_part.as<float>(), 0.f,
                         (float *)nullptr, w.params.as<float>(), m1.as<float>(), m2.as<float>(), lr, c1, c2,
                         (float *)nullptr);
    }
    MLP_TRY(hipGetLastError());
    // the permutation buffer is rewritten next epoch: finish this one first
    MLP_TRY(hipStreamSynchronize(m->st));
  }
  MLP_TRY(hipMemcpyAsync(m->params.data(), w.params.p, sizeof(float) * L.total, hipMemcpyDeviceToHost, m->st));
  w.chunks(m, n, nullptr, m->st);
  const int nch = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(mlp_reduce_kernel<false>, dim3((L.total + 255) / 256), dim3(256), 0, m->st, L.total, nch,
                     w.partial.as<float>(), w.loss_part.as<float>(), 0.5f / (float)n, w.grad.as<float>(),
                     (float *)nullptr, (float *)nullptr, (float *)nullptr, 0.f, 0.f, 0.f, w.loss.as<float>());
  float l = 0;
  MLP_TRY(hipMemcpyAsync(&l, w.loss.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  if (final_loss) *final_loss = l;
  return HMVM_OK;
}

}  // extern "C"
