// hmvm_krylov.hip -- BiCGSTAB on the device with the HIP HMVM as its operator.
//
// Restates the iteration of the reference's HACApK_bicgstab_lfmtx_hyp
// (src_hdump/m_HACApK_solve.f90:306-380): two HMVMs per iteration, vector
// updates and dot products between them.  Every scalar (alpha, zeta, beta,
// the residual norm, the stop state and the iteration count) stays in device
// memory, so one iteration is a fixed sequence of launches with no host
// synchronisation; with use_graph the iteration is captured once into a
// hipGraph and replayed.  Dot products are two-pass (per-block partials, then
// one block) and therefore deterministic: the eager and the graph-replayed
// solve give the same bits.
//
// Stopping: the reference leaves its loop at the top of the first iteration
// whose ||r||/||b|| is below eps (m_HACApK_solve.f90:342).  Here the kernel
// that computes ||r|| sets a device stop word instead, and every later update
// kernel and scalar update returns at once, so the host may launch whole
// chunks of iterations between its read-backs and still get the reference's
// iterate and iteration count.  A zero divisor or a non-finite scalar sets
// the stop word to "breakdown" the same way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "hmvm_hip.h"
#include "hmvm_status.hpp"

namespace {

int hip_fail(const char *what, hipError_t e) {
  char m[512];
  std::snprintf(m, sizeof m, "%s failed: %s", what, hipGetErrorString(e));
  return hmvm::report_error(HMVM_ERR_HIP, m);
}

#define KRY_TRY(expr)                         \
  do {                                        \
    const hipError_t e_ = (expr);             \
    if (e_ != hipSuccess) return hip_fail(#expr, e_); \
  } while (0)

constexpr int kBlock = 256;
constexpr int kDotBlocks = 512;

// device scalars
enum {
  kAlpha = 0, kBeta, kZeta, kZnomOld, kRnorm, kBnorm,
  kStop,   // 0 running, 1 converged, 2 breakdown
  kIters,  // iterations that updated u
  kNumScalars
};

// partial[blockIdx] = (sum a*b, sum c*d) over a grid-stride slice
__global__ __launch_bounds__(kBlock) void dot2_partial(int n, const double *__restrict__ a,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ c,
                                                       const double *__restrict__ d, double *__restrict__ part) {
  __shared__ double s0[kBlock], s1[kBlock];
  double x0 = 0, x1 = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x0 += a[i] * b[i];
    x1 += c[i] * d[i];
  }
  s0[threadIdx.x] = x0;
  s1[threadIdx.x] = x1;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

__device__ __forceinline__ bool below_tol(double rnorm, double bnorm, double tol) {
  return rnorm / (bnorm > 0.0 ? bnorm : 1.0) < tol;
}

// Reduces the partials and updates the scalars:
//   op 0 (start):  bnorm = sqrt(d0)               [d0 = (b,b)]
//   op 1 (rnorm):  rnorm = sqrt(d0)               [d0 = (r,r)], stop if below tol
//   op 2 (alpha):  alpha = d0/d1, znomold = d0    [(s,r), (s,Ap)]
//   op 3 (zeta):   zeta = d0/d1                   [(At,t), (At,At)]
//   op 4 (beta):   beta = alpha/zeta*d0/znomold, rnorm = sqrt(d1), iters += 1,
//                  stop if below tol              [(s,r), (r,r)]
// Ops 2-4 do nothing once the stop word is set; a non-finite alpha, zeta,
// beta or rnorm sets it to 2.
__global__ __launch_bounds__(kBlock) void dot2_finish(int nblk, const double *__restrict__ part, int op,
                                                      double tol, double *__restrict__ S) {
  __shared__ double s0[kBlock], s1[kBlock];
  if (op >= 2 && S[kStop] != 0.0) return;  // uniform over the block
  double x0 = 0, x1 = 0;
  for (int i = threadIdx.x; i < nblk; i += kBlock) {
    x0 += part[2 * i];
    x1 += part[2 * i + 1];
  }
  s0[threadIdx.x] = x0;
  s1[threadIdx.x] = x1;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d0 = s0[0], d1 = s1[0];
    switch (op) {
      case 0: S[kBnorm] = sqrt(d0); break;
      case 1:
        S[kRnorm] = sqrt(d0);
        if (below_tol(S[kRnorm], S[kBnorm], tol)) S[kStop] = 1.0;
        break;
      case 2: {
        const double al = d0 / d1;
        S[kAlpha] = al;
        S[kZnomOld] = d0;
        if (!isfinite(al)) S[kStop] = 2.0;
        break;
      }
      case 3: {
        const double ze = d0 / d1;
        S[kZeta] = ze;
        if (!isfinite(ze)) S[kStop] = 2.0;
        break;
      }
      default: {
        const double be = S[kAlpha] / S[kZeta] * d0 / S[kZnomOld];
        const double rn = sqrt(d1);
        S[kBeta] = be;
        S[kRnorm] = rn;
        S[kIters] += 1.0;
        if (below_tol(rn, S[kBnorm], tol))
          S[kStop] = 1.0;  // the reference leaves the loop before beta is used
        else if (!isfinite(rn) || !isfinite(be))
          S[kStop] = 2.0;
        break;
      }
    }
  }
}

// S = 0 (a kernel rather than a memset, so that no memset node enters a
// captured graph)
__global__ void bicg_scalars_zero(double *__restrict__ S) {
  if (threadIdx.x < kNumScalars) S[threadIdx.x] = 0.0;
}

// r = b - r (r holds A u on entry), shadow = r, p = 0, Ap = 0
__global__ __launch_bounds__(kBlock) void bicg_init(int n, const double *__restrict__ b, double *__restrict__ r,
                                                    double *__restrict__ sh, double *__restrict__ p,
                                                    double *__restrict__ ap) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double v = b[i] - r[i];
  r[i] = v;
  sh[i] = v;
  p[i] = 0.0;
  ap[i] = 0.0;
}

// p = r + beta*(p - zeta*Ap)
__global__ __launch_bounds__(kBlock) void bicg_p(int n, const double *__restrict__ r, double *__restrict__ p,
                                                 const double *__restrict__ ap, const double *__restrict__ S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || S[kStop] != 0.0) return;
  p[i] = r[i] + S[kBeta] * (p[i] - S[kZeta] * ap[i]);
}

// t = r - alpha*Ap
__global__ __launch_bounds__(kBlock) void bicg_t(int n, const double *__restrict__ r, const double *__restrict__ ap,
                                                 double *__restrict__ t, const double *__restrict__ S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || S[kStop] != 0.0) return;
  t[i] = r[i] - S[kAlpha] * ap[i];
}

// u += alpha*p + zeta*t ; r = t - zeta*At
__global__ __launch_bounds__(kBlock) void bicg_update(int n, double *__restrict__ u, const double *__restrict__ p,
                                                      const double *__restrict__ t, const double *__restrict__ at,
                                                      double *__restrict__ r, const double *__restrict__ S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || S[kStop] != 0.0) return;
  const double al = S[kAlpha], ze = S[kZeta];
  u[i] = u[i] + al * p[i] + ze * t[i];
  r[i] = t[i] - ze * at[i];
}

// Device vectors and scalars of one solve: b and u are the caller's (or
// copies of the caller's host vectors), the rest is work space.
struct Solver {
  int n = 0;
  double tol = 0;
  hmvm_plan *plan = nullptr;
  hipStream_t st = nullptr;
  const double *b = nullptr;
  double *u = nullptr, *r = nullptr, *sh = nullptr, *p = nullptr, *ap = nullptr, *t = nullptr, *at = nullptr,
         *part = nullptr, *S = nullptr;
  int nblk() const { return std::min(kDotBlocks, (n + kBlock - 1) / kBlock); }
  dim3 grid() const { return dim3((n + kBlock - 1) / kBlock); }

  void dots(const double *a, const double *bb, const double *c, const double *d, int op) {
    hipLaunchKernelGGL(dot2_partial, dim3(nblk()), dim3(kBlock), 0, st, n, a, bb, c, d, part);
    hipLaunchKernelGGL(dot2_finish, dim3(1), dim3(kBlock), 0, st, nblk(), part, op, tol, S);
  }
  // r = b - A u ; shadow = r ; p = Ap = 0 ; norms (m_HACApK_solve.f90:326-339)
  int start() {
    hipLaunchKernelGGL(bicg_scalars_zero, dim3(1), dim3(64), 0, st, S);
    dots(b, b, b, b, 0);
    int rc = hmvm_plan_apply(plan, u, r, 0, st);
    if (rc) return rc;
    hipLaunchKernelGGL(bicg_init, grid(), dim3(kBlock), 0, st, n, b, r, sh, p, ap);
    dots(r, r, r, r, 1);
    KRY_TRY(hipGetLastError());
    return HMVM_OK;
  }
  // one BiCGSTAB iteration (m_HACApK_solve.f90:341-372)
  int iterate() {
    hipLaunchKernelGGL(bicg_p, grid(), dim3(kBlock), 0, st, n, r, p, ap, S);
    int rc = hmvm_plan_apply(plan, p, ap, 0, st);
    if (rc) return rc;
    dots(sh, r, sh, ap, 2);
    hipLaunchKernelGGL(bicg_t, grid(), dim3(kBlock), 0, st, n, r, ap, t, S);
    rc = hmvm_plan_apply(plan, t, at, 0, st);
    if (rc) return rc;
    dots(at, t, at, at, 3);
    hipLaunchKernelGGL(bicg_update, grid(), dim3(kBlock), 0, st, n, u, p, t, at, r, S);
    dots(sh, r, r, r, 4);
    KRY_TRY(hipGetLastError());
    return HMVM_OK;
  }
  int read_scalars(double *S_host) {
    KRY_TRY(hipMemcpyAsync(S_host, S, sizeof(double) * kNumScalars, hipMemcpyDeviceToHost, st));
    KRY_TRY(hipStreamSynchronize(st));
    return HMVM_OK;
  }
};

// The solve loop shared by the entry points: start, then chunks of
// check_every iterations (eager, or replays of one captured iteration) until
// the stop word is set or max_iter iterations have been launched.  `timed`
// brackets the loop with events.
int solve(Solver &s, int max_iter, int use_graph, int check_every, bool timed, hmvm_krylov_result *res) {
  int rc = s.start();
  if (rc) return rc;
  double S0[kNumScalars];
  if ((rc = s.read_scalars(S0))) return rc;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  if (use_graph && S0[kStop] == 0.0 && max_iter > 0) {
    KRY_TRY(hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal));
    rc = s.iterate();
    const hipError_t e = hipStreamEndCapture(s.st, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc ? rc : hip_fail("hipStreamEndCapture", e);
    }
    const hipError_t ei = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return hip_fail("hipGraphInstantiate", ei);
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (timed) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s.st);
  }
  int launched = 0;
  while (launched < max_iter && S0[kStop] == 0.0) {
    const int chunk = check_every > 0 ? std::min(check_every, max_iter - launched) : max_iter - launched;
    for (int c = 0; c < chunk && !rc; c++) {
      if (gexec) {
        const hipError_t e = hipGraphLaunch(gexec, s.st);
        rc = e == hipSuccess ? HMVM_OK : hip_fail("hipGraphLaunch", e);
      } else
        rc = s.iterate();
    }
    if (rc) break;
    launched += chunk;
    if ((rc = s.read_scalars(S0))) break;
  }
  float ms = 0;
  if (timed) {
    (void)hipEventRecord(e1, s.st);
    if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc) return rc;
  const double bnorm = S0[kBnorm] > 0 ? S0[kBnorm] : 1.0;
  res->iters = (int)S0[kIters];
  res->hmvms = 2 * res->iters + 1;
  res->relres = S0[kRnorm] / bnorm;
  res->ms_loop = ms;
  res->ms_per_iter = launched > 0 ? ms / launched : 0.0;
  res->graph_used = gexec ? 1 : 0;
  if (S0[kStop] == 2.0) {
    char m[256];
    std::snprintf(m, sizeof m, "BiCGSTAB broke down after %d iterations (zero divisor or non-finite scalar)",
                  res->iters);
    return hmvm::report_error(HMVM_ERR_BREAKDOWN, m);
  }
  return HMVM_OK;
}

}  // namespace

extern "C" int hmvm_bicgstab(hmvm_plan *plan, const double *b_host, double *u_host, int max_iter, double tol,
                             int use_graph, int check_every, hmvm_krylov_result *res) {
  hmvm::report_clear();
  if (!plan || !b_host || !u_host || !res || max_iter < 0 || check_every < 0)
    return hmvm::report_error(HMVM_ERR_ARG, "hmvm_bicgstab: NULL argument or negative count");
  std::memset(res, 0, sizeof(*res));
  hmvm_plan_info info;
  int rc = hmvm_plan_get_info(plan, &info);
  if (rc) return rc;
  Solver s;
  s.n = info.nd;
  s.tol = tol;
  s.plan = plan;
  const size_t nb = sizeof(double) * (size_t)s.n;
  // b, u, the six work vectors, the dot partials and the scalars
  struct Buffers {
    double *d[10] = {};
    hipStream_t st = nullptr;
    ~Buffers() {
      if (st) (void)hipStreamSynchronize(st);
      for (double *q : d)
        if (q) (void)hipFree(q);
      if (st) (void)hipStreamDestroy(st);
    }
  } m;
  KRY_TRY(hipStreamCreate(&m.st));
  for (int i = 0; i < 10; i++) {
    const size_t bytes = i < 8 ? nb : i == 8 ? sizeof(double) * 2 * kDotBlocks : sizeof(double) * kNumScalars;
    KRY_TRY(hipMalloc(&m.d[i], bytes));
  }
  s.st = m.st;
  s.b = m.d[0];
  s.u = m.d[1];
  s.r = m.d[2];
  s.sh = m.d[3];
  s.p = m.d[4];
  s.ap = m.d[5];
  s.t = m.d[6];
  s.at = m.d[7];
  s.part = m.d[8];
  s.S = m.d[9];
  KRY_TRY(hipMemcpyAsync(m.d[0], b_host, nb, hipMemcpyHostToDevice, s.st));
  KRY_TRY(hipMemcpyAsync(s.u, u_host, nb, hipMemcpyHostToDevice, s.st));
  rc = solve(s, max_iter, use_graph, check_every, true, res);
  if (rc == HMVM_OK || rc == HMVM_ERR_BREAKDOWN) {
    const hipError_t e = hipMemcpy(u_host, s.u, nb, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail("hipMemcpy (u)", e);
  }
  return rc;
}

// ---- the allocation-free variant ----------------------------------------------
// Work space of hmvm_bicgstab_static: the six work vectors of a solve of at
// most kStaticMaxN unknowns, the dot partials and the scalars, in one array
// of the library's code object (48 MiB + 8 KiB; each device the library runs
// on holds its own copy until the process ends).  Every call on a device
// uses that device's copy: calls that overlap on one device race.
constexpr int kStaticMaxN = HMVM_BICGSTAB_STATIC_MAX_N;
constexpr size_t kScratchDoubles = 6 * (size_t)kStaticMaxN + 2 * kDotBlocks + kNumScalars;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, resolved once per
// device (host code never passes the symbol's own name to a kernel: on the
// host it is only a stand-in).
int scratch_on_current_device(double **out) {
  int dev = -1;
  KRY_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return hmvm::report_error(HMVM_ERR_UNSUPPORTED, "hmvm_bicgstab_static: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    KRY_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return HMVM_OK;
}

}  // namespace

extern "C" int hmvm_bicgstab_static(hmvm_plan *plan, const double *d_b, double *d_u, int max_iter, double tol,
                                    int check_every, void *stream, hmvm_krylov_result *res) {
  hmvm::report_clear();
  if (!plan || !d_b || !d_u || !res || max_iter < 0 || check_every < 0)
    return hmvm::report_error(HMVM_ERR_ARG, "hmvm_bicgstab_static: NULL argument or negative count");
  std::memset(res, 0, sizeof(*res));
  hmvm_plan_info info;
  int rc = hmvm_plan_get_info(plan, &info);
  if (rc) return rc;
  if (info.nd > kStaticMaxN) {
    char m[256];
    std::snprintf(m, sizeof m, "hmvm_bicgstab_static: nd = %d exceeds the fixed work space (at most %d unknowns); "
                  "use hmvm_bicgstab", info.nd, kStaticMaxN);
    return hmvm::report_error(HMVM_ERR_ARG, m);
  }
  double *w = nullptr;
  if ((rc = scratch_on_current_device(&w))) return rc;
  Solver s;
  s.n = info.nd;
  s.tol = tol;
  s.plan = plan;
  s.st = (hipStream_t)stream;
  s.b = d_b;
  s.u = d_u;
  const size_t n = (size_t)s.n;
  s.r = w;
  s.sh = w + n;
  s.p = w + 2 * n;
  s.ap = w + 3 * n;
  s.t = w + 4 * n;
  s.at = w + 5 * n;
  s.part = w + 6 * n;
  s.S = s.part + 2 * kDotBlocks;
  return solve(s, max_iter, 0, check_every, false, res);
}
