// hmvm_krylov.hip -- BiCGSTAB on the device with the HIP HMVM as its operator.
//
// Restates the iteration of the reference's HACApK_bicgstab_lfmtx_hyp
// (src_hdump/m_HACApK_solve.f90:306-380): two HMVMs per iteration, vector
// updates and dot products between them.  Every scalar (alpha, zeta, beta,
// the residual norm) stays in device memory, so one iteration is a fixed
// sequence of launches with no host synchronisation; with use_graph the
// iteration is captured once into a hipGraph and replayed.  Dot products are
// two-pass (per-block partials, then one block) and therefore deterministic:
// the eager and the graph-replayed solve give the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hmvm_hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kDotBlocks = 512;

// device scalars
enum {
  kAlpha = 0, kBeta, kZeta, kZnomOld, kRnorm, kBnorm, kNumScalars = 8
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

// Reduces the partials and updates the scalars:
//   op 0 (start):  bnorm = sqrt(d0)               [d0 = (b,b)]
//   op 1 (rnorm):  rnorm = sqrt(d0)               [d0 = (r,r)]
//   op 2 (alpha):  alpha = d0/d1, znomold = d0    [(s,r), (s,Ap)]
//   op 3 (zeta):   zeta = d0/d1                   [(At,t), (At,At)]
//   op 4 (beta):   beta = alpha/zeta*d0/znomold, rnorm = sqrt(d1)   [(s,r), (r,r)]
__global__ __launch_bounds__(kBlock) void dot2_finish(int nblk, const double *__restrict__ part, int op,
                                                      double *__restrict__ S) {
  __shared__ double s0[kBlock], s1[kBlock];
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
      case 1: S[kRnorm] = sqrt(d0); break;
      case 2: S[kAlpha] = d0 / d1; S[kZnomOld] = d0; break;
      case 3: S[kZeta] = d0 / d1; break;
      default:
        S[kBeta] = S[kAlpha] / S[kZeta] * d0 / S[kZnomOld];
        S[kRnorm] = sqrt(d1);
        break;
    }
  }
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
  if (i >= n) return;
  p[i] = r[i] + S[kBeta] * (p[i] - S[kZeta] * ap[i]);
}

// t = r - alpha*Ap
__global__ __launch_bounds__(kBlock) void bicg_t(int n, const double *__restrict__ r, const double *__restrict__ ap,
                                                 double *__restrict__ t, const double *__restrict__ S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  t[i] = r[i] - S[kAlpha] * ap[i];
}

// u += alpha*p + zeta*t ; r = t - zeta*At
__global__ __launch_bounds__(kBlock) void bicg_update(int n, double *__restrict__ u, const double *__restrict__ p,
                                                      const double *__restrict__ t, const double *__restrict__ at,
                                                      double *__restrict__ r, const double *__restrict__ S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double al = S[kAlpha], ze = S[kZeta];
  u[i] = u[i] + al * p[i] + ze * t[i];
  r[i] = t[i] - ze * at[i];
}

struct Solver {
  int n = 0;
  hmvm_plan *plan = nullptr;
  hipStream_t st = nullptr;
  double *b = nullptr, *u = nullptr, *r = nullptr, *sh = nullptr, *p = nullptr, *ap = nullptr, *t = nullptr,
         *at = nullptr, *part = nullptr, *S = nullptr;
  int nblk() const { return std::min(kDotBlocks, (n + kBlock - 1) / kBlock); }
  dim3 grid() const { return dim3((n + kBlock - 1) / kBlock); }

  void dots(const double *a, const double *bb, const double *c, const double *d, int op) {
    hipLaunchKernelGGL(dot2_partial, dim3(nblk()), dim3(kBlock), 0, st, n, a, bb, c, d, part);
    hipLaunchKernelGGL(dot2_finish, dim3(1), dim3(kBlock), 0, st, nblk(), part, op, S);
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
    return hipGetLastError() == hipSuccess ? HMVM_OK : HMVM_ERR_HIP;
  }
  ~Solver() {
    double *bufs[] = {b, u, r, sh, p, ap, t, at, part, S};
    for (double *q : bufs)
      if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
};

}  // namespace

extern "C" int hmvm_bicgstab(hmvm_plan *plan, const double *b_host, double *u_host, int max_iter, double tol,
                             int use_graph, int check_every, hmvm_krylov_result *res) {
  if (!plan || !b_host || !u_host || !res || max_iter < 0 || check_every < 0) return HMVM_ERR_ARG;
  std::memset(res, 0, sizeof(*res));
  hmvm_plan_info info;
  int rc = hmvm_plan_get_info(plan, &info);
  if (rc) return rc;
  Solver s;
  s.n = info.nd;
  s.plan = plan;
  const size_t nb = sizeof(double) * (size_t)s.n;
  if (hipStreamCreate(&s.st) != hipSuccess) return HMVM_ERR_HIP;
  double **bufs[] = {&s.b, &s.u, &s.r, &s.sh, &s.p, &s.ap, &s.t, &s.at};
  for (double **q : bufs)
    if (hipMalloc(q, nb) != hipSuccess) return HMVM_ERR_HIP;
  if (hipMalloc(&s.part, sizeof(double) * 2 * kDotBlocks) != hipSuccess) return HMVM_ERR_HIP;
  if (hipMalloc(&s.S, sizeof(double) * kNumScalars) != hipSuccess) return HMVM_ERR_HIP;
  if (hipMemcpyAsync(s.b, b_host, nb, hipMemcpyHostToDevice, s.st) != hipSuccess) return HMVM_ERR_HIP;
  if (hipMemcpyAsync(s.u, u_host, nb, hipMemcpyHostToDevice, s.st) != hipSuccess) return HMVM_ERR_HIP;
  {
    const double zero[kNumScalars] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(s.S, zero, sizeof(zero), hipMemcpyHostToDevice, s.st) != hipSuccess) return HMVM_ERR_HIP;
  }
  // r = b - A u ; shadow = r ; p = Ap = 0 ; norms (m_HACApK_solve.f90:326-339)
  s.dots(s.b, s.b, s.b, s.b, 0);
  if ((rc = hmvm_plan_apply(plan, s.u, s.r, 0, s.st))) return rc;
  hipLaunchKernelGGL(bicg_init, s.grid(), dim3(kBlock), 0, s.st, s.n, s.b, s.r, s.sh, s.p, s.ap);
  s.dots(s.r, s.r, s.r, s.r, 1);
  double S0[kNumScalars];
  if (hipMemcpyAsync(S0, s.S, sizeof(S0), hipMemcpyDeviceToHost, s.st) != hipSuccess ||
      hipStreamSynchronize(s.st) != hipSuccess)
    return HMVM_ERR_HIP;
  const double bnorm = S0[kBnorm] > 0 ? S0[kBnorm] : 1.0;
  double relres = S0[kRnorm] / bnorm;

  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  if (use_graph) {
    if (hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal) != hipSuccess) return HMVM_ERR_HIP;
    rc = s.iterate();
    hipError_t e = hipStreamEndCapture(s.st, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc ? rc : HMVM_ERR_HIP;
    }
    if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return HMVM_ERR_HIP;
    }
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipEventRecord(e0, s.st);
  int it = 0;
  while (it < max_iter && !(relres < tol)) {
    const int chunk = check_every > 0 ? std::min(check_every, max_iter - it) : max_iter - it;
    for (int c = 0; c < chunk && !rc; c++) {
      if (use_graph)
        rc = hipGraphLaunch(gexec, s.st) == hipSuccess ? HMVM_OK : HMVM_ERR_HIP;
      else
        rc = s.iterate();
    }
    if (rc) break;
    it += chunk;
    double rn = 0;
    if (hipMemcpyAsync(&rn, s.S + kRnorm, sizeof(double), hipMemcpyDeviceToHost, s.st) != hipSuccess ||
        hipStreamSynchronize(s.st) != hipSuccess) {
      rc = HMVM_ERR_HIP;
      break;
    }
    relres = rn / bnorm;
  }
  hipEventRecord(e1, s.st);
  float ms = 0;
  if (hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc) return rc;
  if (hipMemcpy(u_host, s.u, nb, hipMemcpyDeviceToHost) != hipSuccess) return HMVM_ERR_HIP;
  res->iters = it;
  res->hmvms = 2 * it + 1;
  res->relres = relres;
  res->ms_loop = ms;
  res->ms_per_iter = it > 0 ? ms / it : 0.0;
  res->graph_used = use_graph ? 1 : 0;
  return HMVM_OK;
}
