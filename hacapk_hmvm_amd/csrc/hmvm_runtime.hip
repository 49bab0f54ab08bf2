// hmvm_runtime.hip -- plan upload, launch sequence and the C ABI of the HIP HMVM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hacapk.h"
#include "hmvm_hip.h"
#include "hmvm_kernels.hpp"
#include "hmvm_plan.hpp"

using namespace hmvm;

// ---------------------------------------------------------------------------
// error state
namespace {
thread_local int g_err = HMVM_OK;
thread_local char g_msg[1024] = "ok";

int set_error(int code, const char *fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  if (code != HMVM_OK) std::fprintf(stderr, "hmvm_hip: %s\n", g_msg);
  return code;
}
void clear_error() {
  g_err = HMVM_OK;
  std::snprintf(g_msg, sizeof g_msg, "ok");
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return set_error(HMVM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

bool verbose() {
  static int v = -1;
  if (v < 0) {
    const char *e = std::getenv("HMVM_VERBOSE");
    v = (e && *e && *e != '0') ? 1 : 0;
  }
  return v == 1;
}

template <class T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return HMVM_OK;
  HIP_TRY(hipMalloc((void **)p, sizeof(T) * n));
  return HMVM_OK;
}
template <class T>
int h2d(T *d, const T *h, size_t n) {
  if (n == 0) return HMVM_OK;
  HIP_TRY(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  return HMVM_OK;
}
template <class T>
int upload(T **d, const std::vector<T> &h) {
  int rc = dalloc(d, h.size());
  if (rc) return rc;
  return h2d(*d, h.data(), h.size());
}

int elem_bytes(int storage) { return storage == kF64 ? 8 : storage == kF32 ? 4 : 2; }

}  // namespace

// ---------------------------------------------------------------------------
// the plan object
struct hmvm_plan {
  Plan plan;
  int storage = kF64;
  int block_threads = 256;
  int kernel = 0;  // tile-kernel variant, 0 .. kNumKernelVariants-1
  int mode = 0;    // 0: launches A, [t sum], B, y sum; 1: one persistent launch
  int persist_key = -1, persist_blocks = 0;  // occupancy cache (kernel variant, block size)
  // several right-hand sides: buffers sized for multi_cap columns
  int multi_cap = 0;
  double *d_mslots = nullptr, *d_mt = nullptr, *d_mX = nullptr, *d_mY = nullptr;
  PersistCounters *d_ctr = nullptr;
  int ncu = 0;     // compute units of the device
  int device = 0;
  hipStream_t stream = nullptr;
  void *d_vals = nullptr;
  TileDesc *d_tiles_a = nullptr, *d_tiles_b = nullptr;
  PanelDesc *d_panels = nullptr;
  double *d_t = nullptr, *d_slots = nullptr;  // stage-1 partial slots, then output slots
  int32_t *d_tseg = nullptr, *d_tptr = nullptr, *d_tfac = nullptr;
  int64_t *d_tlist = nullptr;
  int32_t *d_yseg = nullptr, *d_yptr = nullptr;
  int64_t *d_ylist = nullptr;
  float *d_scales = nullptr;
  double *d_x = nullptr, *d_y = nullptr;  // staging for host-pointer calls
  size_t device_bytes = 0;

  ~hmvm_plan() {
    void *ptrs[] = {d_vals,  d_tiles_a, d_tiles_b, d_panels, d_t,      d_slots, d_tseg, d_tptr,
                    d_tfac,  d_tlist,   d_yseg,    d_yptr,   d_ylist,  d_scales, d_x,    d_y,
                    d_ctr,   d_mslots,  d_mt,      d_mX,     d_mY};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Source view of the caller's matrix: descriptor arrays + one contiguous FP64
// value buffer in host memory (rowmat, rowmat_t or a gathered copy of the
// per-leaf arrays).
struct Source {
  LeafArrays la;
  std::vector<int64_t> a1, a2;
  const double *vals = nullptr;
  std::vector<double> gathered;
  std::vector<int> ltmtx, kt, ndl, ndt, nstrtl, nstrtt;
};

int source_from_matrix2(const matrix2 *m, int a2_source, Source *s) {
  if (!m || m->nd < 1 || m->nlf < 1 || !m->ltmtx || !m->kt || !m->ndl || !m->ndt || !m->nstrtl ||
      !m->nstrtt || !m->a1 || !m->a2)
    return set_error(HMVM_ERR_ARG, "matrix2 is NULL or incomplete");
  const double *vals = a2_source == kA2Transposed ? m->rowmat_t : m->rowmat;
  if (!vals) return set_error(HMVM_ERR_ARG, "matrix2 has no %s", a2_source ? "rowmat_t" : "rowmat");
  s->a1.resize(m->nlf);
  s->a2.resize(m->nlf);
  for (int i = 0; i < m->nlf; i++) {
    s->a1[i] = m->a1[i];
    s->a2[i] = m->a2[i];
  }
  LeafArrays &la = s->la;
  la.nd = m->nd;
  la.nlf = m->nlf;
  la.ltmtx = m->ltmtx;
  la.kt = m->kt;
  la.ndl = m->ndl;
  la.ndt = m->ndt;
  la.nstrtl = m->nstrtl;
  la.nstrtt = m->nstrtt;
  la.src_a1 = s->a1.data();
  la.src_a2 = s->a2.data();
  la.a2_source = a2_source;
  la.src_len = m->len;
  s->vals = vals;
  return HMVM_OK;
}

// matrix (array of structs): gather every leaf into one buffer in the packing
// order of the reference matrix2 (src_c/loadmatrix.c:98-128).
int source_from_matrix(const matrix *m, int a2_source, Source *s) {
  if (!m || m->nd < 1 || m->nlf < 1 || !m->submat) return set_error(HMVM_ERR_ARG, "matrix is NULL or incomplete");
  const int nlf = m->nlf;
  s->ltmtx.resize(nlf); s->kt.resize(nlf); s->ndl.resize(nlf); s->ndt.resize(nlf);
  s->nstrtl.resize(nlf); s->nstrtt.resize(nlf); s->a1.resize(nlf); s->a2.resize(nlf);
  int64_t off = 0;
  for (int i = 0; i < nlf; i++) {
    const submatrix &q = m->submat[i];
    s->ltmtx[i] = q.ltmtx; s->kt[i] = q.kt; s->ndl[i] = q.ndl; s->ndt[i] = q.ndt;
    s->nstrtl[i] = q.nstrtl; s->nstrtt[i] = q.nstrtt;
    s->a1[i] = off;
    if (q.ltmtx == 1) {
      if (q.kt < 0 || q.ndl < 0 || q.ndt < 0) return set_error(HMVM_ERR_MATRIX, "leaf %d: negative size", i);
      off += (int64_t)q.kt * q.ndt;
      s->a2[i] = off;
      off += (int64_t)q.kt * q.ndl;
    } else {
      if (q.ndl < 0 || q.ndt < 0) return set_error(HMVM_ERR_MATRIX, "leaf %d: negative size", i);
      s->a2[i] = 0;
      off += (int64_t)q.ndl * q.ndt;
    }
  }
  try {
    s->gathered.resize((size_t)std::max<int64_t>(off, 1));
  } catch (...) {
    return set_error(HMVM_ERR_NOMEM, "cannot gather %lld elements", (long long)off);
  }
  for (int i = 0; i < nlf; i++) {
    const submatrix &q = m->submat[i];
    if (q.ltmtx == 1) {
      const double *a2 = a2_source == kA2Transposed ? q.a2t : q.a2;
      if ((q.kt > 0) && (!q.a1 || !a2)) return set_error(HMVM_ERR_ARG, "leaf %d: missing factor", i);
      if (q.kt > 0) {
        std::memcpy(&s->gathered[s->a1[i]], q.a1, sizeof(double) * (size_t)q.kt * q.ndt);
        std::memcpy(&s->gathered[s->a2[i]], a2, sizeof(double) * (size_t)q.kt * q.ndl);
      }
    } else {
      if (!q.a1) return set_error(HMVM_ERR_ARG, "leaf %d: missing a1", i);
      std::memcpy(&s->gathered[s->a1[i]], q.a1, sizeof(double) * (size_t)q.ndl * q.ndt);
    }
  }
  LeafArrays &la = s->la;
  la.nd = m->nd;
  la.nlf = nlf;
  la.ltmtx = s->ltmtx.data(); la.kt = s->kt.data(); la.ndl = s->ndl.data(); la.ndt = s->ndt.data();
  la.nstrtl = s->nstrtl.data(); la.nstrtt = s->nstrtt.data();
  la.src_a1 = s->a1.data();
  la.src_a2 = s->a2.data();
  la.a2_source = a2_source;
  la.src_len = off;
  s->vals = s->gathered.data();
  return HMVM_OK;
}

PlanOptions plan_options(const hmvm_options *o) {
  PlanOptions po;
  if (o) {
    po.storage = o->storage;
    if (o->w1 > 0) po.w1 = o->w1;
    if (o->w2 > 0) po.w2 = o->w2;
    po.order = o->order;
  }
  return po;
}

static_assert(sizeof(hmvm_tile) == sizeof(TileDesc), "tile record layout");
static_assert(sizeof(hmvm_panel) == sizeof(PanelDesc), "panel record layout");
static_assert(sizeof(hmvm_repack) == sizeof(RepackDesc), "repack record layout");

template <class T>
T *copy_out(const std::vector<T> &v) {
  T *p = (T *)std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

void export_segsum(const SegSum &s, hmvm_segsum *o) {
  o->n = s.n;
  o->nseg = (int)s.ptr.size() - 1;
  o->seg = copy_out(s.seg);
  o->ptr = copy_out(s.ptr);
  o->list = (long long *)copy_out(s.list);
}

template <typename S>
int launch_repack(hmvm_plan *p, const double *d_src, const RepackDesc *d_desc, const RepackChunk *d_chunks,
                  int nchunks, const int32_t *d_chunk_begin, int ndesc, float *d_chunk_max) {
  const bool scaled = p->storage == kBF16;
  if (scaled) {
    hipLaunchKernelGGL(absmax_kernel, dim3(nchunks), dim3(256), 0, p->stream, d_desc, d_chunks, d_src, d_chunk_max);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(scale_kernel, dim3((ndesc + 255) / 256), dim3(256), 0, p->stream, ndesc, d_desc,
                       d_chunk_begin, d_chunk_max, p->d_scales);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((repack_kernel<S, true>), dim3(nchunks), dim3(256), 0, p->stream, d_desc, d_chunks, d_src,
                       p->d_scales, (S *)p->d_vals);
  } else {
    hipLaunchKernelGGL((repack_kernel<S, false>), dim3(nchunks), dim3(256), 0, p->stream, d_desc, d_chunks, d_src,
                       (const float *)nullptr, (S *)p->d_vals);
  }
  HIP_TRY(hipGetLastError());
  return HMVM_OK;
}

int create_plan(Source &src, const hmvm_options *opt, hmvm_plan **out) {
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(HMVM_ERR_NODEVICE, "no HIP device visible");
  const PlanOptions po = plan_options(opt);
  if (po.storage < kF64 || po.storage > kBF16) return set_error(HMVM_ERR_ARG, "unknown storage %d", po.storage);
  std::unique_ptr<hmvm_plan> p(new hmvm_plan());
  p->storage = po.storage;
  if (opt && opt->block_threads > 0) {
    if (opt->block_threads % 64 || opt->block_threads > 256 || opt->block_threads < 64)
      return set_error(HMVM_ERR_ARG, "block_threads must be 64, 128, 192 or 256");
    p->block_threads = opt->block_threads;
  }
  if (opt) {
    if (opt->kernel < 0 || opt->kernel >= kNumKernelVariants)
      return set_error(HMVM_ERR_ARG, "kernel variant must be in [0,%d)", kNumKernelVariants);
    p->kernel = opt->kernel;
    if (opt->mode < 0 || opt->mode > 1) return set_error(HMVM_ERR_ARG, "mode must be 0 or 1");
    p->mode = opt->mode;
  }
  std::string err = build_plan(src.la, po, &p->plan);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "%s", err.c_str());
  err = verify_plan(p->plan, src.la);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "schedule check failed: %s", err.c_str());
  Plan &P = p->plan;
  HIP_TRY(hipGetDevice(&p->device));
  HIP_TRY(hipStreamCreate(&p->stream));
  {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, p->device));
    p->ncu = prop.multiProcessorCount;
  }
  HIP_TRY(hipMalloc(&p->d_ctr, sizeof(PersistCounters)));
  HIP_TRY(hipMemset(p->d_ctr, 0, sizeof(PersistCounters)));

  int rc;
  const size_t eb = (size_t)elem_bytes(p->storage);
  // schedule
  if ((rc = upload(&p->d_tiles_a, P.tiles_a)) || (rc = upload(&p->d_tiles_b, P.tiles_b)) ||
      (rc = upload(&p->d_panels, P.panels)) || (rc = upload(&p->d_tseg, P.tsum.seg)) ||
      (rc = upload(&p->d_tptr, P.tsum.ptr)) || (rc = upload(&p->d_tlist, P.tsum.list)) ||
      (rc = upload(&p->d_yseg, P.ysum.seg)) || (rc = upload(&p->d_yptr, P.ysum.ptr)) ||
      (rc = upload(&p->d_ylist, P.ysum.list)))
    return rc;
  if (p->storage == kBF16) {
    if ((rc = upload(&p->d_tfac, P.factor_of_trow))) return rc;
    if ((rc = dalloc(&p->d_scales, (size_t)P.nfactors))) return rc;
    std::vector<float> ones((size_t)P.nfactors, 1.0f);
    if ((rc = h2d(p->d_scales, ones.data(), ones.size()))) return rc;
  }
  if ((rc = dalloc(&p->d_t, (size_t)P.n_t)) || (rc = dalloc(&p->d_slots, (size_t)(P.n_tp + P.n_z))) ||
      (rc = dalloc(&p->d_x, (size_t)P.nd)) ||
      (rc = dalloc(&p->d_y, (size_t)P.nd)))
    return rc;
  HIP_TRY(hipMalloc(&p->d_vals, eb * (size_t)std::max<int64_t>(P.nvals, 1)));
  // column padding is read (and discarded) by the 16-byte lane loads: keep it zero
  HIP_TRY(hipMemsetAsync(p->d_vals, 0, eb * (size_t)std::max<int64_t>(P.nvals, 1), p->stream));

  // values: upload the caller's buffer once, repack on the GPU
  std::vector<RepackChunk> chunks;
  std::vector<int32_t> chunk_begin;
  chunk_begin.reserve(P.repack.size() + 1);
  for (size_t d = 0; d < P.repack.size(); d++) {
    chunk_begin.push_back((int32_t)chunks.size());
    const int64_t ne = (int64_t)P.repack[d].rows * P.repack[d].cols;
    for (int64_t e0 = 0; e0 < ne; e0 += kRepackChunk) chunks.push_back({(int32_t)d, (int32_t)e0});
  }
  chunk_begin.push_back((int32_t)chunks.size());
  double *d_src = nullptr;
  RepackDesc *d_desc = nullptr;
  RepackChunk *d_chunks = nullptr;
  int32_t *d_chunk_begin = nullptr;
  float *d_chunk_max = nullptr;
  auto cleanup = [&]() {
    if (d_src) (void)hipFree(d_src);
    if (d_desc) (void)hipFree(d_desc);
    if (d_chunks) (void)hipFree(d_chunks);
    if (d_chunk_begin) (void)hipFree(d_chunk_begin);
    if (d_chunk_max) (void)hipFree(d_chunk_max);
  };
  rc = dalloc(&d_src, (size_t)std::max<int64_t>(src.la.src_len, 1));
  if (!rc) rc = h2d(d_src, src.vals, (size_t)src.la.src_len);
  if (!rc) rc = upload(&d_desc, P.repack);
  if (!rc) rc = upload(&d_chunks, chunks);
  if (!rc) rc = upload(&d_chunk_begin, chunk_begin);
  if (!rc && p->storage == kBF16) rc = dalloc(&d_chunk_max, chunks.size());
  if (!rc && !chunks.empty()) {
    const int nch = (int)chunks.size(), nd = (int)P.repack.size();
    if (p->storage == kF64)
      rc = launch_repack<double>(p.get(), d_src, d_desc, d_chunks, nch, d_chunk_begin, nd, d_chunk_max);
    else if (p->storage == kF32)
      rc = launch_repack<float>(p.get(), d_src, d_desc, d_chunks, nch, d_chunk_begin, nd, d_chunk_max);
    else
      rc = launch_repack<uint16_t>(p.get(), d_src, d_desc, d_chunks, nch, d_chunk_begin, nd, d_chunk_max);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = set_error(HMVM_ERR_HIP, "repack failed: %s", hipGetErrorString(e));
  }
  cleanup();
  if (rc) return rc;

  p->device_bytes = eb * (size_t)P.nvals + sizeof(TileDesc) * (P.tiles_a.size() + P.tiles_b.size()) +
                    sizeof(PanelDesc) * P.panels.size() + 8 * (size_t)(P.n_t + P.n_tp + P.n_z + 2 * P.nd) +
                    4 * (P.tsum.seg.size() + P.tsum.ptr.size() + P.ysum.seg.size() + P.ysum.ptr.size()) +
                    8 * (P.tsum.list.size() + P.ysum.list.size());
  if (verbose())
    std::fprintf(stderr,
                 "hmvm_hip: plan nd=%d nlf=%d len=%lld storage=%d tiles=%zu+%zu panels=%zu t=%lld tslots=%lld "
                 "zslots=%lld\n",
                 P.nd, P.nlf, (long long)P.len, p->storage, P.tiles_a.size(), P.tiles_b.size(), P.panels.size(),
                 (long long)P.n_t, (long long)P.n_tp, (long long)P.n_z);
  *out = p.release();
  return HMVM_OK;
}

template <typename S, bool kScaled, int kVar>
void launch_tiles_v(hmvm_plan *p, dim3 grid, dim3 block, hipStream_t st, const TileDesc *d_tiles, int ntiles,
                    const double *x, const double *t, const double *tp, double *out) {
  constexpr int L = (kVar & 1) ? 16 : 8;
  constexpr bool kPipe = (kVar & 2) != 0, kNT = (kVar & 4) != 0;
  hipLaunchKernelGGL((tile_kernel<S, kScaled, L, kPipe, kNT>), grid, block, 0, st, d_tiles, ntiles, p->d_panels,
                     (const S *)p->d_vals, x, t, tp, (const float *)p->d_scales, out);
}

template <typename S, bool kScaled>
void launch_tiles_s(hmvm_plan *p, dim3 grid, dim3 block, hipStream_t st, const TileDesc *d_tiles, int ntiles,
                    const double *x, const double *t, const double *tp, double *out) {
  switch (p->kernel) {
    case 0: launch_tiles_v<S, kScaled, 0>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    case 1: launch_tiles_v<S, kScaled, 1>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    case 2: launch_tiles_v<S, kScaled, 2>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    case 3: launch_tiles_v<S, kScaled, 3>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    case 4: launch_tiles_v<S, kScaled, 4>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    case 5: launch_tiles_v<S, kScaled, 5>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    case 6: launch_tiles_v<S, kScaled, 6>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
    default: launch_tiles_v<S, kScaled, 7>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out); break;
  }
}

template <typename S>
void launch_tiles(hmvm_plan *p, const TileDesc *d_tiles, int ntiles, const double *x, const double *t,
                  const double *tp, double *out, bool scaled, hipStream_t st) {
  if (ntiles == 0) return;
  const int wpb = p->block_threads / 64;
  const dim3 grid((ntiles + wpb - 1) / wpb), block(p->block_threads);
  if (scaled)
    launch_tiles_s<S, true>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out);
  else
    launch_tiles_s<S, false>(p, grid, block, st, d_tiles, ntiles, x, t, tp, out);
}

template <typename S>
void launch_tiles_both(hmvm_plan *p, int which, const double *x, hipStream_t st) {
  const Plan &P = p->plan;
  const bool bf = p->storage == kBF16;
  if (which == 0)
    launch_tiles<S>(p, p->d_tiles_a, (int)P.tiles_a.size(), x, nullptr, nullptr, p->d_slots, bf, st);
  else
    launch_tiles<S>(p, p->d_tiles_b, (int)P.tiles_b.size(), x, p->d_t, p->d_slots, p->d_slots, bf, st);
}

// Launch `phase` of one HMVM: 0 = tiles A (stage 1 + dense), 1 = t segment sum
// (only when some leaf's rank rows straddle two row groups), 2 = tiles B
// (stage 2), 3 = y segment sum.
void launch_phase(hmvm_plan *p, int phase, const double *x, double *y, int accumulate, hipStream_t st) {
  const Plan &P = p->plan;
  const bool bf = p->storage == kBF16;
  if (phase == 0 || phase == 2) {
    const int which = phase == 0 ? 0 : 1;
    if (p->storage == kF64)
      launch_tiles_both<double>(p, which, x, st);
    else if (p->storage == kF32)
      launch_tiles_both<float>(p, which, x, st);
    else
      launch_tiles_both<uint16_t>(p, which, x, st);
  } else if (phase == 1) {
    if (P.t_fused || P.tsum.n == 0) return;
    if (bf)
      hipLaunchKernelGGL(segsum_kernel<true>, dim3((P.tsum.n + 255) / 256), dim3(256), 0, st, P.tsum.n, p->d_tseg,
                         p->d_tptr, p->d_tlist, p->d_slots, p->d_tfac, p->d_scales, p->d_t, 0);
    else
      hipLaunchKernelGGL(segsum_kernel<false>, dim3((P.tsum.n + 255) / 256), dim3(256), 0, st, P.tsum.n, p->d_tseg,
                         p->d_tptr, p->d_tlist, p->d_slots, (const int32_t *)nullptr, (const float *)nullptr,
                         p->d_t, 0);
  } else {
    hipLaunchKernelGGL(segsum_kernel<false>, dim3((P.ysum.n + 255) / 256), dim3(256), 0, st, P.ysum.n, p->d_yseg,
                       p->d_yptr, p->d_ylist, p->d_slots, (const int32_t *)nullptr, (const float *)nullptr, y,
                       accumulate);
  }
}

// ---- several right-hand sides -------------------------------------------------
int reserve_multi(hmvm_plan *p, int nrhs) {
  if (nrhs != 16 && nrhs != 32 && nrhs != 64) return set_error(HMVM_ERR_ARG, "nrhs must be 16, 32 or 64");
  if (p->multi_cap >= nrhs) return HMVM_OK;
  const Plan &P = p->plan;
  double *bufs[4] = {p->d_mslots, p->d_mt, p->d_mX, p->d_mY};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  p->d_mslots = p->d_mt = p->d_mX = p->d_mY = nullptr;
  p->multi_cap = 0;
  int rc;
  if ((rc = dalloc(&p->d_mslots, (size_t)(P.n_tp + P.n_z) * nrhs)) || (rc = dalloc(&p->d_mt, (size_t)P.n_t * nrhs)) ||
      (rc = dalloc(&p->d_mX, (size_t)P.nd * nrhs)) || (rc = dalloc(&p->d_mY, (size_t)P.nd * nrhs)))
    return rc;
  p->multi_cap = nrhs;
  return HMVM_OK;
}

template <typename S, bool kScaled, int NB>
void launch_multi_tiles(hmvm_plan *p, const TileDesc *d_tiles, int ntiles, const double *X, hipStream_t st) {
  if (ntiles == 0) return;
  const int wpb = p->block_threads / 64;
  hipLaunchKernelGGL((tile_multi_kernel<S, kScaled, NB>), dim3((ntiles + wpb - 1) / wpb), dim3(p->block_threads), 0,
                     st, d_tiles, ntiles, p->d_panels, (const S *)p->d_vals, X, p->d_mt, p->d_mslots,
                     (const float *)p->d_scales, p->d_mslots);
}

template <typename S, bool kScaled>
void launch_multi_nb(hmvm_plan *p, const TileDesc *d_tiles, int ntiles, const double *X, int nrhs, hipStream_t st) {
  if (nrhs == 16)
    launch_multi_tiles<S, kScaled, 1>(p, d_tiles, ntiles, X, st);
  else if (nrhs == 32)
    launch_multi_tiles<S, kScaled, 2>(p, d_tiles, ntiles, X, st);
  else
    launch_multi_tiles<S, kScaled, 4>(p, d_tiles, ntiles, X, st);
}

void launch_multi_phase(hmvm_plan *p, int phase, const double *X, double *Y, int nrhs, int accumulate,
                        hipStream_t st) {
  const Plan &P = p->plan;
  const bool bf = p->storage == kBF16;
  if (phase == 0 || phase == 2) {
    const TileDesc *tl = phase == 0 ? p->d_tiles_a : p->d_tiles_b;
    const int nt = (int)(phase == 0 ? P.tiles_a.size() : P.tiles_b.size());
    if (p->storage == kF64)
      launch_multi_nb<double, false>(p, tl, nt, X, nrhs, st);
    else if (p->storage == kF32)
      launch_multi_nb<float, false>(p, tl, nt, X, nrhs, st);
    else
      launch_multi_nb<uint16_t, true>(p, tl, nt, X, nrhs, st);
  } else if (phase == 1) {
    if (P.t_fused || P.tsum.n == 0) return;
    const int64_t tot = (int64_t)P.tsum.n * nrhs;
    if (bf)
      hipLaunchKernelGGL(segsum_multi_kernel<true>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, P.tsum.n,
                         nrhs, p->d_tseg, p->d_tptr, p->d_tlist, p->d_mslots, p->d_tfac, p->d_scales, p->d_mt, 0);
    else
      hipLaunchKernelGGL(segsum_multi_kernel<false>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, P.tsum.n,
                         nrhs, p->d_tseg, p->d_tptr, p->d_tlist, p->d_mslots, (const int32_t *)nullptr,
                         (const float *)nullptr, p->d_mt, 0);
  } else {
    const int64_t tot = (int64_t)P.ysum.n * nrhs;
    hipLaunchKernelGGL(segsum_multi_kernel<false>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, P.ysum.n,
                       nrhs, p->d_yseg, p->d_yptr, p->d_ylist, p->d_mslots, (const int32_t *)nullptr,
                       (const float *)nullptr, Y, accumulate);
  }
}

int apply_multi(hmvm_plan *p, const double *X, double *Y, int nrhs, int accumulate, hipStream_t st) {
  if (p->multi_cap < nrhs) return set_error(HMVM_ERR_ARG, "call hmvm_plan_reserve_multi(plan, %d) first", nrhs);
  for (int ph = 0; ph < 4; ph++) launch_multi_phase(p, ph, X, Y, nrhs, accumulate, st);
  HIP_TRY(hipGetLastError());
  return HMVM_OK;
}

template <typename S, bool kScaled, int kVar>
int launch_persistent_v(hmvm_plan *p, const double *x, double *y, int accumulate, hipStream_t st) {
  constexpr int L = (kVar & 1) ? 16 : 8;
  constexpr bool kPipe = (kVar & 2) != 0, kNT = (kVar & 4) != 0;
  const Plan &P = p->plan;
  auto kfn = persistent_kernel<S, kScaled, L, kPipe, kNT>;
  const int key = kVar * 1024 + p->block_threads;
  if (p->persist_key != key) {  // resident workgroups per CU, queried once per launch shape
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, p->block_threads, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    p->persist_blocks = per_cu * std::max(p->ncu, 1);
    p->persist_key = key;
  }
  const int blocks = p->persist_blocks;
  const int nwaves = blocks * (p->block_threads / 64);
  const int ns1 = P.n_stage1_tiles;
  const int na = (int)P.tiles_a.size();
  const int nt = na + (int)P.tiles_b.size();
  const int nitems = nt + (P.nd + 63) / 64;
  hipLaunchKernelGGL(kfn, dim3(blocks), dim3(p->block_threads), 0, st, p->d_tiles_a, ns1, na, p->d_tiles_b, nt,
                     nitems, nwaves, p->d_panels, (const S *)p->d_vals, x, p->d_t, p->d_slots,
                     (const float *)p->d_scales, p->d_yseg, p->d_yptr, p->d_ylist, P.nd, y, accumulate, p->d_ctr);
  return HMVM_OK;
}

template <typename S, bool kScaled>
int launch_persistent_s(hmvm_plan *p, const double *x, double *y, int accumulate, hipStream_t st) {
  switch (p->kernel) {
    case 0: return launch_persistent_v<S, kScaled, 0>(p, x, y, accumulate, st);
    case 1: return launch_persistent_v<S, kScaled, 1>(p, x, y, accumulate, st);
    case 2: return launch_persistent_v<S, kScaled, 2>(p, x, y, accumulate, st);
    case 3: return launch_persistent_v<S, kScaled, 3>(p, x, y, accumulate, st);
    case 4: return launch_persistent_v<S, kScaled, 4>(p, x, y, accumulate, st);
    case 5: return launch_persistent_v<S, kScaled, 5>(p, x, y, accumulate, st);
    case 6: return launch_persistent_v<S, kScaled, 6>(p, x, y, accumulate, st);
    default: return launch_persistent_v<S, kScaled, 7>(p, x, y, accumulate, st);
  }
}

int launch_persistent(hmvm_plan *p, const double *x, double *y, int accumulate, hipStream_t st) {
  if (p->storage == kF64) return launch_persistent_s<double, false>(p, x, y, accumulate, st);
  if (p->storage == kF32) return launch_persistent_s<float, false>(p, x, y, accumulate, st);
  return launch_persistent_s<uint16_t, true>(p, x, y, accumulate, st);
}

// The persistent mode needs every stage-2 tile to read its leaf's stage-1
// partial slots directly (no materialised t).
bool persistent_ok(const hmvm_plan *p) { return p->mode == 1 && p->plan.t_fused; }

int apply(hmvm_plan *p, const double *d_x, double *d_y, int accumulate, hipStream_t st) {
  if (persistent_ok(p)) {
    launch_persistent(p, d_x, d_y, accumulate, st);
  } else {
    for (int ph = 0; ph < 4; ph++) launch_phase(p, ph, d_x, d_y, accumulate, st);
  }
  HIP_TRY(hipGetLastError());
  return HMVM_OK;
}

void fill_info(const hmvm_plan *p, const Plan &P, int storage, hmvm_plan_info *info) {
  std::memset(info, 0, sizeof(*info));
  info->nd = P.nd;
  info->nlf = P.nlf;
  info->nlowrank = P.nlr;
  info->ndense = P.ndense;
  info->len = P.len;
  info->device_elems = P.nvals;
  info->storage = storage;
  info->elem_bytes = elem_bytes(storage);
  info->ntiles_a = (int)P.tiles_a.size();
  info->ntiles_b = (int)P.tiles_b.size();
  info->npanels = (int)P.panels.size();
  info->ncol_groups = P.ngroups_col;
  info->nrow_groups = P.ngroups_row;
  info->t_fused = P.t_fused ? 1 : 0;
  info->n_t = P.n_t;
  info->n_tslots = P.n_tp;
  info->n_zslots = P.n_z;
  info->elems_a = P.elems_a;
  info->elems_b = P.elems_b;
  info->elems_stage1 = P.elems_stage1;
  info->elems_dense = P.elems_dense;
  info->elems_stage2 = P.elems_stage2;
  info->vec_reads_a = P.vec_a;
  info->vec_reads_b = P.vec_b;
  info->algorithmic_bytes = (long long)info->elem_bytes * P.len + 16LL * P.nd;
  info->device_bytes = p ? (long long)p->device_bytes : 0;
}

// ---------------------------------------------------------------------------
// cache of device copies for the drop-in entry points
struct CacheKey {
  const void *mat;
  int variant;  // 0:1 1:1t 2:2 3:2t
  bool operator<(const CacheKey &o) const { return mat != o.mat ? mat < o.mat : variant < o.variant; }
};
struct CacheEntry {
  hmvm_plan *plan = nullptr;
  // identity of the contents the plan was built from
  int nd = 0, nlf = 0, len = 0;
  const void *data = nullptr;
  const void *first_leaf_a1 = nullptr;
};
std::mutex g_cache_mu;
std::map<CacheKey, CacheEntry> g_cache;

hmvm_plan *cached_plan(const void *mat, int variant, int nd, int nlf, int len, const void *data,
                       const void *first_a1) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_cache.find({mat, variant});
  if (it == g_cache.end()) return nullptr;
  const CacheEntry &e = it->second;
  if (e.nd == nd && e.nlf == nlf && e.len == len && e.data == data && e.first_leaf_a1 == first_a1) return e.plan;
  hmvm_plan_destroy(e.plan);
  g_cache.erase(it);
  return nullptr;
}

void cache_put(const void *mat, int variant, hmvm_plan *p, int nd, int nlf, int len, const void *data,
               const void *first_a1) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  CacheEntry e;
  e.plan = p;
  e.nd = nd;
  e.nlf = nlf;
  e.len = len;
  e.data = data;
  e.first_leaf_a1 = first_a1;
  g_cache[{mat, variant}] = e;
}

const char *kVariantName[4] = {"1", "1t", "2", "2t"};

// Resolves (and on first use builds) the device copy for one variant.
hmvm_plan *plan_for(const matrix *mat, const matrix2 *mat2, int variant) {
  if (variant < 2) {
    if (!mat || !mat->submat) { set_error(HMVM_ERR_ARG, "matrix is NULL"); return nullptr; }
    const void *a1 = mat->nlf > 0 ? (const void *)mat->submat[0].a1 : nullptr;
    hmvm_plan *p = cached_plan(mat, variant, mat->nd, mat->nlf, 0, mat->submat, a1);
    if (p) return p;
    Source s;
    if (source_from_matrix(mat, variant == 1 ? kA2Transposed : kA2RowMajor, &s)) return nullptr;
    if (create_plan(s, nullptr, &p)) return nullptr;
    cache_put(mat, variant, p, mat->nd, mat->nlf, 0, mat->submat, a1);
    return p;
  }
  if (!mat2) { set_error(HMVM_ERR_ARG, "matrix2 is NULL"); return nullptr; }
  const void *data = variant == 3 ? (const void *)mat2->rowmat_t : (const void *)mat2->rowmat;
  hmvm_plan *p = cached_plan(mat2, variant, mat2->nd, mat2->nlf, mat2->len, data, nullptr);
  if (p) return p;
  Source s;
  if (source_from_matrix2(mat2, variant == 3 ? kA2Transposed : kA2RowMajor, &s)) return nullptr;
  if (create_plan(s, nullptr, &p)) return nullptr;
  cache_put(mat2, variant, p, mat2->nd, mat2->nlf, mat2->len, data, nullptr);
  return p;
}

void run_variant(double *v, const matrix *mat, const matrix2 *mat2, const double *b, int variant) {
  clear_error();
  if (!v || !b) { set_error(HMVM_ERR_ARG, "v or b is NULL"); return; }
  hmvm_plan *p = plan_for(mat, mat2, variant);
  if (!p) return;
  hmvm_plan_apply_host(p, b, v, 1);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

int hmvm_hip_last_error(void) { return g_err; }
const char *hmvm_hip_last_error_message(void) { return g_msg; }

void hmvm_options_default(hmvm_options *opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  PlanOptions po;
  opt->storage = po.storage;
  opt->w1 = po.w1;
  opt->w2 = po.w2;
  opt->block_threads = 256;
  opt->kernel = 0;
  opt->order = 0;
  opt->mode = 0;
}

int hmvm_plan_create(const matrix2 *mat2, int a2_source, const hmvm_options *opt, hmvm_plan **out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  int rc = source_from_matrix2(mat2, a2_source, &s);
  if (rc) return rc;
  return create_plan(s, opt, out);
}

int hmvm_plan_create_aos(const matrix *mat, int a2_source, const hmvm_options *opt, hmvm_plan **out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  int rc = source_from_matrix(mat, a2_source, &s);
  if (rc) return rc;
  return create_plan(s, opt, out);
}

void hmvm_plan_destroy(hmvm_plan *plan) { delete plan; }

int hmvm_plan_set_launch(hmvm_plan *plan, int kernel, int block_threads) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  if (kernel < 0 || kernel >= kNumKernelVariants)
    return set_error(HMVM_ERR_ARG, "kernel variant must be in [0,%d)", kNumKernelVariants);
  if (block_threads % 64 || block_threads < 64 || block_threads > 256)
    return set_error(HMVM_ERR_ARG, "block_threads must be 64, 128, 192 or 256");
  plan->kernel = kernel;
  plan->block_threads = block_threads;
  return HMVM_OK;
}

int hmvm_plan_reserve_multi(hmvm_plan *plan, int nrhs) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  return reserve_multi(plan, nrhs);
}

int hmvm_plan_apply_multi(hmvm_plan *plan, const double *d_X, double *d_Y, int nrhs, int accumulate, void *stream) {
  if (!plan || !d_X || !d_Y) return set_error(HMVM_ERR_ARG, "NULL argument");
  if (nrhs != 16 && nrhs != 32 && nrhs != 64) return set_error(HMVM_ERR_ARG, "nrhs must be 16, 32 or 64");
  return apply_multi(plan, d_X, d_Y, nrhs, accumulate, stream ? (hipStream_t)stream : plan->stream);
}

int hmvm_plan_apply_multi_host(hmvm_plan *plan, const double *X, double *Y, int nrhs, int accumulate) {
  if (!plan || !X || !Y) return set_error(HMVM_ERR_ARG, "NULL argument");
  int rc = reserve_multi(plan, nrhs);
  if (rc) return rc;
  const size_t nb = sizeof(double) * (size_t)plan->plan.nd * nrhs;
  HIP_TRY(hipMemcpyAsync(plan->d_mX, X, nb, hipMemcpyHostToDevice, plan->stream));
  if (accumulate) HIP_TRY(hipMemcpyAsync(plan->d_mY, Y, nb, hipMemcpyHostToDevice, plan->stream));
  if ((rc = apply_multi(plan, plan->d_mX, plan->d_mY, nrhs, accumulate, plan->stream))) return rc;
  HIP_TRY(hipMemcpyAsync(Y, plan->d_mY, nb, hipMemcpyDeviceToHost, plan->stream));
  HIP_TRY(hipStreamSynchronize(plan->stream));
  return HMVM_OK;
}

int hmvm_plan_time_multi(hmvm_plan *plan, int nrhs, int warmup, int iters, hmvm_timing *tm) {
  if (!plan || !tm || iters < 1 || warmup < 0) return set_error(HMVM_ERR_ARG, "bad argument");
  int rc = reserve_multi(plan, nrhs);
  if (rc) return rc;
  std::memset(tm, 0, sizeof(*tm));
  hipStream_t st = plan->stream;
  {
    const size_t n = (size_t)plan->plan.nd * nrhs;
    std::vector<double> h(n);
    for (size_t i = 0; i < n; i++) h[i] = std::sin((double)(i + 1));
    HIP_TRY(hipMemcpy(plan->d_mX, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  for (int w = 0; w < warmup; w++)
    if ((rc = apply_multi(plan, plan->d_mX, plan->d_mY, nrhs, 0, st))) return rc;
  hipEvent_t ev[5];
  for (int e = 0; e < 5; e++) HIP_TRY(hipEventCreate(&ev[e]));
  double ph[4] = {0, 0, 0, 0}, total = 0, mn = 1e30;
  for (int it = 0; it < iters; it++) {
    hipEventRecord(ev[0], st);
    for (int phs = 0; phs < 4; phs++) {
      launch_multi_phase(plan, phs, plan->d_mX, plan->d_mY, nrhs, 0, st);
      hipEventRecord(ev[phs + 1], st);
    }
    if (hipEventSynchronize(ev[4]) != hipSuccess) { rc = set_error(HMVM_ERR_HIP, "event sync failed"); break; }
    float ms;
    for (int phs = 0; phs < 4; phs++) {
      hipEventElapsedTime(&ms, ev[phs], ev[phs + 1]);
      ph[phs] += ms;
    }
    hipEventElapsedTime(&ms, ev[0], ev[4]);
    total += ms;
    mn = std::min(mn, (double)ms);
  }
  if (!rc) {
    hipEventRecord(ev[0], st);
    for (int it = 0; it < iters && !rc; it++) rc = apply_multi(plan, plan->d_mX, plan->d_mY, nrhs, 0, st);
    hipEventRecord(ev[1], st);
    float ms = 0;
    if (!rc && hipEventSynchronize(ev[1]) == hipSuccess) hipEventElapsedTime(&ms, ev[0], ev[1]);
    tm->ms_stream = ms / iters;
  }
  for (int e = 0; e < 5; e++) (void)hipEventDestroy(ev[e]);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  tm->iters = iters;
  tm->ms_avg = total / iters;
  tm->ms_min = mn;
  tm->ms_median = tm->ms_avg;
  tm->ms_a = ph[0] / iters;
  tm->ms_tsum = ph[1] / iters;
  tm->ms_b = ph[2] / iters;
  tm->ms_ysum = ph[3] / iters;
  return HMVM_OK;
}

int hmvm_plan_set_mode(hmvm_plan *plan, int mode) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  if (mode < 0 || mode > 1) return set_error(HMVM_ERR_ARG, "mode must be 0 or 1");
  plan->mode = mode;
  return HMVM_OK;
}

int hmvm_plan_fault(hmvm_plan *plan) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  PersistCounters c;
  HIP_TRY(hipStreamSynchronize(plan->stream));
  HIP_TRY(hipMemcpy(&c, plan->d_ctr, sizeof(c), hipMemcpyDeviceToHost));
  return c.fault ? 1 : 0;
}

int hmvm_bandwidth_probe(size_t bytes, int iters, double *read_gbs, double *copy_gbs) {
  clear_error();
  if (bytes < (1u << 20) || iters < 1) return set_error(HMVM_ERR_ARG, "bytes >= 1 MiB and iters >= 1");
  const size_t n4 = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  double *sink = nullptr;
  HIP_TRY(hipMalloc(&a, n4 * sizeof(double4)));
  hipError_t e = hipMalloc(&b, n4 * sizeof(double4));
  if (e == hipSuccess) e = hipMalloc(&sink, sizeof(double));
  if (e == hipSuccess) e = hipMemset(a, 0, n4 * sizeof(double4));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int ncu = 256;
  hipDeviceProp_t prop;
  if (e == hipSuccess && hipGetDeviceProperties(&prop, 0) == hipSuccess) ncu = prop.multiProcessorCount;
  const dim3 grid(ncu * 8), block(256);
  float ms_r = 0, ms_c = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(probe_read_kernel, grid, block, 0, 0, a, n4, sink);  // warm-up
    hipEventRecord(e0, 0);
    for (int i = 0; i < iters; i++) hipLaunchKernelGGL(probe_read_kernel, grid, block, 0, 0, a, n4, sink);
    hipEventRecord(e1, 0);
    e = hipEventSynchronize(e1);
    if (e == hipSuccess) hipEventElapsedTime(&ms_r, e0, e1);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(probe_copy_kernel, grid, block, 0, 0, a, b, n4);
    hipEventRecord(e0, 0);
    for (int i = 0; i < iters; i++) hipLaunchKernelGGL(probe_copy_kernel, grid, block, 0, 0, a, b, n4);
    hipEventRecord(e1, 0);
    e = hipEventSynchronize(e1);
    if (e == hipSuccess) hipEventElapsedTime(&ms_c, e0, e1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (sink) (void)hipFree(sink);
  if (e != hipSuccess) return set_error(HMVM_ERR_HIP, "bandwidth probe: %s", hipGetErrorString(e));
  const double nbytes = (double)n4 * sizeof(double4);
  if (read_gbs) *read_gbs = nbytes * iters / (ms_r * 1e-3) / 1e9;
  if (copy_gbs) *copy_gbs = 2.0 * nbytes * iters / (ms_c * 1e-3) / 1e9;
  return HMVM_OK;
}

int hmvm_plan_get_info(const hmvm_plan *plan, hmvm_plan_info *info) {
  if (!plan || !info) return set_error(HMVM_ERR_ARG, "NULL argument");
  fill_info(plan, plan->plan, plan->storage, info);
  return HMVM_OK;
}

int hmvm_plan_check(const matrix2 *mat2, const hmvm_options *opt, hmvm_plan_info *info) {
  clear_error();
  Source s;
  int rc = source_from_matrix2(mat2, kA2RowMajor, &s);
  if (rc) return rc;
  const PlanOptions po = plan_options(opt);
  Plan P;
  std::string err = build_plan(s.la, po, &P);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "%s", err.c_str());
  err = verify_plan(P, s.la);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "schedule check failed: %s", err.c_str());
  if (info) fill_info(nullptr, P, po.storage, info);
  return HMVM_OK;
}

int hmvm_schedule_export(const matrix2 *mat2, int a2_source, const hmvm_options *opt, hmvm_schedule *out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  std::memset(out, 0, sizeof(*out));
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  int rc = source_from_matrix2(mat2, a2_source, &s);
  if (rc) return rc;
  Plan P;
  std::string err = build_plan(s.la, plan_options(opt), &P);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "%s", err.c_str());
  err = verify_plan(P, s.la);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "schedule check failed: %s", err.c_str());
  out->ntiles_a = (int)P.tiles_a.size();
  out->ntiles_b = (int)P.tiles_b.size();
  out->npanels = (int)P.panels.size();
  out->nrepack = (int)P.repack.size();
  out->t_fused = P.t_fused ? 1 : 0;
  out->nvals = P.nvals;
  out->n_t = P.n_t;
  out->n_tslots = P.n_tp;
  out->n_zslots = P.n_z;
  out->tiles_a = (hmvm_tile *)copy_out(P.tiles_a);
  out->tiles_b = (hmvm_tile *)copy_out(P.tiles_b);
  out->panels = (hmvm_panel *)copy_out(P.panels);
  out->repack = (hmvm_repack *)copy_out(P.repack);
  export_segsum(P.tsum, &out->tsum);
  export_segsum(P.ysum, &out->ysum);
  if (!out->tiles_a || !out->tiles_b || !out->panels || !out->repack || !out->tsum.seg || !out->ysum.seg) {
    hmvm_schedule_release(out);
    return set_error(HMVM_ERR_NOMEM, "schedule export");
  }
  return HMVM_OK;
}

void hmvm_schedule_release(hmvm_schedule *s) {
  if (!s) return;
  void *ptrs[] = {s->tiles_a, s->tiles_b, s->panels, s->repack, s->tsum.seg, s->tsum.ptr, s->tsum.list,
                  s->ysum.seg, s->ysum.ptr, s->ysum.list};
  for (void *p : ptrs) std::free(p);
  std::memset(s, 0, sizeof(*s));
}

int hmvm_plan_apply(hmvm_plan *plan, const double *d_x, double *d_y, int accumulate, void *stream) {
  if (!plan || !d_x || !d_y) return set_error(HMVM_ERR_ARG, "NULL argument");
  hipStream_t st = stream ? (hipStream_t)stream : plan->stream;
  return apply(plan, d_x, d_y, accumulate, st);
}

int hmvm_plan_apply_host(hmvm_plan *plan, const double *x, double *y, int accumulate) {
  if (!plan || !x || !y) return set_error(HMVM_ERR_ARG, "NULL argument");
  const size_t nb = sizeof(double) * (size_t)plan->plan.nd;
  HIP_TRY(hipMemcpyAsync(plan->d_x, x, nb, hipMemcpyHostToDevice, plan->stream));
  if (accumulate) HIP_TRY(hipMemcpyAsync(plan->d_y, y, nb, hipMemcpyHostToDevice, plan->stream));
  int rc = apply(plan, plan->d_x, plan->d_y, accumulate, plan->stream);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(y, plan->d_y, nb, hipMemcpyDeviceToHost, plan->stream));
  HIP_TRY(hipStreamSynchronize(plan->stream));
  return HMVM_OK;
}

int hmvm_plan_bench_apply(hmvm_plan *plan, const double *x, int iters, double *seconds) {
  if (!plan || !x || !seconds || iters < 1) return set_error(HMVM_ERR_ARG, "bad argument");
  hipStream_t st = plan->stream;
  HIP_TRY(hipMemcpyAsync(plan->d_x, x, sizeof(double) * (size_t)plan->plan.nd, hipMemcpyHostToDevice, st));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_error(HMVM_ERR_HIP, "hipEventCreate failed");
  }
  hipEventRecord(e0, st);
  int rc = HMVM_OK;
  for (int i = 0; i < iters && !rc; i++) rc = apply(plan, plan->d_x, plan->d_y, 0, st);
  hipEventRecord(e1, st);
  float ms = 0;
  if (!rc && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *seconds = ms * 1e-3 / iters;
  return HMVM_OK;
}

int hmvm_plan_time(hmvm_plan *plan, int warmup, int iters, hmvm_timing *tm) {
  if (!plan || !tm || iters < 1 || warmup < 0) return set_error(HMVM_ERR_ARG, "bad argument");
  std::memset(tm, 0, sizeof(*tm));
  hipStream_t st = plan->stream;
  const int nd = plan->plan.nd;
  {
    std::vector<double> h((size_t)nd);
    for (int i = 0; i < nd; i++) h[i] = std::sin((double)(i + 1));
    HIP_TRY(hipMemcpy(plan->d_x, h.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  }
  for (int w = 0; w < warmup; w++) {
    int rc = apply(plan, plan->d_x, plan->d_y, 0, st);
    if (rc) return rc;
  }
  hipEvent_t ev[5];
  for (int e = 0; e < 5; e++) HIP_TRY(hipEventCreate(&ev[e]));
  std::vector<double> total;
  double ph[4] = {0, 0, 0, 0};
  int rc = HMVM_OK;
  for (int it = 0; it < iters && !rc; it++) {
    hipEventRecord(ev[0], st);
    if (persistent_ok(plan)) {
      launch_persistent(plan, plan->d_x, plan->d_y, 0, st);
      for (int phs = 0; phs < 4; phs++) hipEventRecord(ev[phs + 1], st);
    } else {
      for (int phs = 0; phs < 4; phs++) {
        launch_phase(plan, phs, plan->d_x, plan->d_y, 0, st);
        hipEventRecord(ev[phs + 1], st);
      }
    }
    if (hipEventSynchronize(ev[4]) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event sync failed");
    float ms;
    for (int phs = 0; phs < 4 && !rc; phs++) {
      hipEventElapsedTime(&ms, ev[phs], ev[phs + 1]);
      ph[phs] += ms;
    }
    hipEventElapsedTime(&ms, ev[0], ev[4]);
    total.push_back(ms);
  }
  // back-to-back throughput: all applies enqueued, one event pair
  if (!rc) {
    hipEventRecord(ev[0], st);
    for (int it = 0; it < iters && !rc; it++) rc = apply(plan, plan->d_x, plan->d_y, 0, st);
    hipEventRecord(ev[1], st);
    if (!rc && hipEventSynchronize(ev[1]) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event sync failed");
    float ms = 0;
    if (!rc) hipEventElapsedTime(&ms, ev[0], ev[1]);
    tm->ms_stream = ms / iters;
  }
  for (int e = 0; e < 5; e++) (void)hipEventDestroy(ev[e]);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  std::vector<double> s = total;
  std::sort(s.begin(), s.end());
  double sum = 0;
  for (double v : total) sum += v;
  tm->iters = iters;
  tm->ms_avg = sum / iters;
  tm->ms_min = s.front();
  tm->ms_median = s[s.size() / 2];
  tm->ms_a = ph[0] / iters;
  tm->ms_tsum = ph[1] / iters;
  tm->ms_b = ph[2] / iters;
  tm->ms_ysum = ph[3] / iters;
  return HMVM_OK;
}

void hmvm_hip_1(double *v, const matrix *mat, const double *b) { run_variant(v, mat, nullptr, b, 0); }
void hmvm_hip_1t(double *v, const matrix *mat, const double *b) { run_variant(v, mat, nullptr, b, 1); }
void hmvm_hip_2(double *v, const matrix2 *mat2, const double *b) { run_variant(v, nullptr, mat2, b, 2); }
void hmvm_hip_2t(double *v, const matrix2 *mat2, const double *b) { run_variant(v, nullptr, mat2, b, 3); }

void hmvm_hip_release(const void *mat) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  for (auto it = g_cache.begin(); it != g_cache.end();) {
    if (it->first.mat == mat) {
      hmvm_plan_destroy(it->second.plan);
      it = g_cache.erase(it);
    } else {
      ++it;
    }
  }
}

void hmvm_hip(const matrix *mat, const matrix2 *mat2, const double *b, int dump_result) {
  clear_error();
  const int nd = mat ? mat->nd : (mat2 ? mat2->nd : 0);
  if (nd < 1 || !b) { set_error(HMVM_ERR_ARG, "no matrix or b"); return; }
  std::vector<double> v((size_t)nd);
  for (int var = 0; var < 4; var++) {
    if ((var < 2 && !mat) || (var >= 2 && !mat2)) continue;
    std::fill(v.begin(), v.end(), 0.0);
    run_variant(v.data(), mat, mat2, b, var);
    if (g_err) return;
    if (dump_result) {
      char fname[64];
      std::snprintf(fname, sizeof fname, "result_hip_%s_d.txt", kVariantName[var]);
      FILE *F = std::fopen(fname, "w");
      if (!F) { set_error(HMVM_ERR_ARG, "cannot write %s", fname); return; }
      for (int i = 0; i < nd; i++) std::fprintf(F, "%.3E\n", v[i]);
      std::fclose(F);
    }
  }
}

void hmvm_hip_bench(const matrix *mat, const matrix2 *mat2, const double *b) {
  clear_error();
  const int L = 10, M = 5;
  const int nd = mat ? mat->nd : (mat2 ? mat2->nd : 0);
  if (nd < 1 || !b) { set_error(HMVM_ERR_ARG, "no matrix or b"); return; }
  for (int var = 0; var < 4; var++) {
    if ((var < 2 && !mat) || (var >= 2 && !mat2)) continue;
    hmvm_plan *p = plan_for(mat, mat2, var);
    if (!p) return;
    if (hipMemcpy(p->d_x, b, sizeof(double) * nd, hipMemcpyHostToDevice) != hipSuccess) {
      set_error(HMVM_ERR_HIP, "upload of b failed");
      return;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      set_error(HMVM_ERR_HIP, "hipEventCreate failed");
      return;
    }
    double dt[L];
    for (int l = 0; l < L; l++) {
      hipEventRecord(e0, p->stream);
      if (apply(p, p->d_x, p->d_y, 0, p->stream)) break;
      hipEventRecord(e1, p->stream);
      hipEventSynchronize(e1);
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      dt[l] = ms * 1e-3;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (g_err) return;
    double mn = 9999.99, mx = 0.0, avg = 0.0;
    for (int l = M; l < L; l++) {
      avg += dt[l];
      mn = std::min(mn, dt[l]);
      mx = std::max(mx, dt[l]);
    }
    avg /= (L - M);
    std::printf("TIME %d hmvm_hip_%s min %e max %e avg %e\n", L, kVariantName[var], mn, mx, avg);
    std::fflush(stdout);
  }
}

int hmvm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
int hmvm_device_alloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(HMVM_ERR_ARG, "NULL argument");
  HIP_TRY(hipMalloc(ptr, bytes));
  return HMVM_OK;
}
int hmvm_device_free(void *ptr) {
  HIP_TRY(hipFree(ptr));
  return HMVM_OK;
}
int hmvm_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return HMVM_OK;
}
int hmvm_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return HMVM_OK;
}
int hmvm_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return HMVM_OK;
}

}  // extern "C"
