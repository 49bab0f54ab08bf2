// hmvm_runtime.hip -- plan upload, launch sequence and the C ABI of the HIP HMVM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <tuple>
#include <vector>

#include "hacapk.h"
#include "hmvm_hip.h"
#include "hmvm_kernels.hpp"
#include "hmvm_plan.hpp"
#include "hmvm_status.hpp"
#include "io_records.hpp"

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

}  // namespace

namespace hmvm {
int report_error(int code, const char *msg) { return set_error(code, "%s", msg); }
void report_clear() { clear_error(); }
}  // namespace hmvm

namespace {

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
int dalloc_bytes(void **p, size_t n) {
  *p = nullptr;
  if (n == 0) return HMVM_OK;
  HIP_TRY(hipMalloc(p, n));
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

// Per-entry [begin, end) list ranges of a segment sum.
std::vector<int2> seg_ranges(const SegSum &g) {
  std::vector<int2> r(g.seg.size());
  for (size_t i = 0; i < g.seg.size(); i++) r[i] = make_int2(g.ptr[g.seg[i]], g.ptr[g.seg[i] + 1]);
  return r;
}

// {first slot, (nch << 16) | stride, t entry, 0} of every summed t entry, in
// entry order: every segment lists its slots at one stride (a rank row's
// partial slots in the column chunks of its row group are one row group
// apart).  Entries marked in `skip` (written through a stage-1 row map) get
// no descriptor.  False when a segment is not at one stride, or its fields do
// not fit.
bool t_descs(const SegSum &g, const std::vector<char> &skip, std::vector<int4> *out) {
  out->clear();
  for (size_t i = 0; i < g.seg.size(); i++) {
    const int s = g.seg[i];
    if (s < 0 || (!skip.empty() && skip[i])) continue;
    const int e0 = g.ptr[s], e1 = g.ptr[s + 1], nch = e1 - e0;
    const int64_t first = nch > 0 ? g.list[e0] + (int64_t)i : 0;
    const int64_t stride = nch > 1 ? g.list[e0 + 1] - g.list[e0] : 0;
    if (nch > 0xFFFF || stride < 0 || stride > 0xFFFF || first < 0 || first > INT32_MAX) return false;
    for (int e = e0 + 1; e < e1; e++)
      if (g.list[e] != g.list[e0] + (int64_t)(e - e0) * stride) return false;
    out->push_back(make_int4((int)first, (int)(((unsigned)nch << 16) | (unsigned)stride), (int)i, 0));
  }
  return true;
}

int elem_bytes(int storage) { return storage == kF64 ? 8 : storage == kF32 ? 4 : 2; }
// Bytes of one slot (SlotOf<S>::T): FP64 for FP64 storage, FP32 otherwise.
size_t slot_bytes(int storage) { return storage == kF64 ? 8 : 4; }
// Bytes of one multi-RHS vector entry (VecOf<S>::T: the storage type).
size_t vec_bytes(int storage) { return (size_t)elem_bytes(storage); }

// N events, destroyed on every return path of the timing functions.
template <int N>
struct EventSet {
  hipEvent_t e[N] = {};
  hipError_t create() {
    for (int i = 0; i < N; i++)
      if (hipError_t err = hipEventCreate(&e[i])) return err;
    return hipSuccess;
  }
  hipEvent_t &operator[](int i) { return e[i]; }
  ~EventSet() {
    for (int i = 0; i < N; i++)
      if (e[i]) (void)hipEventDestroy(e[i]);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// the plan object
struct hmvm_plan {
  Plan plan;
  int storage = kF64;
  int block_threads = 256;
  int kernel = 0;  // tile-kernel variant, 0 .. kNumKernelVariants-1 (launch A, persistent, multi)
  int kernel_b = 0, block_threads_b = 256;  // launch B
  int mode = 0;    // 0: launches on one stream; 1: one persistent launch; 2: dense tiles on a second stream
  hipStream_t stream2 = nullptr;       // dense tiles, concurrent with stage 1 / t sum / stage 2
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool a_mixed = false;  // launch A order is not "stage-1 tiles, then dense" (modes 1 and 2 need that)
  bool has_tp = false;   // some panel sums stage-1 partial slots (kVecTP)
  int persist_key = -1, persist_blocks = 0;  // persistent launch: its shape (kernel variant, block size) and grid
  int tiles_version = 0, persist_version = -1;  // tile lists changed / the dealt lists are of this version
  TileDesc *d_ptiles_a = nullptr, *d_ptiles_b = nullptr;  // persistent launch: tiles dealt to the waves
  int32_t *d_pptr_a = nullptr, *d_pptr_b = nullptr;       // wave w's tiles: [ptr[w], ptr[w+1])
  // several right-hand sides: buffers sized for multi_cap columns
  int multi_cap = 0;
  void *d_mslots = nullptr;  // slot_bytes(storage) per slot
  void *d_mt = nullptr;      // t: vec_bytes(storage) per entry
  void *d_mXc = nullptr;     // X in the operand type (FP32 and BF16 storage)
  TileDesc *d_mtiles_a = nullptr, *d_mtiles_b = nullptr;  // XCD-aware tile layouts of launches A and B
  int mtiles_wpb = 0;                                     // waves per workgroup they were laid out for
  double *d_mX = nullptr, *d_mY = nullptr;
  PersistCounters *d_ctr = nullptr;
  int ncu = 0;     // compute units of the device
  int device = 0;
  hipStream_t stream = nullptr;
  void *d_vals = nullptr;
  TileDesc *d_tiles_a = nullptr, *d_tiles_b = nullptr;
  PanelDesc *d_panels = nullptr;
  double *d_t = nullptr;
  void *d_slots = nullptr;  // stage-1 partial slots, then output slots: slot_bytes(storage) each
  int32_t *d_tfac = nullptr;
  int32_t *d_tmap = nullptr;  // t index of every row of the stage-1 tiles that write t
  double *d_mscale = nullptr;  // BF16: both factor scales of each d_tmap row's leaf
  int64_t *d_ylist = nullptr;
  int2 *d_yrng = nullptr;   // per-entry list ranges of the y sum
  int2 *d_tdesc = nullptr;    // t sum: {first slot, (nch << 16) | stride} per entry
  int n_tsum = 0;             // t entries summed by the one-vector launches
  int4 *d_tdesc_m = nullptr;  // multi-RHS t sum: {first, (nch << 16) | stride, entry, 0} of the
  int n_tsum_m = 0;           // entries no stage-1 row map writes
  float *d_scales = nullptr;
  double *d_x = nullptr, *d_y = nullptr;  // staging for host-pointer calls
  size_t device_bytes = 0;

  ~hmvm_plan() {
    void *ptrs[] = {d_vals,  d_tiles_a, d_tiles_b, d_panels, d_t,      d_slots,
                    d_tfac,  d_tmap,    d_ylist,  d_scales, d_x,      d_y,
                    d_yrng,  d_tdesc,   d_tdesc_m, d_mtiles_a, d_mtiles_b, d_ptiles_a, d_ptiles_b, d_pptr_a, d_pptr_b,
                    d_ctr,   d_mslots,  d_mt,      d_mX,     d_mY,     d_mXc,    d_mscale};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Supplies elements [w0, w1) of a source value buffer that is not one
// contiguous host array (a file being streamed, several matrices of a batch).
// Calls come in increasing w0 order and never overlap.
struct Feeder {
  virtual ~Feeder() = default;
  virtual int fill(double *dst, int64_t w0, int64_t w1) = 0;
};

// Source view of the caller's matrix: descriptor arrays + the FP64 values,
// either one contiguous host buffer (rowmat, rowmat_t or a gathered copy of
// the per-leaf arrays) or a feeder.
struct Source {
  LeafArrays la;
  std::vector<int64_t> a1, a2;
  const double *vals = nullptr;
  Feeder *feeder = nullptr;
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

int source_from_matrix2_64(const matrix2_64 *m, int a2_source, Source *s) {
  if (!m || m->nd < 1 || m->nlf < 1 || !m->ltmtx || !m->kt || !m->ndl || !m->ndt || !m->nstrtl ||
      !m->nstrtt || !m->a1 || !m->a2)
    return set_error(HMVM_ERR_ARG, "matrix2_64 is NULL or incomplete");
  const double *vals = a2_source == kA2Transposed ? m->rowmat_t : m->rowmat;
  if (!vals) return set_error(HMVM_ERR_ARG, "matrix2_64 has no %s", a2_source ? "rowmat_t" : "rowmat");
  s->a1.assign(m->a1, m->a1 + m->nlf);
  s->a2.assign(m->a2, m->a2 + m->nlf);
  LeafArrays &la = s->la;
  la.nd = m->nd;
  la.nlf = m->nlf;
  la.ltmtx = m->ltmtx; la.kt = m->kt; la.ndl = m->ndl; la.ndt = m->ndt;
  la.nstrtl = m->nstrtl; la.nstrtt = m->nstrtt;
  la.src_a1 = s->a1.data();
  la.src_a2 = s->a2.data();
  la.a2_source = a2_source;
  la.src_len = m->len;
  s->vals = vals;
  return HMVM_OK;
}

// Streams the leaf records of a format 1/2/2L file (scanned beforehand) in
// file order; the file order is the packing order of the scan's offsets.
struct FileFeeder : Feeder {
  hacapk_io::Reader rd;
  hacapk_io::LeafStream ls;
  const matrix2_64 *m;
  int leaf = 0;
  FileFeeder(const char *fname, int format, const matrix2_64 *mm) : rd(fname), ls(rd, format), m(mm) {}
  int begin(long long data_pos) { return rd.f ? ls.begin(data_pos) : HACAPK_ERR_OPEN; }
  int fill(double *dst, int64_t w0, int64_t w1) override {
    while (leaf < m->nlf && m->a1[leaf] < w1) {
      const int64_t n = hacapk_io::leaf_len(m->ltmtx[leaf], m->kt[leaf], m->ndl[leaf], m->ndt[leaf]);
      if (m->a1[leaf] < w0 || m->a1[leaf] + n > w1)
        return set_error(HMVM_ERR_MATRIX, "leaf %d does not fit the upload window", leaf);
      int rc = ls.read(m, leaf, dst + (m->a1[leaf] - w0));
      if (rc) return set_error(HMVM_ERR_MATRIX, "reading leaf %d: %s", leaf, hacapk_strerror(rc));
      leaf++;
    }
    return HMVM_OK;
  }
};

// Several matrices laid end to end: segment q holds elements
// [base[q], base[q] + len[q]) of the virtual buffer.
struct ConcatFeeder : Feeder {
  std::vector<const double *> ptr;
  std::vector<int64_t> base, len;
  int fill(double *dst, int64_t w0, int64_t w1) override {
    for (size_t q = 0; q < ptr.size(); q++) {
      const int64_t a = std::max(w0, base[q]), b = std::min(w1, base[q] + len[q]);
      if (a < b) std::memcpy(dst + (a - w0), ptr[q] + (a - base[q]), sizeof(double) * (size_t)(b - a));
    }
    return HMVM_OK;
  }
};

// Block-diagonal union of several matrices: matrix q occupies rows and
// columns [row0[q], row0[q] + nd_q) and its values follow those of matrix q-1.
int source_from_batch(const matrix2 *const *mats, int count, int a2_source, Source *s, ConcatFeeder *feed,
                      std::vector<int> *row0) {
  if (!mats || count < 1) return set_error(HMVM_ERR_ARG, "empty batch");
  int64_t nd = 0, nlf = 0, len = 0;
  for (int q = 0; q < count; q++) {
    const matrix2 *m = mats[q];
    if (!m || m->nd < 1 || m->nlf < 1 || !m->ltmtx || !m->kt || !m->ndl || !m->ndt || !m->nstrtl || !m->nstrtt ||
        !m->a1 || !m->a2)
      return set_error(HMVM_ERR_ARG, "batch matrix %d is NULL or incomplete", q);
    if (!(a2_source == kA2Transposed ? m->rowmat_t : m->rowmat))
      return set_error(HMVM_ERR_ARG, "batch matrix %d has no values", q);
    nd += m->nd;
    nlf += m->nlf;
    len += m->len;
  }
  if (nd > INT32_MAX / 2 || nlf > INT32_MAX / 2) return set_error(HMVM_ERR_ARG, "batch too large");
  const size_t n = (size_t)nlf;
  s->ltmtx.resize(n); s->kt.resize(n); s->ndl.resize(n); s->ndt.resize(n);
  s->nstrtl.resize(n); s->nstrtt.resize(n); s->a1.resize(n); s->a2.resize(n);
  row0->assign((size_t)count + 1, 0);
  int64_t r0 = 0, v0 = 0;
  size_t i = 0;
  for (int q = 0; q < count; q++) {
    const matrix2 *m = mats[q];
    (*row0)[q] = (int)r0;
    for (int j = 0; j < m->nlf; j++, i++) {
      s->ltmtx[i] = m->ltmtx[j]; s->kt[i] = m->kt[j]; s->ndl[i] = m->ndl[j]; s->ndt[i] = m->ndt[j];
      s->nstrtl[i] = m->nstrtl[j] + (int)r0;
      s->nstrtt[i] = m->nstrtt[j] + (int)r0;
      s->a1[i] = m->a1[j] + v0;
      s->a2[i] = m->ltmtx[j] == 1 ? m->a2[j] + v0 : 0;
    }
    feed->ptr.push_back(a2_source == kA2Transposed ? m->rowmat_t : m->rowmat);
    feed->base.push_back(v0);
    feed->len.push_back(m->len);
    r0 += m->nd;
    v0 += m->len;
  }
  (*row0)[count] = (int)r0;
  LeafArrays &la = s->la;
  la.nd = (int)nd;
  la.nlf = (int)nlf;
  la.ltmtx = s->ltmtx.data(); la.kt = s->kt.data(); la.ndl = s->ndl.data(); la.ndt = s->ndt.data();
  la.nstrtl = s->nstrtl.data(); la.nstrtt = s->nstrtt.data();
  la.src_a1 = s->a1.data();
  la.src_a2 = s->a2.data();
  la.a2_source = a2_source;
  la.src_len = len;
  s->feeder = feed;
  return HMVM_OK;
}

// Launch defaults per storage, measured best on MI355X at N = 101,250 (cube,
// eps 1e-4): tile widths in elements, then kernel variant and workgroup size
// of launch A and launch B, and the most rows one tile (one wavefront) may
// cover, for launch A and for the stage-2 tiles of launch B.  Taller tiles
// only pay off for BF16, whose 8-element lane loads leave short stage-2 tiles
// with few lanes busy; FP64 stage-2 tiles run best at 128 rows (one 16-byte
// row group per lane, 96-160 within 0.3%), FP32 at 144-160 (256: 2% slower
// per HMVM), BF16 at 192-320 (256: 2% slower; measured with FP32 slots).
// Options with 0 (widths, workgroup size, rows) or -1 (kernel) take these.
struct LaunchDefaults {
  int w1, w2, kernel_a, bt_a, kernel_b, bt_b, max_rows, max_rows_b, w3;
};
const LaunchDefaults &launch_defaults(int storage) {
  static const LaunchDefaults d[3] = {
      {32768, 32768, 6, 256, 6, 256, 256, 128, 32768},  // FP64
      {32768, 24576, 6, 256, 6, 256, 256, 160, 24576},  // FP32
      {32768, 32768, 6, 256, 6, 256, 512, 192, 32768},  // BF16
  };
  return d[storage >= kF64 && storage <= kBF16 ? storage : 0];
}

// len: elements of the matrix (or of all matrices of a batch).  Default tile
// widths shrink for small matrices so that a launch still has about one tile
// per resident wavefront slot of the device (4096 on MI355X): at N = 4,096
// the default widths would give ~650 tiles, a tenth of the slots.
PlanOptions plan_options(const hmvm_options *o, int64_t len) {
  hmvm_options d;
  if (!o) {
    hmvm_options_default(&d);
    o = &d;
  }
  PlanOptions po;
  po.storage = o->storage;
  const int64_t wfit = std::max<int64_t>(1024, len / 4096);
  po.w1 = o->w1 > 0 ? o->w1 : (int)std::min<int64_t>(launch_defaults(po.storage).w1, wfit);
  po.w2 = o->w2 > 0 ? o->w2 : (int)std::min<int64_t>(launch_defaults(po.storage).w2, wfit);
  po.w3 = o->w3 > 0 ? o->w3 : o->w2 > 0 ? o->w2 : (int)std::min<int64_t>(launch_defaults(po.storage).w3, wfit);
  // an explicit max_rows sets both unless max_rows_b is given; by default
  // stage-2 tiles have their own
  po.max_rows = o->max_rows > 0 ? o->max_rows : launch_defaults(po.storage).max_rows;
  po.max_rows_b = o->max_rows_b > 0 ? o->max_rows_b
                  : o->max_rows > 0 ? o->max_rows
                                    : launch_defaults(po.storage).max_rows_b;
  po.order = o->order;
  // t_mode < 0: materialise t for every leaf.  With t in row-cluster order
  // the stage-2 factors of a cluster merge into one panel, which beats summing
  // the stage-1 partials on the fly in all three storages (measured at
  // N = 101,250) despite the extra launch, and beats t_mode 3 (stage-1 bands
  // writing t directly: thin bands over large clusters re-read x), also in
  // the persistent mode, which runs the t sum as one of its stages.
  po.t_mode = o->t_mode >= 0 ? o->t_mode : 1;
  return po;
}

static_assert(sizeof(hmvm_tile) == sizeof(TileDesc), "tile record layout");
static_assert(HMVM_TILE_ROW_PER_LANE == kTileRowPerLane, "tile flag values");
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

// Repack of one window: `descs` are all descriptors (chunks index them),
// `wdescs`/`chunk_begin` the window's own; chunk maxima are indexed globally
// (the window's chunks start at chunk0).
template <typename S>
int launch_repack(hmvm_plan *p, const double *d_src, const RepackDesc *descs, const RepackDesc *wdescs,
                  const RepackChunk *d_chunks, int nchunks, const int32_t *chunk_begin, int ndesc,
                  float *d_chunk_max, int chunk0) {
  if (p->storage == kBF16) {
    hipLaunchKernelGGL(absmax_kernel, dim3(nchunks), dim3(256), 0, p->stream, descs, d_chunks, d_src,
                       d_chunk_max + chunk0);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(scale_kernel, dim3((ndesc + 255) / 256), dim3(256), 0, p->stream, ndesc, wdescs, chunk_begin,
                       (const float *)d_chunk_max, p->d_scales);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((repack_kernel<S, true>), dim3(nchunks), dim3(256), 0, p->stream, descs, d_chunks, d_src,
                       p->d_scales, (S *)p->d_vals);
  } else {
    hipLaunchKernelGGL((repack_kernel<S, false>), dim3(nchunks), dim3(256), 0, p->stream, descs, d_chunks, d_src,
                       (const float *)nullptr, (S *)p->d_vals);
  }
  HIP_TRY(hipGetLastError());
  return HMVM_OK;
}

// Elements per upload window of the source values: 256 MiB of FP64, or the
// value of HMVM_UPLOAD_WINDOW (elements, >= 1024) when set.
int64_t window_elems() {
  const char *e = std::getenv("HMVM_UPLOAD_WINDOW");
  if (e && *e) {
    const long long v = std::atoll(e);
    if (v >= 1024) return v;
  }
  return int64_t(1) << 25;
}

struct Window {
  int64_t w0, w1;  // source elements [w0, w1)
  int d0, d1;      // repack descriptors (window order)
  int c0, c1;      // repack chunks
};

// Uploads the source values window by window and repacks each window on the
// GPU.  Windows are cut between records (a factor of one leaf; a whole leaf
// for a file), so every repack descriptor reads inside one window; with a
// feeder the host side fills pinned buffers while the GPU repacks the
// previous window.  Device memory beyond the plan: two windows.
int upload_values(hmvm_plan *p, const Source &src) {
  const Plan &P = p->plan;
  const LeafArrays &la = src.la;
  // source records, in source order
  std::vector<std::pair<int64_t, int64_t>> units;
  units.reserve((size_t)la.nlf * 2);
  for (int i = 0; i < la.nlf; i++) {
    if (la.ltmtx[i] == 1) {
      const int64_t n1 = (int64_t)la.kt[i] * la.ndt[i], n2 = (int64_t)la.kt[i] * la.ndl[i];
      if (src.feeder) {
        if (n1 + n2 > 0) units.push_back({la.src_a1[i], la.src_a1[i] + n1 + n2});
      } else {
        if (n1 > 0) units.push_back({la.src_a1[i], la.src_a1[i] + n1});
        if (n2 > 0) units.push_back({la.src_a2[i], la.src_a2[i] + n2});
      }
    } else {
      units.push_back({la.src_a1[i], la.src_a1[i] + (int64_t)la.ndl[i] * la.ndt[i]});
    }
  }
  std::stable_sort(units.begin(), units.end());
  const int64_t kWindowElems = window_elems();
  std::vector<Window> win;
  for (const auto &u : units) {
    if (win.empty() || u.second - win.back().w0 > kWindowElems)
      win.push_back({u.first, u.second, 0, 0, 0, 0});
    else
      win.back().w1 = std::max(win.back().w1, u.second);
  }
  // descriptors -> windows
  const int nrep = (int)P.repack.size();
  std::vector<int> wof(nrep, -1);
  bool fits = true;
  for (int d = 0; d < nrep && fits; d++) {
    const RepackDesc &r = P.repack[d];
    if ((int64_t)r.rows * r.cols == 0) continue;
    const int64_t e = r.src_off + (int64_t)(r.rows - 1) * r.src_rs + (int64_t)(r.cols - 1) * r.src_cs + 1;
    auto it = std::upper_bound(win.begin(), win.end(), r.src_off,
                               [](int64_t v, const Window &w) { return v < w.w0; });
    if (it == win.begin()) { fits = false; break; }
    --it;
    if (e > it->w1) { fits = false; break; }
    wof[d] = (int)(it - win.begin());
  }
  if (!fits) {
    // records overlap or interleave: one window over the whole buffer
    if (src.feeder) return set_error(HMVM_ERR_MATRIX, "source records do not fit the upload windows");
    win.assign(1, Window{0, la.src_len, 0, 0, 0, 0});
    for (int d = 0; d < nrep; d++) wof[d] = (int64_t)P.repack[d].rows * P.repack[d].cols ? 0 : -1;
  }
  // window-ordered, rebased descriptors and their chunks
  std::vector<std::vector<int>> byw(win.size());
  for (int d = 0; d < nrep; d++)
    if (wof[d] >= 0) byw[wof[d]].push_back(d);
  std::vector<RepackDesc> descs;
  std::vector<RepackChunk> chunks;
  std::vector<int32_t> chunk_begin;
  descs.reserve(nrep);
  chunk_begin.reserve(nrep + 1);
  int64_t wmax = 0;
  for (size_t w = 0; w < win.size(); w++) {
    Window &W = win[w];
    W.d0 = (int)descs.size();
    W.c0 = (int)chunks.size();
    for (int d : byw[w]) {
      RepackDesc r = P.repack[d];
      r.src_off -= W.w0;
      chunk_begin.push_back((int32_t)chunks.size());
      const int64_t ne = (int64_t)r.rows * r.cols;
      for (int64_t e0 = 0; e0 < ne; e0 += kRepackChunk) chunks.push_back({(int32_t)descs.size(), (int32_t)e0});
      descs.push_back(r);
    }
    W.d1 = (int)descs.size();
    W.c1 = (int)chunks.size();
    wmax = std::max(wmax, W.w1 - W.w0);
  }
  chunk_begin.push_back((int32_t)chunks.size());

  RepackDesc *d_desc = nullptr;
  RepackChunk *d_chunks = nullptr;
  int32_t *d_chunk_begin = nullptr;
  float *d_chunk_max = nullptr;
  double *d_stage[2] = {nullptr, nullptr};
  double *h_stage[2] = {nullptr, nullptr};
  hipEvent_t copied[2] = {nullptr, nullptr};
  const int nbuf = win.size() > 1 ? 2 : 1;
  auto cleanup = [&]() {
    if (d_desc) (void)hipFree(d_desc);
    if (d_chunks) (void)hipFree(d_chunks);
    if (d_chunk_begin) (void)hipFree(d_chunk_begin);
    if (d_chunk_max) (void)hipFree(d_chunk_max);
    for (int b = 0; b < 2; b++) {
      if (d_stage[b]) (void)hipFree(d_stage[b]);
      if (h_stage[b]) (void)hipHostFree(h_stage[b]);
      if (copied[b]) (void)hipEventDestroy(copied[b]);
    }
  };
  int rc = upload(&d_desc, descs);
  if (!rc) rc = upload(&d_chunks, chunks);
  if (!rc) rc = upload(&d_chunk_begin, chunk_begin);
  if (!rc && p->storage == kBF16) rc = dalloc(&d_chunk_max, chunks.size());
  for (int b = 0; b < nbuf && !rc; b++) {
    rc = dalloc(&d_stage[b], (size_t)std::max<int64_t>(wmax, 1));
    if (!rc && src.feeder) {
      if (hipHostMalloc((void **)&h_stage[b], sizeof(double) * (size_t)std::max<int64_t>(wmax, 1)) != hipSuccess)
        rc = set_error(HMVM_ERR_HIP, "pinned staging buffer of %lld elements", (long long)wmax);
      else if (hipEventCreateWithFlags(&copied[b], hipEventDisableTiming) != hipSuccess)
        rc = set_error(HMVM_ERR_HIP, "hipEventCreate failed");
    }
  }
  for (size_t w = 0; w < win.size() && !rc; w++) {
    const Window &W = win[w];
    const int b = (int)(w % nbuf);
    const size_t n = (size_t)(W.w1 - W.w0);
    if (src.feeder) {
      // the pinned buffer is free once its previous copy has landed
      if (w >= 2 && hipEventSynchronize(copied[b]) != hipSuccess) { rc = set_error(HMVM_ERR_HIP, "staging copy"); break; }
      if ((rc = src.feeder->fill(h_stage[b], W.w0, W.w1))) break;
      if (hipMemcpyAsync(d_stage[b], h_stage[b], sizeof(double) * n, hipMemcpyHostToDevice, p->stream) != hipSuccess ||
          hipEventRecord(copied[b], p->stream) != hipSuccess) {
        rc = set_error(HMVM_ERR_HIP, "window upload failed");
        break;
      }
    } else if (n > 0 && hipMemcpyAsync(d_stage[b], src.vals + W.w0, sizeof(double) * n, hipMemcpyHostToDevice,
                                       p->stream) != hipSuccess) {
      rc = set_error(HMVM_ERR_HIP, "window upload failed");
      break;
    }
    const int nch = W.c1 - W.c0, nd = W.d1 - W.d0;
    if (nch == 0) continue;
    if (p->storage == kF64)
      rc = launch_repack<double>(p, d_stage[b], d_desc, d_desc + W.d0, d_chunks + W.c0, nch, d_chunk_begin + W.d0,
                                 nd, d_chunk_max, W.c0);
    else if (p->storage == kF32)
      rc = launch_repack<float>(p, d_stage[b], d_desc, d_desc + W.d0, d_chunks + W.c0, nch, d_chunk_begin + W.d0,
                                nd, d_chunk_max, W.c0);
    else
      rc = launch_repack<uint16_t>(p, d_stage[b], d_desc, d_desc + W.d0, d_chunks + W.c0, nch,
                                   d_chunk_begin + W.d0, nd, d_chunk_max, W.c0);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = set_error(HMVM_ERR_HIP, "repack failed: %s", hipGetErrorString(e));
  } else {
    (void)hipStreamSynchronize(p->stream);
  }
  cleanup();
  return rc;
}

// Whether launch A's order puts a dense tile before a stage-1 tile.
bool order_mixed(const Plan &P) {
  for (int i = 0; i < P.n_stage1_tiles; i++)
    if (!P.a_stage1[i]) return true;
  return false;
}

// Launch A as "stage-1 tiles, then dense", each part in its current order:
// the layout the persistent and two-stream modes rely on.
int persist_prepare(hmvm_plan *p);

int partition_a(hmvm_plan *p) {
  Plan &P = p->plan;
  if (!order_mixed(P)) {
    p->a_mixed = false;
    return HMVM_OK;
  }
  std::vector<TileDesc> nt;
  std::vector<uint8_t> ns;
  nt.reserve(P.tiles_a.size());
  ns.reserve(P.tiles_a.size());
  for (int part = 1; part >= 0; part--)
    for (size_t i = 0; i < P.tiles_a.size(); i++)
      if (P.a_stage1[i] == part) {
        nt.push_back(P.tiles_a[i]);
        ns.push_back((uint8_t)part);
      }
  HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipMemcpy(p->d_tiles_a, nt.data(), sizeof(TileDesc) * nt.size(), hipMemcpyHostToDevice));
  P.tiles_a.swap(nt);
  P.a_stage1.swap(ns);
  p->a_mixed = false;
  p->tiles_version++;
  return HMVM_OK;
}

int create_plan(Source &src, const hmvm_options *opt, hmvm_plan **out) {
  *out = nullptr;
  hmvm_options dflt;
  if (!opt) {
    hmvm_options_default(&dflt);
    opt = &dflt;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(HMVM_ERR_NODEVICE, "no HIP device visible");
  const PlanOptions po = plan_options(opt, src.la.src_len);
  if (po.storage < kF64 || po.storage > kBF16) return set_error(HMVM_ERR_ARG, "unknown storage %d", po.storage);
  std::unique_ptr<hmvm_plan> p(new hmvm_plan());
  p->storage = po.storage;
  const LaunchDefaults &ld = launch_defaults(po.storage);
  p->kernel = ld.kernel_a;
  p->kernel_b = ld.kernel_b;
  p->block_threads = ld.bt_a;
  p->block_threads_b = ld.bt_b;
  if (opt->block_threads > 0) {
    if (opt->block_threads % 64 || opt->block_threads > 256 || opt->block_threads < 64)
      return set_error(HMVM_ERR_ARG, "block_threads must be 64, 128, 192 or 256");
    p->block_threads = opt->block_threads;
    p->block_threads_b = opt->block_threads;
  }
  if (opt->kernel >= 0) {
    if (opt->kernel >= kNumKernelVariants)
      return set_error(HMVM_ERR_ARG, "kernel variant must be in [0,%d)", kNumKernelVariants);
    p->kernel = opt->kernel;
    p->kernel_b = opt->kernel;
  }
  if (opt->mode < 0 || opt->mode > 2) return set_error(HMVM_ERR_ARG, "mode must be 0, 1 or 2");
  p->mode = opt->mode;
  std::string err = build_plan(src.la, po, &p->plan);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "%s", err.c_str());
  err = verify_plan(p->plan, src.la);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "schedule check failed: %s", err.c_str());
  Plan &P = p->plan;
  for (const PanelDesc &pd : P.panels) p->has_tp = p->has_tp || pd.vec_sel == kVecTP;
  p->a_mixed = order_mixed(P);
  HIP_TRY(hipGetDevice(&p->device));
  HIP_TRY(hipStreamCreate(&p->stream));
  HIP_TRY(hipStreamCreate(&p->stream2));
  HIP_TRY(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
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
      (rc = upload(&p->d_panels, P.panels)) ||
      (rc = upload(&p->d_ylist, P.ysum.list)) || (rc = upload(&p->d_yrng, seg_ranges(P.ysum))))
    return rc;
  {
    std::vector<int4> all, multi;
    if (!t_descs(P.tsum, {}, &all))
      return set_error(HMVM_ERR_UNSUPPORTED, "t sum segment with more than 65535 chunks or an irregular stride");
    if (all.size() != (size_t)P.tsum.n) return set_error(HMVM_ERR_UNSUPPORTED, "t entry without a t sum segment");
    std::vector<int2> one(all.size());  // every entry, in entry order
    for (size_t i = 0; i < all.size(); i++) one[i] = make_int2(all[i].x, all[i].y);
    if ((rc = upload(&p->d_tdesc, one))) return rc;
    p->n_tsum = (int)one.size();
    if (!P.t_direct && !P.tmap.empty()) {  // multi-RHS: the entries the row maps write are not summed
      std::vector<char> mapped((size_t)P.n_t, 0);
      for (int32_t e : P.tmap) mapped[(size_t)e] = 1;
      (void)t_descs(P.tsum, mapped, &multi);  // a subset of `all`, which passed
    } else {
      multi.swap(all);
    }
    if ((rc = upload(&p->d_tdesc_m, multi))) return rc;
    p->n_tsum_m = (int)multi.size();
  }
  if (p->mode != 0 && (rc = partition_a(p.get()))) return rc;
  if (!P.tmap.empty() && (rc = upload(&p->d_tmap, P.tmap))) return rc;
  if (p->storage == kBF16) {
    if ((rc = upload(&p->d_tfac, P.factor_of_trow))) return rc;
    if ((rc = dalloc(&p->d_scales, (size_t)P.nfactors))) return rc;
    std::vector<float> ones((size_t)P.nfactors, 1.0f);
    if ((rc = h2d(p->d_scales, ones.data(), ones.size()))) return rc;
  }
  if ((rc = dalloc(&p->d_t, (size_t)P.n_t)) || (rc = dalloc_bytes(&p->d_slots, slot_bytes(p->storage) * (size_t)(P.n_tp + P.n_z))) ||
      (rc = dalloc(&p->d_x, (size_t)P.nd)) ||
      (rc = dalloc(&p->d_y, (size_t)P.nd)))
    return rc;
  HIP_TRY(hipMalloc(&p->d_vals, eb * (size_t)std::max<int64_t>(P.nvals, 1)));
  // column padding is read (and discarded) by the 16-byte lane loads: keep it zero
  HIP_TRY(hipMemsetAsync(p->d_vals, 0, eb * (size_t)std::max<int64_t>(P.nvals, 1), p->stream));

  // values: streamed to the GPU in windows and repacked there
  if ((rc = upload_values(p.get(), src))) return rc;
  if (p->storage == kBF16 && !P.tmap.empty()) {
    if ((rc = dalloc(&p->d_mscale, P.tmap.size()))) return rc;
    hipLaunchKernelGGL(map_scale_kernel, dim3((unsigned)((P.tmap.size() + 255) / 256)), dim3(256), 0, p->stream,
                       (int)P.tmap.size(), p->d_tmap, p->d_tfac, (const float *)p->d_scales, p->d_mscale);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(p->stream));
  }
  if (p->mode == 1 && (rc = persist_prepare(p.get()))) return rc;

  p->device_bytes = eb * (size_t)P.nvals + sizeof(TileDesc) * (P.tiles_a.size() + P.tiles_b.size()) +
                    sizeof(PanelDesc) * P.panels.size() + 8 * (size_t)(P.n_t + 2 * P.nd) +
                    slot_bytes(p->storage) * (size_t)(P.n_tp + P.n_z) +
                    4 * (size_t)P.n_t + (p->d_mscale ? 12 : 4) * P.tmap.size() +
                    sizeof(int2) * (size_t)p->n_tsum + sizeof(int4) * (size_t)p->n_tsum_m +
                    8 * (P.ysum.list.size() + P.ysum.seg.size());
  if (verbose())
    std::fprintf(stderr,
                 "hmvm_hip: plan nd=%d nlf=%d len=%lld storage=%d tiles=%zu+%zu panels=%zu t=%lld tslots=%lld "
                 "zslots=%lld\n",
                 P.nd, P.nlf, (long long)P.len, p->storage, P.tiles_a.size(), P.tiles_b.size(), P.panels.size(),
                 (long long)P.n_t, (long long)P.n_tp, (long long)P.n_z);
  *out = p.release();
  return HMVM_OK;
}

// Where the tiles of one launch store their rows: the slots, and for
// stage-1 tiles with a row map the plan's t (or the multi-RHS T).
TileOut tile_out(const hmvm_plan *p, void *out, void *tout) {
  return TileOut{out, p->d_tmap, tout, p->d_mscale};
}

// kind: unused by the one-wave-per-tile launch (kept so every launch names
// its part of the HMVM: 0 A, 1 B, 2 stage-1 part of A, 3 dense part)
template <typename S, bool kScaled, int kVar>
void launch_tiles_v(hmvm_plan *p, int kind, int bt, hipStream_t st, const TileDesc *d_tiles, int ntiles,
                    const double *x, const double *t, const void *tp, void *out) {
  (void)kind;
  constexpr int L = 8;
  constexpr bool kBV = (kVar & 1) != 0;
  constexpr bool kPipe = (kVar & 2) != 0, kNT = (kVar & 4) != 0;
  const int wpb = bt / 64;
  const size_t lds = sizeof(double) * wpb * lds_wave_doubles<S>();
  hipLaunchKernelGGL((tile_kernel<S, kScaled, L, kPipe, kNT, kBV>), dim3((ntiles + wpb - 1) / wpb), dim3(bt), lds, st,
                     d_tiles, ntiles, p->d_panels, (const S *)p->d_vals, x, t,
                     (const typename SlotOf<S>::T *)tp, (const float *)p->d_scales, tile_out(p, out, p->d_t));
}

template <typename S, bool kScaled>
void launch_tiles_s(hmvm_plan *p, int kern, int kind, int bt, hipStream_t st, const TileDesc *d_tiles, int ntiles,
                    const double *x, const double *t, const void *tp, void *out) {
#define HMVM_TILE_CASE(v) \
  case v: launch_tiles_v<S, kScaled, v>(p, kind, bt, st, d_tiles, ntiles, x, t, tp, out); break;
  switch (kern) {
    HMVM_TILE_CASE(0) HMVM_TILE_CASE(1) HMVM_TILE_CASE(2) HMVM_TILE_CASE(3)
    HMVM_TILE_CASE(4) HMVM_TILE_CASE(5) HMVM_TILE_CASE(6)
    default: launch_tiles_v<S, kScaled, 7>(p, kind, bt, st, d_tiles, ntiles, x, t, tp, out); break;
  }
#undef HMVM_TILE_CASE
}

template <typename S>
void launch_tiles(hmvm_plan *p, int which, int kind, const TileDesc *d_tiles, int ntiles, const double *x,
                  const double *t, const void *tp, void *out, bool scaled, hipStream_t st) {
  if (ntiles == 0) return;
  const int bt = which ? p->block_threads_b : p->block_threads;
  const int kern = which ? p->kernel_b : p->kernel;
  if (scaled)
    launch_tiles_s<S, true>(p, kern, kind, bt, st, d_tiles, ntiles, x, t, tp, out);
  else
    launch_tiles_s<S, false>(p, kern, kind, bt, st, d_tiles, ntiles, x, t, tp, out);
}

// which: 0 launch A (all), 1 launch B, 2 stage-1 tiles of A only, 3 dense tiles of A only
template <typename S>
void launch_tiles_both(hmvm_plan *p, int which, const double *x, hipStream_t st) {
  const Plan &P = p->plan;
  const bool bf = p->storage == kBF16;
  const int na = (int)P.tiles_a.size(), n1 = P.n_stage1_tiles;
  if (which == 0)
    launch_tiles<S>(p, 0, 0, p->d_tiles_a, na, x, nullptr, nullptr, p->d_slots, bf, st);
  else if (which == 2)
    launch_tiles<S>(p, 0, 2, p->d_tiles_a, n1, x, nullptr, nullptr, p->d_slots, bf, st);
  else if (which == 3)
    launch_tiles<S>(p, 0, 3, p->d_tiles_a + n1, na - n1, x, nullptr, nullptr, p->d_slots, bf, st);
  else
    launch_tiles<S>(p, 1, 1, p->d_tiles_b, (int)P.tiles_b.size(), x, p->d_t, p->d_slots, p->d_slots, bf, st);
}

void launch_tiles_part(hmvm_plan *p, int which, const double *x, hipStream_t st) {
  if (p->storage == kF64)
    launch_tiles_both<double>(p, which, x, st);
  else if (p->storage == kF32)
    launch_tiles_both<float>(p, which, x, st);
  else
    launch_tiles_both<uint16_t>(p, which, x, st);
}

// ---- per-tile timing (cost-model training data) -----------------------------
template <typename S, bool kScaled, int kVar>
void launch_timed_v(hmvm_plan *p, int bt, const TileDesc *d_tiles, int ntiles, const double *x, const double *t,
                    const void *tp, void *out, long long *stamps, hipStream_t st) {
  constexpr int L = 8;
  constexpr bool kBV = (kVar & 1) != 0;
  constexpr bool kPipe = (kVar & 2) != 0, kNT = (kVar & 4) != 0;
  const int wpb = bt / 64;
  const size_t lds = sizeof(double) * wpb * lds_wave_doubles<S>();
  hipLaunchKernelGGL((tile_kernel_timed<S, kScaled, L, kPipe, kNT, kBV>), dim3((ntiles + wpb - 1) / wpb),
                     dim3(bt), lds, st, d_tiles, ntiles, p->d_panels, (const S *)p->d_vals, x, t,
                     (const typename SlotOf<S>::T *)tp, (const float *)p->d_scales, tile_out(p, out, p->d_t), stamps);
}

template <typename S, bool kScaled>
void launch_timed_s(hmvm_plan *p, int which, const TileDesc *d_tiles, int ntiles, const double *x, const double *t,
                    const void *tp, void *out, long long *stamps, hipStream_t st) {
  const int bt = which ? p->block_threads_b : p->block_threads;
  switch (which ? p->kernel_b : p->kernel) {
    case 0: launch_timed_v<S, kScaled, 0>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    case 1: launch_timed_v<S, kScaled, 1>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    case 2: launch_timed_v<S, kScaled, 2>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    case 3: launch_timed_v<S, kScaled, 3>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    case 4: launch_timed_v<S, kScaled, 4>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    case 5: launch_timed_v<S, kScaled, 5>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    case 6: launch_timed_v<S, kScaled, 6>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
    default: launch_timed_v<S, kScaled, 7>(p, bt, d_tiles, ntiles, x, t, tp, out, stamps, st); break;
  }
}

// Timed launch of the tiles of launch A (which = 0) or B (which = 1).
void launch_timed(hmvm_plan *p, int which, const double *x, long long *stamps, hipStream_t st) {
  const Plan &P = p->plan;
  const TileDesc *tl = which == 0 ? p->d_tiles_a : p->d_tiles_b;
  const int nt = (int)(which == 0 ? P.tiles_a.size() : P.tiles_b.size());
  if (nt == 0) return;
  const double *t = which == 0 ? nullptr : p->d_t;
  const void *tp = which == 0 ? nullptr : p->d_slots;
  if (p->storage == kF64)
    launch_timed_s<double, false>(p, which, tl, nt, x, t, tp, p->d_slots, stamps, st);
  else if (p->storage == kF32)
    launch_timed_s<float, false>(p, which, tl, nt, x, t, tp, p->d_slots, stamps, st);
  else
    launch_timed_s<uint16_t, true>(p, which, tl, nt, x, t, tp, p->d_slots, stamps, st);
}

// t segment sum (phase 1) or y segment sum (phase 3) over slots of type P.
template <typename P>
void launch_segsum(hmvm_plan *p, int phase, double *y, int accumulate, hipStream_t st) {
  const Plan &P_ = p->plan;
  const P *slots = (const P *)p->d_slots;
  if (phase == 1) {
    if (P_.t_fused || p->n_tsum == 0) return;
    const dim3 grid((p->n_tsum + 255) / 256);
    if (p->storage == kBF16)
      hipLaunchKernelGGL((tsum_kernel<true, P>), grid, dim3(256), 0, st, p->n_tsum, p->d_tdesc, slots, p->d_tfac,
                         p->d_scales, p->d_t);
    else
      hipLaunchKernelGGL((tsum_kernel<false, P>), grid, dim3(256), 0, st, p->n_tsum, p->d_tdesc, slots,
                         (const int32_t *)nullptr, (const float *)nullptr, p->d_t);
  } else {
    hipLaunchKernelGGL((ysum_kernel<P>), dim3((P_.ysum.n + 255) / 256), dim3(256), 0, st, P_.ysum.n, p->d_yrng,
                       p->d_ylist, slots, y, accumulate);
  }
}

// Launch `phase` of one HMVM: 0 = tiles A (stage 1 + dense), 1 = t segment sum
// (only when some leaf's rank rows straddle two row groups), 2 = tiles B
// (stage 2), 3 = y segment sum.
void launch_phase(hmvm_plan *p, int phase, const double *x, double *y, int accumulate, hipStream_t st) {
  if (phase == 0 || phase == 2) {
    const int which = phase == 0 ? 0 : 1;
    if (p->storage == kF64)
      launch_tiles_both<double>(p, which, x, st);
    else if (p->storage == kF32)
      launch_tiles_both<float>(p, which, x, st);
    else
      launch_tiles_both<uint16_t>(p, which, x, st);
  } else if (p->storage == kF64) {
    launch_segsum<double>(p, phase, y, accumulate, st);
  } else {
    launch_segsum<float>(p, phase, y, accumulate, st);
  }
}

// ---- several right-hand sides -------------------------------------------------
// Workgroup w of a launch runs on XCD w % kXcds: MI355X hands the workgroups
// of a grid to its 8 XCDs round robin, and each XCD has its own L2.  A
// multi-RHS tile reads a slice of X (launch A) or t (launch B) of nrhs
// entries per column, which other tiles over the same columns read again.
// Sorted by the start of that slice, the tiles are cut into kXcds runs of
// equal work (elements), and laid out so that the workgroups of XCD k take
// run k's tiles, largest first: each XCD then reads about one kXcds-th of the
// vector, which its L2 holds, instead of all of it.  A workgroup whose run is
// used up takes tiles of the run with the most left.
std::vector<TileDesc> xcd_layout(const std::vector<TileDesc> &tiles, int wpb) {
  const size_t n = tiles.size();
  std::vector<size_t> idx(n);
  for (size_t i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t a, size_t b) { return tiles[a].first.vec_off < tiles[b].first.vec_off; });
  auto work = [&](size_t i) { return (double)tiles[i].nrows * (double)tiles[i].ncols; };
  double total = 0;
  for (size_t i = 0; i < n; i++) total += work(i);
  std::vector<std::vector<size_t>> run(kXcds);
  double cum = 0;
  for (size_t i : idx) {
    const int r = total > 0 ? std::min(kXcds - 1, (int)((cum + 0.5 * work(i)) * kXcds / total)) : 0;
    run[r].push_back(i);
    cum += work(i);
  }
  for (auto &r : run)
    std::stable_sort(r.begin(), r.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
  std::vector<TileDesc> out;
  out.reserve(n);
  std::vector<size_t> pos(kXcds, 0);
  for (int w = 0; out.size() < n; w++) {
    for (int k = 0; k < wpb && out.size() < n; k++) {
      int r = w % kXcds;
      if (pos[r] == run[r].size()) {  // run used up: the run with the most tiles left
        for (int q = 0; q < kXcds; q++)
          if (run[q].size() - pos[q] > run[r].size() - pos[r]) r = q;
      }
      out.push_back(tiles[run[r][pos[r]++]]);
    }
  }
  return out;
}

// Uploads the XCD-aware layouts of both launches for the plan's current
// workgroup size (again when that changes).
int multi_layout(hmvm_plan *p) {
  const int wpb = p->block_threads / 64;
  if (p->mtiles_wpb == wpb && p->d_mtiles_a) return HMVM_OK;
  for (TileDesc **d : {&p->d_mtiles_a, &p->d_mtiles_b})
    if (*d) {
      (void)hipFree(*d);
      *d = nullptr;
    }
  p->mtiles_wpb = 0;
  // the multi-RHS launches send the rows of stage-1 tiles with a
  // multi-RHS row map to t
  std::vector<TileDesc> ta = p->plan.tiles_a;
  for (TileDesc &t : ta) t.map_off = t.map_off_multi;
  int rc;
  if ((rc = upload(&p->d_mtiles_a, xcd_layout(ta, wpb))) ||
      (rc = upload(&p->d_mtiles_b, xcd_layout(p->plan.tiles_b, wpb))))
    return rc;
  p->mtiles_wpb = wpb;
  return HMVM_OK;
}

int reserve_multi(hmvm_plan *p, int nrhs) {
  if (nrhs != 16 && nrhs != 32 && nrhs != 64) return set_error(HMVM_ERR_ARG, "nrhs must be 16, 32 or 64");
  int rc = multi_layout(p);
  if (rc) return rc;
  if (p->multi_cap >= nrhs) return HMVM_OK;
  const Plan &P = p->plan;
  void *bufs[5] = {p->d_mslots, p->d_mt, p->d_mX, p->d_mY, p->d_mXc};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  p->d_mslots = p->d_mt = nullptr;
  p->d_mX = p->d_mY = nullptr;
  p->d_mXc = nullptr;
  p->multi_cap = 0;
  if ((rc = dalloc_bytes(&p->d_mslots, slot_bytes(p->storage) * (size_t)(P.n_tp + P.n_z) * nrhs)) ||
      (rc = dalloc_bytes(&p->d_mt, vec_bytes(p->storage) * (size_t)P.n_t * nrhs)) ||
      (rc = dalloc(&p->d_mX, (size_t)P.nd * nrhs)) || (rc = dalloc(&p->d_mY, (size_t)P.nd * nrhs)))
    return rc;
  if (p->storage != kF64 && (rc = dalloc_bytes(&p->d_mXc, vec_bytes(p->storage) * (size_t)P.nd * nrhs))) return rc;
  p->multi_cap = nrhs;
  return HMVM_OK;
}

// T: the t read by stage-2 panels (launch B; NULL in launch A, whose
// stage-1 tiles with a row map write it instead)
template <typename S, bool kScaled, int NB>
void launch_multi_tiles(hmvm_plan *p, const TileDesc *d_tiles, int ntiles, const void *Xv, const void *Tv,
                        hipStream_t st) {
  if (ntiles == 0) return;
  const int wpb = p->block_threads / 64;
  const TileOut o = tile_out(p, p->d_mslots, p->d_mt);
  const auto *TP = (const typename SlotOf<S>::T *)p->d_mslots;
  const auto *X = (const typename VecOf<S>::T *)Xv;
  const auto *T = (const typename VecOf<S>::T *)Tv;
  if (p->has_tp)
    hipLaunchKernelGGL((tile_multi_kernel<S, kScaled, NB, true>), dim3((ntiles + wpb - 1) / wpb),
                       dim3(p->block_threads), 0, st, d_tiles, ntiles, p->d_panels, (const S *)p->d_vals, X, T,
                       TP, (const float *)p->d_scales, o);
  else
    hipLaunchKernelGGL((tile_multi_kernel<S, kScaled, NB, false>), dim3((ntiles + wpb - 1) / wpb),
                       dim3(p->block_threads), 0, st, d_tiles, ntiles, p->d_panels, (const S *)p->d_vals, X, T,
                       TP, (const float *)p->d_scales, o);
}

template <typename S, bool kScaled>
void launch_multi_nb(hmvm_plan *p, const TileDesc *d_tiles, int ntiles, const void *X, const void *T, int nrhs,
                     hipStream_t st) {
  if (nrhs == 16)
    launch_multi_tiles<S, kScaled, 1>(p, d_tiles, ntiles, X, T, st);
  else if (nrhs == 32)
    launch_multi_tiles<S, kScaled, 2>(p, d_tiles, ntiles, X, T, st);
  else
    launch_multi_tiles<S, kScaled, 4>(p, d_tiles, ntiles, X, T, st);
}

// Multi-RHS t (phase 1) or y (phase 3) segment sum over slots of type P.
template <typename P>
void launch_segsum_multi(hmvm_plan *p, int phase, double *Y, int nrhs, int accumulate, hipStream_t st) {
  const Plan &P_ = p->plan;
  const P *slots = (const P *)p->d_mslots;
  if (phase == 1) {
    // 16-byte loads and stores (slots and t are the plan's own allocations)
    const int4 *desc = p->d_tdesc_m;
    const int n = p->n_tsum_m;
    if (P_.t_fused || n == 0) return;
    // t in the multi-RHS operand type (VecOf: FP64, FP32 or BF16 bits)
    const int64_t tw = (int64_t)n * (nrhs / (16 / (int)sizeof(P)));
    const dim3 grid((unsigned)((tw + 255) / 256));
    P *T = (P *)p->d_mt;
    if (p->storage == kBF16)
      hipLaunchKernelGGL((tsum_multi_kernel<true, P, uint16_t>), grid, dim3(256), 0, st, n, nrhs, desc, slots,
                         p->d_tfac, p->d_scales, (uint16_t *)p->d_mt);
    else
      hipLaunchKernelGGL((tsum_multi_kernel<false, P, P>), grid, dim3(256), 0, st, n, nrhs, desc,
                         slots, (const int32_t *)nullptr, (const float *)nullptr, T);
  } else if (((uintptr_t)Y & 15) == 0) {  // 16-byte slot loads where the caller's Y is 16-byte aligned
    constexpr int W = 16 / sizeof(P);
    const int64_t tot = (int64_t)P_.ysum.n * (nrhs / W);
    hipLaunchKernelGGL((ysum_multi_kernel<W, P>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, P_.ysum.n,
                       nrhs, p->d_yrng, p->d_ylist, slots, Y, accumulate);
  } else {
    const int64_t tot = (int64_t)P_.ysum.n * nrhs;
    hipLaunchKernelGGL((ysum_multi_kernel<1, P>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, P_.ysum.n,
                       nrhs, p->d_yrng, p->d_ylist, slots, Y, accumulate);
  }
}

void launch_multi_phase(hmvm_plan *p, int phase, const double *X, double *Y, int nrhs, int accumulate,
                        hipStream_t st) {
  const Plan &P = p->plan;
  if (phase == 0 || phase == 2) {
    const TileDesc *tl = phase == 0 ? p->d_mtiles_a : p->d_mtiles_b;
    const int nt = (int)(phase == 0 ? P.tiles_a.size() : P.tiles_b.size());
    const void *T = phase == 0 ? nullptr : p->d_mt;
    if (p->storage != kF64) {
      // the tiles read X and t in the operand type: a copy of X made here,
      // before launch A, and a t the t sum writes so
      if (phase == 0) {
        const size_t n = (size_t)P.nd * nrhs;
        const unsigned blocks = (unsigned)std::min<size_t>((n / 2 + 255) / 256, 8192);
        const bool a16 = ((uintptr_t)X & 15) == 0;
        if (p->storage == kF32 && a16)
          hipLaunchKernelGGL((to_vec_kernel<2, float>), dim3(blocks), dim3(256), 0, st, X, (float *)p->d_mXc, n);
        else if (p->storage == kF32)
          hipLaunchKernelGGL((to_vec_kernel<1, float>), dim3(2 * blocks), dim3(256), 0, st, X, (float *)p->d_mXc, n);
        else if (a16)
          hipLaunchKernelGGL((to_vec_kernel<2, uint16_t>), dim3(blocks), dim3(256), 0, st, X, (uint16_t *)p->d_mXc,
                             n);
        else
          hipLaunchKernelGGL((to_vec_kernel<1, uint16_t>), dim3(2 * blocks), dim3(256), 0, st, X,
                             (uint16_t *)p->d_mXc, n);
      }
    }
    if (p->storage == kF64 && nrhs > 16 && ((uintptr_t)X & 15) != 0) {
      // the tiles load two or four right-hand sides at once (16-byte loads):
      // an X that is not 16-byte aligned is first copied to the plan's own
      if (phase == 0)
        (void)hipMemcpyAsync(p->d_mX, X, sizeof(double) * (size_t)P.nd * nrhs, hipMemcpyDeviceToDevice, st);
      X = p->d_mX;
    }
    const void *Xv = p->storage == kF64 ? (const void *)X : (const void *)p->d_mXc;
    if (p->storage == kF64)
      launch_multi_nb<double, false>(p, tl, nt, Xv, T, nrhs, st);
    else if (p->storage == kF32)
      launch_multi_nb<float, false>(p, tl, nt, Xv, T, nrhs, st);
    else
      launch_multi_nb<uint16_t, true>(p, tl, nt, Xv, T, nrhs, st);
  } else if (p->storage == kF64) {
    launch_segsum_multi<double>(p, phase, Y, nrhs, accumulate, st);
  } else {
    launch_segsum_multi<float>(p, phase, Y, nrhs, accumulate, st);
  }
}

int apply_multi(hmvm_plan *p, const double *X, double *Y, int nrhs, int accumulate, hipStream_t st) {
  if (p->multi_cap < nrhs) return set_error(HMVM_ERR_ARG, "call hmvm_plan_reserve_multi(plan, %d) first", nrhs);
  int rc = multi_layout(p);  // no-op unless the workgroup size changed since hmvm_plan_reserve_multi
  if (rc) return rc;
  for (int ph = 0; ph < 4; ph++) launch_multi_phase(p, ph, X, Y, nrhs, accumulate, st);
  HIP_TRY(hipGetLastError());
  return HMVM_OK;
}

// ---- persistent launch ---------------------------------------------------------
// A tile's fixed cost in element-equivalents (descriptor, panel table, first
// round trip), for dealing tiles to waves when their durations could not be
// measured.
constexpr double kTileOverheadElems = 1024;

// Per-tile cost of launch A (which = 0) or B (1): the tile's duration in the
// launches, measured by the clock-stamping tile kernel (two timed runs);
// elements plus a fixed overhead if that fails.  Dealt by element counts
// alone, the waves of the persistent launch finished their launch-A tiles up
// to 60 us apart at N=101,250 (FP64: 175 against 235 us).
std::vector<double> tile_costs(hmvm_plan *p, int which) {
  const std::vector<TileDesc> &tl = which == 0 ? p->plan.tiles_a : p->plan.tiles_b;
  std::vector<double> c(tl.size()), s(tl.size());
  double lms = 0;
  if (!tl.empty() && hmvm_plan_profile_tiles(p, which, 2, c.data(), s.data(), &lms) == HMVM_OK) return c;
  for (size_t i = 0; i < tl.size(); i++) c[i] = (double)tl[i].nrows * (double)tl[i].ncols + kTileOverheadElems;
  return c;
}

// Deals a launch's tiles to nw waves, longest first, each to the wave with
// the least work so far (what the hardware dispatcher does with one tile per
// wave); wave w's tiles, in the order dealt, are out[ptr[w] .. ptr[w+1]).
void deal_tiles(const std::vector<TileDesc> &tiles, const std::vector<double> &costs, int nw,
                std::vector<TileDesc> *out, std::vector<int32_t> *ptr) {
  const size_t n = tiles.size();
  std::vector<size_t> idx(n);
  for (size_t i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return costs[a] > costs[b]; });
  std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>, std::greater<>> load;
  for (int w = 0; w < nw; w++) load.push({0.0, w});
  std::vector<std::vector<size_t>> per((size_t)nw);
  for (size_t i : idx) {
    const std::pair<double, int> top = load.top();
    load.pop();
    per[(size_t)top.second].push_back(i);
    load.push({top.first + costs[i], top.second});
  }
  out->clear();
  out->reserve(n);
  ptr->assign((size_t)nw + 1, 0);
  for (int w = 0; w < nw; w++) {
    (*ptr)[(size_t)w] = (int32_t)out->size();
    for (size_t i : per[(size_t)w]) out->push_back(tiles[i]);
  }
  (*ptr)[(size_t)nw] = (int32_t)out->size();
}

// (four loads in flight per lane group rather than the tile launches' eight:
// the stage loop around the tile code then fits 128 registers, 4 waves per
// SIMD; with eight it spilled)
template <typename S, bool kScaled, int kVar>
auto persistent_fn() {
  constexpr int L = 4;
  return persistent_kernel<S, kScaled, L, (kVar & 2) != 0, (kVar & 4) != 0, (kVar & 1) != 0>;
}

// Workgroups of the persistent launch: as many as the device holds at once
// (the grid barriers need every workgroup resident).
template <typename S, bool kScaled, int kVar>
int persistent_blocks_v(hmvm_plan *p) {
  const size_t lds = sizeof(double) * (p->block_threads / 64) * lds_wave_doubles<S>();
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_fn<S, kScaled, kVar>(), p->block_threads,
                                                   lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * std::max(p->ncu, 1);
}

template <typename S, bool kScaled>
int persistent_blocks_s(hmvm_plan *p) {
  switch (p->kernel) {
    case 0: return persistent_blocks_v<S, kScaled, 0>(p);
    case 1: return persistent_blocks_v<S, kScaled, 1>(p);
    case 2: return persistent_blocks_v<S, kScaled, 2>(p);
    case 3: return persistent_blocks_v<S, kScaled, 3>(p);
    case 4: return persistent_blocks_v<S, kScaled, 4>(p);
    case 5: return persistent_blocks_v<S, kScaled, 5>(p);
    case 6: return persistent_blocks_v<S, kScaled, 6>(p);
    default: return persistent_blocks_v<S, kScaled, 7>(p);
  }
}

// Deals both launches' tiles to the persistent launch's waves (again when the
// launch shape or the tile lists changed since).
int persist_prepare(hmvm_plan *p) {
  const int key = p->kernel * 1024 + p->block_threads;
  if (p->persist_key == key && p->persist_version == p->tiles_version && p->d_ptiles_a) return HMVM_OK;
  const int blocks = p->storage == kF64   ? persistent_blocks_s<double, false>(p)
                     : p->storage == kF32 ? persistent_blocks_s<float, false>(p)
                                          : persistent_blocks_s<uint16_t, true>(p);
  const int nw = blocks * (p->block_threads / 64);
  for (void *d : {(void *)p->d_ptiles_a, (void *)p->d_ptiles_b, (void *)p->d_pptr_a, (void *)p->d_pptr_b})
    if (d) (void)hipFree(d);
  p->d_ptiles_a = p->d_ptiles_b = nullptr;
  p->d_pptr_a = p->d_pptr_b = nullptr;
  p->persist_key = -1;
  std::vector<TileDesc> ta, tb;
  std::vector<int32_t> pa, pb;
  deal_tiles(p->plan.tiles_a, tile_costs(p, 0), nw, &ta, &pa);
  deal_tiles(p->plan.tiles_b, tile_costs(p, 1), nw, &tb, &pb);
  int rc;
  if ((rc = upload(&p->d_ptiles_a, ta)) || (rc = upload(&p->d_ptiles_b, tb)) || (rc = upload(&p->d_pptr_a, pa)) ||
      (rc = upload(&p->d_pptr_b, pb)))
    return rc;
  p->persist_blocks = blocks;
  p->persist_key = key;
  p->persist_version = p->tiles_version;
  if (verbose())
    std::fprintf(stderr, "hmvm_hip: persistent launch %d workgroups of %d threads (%d CUs), %zu + %zu tiles\n",
                 blocks, p->block_threads, p->ncu, p->plan.tiles_a.size(), p->plan.tiles_b.size());
  return HMVM_OK;
}

template <typename S, bool kScaled, int kVar>
void launch_persistent_v(hmvm_plan *p, const double *x, double *y, int accumulate, hipStream_t st) {
  const Plan &P = p->plan;
  const size_t lds = sizeof(double) * (p->block_threads / 64) * lds_wave_doubles<S>();
  const int n_t = P.t_fused || P.t_direct ? 0 : p->n_tsum;
  hipLaunchKernelGGL((persistent_fn<S, kScaled, kVar>()), dim3(p->persist_blocks), dim3(p->block_threads), lds, st,
                     p->d_ptiles_a, p->d_pptr_a, p->d_ptiles_b, p->d_pptr_b, p->d_panels, (const S *)p->d_vals, x,
                     p->d_t, (typename SlotOf<S>::T *)p->d_slots, (const float *)p->d_scales, p->d_tmap, p->d_mscale, p->d_tfac,
                     n_t, p->d_tdesc, P.nd, p->d_yrng, p->d_ylist, y, accumulate, p->d_ctr);
}

template <typename S, bool kScaled>
void launch_persistent_s(hmvm_plan *p, const double *x, double *y, int accumulate, hipStream_t st) {
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

// The tile lists are dealt when mode 1 is selected and whenever the launch
// shape or the tiles change; dealing here (a synchronous upload) only
// happens if a caller changed them behind the plan's back.
int launch_persistent(hmvm_plan *p, const double *x, double *y, int accumulate, hipStream_t st) {
  const int rc = persist_prepare(p);
  if (rc) return rc;
  if (p->storage == kF64)
    launch_persistent_s<double, false>(p, x, y, accumulate, st);
  else if (p->storage == kF32)
    launch_persistent_s<float, false>(p, x, y, accumulate, st);
  else
    launch_persistent_s<uint16_t, true>(p, x, y, accumulate, st);
  return HMVM_OK;
}

// Mode 1 runs the persistent launch (any t_mode: it sums t itself when the
// plan materialises t).
bool persistent_ok(const hmvm_plan *p) { return p->mode == 1; }

// Dense tiles depend only on x and write their own output slots, so in mode 2
// they run on a second stream, concurrently with the stage-1 tiles, the t sum
// and the stage-2 tiles; the y sum waits for both.  The fork waits for
// everything earlier on the caller's stream (including a previous HMVM's y
// sum, which reads the same slots).  Measured on MI355X it does not beat one
// stream (the launches already fill the device), so it is not the default.
bool overlap_ok(const hmvm_plan *p) {
  const Plan &P = p->plan;
  return p->mode == 2 && !p->a_mixed && P.n_stage1_tiles > 0 && (int)P.tiles_a.size() > P.n_stage1_tiles;
}

int apply(hmvm_plan *p, const double *d_x, double *d_y, int accumulate, hipStream_t st) {
  if (persistent_ok(p)) {
    const int rc = launch_persistent(p, d_x, d_y, accumulate, st);
    if (rc) return rc;
  } else if (overlap_ok(p)) {
    HIP_TRY(hipEventRecord(p->ev_fork, st));
    HIP_TRY(hipStreamWaitEvent(p->stream2, p->ev_fork, 0));
    launch_tiles_part(p, 3, d_x, p->stream2);
    HIP_TRY(hipEventRecord(p->ev_join, p->stream2));
    launch_tiles_part(p, 2, d_x, st);
    launch_phase(p, 1, d_x, d_y, accumulate, st);
    launch_phase(p, 2, d_x, d_y, accumulate, st);
    HIP_TRY(hipStreamWaitEvent(st, p->ev_join, 0));
    launch_phase(p, 3, d_x, d_y, accumulate, st);
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
  info->mode = !p ? 0 : persistent_ok(p) ? 1 : overlap_ok(p) ? 2 : 0;
  info->t_direct = P.t_direct ? 1 : 0;
}

// Reads and clears the fault word of the persistent launches (after the
// plan's stream has drained).
int take_fault(hmvm_plan *p, int *fault) {
  *fault = 0;
  HIP_TRY(hipStreamSynchronize(p->stream));
  int f = 0;
  HIP_TRY(hipMemcpy(&f, &p->d_ctr->fault, sizeof(int), hipMemcpyDeviceToHost));
  if (f) {
    const int zero = 0;
    HIP_TRY(hipMemcpy(&p->d_ctr->fault, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  *fault = f;
  return HMVM_OK;
}

// For the blocking entry points, after their synchronisation: an error when
// a persistent launch of this plan gave up a bounded wait.
int check_wait(hmvm_plan *p) {
  if (!persistent_ok(p)) return HMVM_OK;
  int f = 0;
  const int rc = take_fault(p, &f);
  if (rc) return rc;
  if (!f) return HMVM_OK;
  int seen = 0;  // 16 * barrier + XCDs through when a barrier gave up
  HIP_TRY(hipMemcpy(&seen, &p->d_ctr->seen, sizeof(int), hipMemcpyDeviceToHost));
  return set_error(HMVM_ERR_WAIT,
                   "persistent launch of %d workgroups: a bounded wait gave up (barrier %d, %d XCDs through); the "
                   "results are invalid",
                   p->persist_blocks, seen / 16, seen % 16);
}

// ---------------------------------------------------------------------------
// cache of device copies for the drop-in entry points
struct CacheKey {
  const void *mat;
  int variant;  // 0:1 1:1t 2:2 3:2t 4:2 (int64 form) 5:2t (int64 form)
  bool operator<(const CacheKey &o) const { return mat != o.mat ? mat < o.mat : variant < o.variant; }
};
struct CacheEntry {
  hmvm_plan *plan = nullptr;
  // identity of the contents the plan was built from
  int nd = 0, nlf = 0;
  int64_t len = 0;
  const void *data = nullptr;
  const void *first_leaf_a1 = nullptr;
};
std::mutex g_cache_mu;
std::map<CacheKey, CacheEntry> g_cache;

hmvm_plan *cached_plan(const void *mat, int variant, int nd, int nlf, int64_t len, const void *data,
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

void cache_put(const void *mat, int variant, hmvm_plan *p, int nd, int nlf, int64_t len, const void *data,
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

hmvm_plan *plan_for64(const matrix2_64 *m, int variant) {
  if (!m) { set_error(HMVM_ERR_ARG, "matrix2_64 is NULL"); return nullptr; }
  const void *data = variant == 5 ? (const void *)m->rowmat_t : (const void *)m->rowmat;
  hmvm_plan *p = cached_plan(m, variant, m->nd, m->nlf, m->len, data, nullptr);
  if (p) return p;
  Source s;
  if (source_from_matrix2_64(m, variant == 5 ? kA2Transposed : kA2RowMajor, &s)) return nullptr;
  if (create_plan(s, nullptr, &p)) return nullptr;
  cache_put(m, variant, p, m->nd, m->nlf, m->len, data, nullptr);
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
  // measured best on MI355X for N = 101,250 (cube, eps 1e-4) in all three
  // storages: pipelined non-temporal loads, largest tiles first
  opt->storage = kF64;
  opt->w1 = 0;             // by storage (launch_defaults)
  opt->w2 = 0;
  opt->block_threads = 0;  // by storage, per launch
  opt->kernel = -1;        // by storage, per launch
  opt->order = 1;
  opt->mode = 0;
  opt->t_mode = -1;  // by storage
  opt->max_rows = 0;  // by storage
  opt->max_rows_b = 0;
  opt->w3 = 0;
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

int hmvm_plan_create64(const matrix2_64 *mat2, int a2_source, const hmvm_options *opt, hmvm_plan **out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  int rc = source_from_matrix2_64(mat2, a2_source, &s);
  if (rc) return rc;
  return create_plan(s, opt, out);
}

int hmvm_plan_create_file(const char *fname, const hmvm_options *opt, hmvm_plan **out) {
  clear_error();
  if (!out || !fname) return set_error(HMVM_ERR_ARG, "NULL argument");
  *out = nullptr;
  matrix2_64 m;
  long long data_pos = -1;
  int format = 0;
  int rc = hacapk_scan_file(fname, &m, &data_pos, &format);
  if (rc) return set_error(HMVM_ERR_MATRIX, "%s: %s", fname, hacapk_strerror(rc));
  Source s;
  s.a1.assign(m.a1, m.a1 + m.nlf);
  s.a2.assign(m.a2, m.a2 + m.nlf);
  LeafArrays &la = s.la;
  la.nd = m.nd;
  la.nlf = m.nlf;
  la.ltmtx = m.ltmtx; la.kt = m.kt; la.ndl = m.ndl; la.ndt = m.ndt;
  la.nstrtl = m.nstrtl; la.nstrtt = m.nstrtt;
  la.src_a1 = s.a1.data();
  la.src_a2 = s.a2.data();
  la.a2_source = kA2RowMajor;
  la.src_len = m.len;
  {
    FileFeeder feed(fname, format, &m);
    rc = feed.begin(data_pos);
    if (rc) {
      rc = set_error(HMVM_ERR_MATRIX, "%s: %s", fname, hacapk_strerror(rc));
    } else {
      s.feeder = &feed;
      rc = create_plan(s, opt, out);
    }
  }
  freeHmatrix64(&m);
  return rc;
}

int hmvm_plan_create_batch(const matrix2 *const *mats, int count, int a2_source, const hmvm_options *opt,
                           hmvm_plan **out, int *row_offsets) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  ConcatFeeder feed;
  std::vector<int> row0;
  int rc = source_from_batch(mats, count, a2_source, &s, &feed, &row0);
  if (rc) return rc;
  rc = create_plan(s, opt, out);
  if (!rc && row_offsets) std::memcpy(row_offsets, row0.data(), sizeof(int) * row0.size());
  return rc;
}

void hmvm_plan_destroy(hmvm_plan *plan) { delete plan; }

int hmvm_plan_set_launch(hmvm_plan *plan, int kernel, int block_threads) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  if (kernel < 0 || kernel >= kNumKernelVariants)
    return set_error(HMVM_ERR_ARG, "kernel variant must be in [0,%d)", kNumKernelVariants);
  if (block_threads % 64 || block_threads < 64 || block_threads > 256)
    return set_error(HMVM_ERR_ARG, "block_threads must be 64, 128, 192 or 256");
  plan->kernel = plan->kernel_b = kernel;
  plan->block_threads = plan->block_threads_b = block_threads;
  return plan->mode == 1 ? persist_prepare(plan) : HMVM_OK;
}

int hmvm_plan_reset_launch(hmvm_plan *plan) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  const LaunchDefaults &ld = launch_defaults(plan->storage);
  plan->kernel = ld.kernel_a;
  plan->kernel_b = ld.kernel_b;
  plan->block_threads = ld.bt_a;
  plan->block_threads_b = ld.bt_b;
  return plan->mode == 1 ? persist_prepare(plan) : HMVM_OK;
}

int hmvm_plan_set_phase_launch(hmvm_plan *plan, int which, int kernel, int block_threads) {
  if (!plan || (which != 0 && which != 1)) return set_error(HMVM_ERR_ARG, "bad argument");
  if (kernel < 0 || kernel >= kNumKernelVariants)
    return set_error(HMVM_ERR_ARG, "kernel variant must be in [0,%d)", kNumKernelVariants);
  if (block_threads % 64 || block_threads < 64 || block_threads > 256)
    return set_error(HMVM_ERR_ARG, "block_threads must be 64, 128, 192 or 256");
  if (which == 0) {
    plan->kernel = kernel;
    plan->block_threads = block_threads;
  } else {
    plan->kernel_b = kernel;
    plan->block_threads_b = block_threads;
  }
  return plan->mode == 1 ? persist_prepare(plan) : HMVM_OK;
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
  EventSet<5> ev;
  HIP_TRY(ev.create());
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
  if (mode < 0 || mode > 2) return set_error(HMVM_ERR_ARG, "mode must be 0, 1 or 2");
  if (mode != 0) {
    const int rc = partition_a(plan);
    if (rc) return rc;
  }
  if (mode == 1) {
    const int rc = persist_prepare(plan);
    if (rc) return rc;
  }
  plan->mode = mode;
  return HMVM_OK;
}

int hmvm_plan_fault(hmvm_plan *plan) {
  if (!plan) return set_error(HMVM_ERR_ARG, "NULL plan");
  int f = 0;
  const int rc = take_fault(plan, &f);
  return rc ? rc : f;
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
  const PlanOptions po = plan_options(opt, s.la.src_len);
  Plan P;
  std::string err = build_plan(s.la, po, &P);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "%s", err.c_str());
  err = verify_plan(P, s.la);
  if (!err.empty()) return set_error(HMVM_ERR_MATRIX, "schedule check failed: %s", err.c_str());
  if (info) fill_info(nullptr, P, po.storage, info);
  return HMVM_OK;
}

}  // extern "C"

namespace {
int export_schedule(const Source &s, const hmvm_options *opt, hmvm_schedule *out) {
  Plan P;
  std::string err = build_plan(s.la, plan_options(opt, s.la.src_len), &P);
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
  out->t_direct = P.t_direct ? 1 : 0;
  out->ntmap = (long long)P.tmap.size();
  out->tmap = copy_out(P.tmap);
  if (!out->tiles_a || !out->tiles_b || !out->panels || !out->repack || !out->tsum.seg || !out->ysum.seg ||
      !out->tmap) {
    hmvm_schedule_release(out);
    return set_error(HMVM_ERR_NOMEM, "schedule export");
  }
  return HMVM_OK;
}
}  // namespace

extern "C" {

int hmvm_schedule_export(const matrix2 *mat2, int a2_source, const hmvm_options *opt, hmvm_schedule *out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  std::memset(out, 0, sizeof(*out));
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  int rc = source_from_matrix2(mat2, a2_source, &s);
  if (rc) return rc;
  return export_schedule(s, opt, out);
}

int hmvm_schedule_export64(const matrix2_64 *mat2, int a2_source, const hmvm_options *opt, hmvm_schedule *out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  std::memset(out, 0, sizeof(*out));
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  int rc = source_from_matrix2_64(mat2, a2_source, &s);
  if (rc) return rc;
  return export_schedule(s, opt, out);
}

int hmvm_schedule_export_batch(const matrix2 *const *mats, int count, int a2_source, const hmvm_options *opt,
                               hmvm_schedule *out) {
  clear_error();
  if (!out) return set_error(HMVM_ERR_ARG, "out is NULL");
  std::memset(out, 0, sizeof(*out));
  if (a2_source != kA2RowMajor && a2_source != kA2Transposed) return set_error(HMVM_ERR_ARG, "bad a2_source");
  Source s;
  ConcatFeeder feed;
  std::vector<int> row0;
  int rc = source_from_batch(mats, count, a2_source, &s, &feed, &row0);
  if (rc) return rc;
  return export_schedule(s, opt, out);
}

void hmvm_schedule_release(hmvm_schedule *s) {
  if (!s) return;
  void *ptrs[] = {s->tiles_a, s->tiles_b, s->panels, s->repack, s->tsum.seg, s->tsum.ptr, s->tsum.list,
                  s->ysum.seg, s->ysum.ptr, s->ysum.list, s->tmap};
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
  return check_wait(plan);
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
  return check_wait(plan);
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
  EventSet<5> ev;
  HIP_TRY(ev.create());
  std::vector<double> total;
  double ph[4] = {0, 0, 0, 0};
  int rc = HMVM_OK;
  // single applies: one event pair around each (nothing else in between)
  for (int it = 0; it < iters && !rc; it++) {
    hipEventRecord(ev[0], st);
    if ((rc = apply(plan, plan->d_x, plan->d_y, 0, st))) break;
    hipEventRecord(ev[4], st);
    if (hipEventSynchronize(ev[4]) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event sync failed");
    float ms = 0;
    if (!rc) hipEventElapsedTime(&ms, ev[0], ev[4]);
    total.push_back(ms);
  }
  // per-launch breakdown: a further iters applies with an event after every launch
  for (int it = 0; it < iters && !rc; it++) {
    hipEventRecord(ev[0], st);
    if (persistent_ok(plan)) {
      if ((rc = launch_persistent(plan, plan->d_x, plan->d_y, 0, st))) break;
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
  return check_wait(plan);
}

int hmvm_plan_time_flushed(hmvm_plan *plan, long long flush_bytes, int iters, hmvm_timing *tm) {
  if (!plan || !tm || iters < 1 || flush_bytes < 0) return set_error(HMVM_ERR_ARG, "bad argument");
  std::memset(tm, 0, sizeof(*tm));
  hipStream_t st = plan->stream;
  const int nd = plan->plan.nd;
  {
    std::vector<double> h((size_t)nd);
    for (int i = 0; i < nd; i++) h[i] = std::sin((double)(i + 1));
    HIP_TRY(hipMemcpy(plan->d_x, h.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  }
  const size_t n4 = (size_t)flush_bytes / sizeof(double4);
  struct Buf {  // released on every return path
    double4 *p = nullptr;
    ~Buf() {
      if (p) (void)hipFree(p);
    }
  } scratch;
  if (n4) HIP_TRY(hipMalloc(&scratch.p, n4 * sizeof(double4)));
  EventSet<2> ev;
  HIP_TRY(ev.create());
  double4 *buf = scratch.p;
  std::vector<double> total;
  int rc = apply(plan, plan->d_x, plan->d_y, 0, st);  // one untimed apply
  for (int it = 0; it < iters && !rc; it++) {
    if (n4)
      hipLaunchKernelGGL(flush_kernel, dim3(4096), dim3(256), 0, st, buf, n4, (double)it);
    hipEventRecord(ev[0], st);
    rc = apply(plan, plan->d_x, plan->d_y, 0, st);
    hipEventRecord(ev[1], st);
    if (!rc && hipEventSynchronize(ev[1]) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event sync failed");
    float ms = 0;
    if (!rc) hipEventElapsedTime(&ms, ev[0], ev[1]);
    total.push_back(ms);
  }
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
  tm->ms_stream = tm->ms_avg;
  return check_wait(plan);
}

// ---- hipGraph form of the apply -------------------------------------------
struct hmvm_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~hmvm_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

int hmvm_plan_graph_create(hmvm_plan *plan, const double *d_x, double *d_y, int accumulate, int count,
                           hmvm_graph **out) {
  if (!plan || !d_x || !d_y || !out || count < 1) return set_error(HMVM_ERR_ARG, "bad argument");
  *out = nullptr;
  std::unique_ptr<hmvm_graph> g(new hmvm_graph());
  hipStream_t st = plan->stream;
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = HMVM_OK;
  for (int i = 0; i < count && !rc; i++) rc = apply(plan, d_x, d_y, accumulate, st);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  g->graph = graph;
  if (rc) return rc;
  if (e != hipSuccess || !graph) return set_error(HMVM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return HMVM_OK;
}

int hmvm_graph_launch(hmvm_graph *graph, void *stream) {
  if (!graph || !graph->exec) return set_error(HMVM_ERR_ARG, "NULL graph");
  HIP_TRY(hipGraphLaunch(graph->exec, (hipStream_t)stream));
  return HMVM_OK;
}

void hmvm_graph_destroy(hmvm_graph *graph) { delete graph; }

int hmvm_plan_time_graph(hmvm_plan *plan, int count, int warmup, int iters, hmvm_timing *tm) {
  if (!plan || !tm || iters < 1 || warmup < 0 || count < 1) return set_error(HMVM_ERR_ARG, "bad argument");
  std::memset(tm, 0, sizeof(*tm));
  const int nd = plan->plan.nd;
  {
    std::vector<double> h((size_t)nd);
    for (int i = 0; i < nd; i++) h[i] = std::sin((double)(i + 1));
    HIP_TRY(hipMemcpy(plan->d_x, h.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  }
  hmvm_graph *g = nullptr;
  int rc = hmvm_plan_graph_create(plan, plan->d_x, plan->d_y, 0, count, &g);
  if (rc) return rc;
  hipStream_t st = plan->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<double> t;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event");
  for (int w = 0; w < warmup && !rc; w++) rc = hmvm_graph_launch(g, st);
  for (int it = 0; it < iters && !rc; it++) {
    hipEventRecord(e0, st);
    rc = hmvm_graph_launch(g, st);
    hipEventRecord(e1, st);
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event sync failed");
    float ms = 0;
    if (!rc) hipEventElapsedTime(&ms, e0, e1);
    t.push_back(ms / count);
  }
  if (!rc) {
    hipEventRecord(e0, st);
    for (int it = 0; it < iters && !rc; it++) rc = hmvm_graph_launch(g, st);
    hipEventRecord(e1, st);
    float ms = 0;
    if (!rc && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
    tm->ms_stream = ms / ((double)iters * count);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  hmvm_graph_destroy(g);
  if (rc) return rc;
  std::vector<double> srt = t;
  std::sort(srt.begin(), srt.end());
  double sum = 0;
  for (double v : t) sum += v;
  tm->iters = iters;
  tm->ms_avg = sum / iters;
  tm->ms_min = srt.front();
  tm->ms_median = srt[srt.size() / 2];
  return check_wait(plan);
}

int hmvm_plan_profile_tiles(hmvm_plan *plan, int which, int iters, double *tile_us, double *start_us,
                            double *launch_ms) {
  if (!plan || (which != 0 && which != 1) || iters < 1 || !tile_us) return set_error(HMVM_ERR_ARG, "bad argument");
  const Plan &P = plan->plan;
  const size_t nt = which == 0 ? P.tiles_a.size() : P.tiles_b.size();
  if (nt == 0) {
    if (launch_ms) *launch_ms = 0;
    return HMVM_OK;
  }
  int rate_khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, plan->device));
  if (rate_khz <= 0) return set_error(HMVM_ERR_UNSUPPORTED, "device wall clock rate unknown");
  const double us_per_tick = 1e3 / rate_khz;
  const int nd = plan->plan.nd;
  {
    std::vector<double> h((size_t)nd);
    for (int i = 0; i < nd; i++) h[i] = std::sin((double)(i + 1));
    HIP_TRY(hipMemcpy(plan->d_x, h.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  }
  long long *d_stamps = nullptr;
  int rc = dalloc(&d_stamps, 2 * nt);
  if (rc) return rc;
  // cache eviction before every timed launch: 512 MB, twice the last-level cache
  const size_t n4 = ((size_t)512 << 20) / sizeof(double4);
  double4 *d_flush = nullptr;
  if (hipMalloc(&d_flush, n4 * sizeof(double4)) != hipSuccess) {
    (void)hipFree(d_stamps);
    return set_error(HMVM_ERR_HIP, "hipMalloc of the 512 MB flush buffer failed");
  }
  std::vector<long long> h(2 * nt);
  std::vector<double> acc(nt, 0.0), acc0(nt, 0.0);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double lms = 0;
  hipStream_t st = plan->stream;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = set_error(HMVM_ERR_HIP, "event");
  for (int it = 0; it <= iters && !rc; it++) {  // iteration 0 warms up
    hipLaunchKernelGGL(flush_kernel, dim3(4096), dim3(256), 0, st, d_flush, n4, (double)it);
    if (which == 1) {
      launch_phase(plan, 0, plan->d_x, plan->d_y, 0, st);
      launch_phase(plan, 1, plan->d_x, plan->d_y, 0, st);
    }
    hipEventRecord(e0, st);
    launch_timed(plan, which, plan->d_x, d_stamps, st);
    hipEventRecord(e1, st);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = set_error(HMVM_ERR_HIP, "timed launch failed");
      break;
    }
    if (hipMemcpy(h.data(), d_stamps, sizeof(long long) * 2 * nt, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = set_error(HMVM_ERR_HIP, "stamp copy failed");
      break;
    }
    if (it == 0) continue;
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    lms += ms;
    long long t_min = h[0];
    for (size_t i = 0; i < nt; i++) t_min = std::min(t_min, h[2 * i]);
    for (size_t i = 0; i < nt; i++) {
      acc[i] += (double)(h[2 * i + 1] - h[2 * i]) * us_per_tick;
      acc0[i] += (double)(h[2 * i] - t_min) * us_per_tick;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(d_stamps);
  (void)hipFree(d_flush);
  if (rc) return rc;
  for (size_t i = 0; i < nt; i++) {
    tile_us[i] = acc[i] / iters;
    if (start_us) start_us[i] = acc0[i] / iters;
  }
  if (launch_ms) *launch_ms = lms / iters;
  return HMVM_OK;
}

}  // extern "C"

namespace {
template <typename S, bool kScaled, int kVar>
const void *tile_fn() {
  constexpr int L = 8;
  constexpr bool kBV = (kVar & 1) != 0;
  constexpr bool kPipe = (kVar & 2) != 0, kNT = (kVar & 4) != 0;
  return (const void *)tile_kernel<S, kScaled, L, kPipe, kNT, kBV>;
}
template <typename S, bool kScaled>
const void *tile_fn_v(int kernel) {
  switch (kernel) {
    case 0: return tile_fn<S, kScaled, 0>();
    case 1: return tile_fn<S, kScaled, 1>();
    case 2: return tile_fn<S, kScaled, 2>();
    case 3: return tile_fn<S, kScaled, 3>();
    case 4: return tile_fn<S, kScaled, 4>();
    case 5: return tile_fn<S, kScaled, 5>();
    case 6: return tile_fn<S, kScaled, 6>();
    default: return tile_fn<S, kScaled, 7>();
  }
}
}  // namespace

extern "C" {

int hmvm_tile_occupancy(int storage, int kernel, int block_threads, int *waves) {
  if (!waves || storage < kF64 || storage > kBF16 || kernel < 0 || kernel >= kNumKernelVariants ||
      block_threads % 64 || block_threads < 64 || block_threads > 256)
    return set_error(HMVM_ERR_ARG, "bad argument");
  // cached per (device, storage, variant, workgroup size): the features of
  // every tile of a launch need it
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, int> cache;
  int dev0 = 0;
  HIP_TRY(hipGetDevice(&dev0));
  const auto key = std::make_tuple(dev0, storage, kernel, block_threads);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *waves = it->second;
      return HMVM_OK;
    }
  }
  const void *fn = storage == kF64   ? tile_fn_v<double, false>(kernel)
                   : storage == kF32 ? tile_fn_v<float, false>(kernel)
                                     : tile_fn_v<uint16_t, true>(kernel);
  const int wave_doubles = storage == kF64   ? lds_wave_doubles<double>()
                           : storage == kF32 ? lds_wave_doubles<float>()
                                             : lds_wave_doubles<uint16_t>();
  const size_t lds = sizeof(double) * (block_threads / 64) * (size_t)wave_doubles;
  int per_cu = 0, dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block_threads, lds));
  *waves = per_cu * (block_threads / 64) * prop.multiProcessorCount;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = *waves;
  return HMVM_OK;
}

int hmvm_plan_tile_features(const hmvm_plan *plan, int which, int nrhs, float *feat) {
  if (!plan || (which != 0 && which != 1) || !feat) return set_error(HMVM_ERR_ARG, "bad argument");
  const Plan &P = plan->plan;
  const std::vector<TileDesc> &tl = which == 0 ? P.tiles_a : P.tiles_b;
  const int V = P.lane_elems;
  const float lb = std::log2((float)elem_bytes(plan->storage));
  const int kern = which ? plan->kernel_b : plan->kernel;
  const int bt = which ? plan->block_threads_b : plan->block_threads;
  int slots = 0;
  if (int rc = hmvm_tile_occupancy(plan->storage, kern, bt, &slots)) return rc;
  slots = std::max(slots, 1);
  const float f_sat = std::log2((float)std::min<size_t>(tl.size(), (size_t)slots));
  const float f_slots = std::log2((float)slots);
  for (size_t i = 0; i < tl.size(); i++) {
    const TileDesc &td = tl[i];
    // stage-1 panels carry no factor scale (factor -1); dense panels do
    const int kind = which == 1 ? 2 : (P.panels[td.panel0].factor < 0 ? 0 : 1);
    int64_t cols = 0;
    double nch = 0;
    for (int q = 0; q < td.npanel; q++) {
      const PanelDesc &pd = P.panels[td.panel0 + q];
      cols += pd.ncols;
      if (pd.vec_sel == kVecTP) nch += (double)(pd.vec_chunks >> 16) * pd.ncols;
    }
    const int shift = (int)(P.panels[td.panel0].val_off % V);
    const int lpc = (td.nrows + shift + V - 1) / V;
    float *f = feat + i * HMVM_NFEAT;
    f[0] = kind == 0; f[1] = kind == 1; f[2] = kind == 2;
    f[3] = std::log2((float)td.nrows);
    f[4] = std::log2(1.0f + (float)cols);
    f[5] = std::log2((float)td.npanel);
    f[6] = std::log2((float)lpc);
    f[7] = std::log2(1.0f + (float)lpc * V * (float)cols);
    f[8] = std::log2(1.0f + (float)(cols ? nch / cols : 0.0));
    f[9] = lb;
    f[10] = (float)(kern & 1);
    f[11] = (float)((kern >> 1) & 1);
    f[12] = (float)((kern >> 2) & 1);
    f[13] = std::log2((float)(bt / 64));
    f[14] = std::log2((float)std::max(nrhs, 1));
    f[15] = f_sat;
    f[16] = f_slots;
    f[17] = (lpc > 32 || (td.flags & kTileRowPerLane)) ? 1.0f : 0.0f;
  }
  return HMVM_OK;
}

int hmvm_plan_set_tile_flags(hmvm_plan *plan, int which, const int *flags) {
  if (!plan || (which != 0 && which != 1) || !flags) return set_error(HMVM_ERR_ARG, "bad argument");
  Plan &P = plan->plan;
  std::vector<TileDesc> &tl = which == 0 ? P.tiles_a : P.tiles_b;
  std::vector<TileDesc> nt = tl;
  for (size_t i = 0; i < nt.size(); i++) {
    if (flags[i] & ~HMVM_TILE_ROW_PER_LANE) return set_error(HMVM_ERR_ARG, "unknown tile flag %d", flags[i]);
    nt[i].flags = flags[i];
  }
  TileDesc *d = which == 0 ? plan->d_tiles_a : plan->d_tiles_b;
  HIP_TRY(hipStreamSynchronize(plan->stream));
  HIP_TRY(hipMemcpy(d, nt.data(), sizeof(TileDesc) * nt.size(), hipMemcpyHostToDevice));
  tl.swap(nt);
  plan->tiles_version++;
  return plan->mode == 1 ? persist_prepare(plan) : HMVM_OK;
}

int hmvm_plan_get_tile_flags(const hmvm_plan *plan, int which, int *flags) {
  if (!plan || (which != 0 && which != 1) || !flags) return set_error(HMVM_ERR_ARG, "bad argument");
  const std::vector<TileDesc> &tl = which == 0 ? plan->plan.tiles_a : plan->plan.tiles_b;
  for (size_t i = 0; i < tl.size(); i++) flags[i] = tl[i].flags;
  return HMVM_OK;
}

int hmvm_plan_set_tile_order(hmvm_plan *plan, int which, const int *perm) {
  if (!plan || (which != 0 && which != 1) || !perm) return set_error(HMVM_ERR_ARG, "bad argument");
  Plan &P = plan->plan;
  std::vector<TileDesc> &tl = which == 0 ? P.tiles_a : P.tiles_b;
  const int n = (int)tl.size();
  std::vector<char> seen((size_t)n, 0);
  for (int i = 0; i < n; i++) {
    if (perm[i] < 0 || perm[i] >= n || seen[perm[i]]) return set_error(HMVM_ERR_ARG, "not a permutation");
    seen[perm[i]] = 1;
  }
  std::vector<TileDesc> nt((size_t)n);
  for (int i = 0; i < n; i++) nt[i] = tl[perm[i]];
  TileDesc *d = which == 0 ? plan->d_tiles_a : plan->d_tiles_b;
  HIP_TRY(hipStreamSynchronize(plan->stream));
  HIP_TRY(hipMemcpy(d, nt.data(), sizeof(TileDesc) * (size_t)n, hipMemcpyHostToDevice));
  tl.swap(nt);
  if (which == 0) {
    std::vector<uint8_t> ns((size_t)n);
    for (int i = 0; i < n; i++) ns[i] = P.a_stage1[perm[i]];
    P.a_stage1.swap(ns);
    plan->a_mixed = order_mixed(P);
  }
  plan->tiles_version++;
  return plan->mode == 1 ? persist_prepare(plan) : HMVM_OK;
}

void hmvm_hip_1(double *v, const matrix *mat, const double *b) { run_variant(v, mat, nullptr, b, 0); }
void hmvm_hip_1t(double *v, const matrix *mat, const double *b) { run_variant(v, mat, nullptr, b, 1); }
void hmvm_hip_2(double *v, const matrix2 *mat2, const double *b) { run_variant(v, nullptr, mat2, b, 2); }
void hmvm_hip_2t(double *v, const matrix2 *mat2, const double *b) { run_variant(v, nullptr, mat2, b, 3); }

void hmvm_hip_2_64(double *v, const matrix2_64 *mat2, const double *b) {
  clear_error();
  if (!v || !b) { set_error(HMVM_ERR_ARG, "v or b is NULL"); return; }
  hmvm_plan *p = plan_for64(mat2, 4);
  if (p) hmvm_plan_apply_host(p, b, v, 1);
}
void hmvm_hip_2t_64(double *v, const matrix2_64 *mat2, const double *b) {
  clear_error();
  if (!v || !b) { set_error(HMVM_ERR_ARG, "v or b is NULL"); return; }
  hmvm_plan *p = plan_for64(mat2, 5);
  if (p) hmvm_plan_apply_host(p, b, v, 1);
}

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
