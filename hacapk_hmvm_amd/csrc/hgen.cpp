// hgen.cpp -- geometry, cluster tree, block tree and ACA fill for the inputs
// of the HMVM path.  The construction restates HACApK 1.0 (reference
// src_hdump/m_HACApK_base.f90); line references are given per routine.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hacapk.h"
#include "hgen.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// splitmix64: a portable, seedable generator so the jitter is identical on
// every machine.
struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

struct Cluster {
  int nstrt;  // 0-based first index in cluster order
  int nsize;
  int son[2];
  int nson;
  double bmin[3], bmax[3];
  double width;  // bounding-box diagonal (HACApK zwdth)
};

struct Leaf {
  int ltmtx, nstrtl, nstrtt, ndl, ndt;  // nstrt 1-based as in the dump format
  int kt;
  std::vector<double> a1, a2;  // a1: kt x ndt (or ndl x ndt), a2: kt x ndl, row-major
};

struct Problem {
  hgen_params p;
  std::vector<double> pts;  // 3*n, cluster order
  double diag = 0;
  std::vector<Cluster> cl;

  // Kernel entry between cluster-order indices i and j.
  inline double entry(int i, int j) const {
    if (i == j) return diag;
    const double dx = pts[3 * i] - pts[3 * j], dy = pts[3 * i + 1] - pts[3 * j + 1],
                 dz = pts[3 * i + 2] - pts[3 * j + 2];
    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    return r > 0 ? 1.0 / (4.0 * kPi * r) : diag;
  }
};

// Fibonacci lattice on the unit sphere; mean spacing sqrt(4*pi/n).
void sphere_points(int n, std::vector<double> &xyz, double &h) {
  xyz.resize(3 * (size_t)n);
  const double ga = kPi * (3.0 - std::sqrt(5.0));
  for (int i = 0; i < n; i++) {
    const double z = 1.0 - (2.0 * i + 1.0) / n;
    const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    const double phi = ga * i;
    xyz[3 * i] = r * std::cos(phi);
    xyz[3 * i + 1] = r * std::sin(phi);
    xyz[3 * i + 2] = z;
  }
  h = std::sqrt(4.0 * kPi / n);
}

// Grid on the six faces of [0,1]^3 with uniform jitter of +-jitter*spacing.
// Face f fixes axis f/2 at (f%2); each face gets an a x b grid with a*b as
// close to square as the point count allows (101250 -> 6 x 135 x 125).
void cube_points(int n, double jitter, uint64_t seed, std::vector<double> &xyz, double &h) {
  xyz.resize(3 * (size_t)n);
  SplitMix rng(seed);
  int idx = 0;
  for (int f = 0; f < 6; f++) {
    const int m = n / 6 + (f < n % 6 ? 1 : 0);
    if (m == 0) continue;
    int a = (int)std::ceil(std::sqrt((double)m)), b = (m + a - 1) / a;
    for (int d = (int)std::floor(std::sqrt((double)m)); d >= 1; d--) {
      if (m % d == 0) {
        const int e = m / d;
        if (std::max(d, e) <= 2 * std::min(d, e)) { a = std::max(d, e); b = std::min(d, e); }
        break;
      }
    }
    const int axis = f / 2;
    const double side = (double)(f % 2);
    for (int c = 0; c < m; c++) {
      const int ia = c % a, ib = c / a;
      double u = (ia + 0.5) / a + jitter * (2.0 * rng.uniform() - 1.0) / a;
      double v = (ib + 0.5) / b + jitter * (2.0 * rng.uniform() - 1.0) / b;
      u = std::min(1.0, std::max(0.0, u));
      v = std::min(1.0, std::max(0.0, v));
      double q[3];
      q[axis] = side;
      q[(axis + 1) % 3] = u;
      q[(axis + 2) % 3] = v;
      xyz[3 * idx] = q[0];
      xyz[3 * idx + 1] = q[1];
      xyz[3 * idx + 2] = q[2];
      idx++;
    }
  }
  h = std::sqrt(6.0 / n);
}

// Recursive bisection (reference m_HACApK_base.f90:1009-1054): split along
// the largest extent at the midpoint; points with coordinate <= mid go to the
// first son.  Leaves hold at most `leaf` points.
int build_tree(std::vector<Cluster> &cl, const std::vector<double> &xyz, std::vector<int> &lod,
               int nstrt, int nsize, int leaf) {
  const int id = (int)cl.size();
  cl.push_back(Cluster());
  cl[id].nstrt = nstrt;
  cl[id].nsize = nsize;
  cl[id].nson = 0;
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = nstrt; i < nstrt + nsize; i++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], xyz[3 * lod[i] + d]);
      hi[d] = std::max(hi[d], xyz[3 * lod[i] + d]);
    }
  for (int d = 0; d < 3; d++) {
    cl[id].bmin[d] = lo[d];
    cl[id].bmax[d] = hi[d];
  }
  double w2 = 0;
  for (int d = 0; d < 3; d++) w2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
  cl[id].width = std::sqrt(w2);
  if (nsize <= leaf) return id;

  int cut = 0;
  for (int d = 1; d < 3; d++)
    if (hi[d] - lo[d] > hi[cut] - lo[cut]) cut = d;
  int nl;
  if (hi[cut] > lo[cut]) {
    const double mid = 0.5 * (hi[cut] + lo[cut]);
    auto it = std::stable_partition(lod.begin() + nstrt, lod.begin() + nstrt + nsize,
                                    [&](int q) { return xyz[3 * q + cut] <= mid; });
    nl = (int)(it - (lod.begin() + nstrt));
  } else {
    nl = nsize / 2;  // coincident points: split by count
  }
  const int s0 = build_tree(cl, xyz, lod, nstrt, nl, leaf);
  const int s1 = build_tree(cl, xyz, lod, nstrt + nl, nsize - nl, leaf);
  cl[id].son[0] = s0;
  cl[id].son[1] = s1;
  cl[id].nson = 2;
  return id;
}

int make_problem(const hgen_params *p, Problem &pb) {
  if (!p || p->n < 2 || p->leaf < 1 || p->eta <= 0 || p->eps <= 0 || p->kmax < 1) return HACAPK_ERR_ARG;
  pb.p = *p;
  std::vector<double> xyz;
  double h = 0;
  if (p->geometry == HGEN_SPHERE)
    sphere_points(p->n, xyz, h);
  else if (p->geometry == HGEN_CUBE)
    cube_points(p->n, p->jitter, p->seed, xyz, h);
  else
    return HACAPK_ERR_ARG;
  pb.diag = 1.0 / (4.0 * kPi * h);
  std::vector<int> lod(p->n);
  for (int i = 0; i < p->n; i++) lod[i] = i;
  pb.cl.reserve(4 * (size_t)p->n / std::max(1, p->leaf) + 16);
  build_tree(pb.cl, xyz, lod, 0, p->n, p->leaf);
  pb.pts.resize(3 * (size_t)p->n);
  for (int i = 0; i < p->n; i++)
    for (int d = 0; d < 3; d++) pb.pts[3 * i + d] = xyz[3 * lod[i] + d];
  return HACAPK_OK;
}

// Block tree (reference m_HACApK_base.f90:845-891).
void build_blocks(const Problem &pb, int tl, int tt, std::vector<Leaf> &out) {
  const Cluster &L = pb.cl[tl], &T = pb.cl[tt];
  const int ndl = L.nsize, ndt = T.nsize;
  const int nleaf = pb.p.leaf + 1;
  double zs = 0;
  for (int d = 0; d < 3; d++) {
    if (L.bmax[d] < T.bmin[d])
      zs += (T.bmin[d] - L.bmax[d]) * (T.bmin[d] - L.bmax[d]);
    else if (T.bmax[d] < L.bmin[d])
      zs += (L.bmin[d] - T.bmax[d]) * (L.bmin[d] - T.bmax[d]);
  }
  const double dist = std::sqrt(zs);
  const double eta = pb.p.eta;
  if ((L.width <= eta * dist || T.width <= eta * dist) && ndl >= nleaf && ndt >= nleaf) {
    Leaf lf;
    lf.ltmtx = 1;
    lf.nstrtl = L.nstrt + 1;
    lf.nstrtt = T.nstrt + 1;
    lf.ndl = ndl;
    lf.ndt = ndt;
    lf.kt = 0;
    out.push_back(std::move(lf));
  } else if (L.nson == 0 || T.nson == 0 || ndl <= nleaf || ndt <= nleaf) {
    Leaf lf;
    lf.ltmtx = 2;
    lf.nstrtl = L.nstrt + 1;
    lf.nstrtt = T.nstrt + 1;
    lf.ndl = ndl;
    lf.ndt = ndt;
    lf.kt = 0;
    out.push_back(std::move(lf));
  } else {
    for (int a = 0; a < L.nson; a++)
      for (int b = 0; b < T.nson; b++) build_blocks(pb, L.son[a], T.son[b], out);
  }
}

double norm2(const double *v, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += v[i] * v[i];
  return std::sqrt(s);
}

// ACA with partial pivoting, following HACApK_aca (m_HACApK_base.f90:494-577)
// with param(61)=1: the stopping quantity is |u_k||v_k| relative to the first
// rank-1 term, and the iteration also stops when both pivots fall below the
// absolute floor aca_eps.  Produces zaa (ndl x kt, column k at zaa[k*ndl]) and
// zab (ndt x kt, column k at zab[k*ndt]); the block is zaa * zab^T.
int aca(const Problem &pb, int r0, int ndl, int c0, int ndt, int nstrtl, int nstrtt,
        std::vector<double> &zaa, std::vector<double> &zab) {
  const double eps = pb.p.eps;
  const double aca_eps = pb.p.aca_eps_factor * eps;
  const int kstop = std::min(pb.p.kmax, std::min(ndl, ndt));
  std::vector<char> rmask(ndl, 0), cmask(ndt, 0);
  zaa.clear();
  zab.clear();
  int ist = (nstrtl > nstrtt) ? 0 : ndl - 1;
  double znrm = 1.0;
  int k = 1;
  for (;;) {
    zaa.resize((size_t)k * ndl);
    zab.resize((size_t)k * ndt);
    double *prow = &zab[(size_t)(k - 1) * ndt];
    double *pcol = &zaa[(size_t)(k - 1) * ndl];
    // residual row ist
    for (int it = 0; it < ndt; it++) {
      if (cmask[it]) { prow[it] = 0; continue; }
      double v = pb.entry(r0 + ist, c0 + it);
      for (int l = 0; l < k - 1; l++) v -= zaa[(size_t)l * ndl + ist] * zab[(size_t)l * ndt + it];
      prow[it] = v;
    }
    int jst = -1;
    double row_max = 0;
    for (int it = 0; it < ndt; it++)
      if (!cmask[it] && std::fabs(prow[it]) > row_max) { row_max = std::fabs(prow[it]); jst = it; }
    if (jst < 0) {  // residual row vanished: keep a zero term and stop
      std::fill(prow, prow + ndt, 0.0);
      std::fill(pcol, pcol + ndl, 0.0);
      break;
    }
    const double inv = 1.0 / prow[jst];
    for (int it = 0; it < ndt; it++) prow[it] *= inv;
    // residual column jst
    for (int il = 0; il < ndl; il++) {
      if (rmask[il]) { pcol[il] = 0; continue; }
      double v = pb.entry(r0 + il, c0 + jst);
      for (int l = 0; l < k - 1; l++) v -= zab[(size_t)l * ndt + jst] * zaa[(size_t)l * ndl + il];
      pcol[il] = v;
    }
    rmask[ist] = 1;
    int istn = -1;
    double col_max = 0;
    for (int il = 0; il < ndl; il++)
      if (!rmask[il] && std::fabs(pcol[il]) > col_max) { col_max = std::fabs(pcol[il]); istn = il; }
    cmask[jst] = 1;
    if (istn >= 0) ist = istn;
    if (row_max < aca_eps && col_max < aca_eps) break;
    double zeps = norm2(pcol, ndl) * norm2(prow, ndt);
    if (k == 1) znrm = zeps > 0 ? zeps : 1.0;
    zeps /= znrm;
    if (zeps < eps || k == kstop || istn < 0) break;
    k++;
  }
  return k;
}

void fill_leaf(const Problem &pb, Leaf &lf) {
  const int r0 = lf.nstrtl - 1, c0 = lf.nstrtt - 1;
  if (lf.ltmtx == 2) {
    lf.a1.resize((size_t)lf.ndl * lf.ndt);
    for (int il = 0; il < lf.ndl; il++)
      for (int it = 0; it < lf.ndt; it++) lf.a1[(size_t)il * lf.ndt + it] = pb.entry(r0 + il, c0 + it);
    lf.kt = 0;
  } else {
    std::vector<double> zaa, zab;
    lf.kt = aca(pb, r0, lf.ndl, c0, lf.ndt, lf.nstrtl, lf.nstrtt, zaa, zab);
    // Fortran a1(ndt,kt) = zab, a2(ndl,kt) = zaa: the same memory order.
    lf.a1.assign(zab.begin(), zab.begin() + (size_t)lf.kt * lf.ndt);
    lf.a2.assign(zaa.begin(), zaa.begin() + (size_t)lf.kt * lf.ndl);
  }
}

// Packs the filled leaves in leaf order as reference src_c/loadmatrix.c:88-128
// does (a1 part, then a2 = a1 + kt*ndt; rowmat_t holds a2t).  M is matrix2
// (int offsets, rowmat_t always built) or matrix2_64 (rowmat_t optional).
template <class M, class Off>
int pack_leaves(const Problem &pb, std::vector<Leaf> &leaves, M *m, bool with_t, int64_t max_len) {
  std::memset(m, 0, sizeof(*m));
  const int nlf = (int)leaves.size();
  int64_t len = 0;
  int ktmax = 0;
  for (const Leaf &lf : leaves) {
    if (lf.ltmtx == 1) {
      len += (int64_t)lf.kt * (lf.ndl + lf.ndt);
      ktmax = std::max(ktmax, lf.kt);
    } else {
      len += (int64_t)lf.ndl * lf.ndt;
    }
  }
  if (len > max_len) return HACAPK_ERR_OVERFLOW;
  m->nd = pb.p.n;
  m->nlf = nlf;
  m->ktmax = ktmax;
  m->len = len;
  auto ia = [&]() { return (int *)std::malloc(sizeof(int) * (size_t)nlf); };
  auto oa = [&]() { return (Off *)std::malloc(sizeof(Off) * (size_t)nlf); };
  m->ltmtx = ia(); m->kt = ia(); m->ndl = ia(); m->ndt = ia();
  m->nstrtl = ia(); m->nstrtt = ia(); m->a1 = oa(); m->a2 = oa();
  m->rowmat = (double *)std::malloc(sizeof(double) * (size_t)std::max<int64_t>(len, 1));
  if (with_t) m->rowmat_t = (double *)std::malloc(sizeof(double) * (size_t)std::max<int64_t>(len, 1));
  if (!m->ltmtx || !m->kt || !m->ndl || !m->ndt || !m->nstrtl || !m->nstrtt || !m->a1 || !m->a2 ||
      !m->rowmat || (with_t && !m->rowmat_t))
    return HACAPK_ERR_NOMEM;
  int64_t off = 0;
  for (int i = 0; i < nlf; i++) {
    Leaf &lf = leaves[i];
    m->ltmtx[i] = lf.ltmtx;
    m->kt[i] = lf.kt;
    m->ndl[i] = lf.ndl;
    m->ndt[i] = lf.ndt;
    m->nstrtl[i] = lf.nstrtl;
    m->nstrtt[i] = lf.nstrtt;
    m->a1[i] = (Off)off;
    if (lf.ltmtx == 1) {
      const size_t n1 = (size_t)lf.kt * lf.ndt, n2 = (size_t)lf.kt * lf.ndl;
      std::memcpy(m->rowmat + off, lf.a1.data(), sizeof(double) * n1);
      if (with_t) std::memcpy(m->rowmat_t + off, lf.a1.data(), sizeof(double) * n1);
      off += (int64_t)n1;
      m->a2[i] = (Off)off;
      std::memcpy(m->rowmat + off, lf.a2.data(), sizeof(double) * n2);
      if (with_t)
        for (int k = 0; k < lf.kt; k++)
          for (int il = 0; il < lf.ndl; il++)
            m->rowmat_t[off + k + (int64_t)il * lf.kt] = lf.a2[(size_t)k * lf.ndl + il];
      off += (int64_t)n2;
    } else {
      const size_t n1 = (size_t)lf.ndl * lf.ndt;
      std::memcpy(m->rowmat + off, lf.a1.data(), sizeof(double) * n1);
      if (with_t) std::memcpy(m->rowmat_t + off, lf.a1.data(), sizeof(double) * n1);
      off += (int64_t)n1;
      m->a2[i] = 0;
    }
    std::vector<double>().swap(lf.a1);
    std::vector<double>().swap(lf.a2);
  }
  return HACAPK_OK;
}

// Cluster tree, block tree, leaf sort and the parallel fill
// (largest leaves first for the dynamic schedule).
int build_leaves(const hgen_params *p, Problem &pb, std::vector<Leaf> &leaves) {
  int rc = make_problem(p, pb);
  if (rc) return rc;
  build_blocks(pb, 0, 0, leaves);
  std::sort(leaves.begin(), leaves.end(), [](const Leaf &a, const Leaf &b) {
    return a.nstrtl != b.nstrtl ? a.nstrtl < b.nstrtl : a.nstrtt < b.nstrtt;
  });
  std::vector<int> order(leaves.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    const int64_t wa = (int64_t)leaves[a].ndl * leaves[a].ndt, wb = (int64_t)leaves[b].ndl * leaves[b].ndt;
    return wa != wb ? wa > wb : a < b;
  });
  const int nl = (int)leaves.size();
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 0; q < nl; q++) fill_leaf(pb, leaves[order[q]]);
  return HACAPK_OK;
}

// ||A - H||_F / ||A||_F over every stride-th leaf (HACApK_accuracy_leafmtx,
// m_HACApK_base.f90:407-466), entries of A recomputed.
template <class M>
double accuracy(const Problem &pb, const M *m, int stride) {
  if (stride < 1) stride = 1;
  double num = 0, den = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : num, den)
  for (int i = 0; i < m->nlf; i += stride) {
    const int ndl = m->ndl[i], ndt = m->ndt[i], r0 = m->nstrtl[i] - 1, c0 = m->nstrtt[i] - 1;
    const double *a1 = m->rowmat + m->a1[i];
    for (int il = 0; il < ndl; il++)
      for (int it = 0; it < ndt; it++) {
        const double a = pb.entry(r0 + il, c0 + it);
        double h;
        if (m->ltmtx[i] == 1) {
          const double *a2 = m->rowmat + m->a2[i];
          h = 0;
          for (int k = 0; k < m->kt[i]; k++) h += a2[(size_t)k * ndl + il] * a1[(size_t)k * ndt + it];
        } else {
          h = a1[(size_t)il * ndt + it];
        }
        num += (a - h) * (a - h);
        den += a * a;
      }
  }
  return den > 0 ? std::sqrt(num / den) : 0.0;
}

}  // namespace

extern "C" {

void hgen_default_params(hgen_params *p, int n) {
  if (!p) return;
  p->n = n;
  p->geometry = HGEN_SPHERE;
  p->leaf = 32;
  p->kmax = 1000;
  p->eta = 2.0;
  p->eps = 1e-4;
  p->aca_eps_factor = 1e-3;
  p->jitter = 0.25;
  p->seed = 20261202ull;
}

int hgen_generate(const hgen_params *p, matrix2 *mat2) {
  if (!mat2) return HACAPK_ERR_ARG;
  Problem pb;
  std::vector<Leaf> leaves;
  int rc = build_leaves(p, pb, leaves);
  if (!rc) rc = pack_leaves<matrix2, int>(pb, leaves, mat2, true, INT_MAX);
  if (rc) freeHmatrix(nullptr, mat2);
  return rc;
}

int hgen_generate64(const hgen_params *p, matrix2_64 *mat2, int flags) {
  if (!mat2) return HACAPK_ERR_ARG;
  Problem pb;
  std::vector<Leaf> leaves;
  int rc = build_leaves(p, pb, leaves);
  if (!rc) rc = pack_leaves<matrix2_64, long long>(pb, leaves, mat2, (flags & HACAPK_WITH_ROWMAT_T) != 0, INT64_MAX);
  if (rc) freeHmatrix64(mat2);
  return rc;
}

int hgen_matrix_from_matrix2(const matrix2 *m, matrix *mat) {
  if (!m || !mat || !m->rowmat) return HACAPK_ERR_ARG;
  std::memset(mat, 0, sizeof(*mat));
  mat->nd = m->nd;
  mat->nlf = m->nlf;
  mat->ktmax = m->ktmax;
  mat->submat = (submatrix *)std::calloc((size_t)m->nlf, sizeof(submatrix));
  if (!mat->submat) return HACAPK_ERR_NOMEM;
  for (int i = 0; i < m->nlf; i++) {
    submatrix &s = mat->submat[i];
    s.ltmtx = m->ltmtx[i];
    s.kt = m->kt[i];
    s.ndl = m->ndl[i];
    s.ndt = m->ndt[i];
    s.nstrtl = m->nstrtl[i];
    s.nstrtt = m->nstrtt[i];
    if (s.ltmtx == 1) {
      const size_t n1 = (size_t)s.kt * s.ndt, n2 = (size_t)s.kt * s.ndl;
      s.a1 = (double *)std::malloc(sizeof(double) * std::max<size_t>(n1, 1));
      s.a2 = (double *)std::malloc(sizeof(double) * std::max<size_t>(n2, 1));
      s.a2t = (double *)std::malloc(sizeof(double) * std::max<size_t>(n2, 1));
      if (!s.a1 || !s.a2 || !s.a2t) { freeHmatrix(mat, nullptr); return HACAPK_ERR_NOMEM; }
      std::memcpy(s.a1, m->rowmat + m->a1[i], sizeof(double) * n1);
      std::memcpy(s.a2, m->rowmat + m->a2[i], sizeof(double) * n2);
      std::memcpy(s.a2t, m->rowmat_t + m->a2[i], sizeof(double) * n2);
    } else {
      const size_t n1 = (size_t)s.ndl * s.ndt;
      s.a1 = (double *)std::malloc(sizeof(double) * std::max<size_t>(n1, 1));
      if (!s.a1) { freeHmatrix(mat, nullptr); return HACAPK_ERR_NOMEM; }
      std::memcpy(s.a1, m->rowmat + m->a1[i], sizeof(double) * n1);
    }
  }
  return HACAPK_OK;
}

int hgen_points(const hgen_params *p, double *xyz, double *diag) {
  Problem pb;
  int rc = make_problem(p, pb);
  if (rc) return rc;
  if (xyz) std::memcpy(xyz, pb.pts.data(), sizeof(double) * pb.pts.size());
  if (diag) *diag = pb.diag;
  return HACAPK_OK;
}

int hgen_kernel_block(const hgen_params *p, int i0, int nr, int j0, int nc, double *out) {
  Problem pb;
  int rc = make_problem(p, pb);
  if (rc) return rc;
  if (!out || i0 < 0 || j0 < 0 || nr < 0 || nc < 0 || i0 + nr > p->n || j0 + nc > p->n) return HACAPK_ERR_ARG;
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) out[(size_t)r * nc + c] = pb.entry(i0 + r, j0 + c);
  return HACAPK_OK;
}

double hgen_accuracy(const hgen_params *p, const matrix2 *m, int stride) {
  Problem pb;
  if (!m || make_problem(p, pb)) return -1.0;
  return accuracy(pb, m, stride);
}

double hgen_accuracy64(const hgen_params *p, const matrix2_64 *m, int stride) {
  Problem pb;
  if (!m || make_problem(p, pb)) return -1.0;
  return accuracy(pb, m, stride);
}

}  // extern "C"
