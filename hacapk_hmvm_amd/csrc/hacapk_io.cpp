// hacapk_io.cpp -- readers and writers for the HACApK H-matrix dump formats.
//
// Format (reference m_HACApK_hdump.f90:78-180; reader src_c/loadmatrix.c:8-232):
// Fortran unformatted sequential records, each `[int32 n] payload [int32 n]`.
//   format 1: {nd,nlf,ktmax}; per leaf {ltmtx,ndl,ndt,nstrtl,nstrtt,kt}, then
//             a1 (kt*ndt doubles) and a2 (kt*ndl doubles) for a low-rank leaf
//             or a1 (ndl*ndt doubles) for a dense leaf.
//   format 2: {nd,nlf,ktmax,len}; six int[nlf] records (ltmtx, ndl, ndt,
//             nstrtl, nstrtt, kt); then the leaf data records as in format 1.
//
// The structures built here are field-for-field what the reference loaders
// build (same offsets, same a2t transpose, same rowmat/rowmat_t packing), but
// every record marker, size and index is validated and errors are returned as
// codes instead of being read through silently.
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hacapk.h"
#include "io_records.hpp"

using namespace hacapk_io;

namespace {

template <class P>
P *alloc_n(int64_t n) {
  return (P *)std::malloc(sizeof(P) * (size_t)(n > 0 ? n : 1));
}

void zero_matrix2(matrix2 *m) { std::memset(m, 0, sizeof(*m)); }

int alloc_index_arrays(matrix2 *m, int nlf) {
  m->ltmtx = alloc_n<int>(nlf);
  m->kt = alloc_n<int>(nlf);
  m->ndl = alloc_n<int>(nlf);
  m->ndt = alloc_n<int>(nlf);
  m->nstrtl = alloc_n<int>(nlf);
  m->nstrtt = alloc_n<int>(nlf);
  m->a1 = alloc_n<int>(nlf);
  m->a2 = alloc_n<int>(nlf);
  if (!m->ltmtx || !m->kt || !m->ndl || !m->ndt || !m->nstrtl || !m->nstrtt || !m->a1 || !m->a2)
    return HACAPK_ERR_NOMEM;
  return HACAPK_OK;
}

}  // namespace

extern "C" {

const char *hacapk_strerror(int code) {
  switch (code) {
    case HACAPK_OK: return "ok";
    case HACAPK_ERR_OPEN: return "cannot open file";
    case HACAPK_ERR_TRUNCATED: return "file truncated inside a record";
    case HACAPK_ERR_MARKER: return "record marker does not match the payload size";
    case HACAPK_ERR_HEADER: return "header values out of range";
    case HACAPK_ERR_LEAF: return "leaf descriptor out of range";
    case HACAPK_ERR_OVERFLOW: return "element count exceeds the int32 len of matrix2";
    case HACAPK_ERR_NOMEM: return "out of host memory";
    case HACAPK_ERR_LEN: return "stored len disagrees with the leaf sizes";
    case HACAPK_ERR_ARG: return "invalid argument";
    default: return "unknown error";
  }
}

void freeHmatrix(matrix *mat, matrix2 *mat2) {
  if (mat && mat->submat) {
    for (int i = 0; i < mat->nlf; i++) {
      std::free(mat->submat[i].a1);
      std::free(mat->submat[i].a2);
      std::free(mat->submat[i].a2t);
    }
    std::free(mat->submat);
    mat->submat = nullptr;
  }
  if (mat2) {
    std::free(mat2->ltmtx);
    std::free(mat2->kt);
    std::free(mat2->ndl);
    std::free(mat2->ndt);
    std::free(mat2->nstrtl);
    std::free(mat2->nstrtt);
    std::free(mat2->a1);
    std::free(mat2->a2);
    std::free(mat2->rowmat);
    std::free(mat2->rowmat_t);
    zero_matrix2(mat2);
  }
}

// Format 1 -> matrix (per-leaf arrays incl. a2t) + matrix2 (packed), following
// the construction order of reference src_c/loadmatrix.c:8-132.
int loadHmatrix(const char *fname, matrix *mat, matrix2 *mat2) {
  if (!fname || !mat || !mat2) return HACAPK_ERR_ARG;
  std::memset(mat, 0, sizeof(*mat));
  zero_matrix2(mat2);
  Reader rd(fname);
  if (!rd.f) {
    std::fprintf(stderr, "fopen %s failed\n", fname);
    return HACAPK_ERR_OPEN;
  }
  int hdr[3];
  int rc = rd.record(hdr, sizeof(hdr));
  if (rc) return rc;
  const int nd = hdr[0], nlf = hdr[1], ktmax = hdr[2];
  if (nd < 1 || nlf < 1 || ktmax < 0) return HACAPK_ERR_HEADER;
  mat->nd = nd;
  mat->nlf = nlf;
  mat->ktmax = ktmax;
  mat->submat = (submatrix *)std::calloc((size_t)nlf, sizeof(submatrix));
  if (!mat->submat) return HACAPK_ERR_NOMEM;

  int64_t len = 0;
  for (int i = 0; i < nlf; i++) {
    int d[6];  // ltmtx, ndl, ndt, nstrtl, nstrtt, kt (file order)
    if ((rc = rd.record(d, sizeof(d)))) goto fail;
    {
      submatrix &s = mat->submat[i];
      s.ltmtx = d[0];
      s.ndl = d[1];
      s.ndt = d[2];
      s.nstrtl = d[3];
      s.nstrtt = d[4];
      s.kt = d[5];
      if ((rc = check_leaf(nd, ktmax, s.ltmtx, s.kt, s.ndl, s.ndt, s.nstrtl, s.nstrtt))) goto fail;
      if (s.ltmtx == 1) {
        const int64_t n1 = (int64_t)s.kt * s.ndt, n2 = (int64_t)s.kt * s.ndl;
        s.a1 = alloc_n<double>(n1);
        s.a2 = alloc_n<double>(n2);
        s.a2t = alloc_n<double>(n2);
        if (!s.a1 || !s.a2 || !s.a2t) { rc = HACAPK_ERR_NOMEM; goto fail; }
        if ((rc = rd.record(s.a1, sizeof(double) * n1))) goto fail;
        if ((rc = rd.record(s.a2, sizeof(double) * n2))) goto fail;
        transpose_a2(s.a2, s.a2t, s.kt, s.ndl);
        len += n1 + n2;
      } else {
        const int64_t n1 = (int64_t)s.ndl * s.ndt;
        s.a1 = alloc_n<double>(n1);
        if (!s.a1) { rc = HACAPK_ERR_NOMEM; goto fail; }
        if ((rc = rd.record(s.a1, sizeof(double) * n1))) goto fail;
        len += n1;
      }
    }
  }
  if (len > INT_MAX) { rc = HACAPK_ERR_OVERFLOW; goto fail; }

  // Pack matrix2: leaf data in leaf order, rowmat_t holds a2t instead of a2.
  mat2->nd = nd;
  mat2->nlf = nlf;
  mat2->ktmax = ktmax;
  if ((rc = alloc_index_arrays(mat2, nlf))) goto fail;
  mat2->rowmat = alloc_n<double>(len);
  mat2->rowmat_t = alloc_n<double>(len);
  if (!mat2->rowmat || !mat2->rowmat_t) { rc = HACAPK_ERR_NOMEM; goto fail; }
  {
    int64_t off = 0;
    for (int i = 0; i < nlf; i++) {
      const submatrix &s = mat->submat[i];
      mat2->ltmtx[i] = s.ltmtx;
      mat2->kt[i] = s.kt;
      mat2->ndl[i] = s.ndl;
      mat2->ndt[i] = s.ndt;
      mat2->nstrtl[i] = s.nstrtl;
      mat2->nstrtt[i] = s.nstrtt;
      mat2->a1[i] = (int)off;
      if (s.ltmtx == 1) {
        const size_t n1 = (size_t)s.kt * s.ndt, n2 = (size_t)s.kt * s.ndl;
        std::memcpy(mat2->rowmat + off, s.a1, sizeof(double) * n1);
        std::memcpy(mat2->rowmat_t + off, s.a1, sizeof(double) * n1);
        off += (int64_t)n1;
        mat2->a2[i] = (int)off;
        std::memcpy(mat2->rowmat + off, s.a2, sizeof(double) * n2);
        std::memcpy(mat2->rowmat_t + off, s.a2t, sizeof(double) * n2);
        off += (int64_t)n2;
      } else {
        const size_t n1 = (size_t)s.ndl * s.ndt;
        std::memcpy(mat2->rowmat + off, s.a1, sizeof(double) * n1);
        std::memcpy(mat2->rowmat_t + off, s.a1, sizeof(double) * n1);
        off += (int64_t)n1;
        mat2->a2[i] = 0;
      }
    }
    mat2->len = (int)off;
  }
  return HACAPK_OK;

fail:
  freeHmatrix(mat, mat2);
  return rc;
}

// Format 2 -> matrix2, following reference src_c/loadmatrix.c:134-232.
int loadHmatrix2(const char *fname, matrix2 *mat2) {
  if (!fname || !mat2) return HACAPK_ERR_ARG;
  zero_matrix2(mat2);
  Reader rd(fname);
  if (!rd.f) {
    std::fprintf(stderr, "fopen %s failed\n", fname);
    return HACAPK_ERR_OPEN;
  }
  int hdr[4];
  int rc = rd.record(hdr, sizeof(hdr));
  if (rc) return rc;
  const int nd = hdr[0], nlf = hdr[1], ktmax = hdr[2], len = hdr[3];
  if (nd < 1 || nlf < 1 || ktmax < 0 || len < 0) return HACAPK_ERR_HEADER;
  mat2->nd = nd;
  mat2->nlf = nlf;
  mat2->ktmax = ktmax;
  mat2->len = len;
  if ((rc = alloc_index_arrays(mat2, nlf))) goto fail;
  {
    int *cols[6] = {mat2->ltmtx, mat2->ndl, mat2->ndt, mat2->nstrtl, mat2->nstrtt, mat2->kt};
    for (int c = 0; c < 6; c++)
      if ((rc = rd.record(cols[c], sizeof(int) * (size_t)nlf))) goto fail;
  }
  {
    int64_t total = 0;
    for (int i = 0; i < nlf; i++) {
      if ((rc = check_leaf(nd, ktmax, mat2->ltmtx[i], mat2->kt[i], mat2->ndl[i], mat2->ndt[i],
                           mat2->nstrtl[i], mat2->nstrtt[i])))
        goto fail;
      total += leaf_len(mat2->ltmtx[i], mat2->kt[i], mat2->ndl[i], mat2->ndt[i]);
    }
    if (total > INT_MAX) { rc = HACAPK_ERR_OVERFLOW; goto fail; }
    if (total != len) { rc = HACAPK_ERR_LEN; goto fail; }
  }
  mat2->rowmat = alloc_n<double>(len);
  mat2->rowmat_t = alloc_n<double>(len);
  if (!mat2->rowmat || !mat2->rowmat_t) { rc = HACAPK_ERR_NOMEM; goto fail; }
  {
    int64_t off = 0;
    for (int i = 0; i < nlf; i++) {
      const int kt = mat2->kt[i], ndl = mat2->ndl[i], ndt = mat2->ndt[i];
      mat2->a1[i] = (int)off;
      if (mat2->ltmtx[i] == 1) {
        const size_t n1 = (size_t)kt * ndt, n2 = (size_t)kt * ndl;
        if ((rc = rd.record(mat2->rowmat + off, sizeof(double) * n1))) goto fail;
        std::memcpy(mat2->rowmat_t + off, mat2->rowmat + off, sizeof(double) * n1);
        off += (int64_t)n1;
        mat2->a2[i] = (int)off;
        if ((rc = rd.record(mat2->rowmat + off, sizeof(double) * n2))) goto fail;
        transpose_a2(mat2->rowmat + off, mat2->rowmat_t + off, kt, ndl);
        off += (int64_t)n2;
      } else {
        const size_t n1 = (size_t)ndl * ndt;
        if ((rc = rd.record(mat2->rowmat + off, sizeof(double) * n1))) goto fail;
        std::memcpy(mat2->rowmat_t + off, mat2->rowmat + off, sizeof(double) * n1);
        off += (int64_t)n1;
        mat2->a2[i] = 0;
      }
    }
  }
  return HACAPK_OK;

fail:
  freeHmatrix(nullptr, mat2);
  return rc;
}

// All-dense synthetic matrix, following reference src_c/loadmatrix.c:240-323:
// leaf y*n+x covers rows ndl*x.. and columns ndt*y.., values d*(rand()%1000)/1000
// with d=100 after srand(0).
int dummymatrix(matrix *mat, matrix2 *mat2, int ndl, int ndt, int n) {
  if (!mat || !mat2 || ndl < 1 || ndt < 1 || n < 1) return HACAPK_ERR_ARG;
  std::memset(mat, 0, sizeof(*mat));
  zero_matrix2(mat2);
  const int64_t total = (int64_t)n * n * ndl * ndt;
  if (total > INT_MAX || (int64_t)n * n > INT_MAX) return HACAPK_ERR_OVERFLOW;
  mat->nd = n * ndl;
  mat->nlf = n * n;
  mat->ktmax = 0;
  mat->submat = (submatrix *)std::calloc((size_t)mat->nlf, sizeof(submatrix));
  if (!mat->submat) return HACAPK_ERR_NOMEM;
  std::srand(0);
  for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++) {
      submatrix &s = mat->submat[y * n + x];
      s.ltmtx = 2;
      s.kt = 0;
      s.ndl = ndl;
      s.ndt = ndt;
      s.nstrtl = ndl * x + 1;
      s.nstrtt = ndt * y + 1;
      s.a1 = alloc_n<double>((int64_t)ndl * ndt);
      if (!s.a1) { freeHmatrix(mat, mat2); return HACAPK_ERR_NOMEM; }
      for (int z = 0; z < ndl * ndt; z++) s.a1[z] = 100.0 * (std::rand() % 1000) / 1000.0;
    }
  mat2->nd = mat->nd;
  mat2->nlf = mat->nlf;
  mat2->ktmax = 0;
  if (alloc_index_arrays(mat2, mat->nlf)) { freeHmatrix(mat, mat2); return HACAPK_ERR_NOMEM; }
  mat2->rowmat = alloc_n<double>(total);
  mat2->rowmat_t = alloc_n<double>(total);
  if (!mat2->rowmat || !mat2->rowmat_t) { freeHmatrix(mat, mat2); return HACAPK_ERR_NOMEM; }
  int64_t off = 0;
  for (int i = 0; i < mat->nlf; i++) {
    const submatrix &s = mat->submat[i];
    mat2->ltmtx[i] = s.ltmtx;
    mat2->kt[i] = s.kt;
    mat2->ndl[i] = s.ndl;
    mat2->ndt[i] = s.ndt;
    mat2->nstrtl[i] = s.nstrtl;
    mat2->nstrtt[i] = s.nstrtt;
    mat2->a1[i] = (int)off;
    mat2->a2[i] = 0;
    std::memcpy(mat2->rowmat + off, s.a1, sizeof(double) * (size_t)ndl * ndt);
    std::memcpy(mat2->rowmat_t + off, s.a1, sizeof(double) * (size_t)ndl * ndt);
    off += (int64_t)ndl * ndt;
  }
  mat2->len = (int)off;
  return HACAPK_OK;
}

// matrix2 -> format 1, record for record as reference m_HACApK_hdump.f90:93-105.
int writeHmatrix1(const char *fname, const matrix2 *m) {
  if (!fname || !m || !m->rowmat) return HACAPK_ERR_ARG;
  Writer w(fname);
  if (!w.f) return HACAPK_ERR_OPEN;
  int hdr[3] = {m->nd, m->nlf, m->ktmax};
  w.record(hdr, sizeof(hdr));
  for (int i = 0; i < m->nlf; i++) {
    int d[6] = {m->ltmtx[i], m->ndl[i], m->ndt[i], m->nstrtl[i], m->nstrtt[i], m->kt[i]};
    w.record(d, sizeof(d));
    if (m->ltmtx[i] == 1) {
      w.record(m->rowmat + m->a1[i], sizeof(double) * (size_t)m->kt[i] * m->ndt[i]);
      w.record(m->rowmat + m->a2[i], sizeof(double) * (size_t)m->kt[i] * m->ndl[i]);
    } else {
      w.record(m->rowmat + m->a1[i], sizeof(double) * (size_t)m->ndl[i] * m->ndt[i]);
    }
  }
  return w.close();
}

// matrix2 -> format 2, as reference m_HACApK_hdump.f90:114-180.
int writeHmatrix2(const char *fname, const matrix2 *m) {
  if (!fname || !m || !m->rowmat) return HACAPK_ERR_ARG;
  Writer w(fname);
  if (!w.f) return HACAPK_ERR_OPEN;
  int hdr[4] = {m->nd, m->nlf, m->ktmax, m->len};
  w.record(hdr, sizeof(hdr));
  const int *cols[6] = {m->ltmtx, m->ndl, m->ndt, m->nstrtl, m->nstrtt, m->kt};
  for (int c = 0; c < 6; c++) w.record(cols[c], sizeof(int) * (size_t)m->nlf);
  for (int i = 0; i < m->nlf; i++) {
    if (m->ltmtx[i] == 1) {
      w.record(m->rowmat + m->a1[i], sizeof(double) * (size_t)m->kt[i] * m->ndt[i]);
      w.record(m->rowmat + m->a2[i], sizeof(double) * (size_t)m->kt[i] * m->ndl[i]);
    } else {
      w.record(m->rowmat + m->a1[i], sizeof(double) * (size_t)m->ndl[i] * m->ndt[i]);
    }
  }
  return w.close();
}

}  // extern "C"
