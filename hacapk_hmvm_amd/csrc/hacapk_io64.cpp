// hacapk_io64.cpp -- the int64 form of matrix2 (matrix2_64) and its files.
//
// The reference's matrix2 and format-2 header keep the element count in int32
// (src_c/hacapk.h:22,26; m_HACApK_hdump.f90:119,143).  matrix2_64 widens len
// and the a1/a2 offsets to int64 and keeps everything else: the same leaf
// arrays, the same packing order as src_c/loadmatrix.c:88-128 (leaf data
// concatenated in leaf order, a1 part first, a2 = a1 + kt*ndt), and the same
// rowmat_t convention (the a2 part stored as a2t).  "Format 2L" is format 2
// with a 20-byte first record {nd, nlf, ktmax, int64 len}.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hacapk.h"
#include "io_records.hpp"

using namespace hacapk_io;

namespace {

template <class P>
P *alloc_n(int64_t n) {
  return (P *)std::malloc(sizeof(P) * (size_t)(n > 0 ? n : 1));
}

int alloc_index64(matrix2_64 *m, int nlf) {
  m->ltmtx = alloc_n<int>(nlf);
  m->kt = alloc_n<int>(nlf);
  m->ndl = alloc_n<int>(nlf);
  m->ndt = alloc_n<int>(nlf);
  m->nstrtl = alloc_n<int>(nlf);
  m->nstrtt = alloc_n<int>(nlf);
  m->a1 = alloc_n<long long>(nlf);
  m->a2 = alloc_n<long long>(nlf);
  if (!m->ltmtx || !m->kt || !m->ndl || !m->ndt || !m->nstrtl || !m->nstrtt || !m->a1 || !m->a2)
    return HACAPK_ERR_NOMEM;
  return HACAPK_OK;
}

// Offsets of the packed layout; checks every leaf; returns the element count
// or a negative error code.
long long pack_offsets(matrix2_64 *m) {
  long long off = 0;
  int ktmax = 0;
  for (int i = 0; i < m->nlf; i++) {
    int rc = check_leaf(m->nd, INT_MAX, m->ltmtx[i], m->kt[i], m->ndl[i], m->ndt[i], m->nstrtl[i], m->nstrtt[i]);
    if (rc) return rc;
    m->a1[i] = off;
    if (m->ltmtx[i] == 1) {
      ktmax = std::max(ktmax, m->kt[i]);
      off += (long long)m->kt[i] * m->ndt[i];
      m->a2[i] = off;
      off += (long long)m->kt[i] * m->ndl[i];
    } else {
      m->a2[i] = 0;
      off += (long long)m->ndl[i] * m->ndt[i];
    }
  }
  if (ktmax > m->ktmax) return HACAPK_ERR_LEAF;
  return off;
}

int scan(Reader &rd, matrix2_64 *m, long long *data_pos, int *format) {
  int32_t m0 = 0;
  int rc = rd.peek_marker(&m0);
  if (rc) return rc;
  if (m0 == 12) {  // format 1: {nd, nlf, ktmax}, then per leaf descriptor + data
    int hdr[3];
    if ((rc = rd.record(hdr, sizeof(hdr)))) return rc;
    m->nd = hdr[0]; m->nlf = hdr[1]; m->ktmax = hdr[2];
    if (m->nd < 1 || m->nlf < 1 || m->ktmax < 0) return HACAPK_ERR_HEADER;
    if ((rc = alloc_index64(m, m->nlf))) return rc;
    for (int i = 0; i < m->nlf; i++) {
      int d[6];  // ltmtx, ndl, ndt, nstrtl, nstrtt, kt (file order)
      if ((rc = rd.record(d, sizeof(d)))) return rc;
      m->ltmtx[i] = d[0]; m->ndl[i] = d[1]; m->ndt[i] = d[2];
      m->nstrtl[i] = d[3]; m->nstrtt[i] = d[4]; m->kt[i] = d[5];
      if ((rc = check_leaf(m->nd, m->ktmax, d[0], d[5], d[1], d[2], d[3], d[4]))) return rc;
      if (d[0] == 1) {
        if ((rc = rd.skip(sizeof(double) * (size_t)d[5] * d[2]))) return rc;
        if ((rc = rd.skip(sizeof(double) * (size_t)d[5] * d[1]))) return rc;
      } else {
        if ((rc = rd.skip(sizeof(double) * (size_t)d[1] * d[2]))) return rc;
      }
    }
    const long long len = pack_offsets(m);
    if (len < 0) return (int)len;
    m->len = len;
    *data_pos = -1;
    *format = kFormat1;
    return HACAPK_OK;
  }
  if (m0 != 16 && m0 != 20) return HACAPK_ERR_MARKER;
  unsigned char hdr[20];
  if ((rc = rd.record(hdr, (size_t)m0))) return rc;
  int h3[3];
  std::memcpy(h3, hdr, 12);
  long long len;
  if (m0 == 16) {
    int l32;
    std::memcpy(&l32, hdr + 12, 4);
    len = l32;
  } else {
    int64_t l64;
    std::memcpy(&l64, hdr + 12, 8);
    len = (long long)l64;
  }
  m->nd = h3[0]; m->nlf = h3[1]; m->ktmax = h3[2];
  if (m->nd < 1 || m->nlf < 1 || m->ktmax < 0 || len < 0) return HACAPK_ERR_HEADER;
  if ((rc = alloc_index64(m, m->nlf))) return rc;
  int *cols[6] = {m->ltmtx, m->ndl, m->ndt, m->nstrtl, m->nstrtt, m->kt};
  for (int c = 0; c < 6; c++)
    if ((rc = rd.record(cols[c], sizeof(int) * (size_t)m->nlf))) return rc;
  for (int i = 0; i < m->nlf; i++)
    if ((rc = check_leaf(m->nd, m->ktmax, m->ltmtx[i], m->kt[i], m->ndl[i], m->ndt[i], m->nstrtl[i], m->nstrtt[i])))
      return rc;
  const long long total = pack_offsets(m);
  if (total < 0) return (int)total;
  if (total != len) return HACAPK_ERR_LEN;
  m->len = len;
  *data_pos = rd.tell();
  *format = m0 == 16 ? kFormat2 : kFormat2L;
  return HACAPK_OK;
}

void write_leaf_data(Writer &w, const matrix2_64 *m, int i) {
  if (m->ltmtx[i] == 1) {
    w.record(m->rowmat + m->a1[i], sizeof(double) * (size_t)m->kt[i] * m->ndt[i]);
    w.record(m->rowmat + m->a2[i], sizeof(double) * (size_t)m->kt[i] * m->ndl[i]);
  } else {
    w.record(m->rowmat + m->a1[i], sizeof(double) * (size_t)m->ndl[i] * m->ndt[i]);
  }
}

}  // namespace

extern "C" {

void freeHmatrix64(matrix2_64 *m) {
  if (!m) return;
  std::free(m->ltmtx); std::free(m->kt); std::free(m->ndl); std::free(m->ndt);
  std::free(m->nstrtl); std::free(m->nstrtt); std::free(m->a1); std::free(m->a2);
  std::free(m->rowmat); std::free(m->rowmat_t);
  std::memset(m, 0, sizeof(*m));
}

int hacapk_scan_file(const char *fname, matrix2_64 *m, long long *data_pos, int *format) {
  if (!fname || !m || !data_pos || !format) return HACAPK_ERR_ARG;
  std::memset(m, 0, sizeof(*m));
  Reader rd(fname);
  if (!rd.f) {
    std::fprintf(stderr, "fopen %s failed\n", fname);
    return HACAPK_ERR_OPEN;
  }
  int rc = scan(rd, m, data_pos, format);
  if (rc) freeHmatrix64(m);
  return rc;
}

int loadHmatrix_64(const char *fname, matrix2_64 *m, int flags) {
  if (!fname || !m) return HACAPK_ERR_ARG;
  std::memset(m, 0, sizeof(*m));
  Reader rd(fname);
  if (!rd.f) {
    std::fprintf(stderr, "fopen %s failed\n", fname);
    return HACAPK_ERR_OPEN;
  }
  long long data_pos = -1;
  int format = 0;
  int rc = scan(rd, m, &data_pos, &format);
  if (!rc) {
    m->rowmat = alloc_n<double>(m->len);
    if (flags & HACAPK_WITH_ROWMAT_T) m->rowmat_t = alloc_n<double>(m->len);
    if (!m->rowmat || ((flags & HACAPK_WITH_ROWMAT_T) && !m->rowmat_t)) rc = HACAPK_ERR_NOMEM;
  }
  if (!rc) {
    LeafStream ls(rd, format);
    rc = ls.begin(data_pos);
    for (int i = 0; !rc && i < m->nlf; i++) {
      rc = ls.read(m, i, m->rowmat + m->a1[i]);
      if (!rc && m->rowmat_t) {
        const int kt = m->kt[i], ndl = m->ndl[i], ndt = m->ndt[i];
        if (m->ltmtx[i] == 1) {
          std::memcpy(m->rowmat_t + m->a1[i], m->rowmat + m->a1[i], sizeof(double) * (size_t)kt * ndt);
          transpose_a2(m->rowmat + m->a2[i], m->rowmat_t + m->a2[i], kt, ndl);
        } else {
          std::memcpy(m->rowmat_t + m->a1[i], m->rowmat + m->a1[i], sizeof(double) * (size_t)ndl * ndt);
        }
      }
    }
  }
  if (rc) freeHmatrix64(m);
  return rc;
}

int writeHmatrix2_64(const char *fname, const matrix2_64 *m) {
  if (!fname || !m || !m->rowmat) return HACAPK_ERR_ARG;
  Writer w(fname);
  if (!w.f) return HACAPK_ERR_OPEN;
  unsigned char hdr[20];
  const int h3[3] = {m->nd, m->nlf, m->ktmax};
  const int64_t len = m->len;
  std::memcpy(hdr, h3, 12);
  std::memcpy(hdr + 12, &len, 8);
  w.record(hdr, sizeof(hdr));
  const int *cols[6] = {m->ltmtx, m->ndl, m->ndt, m->nstrtl, m->nstrtt, m->kt};
  for (int c = 0; c < 6; c++) w.record(cols[c], sizeof(int) * (size_t)m->nlf);
  for (int i = 0; i < m->nlf; i++) write_leaf_data(w, m, i);
  return w.close();
}

int writeHmatrix1_64(const char *fname, const matrix2_64 *m) {
  if (!fname || !m || !m->rowmat) return HACAPK_ERR_ARG;
  Writer w(fname);
  if (!w.f) return HACAPK_ERR_OPEN;
  const int hdr[3] = {m->nd, m->nlf, m->ktmax};
  w.record(hdr, sizeof(hdr));
  for (int i = 0; i < m->nlf; i++) {
    const int d[6] = {m->ltmtx[i], m->ndl[i], m->ndt[i], m->nstrtl[i], m->nstrtt[i], m->kt[i]};
    w.record(d, sizeof(d));
    write_leaf_data(w, m, i);
  }
  return w.close();
}

int hacapk_widen_matrix2(const matrix2 *s, matrix2_64 *d) {
  if (!s || !d || !s->rowmat || s->nlf < 1 || s->len < 0) return HACAPK_ERR_ARG;
  std::memset(d, 0, sizeof(*d));
  d->nd = s->nd;
  d->nlf = s->nlf;
  d->ktmax = s->ktmax;
  d->len = s->len;
  int rc = alloc_index64(d, s->nlf);
  if (!rc) {
    d->rowmat = alloc_n<double>(s->len);
    if (s->rowmat_t) d->rowmat_t = alloc_n<double>(s->len);
    if (!d->rowmat || (s->rowmat_t && !d->rowmat_t)) rc = HACAPK_ERR_NOMEM;
  }
  if (rc) {
    freeHmatrix64(d);
    return rc;
  }
  const size_t ni = sizeof(int) * (size_t)s->nlf;
  std::memcpy(d->ltmtx, s->ltmtx, ni); std::memcpy(d->kt, s->kt, ni);
  std::memcpy(d->ndl, s->ndl, ni); std::memcpy(d->ndt, s->ndt, ni);
  std::memcpy(d->nstrtl, s->nstrtl, ni); std::memcpy(d->nstrtt, s->nstrtt, ni);
  for (int i = 0; i < s->nlf; i++) {
    d->a1[i] = s->a1[i];
    d->a2[i] = s->a2[i];
  }
  std::memcpy(d->rowmat, s->rowmat, sizeof(double) * (size_t)s->len);
  if (s->rowmat_t) std::memcpy(d->rowmat_t, s->rowmat_t, sizeof(double) * (size_t)s->len);
  return HACAPK_OK;
}

}  // extern "C"
