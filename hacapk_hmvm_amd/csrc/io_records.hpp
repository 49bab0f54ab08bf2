// io_records.hpp -- Fortran unformatted sequential records, as written by the
// reference dump (m_HACApK_hdump.f90:78-180): `[int32 n] payload [int32 n]`.
// Shared by the in-memory loaders, the int64 loaders and the streaming plan
// builder.
#pragma once

#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "hacapk.h"

namespace hacapk_io {

struct Reader {
  FILE *f = nullptr;
  explicit Reader(const char *name) { f = std::fopen(name, "rb"); }
  ~Reader() {
    if (f) std::fclose(f);
  }
  Reader(const Reader &) = delete;
  Reader &operator=(const Reader &) = delete;

  // Reads one record whose payload must be exactly `bytes` long.
  int record(void *dst, size_t bytes) {
    int32_t m0 = 0, m1 = 0;
    if (std::fread(&m0, 4, 1, f) != 1) return HACAPK_ERR_TRUNCATED;
    if (m0 < 0 || (size_t)(uint32_t)m0 != bytes) return HACAPK_ERR_MARKER;
    if (bytes && std::fread(dst, 1, bytes, f) != bytes) return HACAPK_ERR_TRUNCATED;
    if (std::fread(&m1, 4, 1, f) != 1) return HACAPK_ERR_TRUNCATED;
    if (m1 != m0) return HACAPK_ERR_MARKER;
    return HACAPK_OK;
  }
  // Skips one record whose payload must be exactly `bytes` long; the closing
  // marker is still checked.
  int skip(size_t bytes) {
    int32_t m0 = 0, m1 = 0;
    if (std::fread(&m0, 4, 1, f) != 1) return HACAPK_ERR_TRUNCATED;
    if (m0 < 0 || (size_t)(uint32_t)m0 != bytes) return HACAPK_ERR_MARKER;
    if (fseeko(f, (off_t)bytes, SEEK_CUR) != 0) return HACAPK_ERR_TRUNCATED;
    if (std::fread(&m1, 4, 1, f) != 1) return HACAPK_ERR_TRUNCATED;
    if (m1 != m0) return HACAPK_ERR_MARKER;
    return HACAPK_OK;
  }
  // Reads the opening marker of the next record and steps back over it.
  int peek_marker(int32_t *m) {
    if (std::fread(m, 4, 1, f) != 1) return HACAPK_ERR_TRUNCATED;
    if (fseeko(f, -4, SEEK_CUR) != 0) return HACAPK_ERR_TRUNCATED;
    return HACAPK_OK;
  }
  long long tell() { return (long long)ftello(f); }
  int seek(long long pos) { return fseeko(f, (off_t)pos, SEEK_SET) == 0 ? HACAPK_OK : HACAPK_ERR_TRUNCATED; }
};

struct Writer {
  FILE *f = nullptr;
  bool ok = true;
  explicit Writer(const char *name) { f = std::fopen(name, "wb"); }
  ~Writer() {
    if (f) std::fclose(f);
  }
  Writer(const Writer &) = delete;
  Writer &operator=(const Writer &) = delete;
  void record(const void *src, size_t bytes) {
    if (bytes > (size_t)INT32_MAX) { ok = false; return; }
    int32_t m = (int32_t)bytes;
    ok = ok && std::fwrite(&m, 4, 1, f) == 1;
    if (bytes) ok = ok && std::fwrite(src, 1, bytes, f) == bytes;
    ok = ok && std::fwrite(&m, 4, 1, f) == 1;
  }
  int close() {
    int rc = std::fclose(f);
    f = nullptr;
    return (ok && rc == 0) ? HACAPK_OK : HACAPK_ERR_OPEN;
  }
};

// Element count of one leaf (reference src_c/loadmatrix.c:70,84).
inline int64_t leaf_len(int ltmtx, int kt, int ndl, int ndt) {
  return ltmtx == 1 ? (int64_t)kt * ndt + (int64_t)kt * ndl : (int64_t)ndl * ndt;
}

inline int check_leaf(int nd, int ktmax, int ltmtx, int kt, int ndl, int ndt, int nstrtl, int nstrtt) {
  if (ltmtx != 1 && ltmtx != 2) return HACAPK_ERR_LEAF;
  if (ndl < 1 || ndt < 1) return HACAPK_ERR_LEAF;
  if (nstrtl < 1 || nstrtt < 1) return HACAPK_ERR_LEAF;
  if ((int64_t)nstrtl + ndl - 1 > nd || (int64_t)nstrtt + ndt - 1 > nd) return HACAPK_ERR_LEAF;
  if (ltmtx == 1 && (kt < 0 || kt > ktmax)) return HACAPK_ERR_LEAF;
  return HACAPK_OK;
}

// a2 (kt x ndl row-major) -> a2t (ndl x kt row-major).
inline void transpose_a2(const double *a2, double *a2t, int kt, int ndl) {
  for (int k = 0; k < kt; k++)
    for (int il = 0; il < ndl; il++) a2t[k + (int64_t)il * kt] = a2[(int64_t)k * ndl + il];
}

// File layouts the readers accept.
enum Format : int { kFormat1 = 1, kFormat2 = 2, kFormat2L = 3 };

// Sequential reader of the leaf data records of a scanned file.  For format 1
// every leaf's descriptor record precedes its data and is checked against the
// scanned descriptors; for formats 2/2L the data records follow each other.
struct LeafStream {
  Reader &rd;
  int format;
  LeafStream(Reader &r, int fmt) : rd(r), format(fmt) {}
  // Positions the reader at leaf 0 (`data_pos` from the scan for 2/2L).
  int begin(long long data_pos) { return rd.seek(format == kFormat1 ? 20 : data_pos); }
  // Reads leaf i's values into dst: a1 part, then the a2 part of a low-rank
  // leaf (kt*ndt then kt*ndl elements; ndl*ndt for a dense leaf).
  int read(const matrix2_64 *m, int i, double *dst) {
    const int ltmtx = m->ltmtx[i], kt = m->kt[i], ndl = m->ndl[i], ndt = m->ndt[i];
    if (format == kFormat1) {
      int d[6];
      int rc = rd.record(d, sizeof(d));
      if (rc) return rc;
      if (d[0] != ltmtx || d[1] != ndl || d[2] != ndt || d[3] != m->nstrtl[i] || d[4] != m->nstrtt[i] ||
          (ltmtx == 1 && d[5] != kt))
        return HACAPK_ERR_LEAF;
    }
    if (ltmtx == 1) {
      const size_t n1 = (size_t)kt * ndt, n2 = (size_t)kt * ndl;
      int rc = rd.record(dst, sizeof(double) * n1);
      if (rc) return rc;
      return rd.record(dst + n1, sizeof(double) * n2);
    }
    return rd.record(dst, sizeof(double) * (size_t)ndl * ndt);
  }
};

}  // namespace hacapk_io
