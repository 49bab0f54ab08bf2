/*
 * hacapk.h -- H-matrix containers and loaders for the MI355X HMVM library.
 *
 * The three structs below are the drop-in boundary: their field order, types
 * and therefore their x86-64 SysV byte layout are identical to the C structs
 * of the reference benchmark (src_hmvm/src_c/hacapk.h:5-28), so a program
 * written against the reference's loader can hand its `matrix`/`matrix2`
 * straight to this library and vice versa.
 *
 * Index conventions (reference m_HACApK_base.f90:51, src_c/loadmatrix.c:41-128):
 *   ltmtx  1 = low-rank leaf  A = a2 * a1^T,  2 = dense leaf
 *   nstrtl/nstrtt are 1-based row/column starts of the leaf
 *   low-rank a1 is kt x ndt row-major (a1[k*ndt + it]),
 *            a2 is kt x ndl row-major (a2[k*ndl + il]),
 *            a2t is ndl x kt row-major (a2t[il*kt + k])
 *   dense    a1 is ndl x ndt row-major (a1[il*ndt + it])
 *   matrix2 packs every leaf into rowmat (and rowmat_t, where the a2 part is
 *   stored as a2t) in leaf order; a1[i]/a2[i] are element offsets of leaf i.
 */
#ifndef HACAPK_AMD_HACAPK_H
#define HACAPK_AMD_HACAPK_H

#ifdef __cplusplus
extern "C" {
#endif

/* One leaf of the H-matrix (reference src_c/hacapk.h:5-8). */
typedef struct submatrix {
  int ltmtx, kt, ndl, ndt, nstrtl, nstrtt;
  double *a1, *a2, *a2t;
} submatrix;

/* Leaf list, array-of-structs form (reference src_c/hacapk.h:9-17).
 * nleavs, b1..b3 and e1..e3 are carried but unused, as in the reference. */
typedef struct matrix {
  int nd;
  int nlf;
  int ktmax;
  int *nleavs;
  submatrix *submat;
  int b1, b2, b3;
  int e1, e2, e3;
} matrix;

/* Leaf list, struct-of-arrays form (reference src_c/hacapk.h:19-28). */
typedef struct matrix2 {
  int nd;
  int nlf;
  int ktmax;
  int len;
  int *nleavs;
  int *ltmtx, *kt, *ndl, *ndt, *nstrtl, *nstrtt;
  int *a1, *a2;              /* element offsets into rowmat / rowmat_t */
  double *rowmat, *rowmat_t; /* leaf data, a2 stored as a2 / as a2t   */
} matrix2;

/* Error codes returned by the loaders and writers (the reference returns -1
 * only when fopen fails and reads damaged files silently). */
enum {
  HACAPK_OK = 0,
  HACAPK_ERR_OPEN = -1,        /* fopen failed (same value as the reference) */
  HACAPK_ERR_TRUNCATED = -2,   /* file ended inside a record                 */
  HACAPK_ERR_MARKER = -3,      /* record marker does not match payload size  */
  HACAPK_ERR_HEADER = -4,      /* nd/nlf/ktmax/len out of range              */
  HACAPK_ERR_LEAF = -5,        /* leaf outside the matrix, bad ltmtx, kt>ktmax */
  HACAPK_ERR_OVERFLOW = -6,    /* element count does not fit the int32 len   */
  HACAPK_ERR_NOMEM = -7,       /* malloc failed                              */
  HACAPK_ERR_LEN = -8,         /* stored len differs from the leaf sizes     */
  HACAPK_ERR_ARG = -9          /* NULL or otherwise invalid argument         */
};

/* Format 1 (hmatrix_1.bin) -> matrix + matrix2, as reference
 * src_c/loadmatrix.c:8-132.  Returns HACAPK_OK or a negative error code. */
int loadHmatrix(const char *fname, matrix *mat, matrix2 *mat2);

/* Format 2 (hmatrix_2.bin) -> matrix2, as reference src_c/loadmatrix.c:134-232. */
int loadHmatrix2(const char *fname, matrix2 *mat2);

/* Synthetic all-dense n x n grid of ndl x ndt blocks, as reference
 * src_c/loadmatrix.c:240-323 (values d*(rand()%1000)/1000 after srand(0)). */
int dummymatrix(matrix *mat, matrix2 *mat2, int ndl, int ndt, int n);

/* Release everything the loaders allocated (the reference never frees). */
void freeHmatrix(matrix *mat, matrix2 *mat2);

/* Writers for the two on-disk formats (reference m_HACApK_hdump.f90:78-180):
 * Fortran unformatted sequential records with 4-byte little-endian markers. */
int writeHmatrix1(const char *fname, const matrix2 *mat2);
int writeHmatrix2(const char *fname, const matrix2 *mat2);

/* Text of an error code. */
const char *hacapk_strerror(int code);

/* ------------------------------------------------------------------------
 * int64 extension for matrices with more than INT_MAX stored elements.
 *
 * The reference's matrix2 keeps len and the a1/a2 offsets in int
 * (src_c/hacapk.h:22,26), and its format-2 header stores len as int32
 * (m_HACApK_hdump.f90:119,143), so neither can describe a matrix of more
 * than 2^31-1 elements (N = 2^20 with eps = 1e-6 holds about 5e9).
 * matrix2_64 is matrix2 with int64 len and offsets; rowmat_t is optional.
 *
 * On disk, "format 2L" is format 2 with a 20-byte first record
 * {nd, nlf, ktmax, int64 len}; every other record is as in format 2 (no
 * single leaf record comes near 2 GiB).  The readers below accept formats 1,
 * 2 and 2L; they tell 2 from 2L by the length of the first record.
 * ------------------------------------------------------------------------ */
typedef struct matrix2_64 {
  int nd;
  int nlf;
  int ktmax;
  long long len;
  int *ltmtx, *kt, *ndl, *ndt, *nstrtl, *nstrtt;
  long long *a1, *a2;        /* element offsets into rowmat / rowmat_t        */
  double *rowmat, *rowmat_t; /* rowmat_t is NULL unless HACAPK_WITH_ROWMAT_T */
} matrix2_64;

enum { HACAPK_WITH_ROWMAT_T = 1 };

/* Format 1, 2 or 2L -> matrix2_64 (format detected from the first record). */
int loadHmatrix_64(const char *fname, matrix2_64 *mat2, int flags);
/* Format 2L (always the 20-byte header) and format 1, from a matrix2_64. */
int writeHmatrix2_64(const char *fname, const matrix2_64 *mat2);
int writeHmatrix1_64(const char *fname, const matrix2_64 *mat2);
/* Deep copy of a matrix2 into the int64 form (rowmat_t copied if present). */
int hacapk_widen_matrix2(const matrix2 *src, matrix2_64 *dst);
void freeHmatrix64(matrix2_64 *mat2);

/* Leaf descriptors of a format 1/2/2L file without its values: fills the
 * index arrays, ktmax, len and offsets; rowmat and rowmat_t stay NULL.
 * *data_pos receives the file position of the first data record when the
 * file is format 2/2L (-1 for format 1, whose data records are interleaved
 * with the leaf records). */
int hacapk_scan_file(const char *fname, matrix2_64 *mat2, long long *data_pos, int *format);

#ifdef __cplusplus
}
#endif

#endif /* HACAPK_AMD_HACAPK_H */
