// hacapk.hpp -- C++ (templated) H-matrix containers, layout-identical to the
// reference's C++ benchmark structs (src_hmvm/src_cpp/hacapk.h:5-41): the
// first fields of matrix2<double> are byte-for-byte the C matrix2 of
// hacapk.h, followed by the approx/dense leaf lists the reference GPU code
// fills.  Do not include together with hacapk.h (the C structs share names).
#ifndef HACAPK_AMD_HACAPK_HPP
#define HACAPK_AMD_HACAPK_HPP

template <class T>
struct submatrix {
  int ltmtx, kt, ndl, ndt, nstrtl, nstrtt;
  T *a1, *a2, *a2t;
};

template <class T>
struct matrix {
  int nd;
  int nlf;
  int ktmax;
  int *nleavs;
  submatrix<T> *submat;
  int b1, b2, b3;
  int e1, e2, e3;
};

template <class T>
struct matrix2 {
  int nd;
  int nlf;
  int ktmax;
  int len;
  int *nleavs;
  int *ltmtx, *kt, *ndl, *ndt, *nstrtl, *nstrtt;
  int *a1, *a2;          // element offsets into rowmat / rowmat_t
  T *rowmat, *rowmat_t;  // leaf data
  int napprox, ndense;   // low-rank / dense leaf lists (filled on demand)
  int *approx, *dense;
};

// Loaders (src_cpp/loadmatrix.cpp:7-250): same files and semantics as the C
// loaders of hacapk.h; return 0 or a negative error code.
int loadHmatrix(const char *fname, matrix<double> *mat, matrix2<double> *mat2);
int loadHmatrix2(const char *fname, matrix2<double> *mat2);
// FP64 -> FP32 copy of every factor, offsets kept (src_cpp/loadmatrix.cpp:252-342).
int convertD2F(matrix<float> *matF, matrix2<float> *matF2, const matrix<double> *matD,
               const matrix2<double> *matD2);
// Releases what the loaders / convertD2F allocated.
void freeHmatrix(matrix<double> *mat, matrix2<double> *mat2);
void freeHmatrix(matrix<float> *mat, matrix2<float> *mat2);

#endif  // HACAPK_AMD_HACAPK_HPP
