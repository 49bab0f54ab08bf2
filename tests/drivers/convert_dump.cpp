// convertD2F check: loads an H-matrix file with the C++ loader, converts it
// with convertD2F (the FP32 matrix of the reference's src_cpp/loadmatrix.cpp:
// 252-342) and writes, array by array, the source doubles followed by the
// converted floats, so a test can check every element of the conversion.
//   usage: convert_dump hmatrix_1.bin out.bin
// out.bin: int32 count, then count records of
//   int64 n, n doubles (source), n floats (converted), int32 index check
// for rowmat, rowmat_t and, per leaf, a1, a2, a2t (a2/a2t only for low-rank
// leaves); the index check is 1 when the converted leaf descriptors and
// offsets equal the source's.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "hacapk.hpp"

namespace {

struct Out {
  FILE *f;
  int count = 0;
  void rec(const double *d, const float *s, int64_t n, int same) {
    std::fwrite(&n, sizeof n, 1, f);
    std::fwrite(d, sizeof(double), (size_t)n, f);
    std::fwrite(s, sizeof(float), (size_t)n, f);
    std::fwrite(&same, sizeof same, 1, f);
    count++;
  }
};

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::printf("usage: %s hmatrix_1.bin out.bin\n", argv[0]);
    return 2;
  }
  matrix<double> matD;
  matrix2<double> matD2;
  if (loadHmatrix(argv[1], &matD, &matD2)) return 3;
  matrix<float> matF;
  matrix2<float> matF2;
  if (convertD2F(&matF, &matF2, &matD, &matD2)) return 4;
  FILE *f = std::fopen(argv[2], "wb");
  if (!f) return 5;
  Out o{f};
  int placeholder = 0;
  std::fwrite(&placeholder, sizeof placeholder, 1, f);
  int same2 = matF2.nd == matD2.nd && matF2.nlf == matD2.nlf && matF2.ktmax == matD2.ktmax && matF2.len == matD2.len;
  for (int i = 0; i < matD2.nlf && same2; i++)
    same2 = matF2.ltmtx[i] == matD2.ltmtx[i] && matF2.kt[i] == matD2.kt[i] && matF2.ndl[i] == matD2.ndl[i] &&
            matF2.ndt[i] == matD2.ndt[i] && matF2.nstrtl[i] == matD2.nstrtl[i] &&
            matF2.nstrtt[i] == matD2.nstrtt[i] && matF2.a1[i] == matD2.a1[i] && matF2.a2[i] == matD2.a2[i];
  o.rec(matD2.rowmat, matF2.rowmat, matD2.len, same2);
  o.rec(matD2.rowmat_t, matF2.rowmat_t, matD2.len, same2);
  for (int i = 0; i < matD.nlf; i++) {
    const submatrix<double> &d = matD.submat[i];
    const submatrix<float> &s = matF.submat[i];
    const int same = s.ltmtx == d.ltmtx && s.kt == d.kt && s.ndl == d.ndl && s.ndt == d.ndt &&
                     s.nstrtl == d.nstrtl && s.nstrtt == d.nstrtt;
    if (d.ltmtx == 1) {
      o.rec(d.a1, s.a1, (int64_t)d.kt * d.ndt, same);
      o.rec(d.a2, s.a2, (int64_t)d.kt * d.ndl, same);
      o.rec(d.a2t, s.a2t, (int64_t)d.kt * d.ndl, same);
    } else {
      o.rec(d.a1, s.a1, (int64_t)d.ndl * d.ndt, same);
    }
  }
  std::fseek(f, 0, SEEK_SET);
  std::fwrite(&o.count, sizeof o.count, 1, f);
  std::fclose(f);
  freeHmatrix(&matD, &matD2);
  freeHmatrix(&matF, &matF2);
  return 0;
}
