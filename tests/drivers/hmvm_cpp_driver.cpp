// Drop-in check in C++: the reference sweep driver flow (src_hmvm/src_cpp/hmvm1.cpp:12-147)
// with hmvm_cuda1<T> replaced by hmvm_hip<T>, FP64 then FP32 (convertD2F).
//   usage: hmvm_cpp_driver hmatrix_1.bin [dump nbench]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "hacapk.hpp"
#include "hmvm_hip.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::printf("usage: %s filename [dump(0,1) nbench]\n", argv[0]);
    return -1;
  }
  const int dump_result = argc >= 3 ? std::atoi(argv[2]) : 0;
  const int nbench = argc >= 4 ? std::atoi(argv[3]) : 5;
  matrix<double> matD;
  matrix2<double> matD2;
  if (loadHmatrix(argv[1], &matD, &matD2)) return -1;
  const int nd = matD.nd;
  double *b = new double[nd];
  for (int i = 0; i < nd; i++) b[i] = std::sin((double)(i + 1));
  for (int k = -1; k < 8; k += 3) hmvm_hip<double>(&matD2, b, k, dump_result, nbench);
  matrix<float> matF;
  matrix2<float> matF2;
  if (convertD2F(&matF, &matF2, &matD, &matD2)) return -2;
  float *fb = new float[nd];
  for (int i = 0; i < nd; i++) fb[i] = std::sin((float)(i + 1));
  hmvm_hip<float>(&matF2, fb, -1, dump_result, nbench);
  hmvm_hip_release_cpp(&matD2);
  hmvm_hip_release_cpp(&matF2);
  freeHmatrix(&matD, &matD2);
  freeHmatrix(&matF, &matF2);
  delete[] b;
  delete[] fb;
  return 0;
}
