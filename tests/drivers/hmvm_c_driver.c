/* Drop-in check in C: the reference driver flow (src_hmvm/src_c/hmvm1.c:8-62)
 * with hmvm_seq/hmvm_omp replaced by the HIP entry points.
 *   usage: hmvm_c_driver hmatrix_1.bin [dump_result] */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hmvm_hip.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    printf("usage: %s filename [dump_result]\n", argv[0]);
    return -1;
  }
  int dump_result = argc >= 3 ? atoi(argv[2]) : 0;
  matrix mat;
  matrix2 mat2;
  int ret = loadHmatrix(argv[1], &mat, &mat2);
  if (ret) {
    printf("loadHmatrix failed: %s\n", hacapk_strerror(ret));
    return -1;
  }
  int nd = mat.nd;
  double *b = (double *)malloc(sizeof(double) * nd);
  for (int i = 0; i < nd; i++) b[i] = sin((double)(i + 1));
  hmvm_hip(&mat, &mat2, b, dump_result);
  if (hmvm_hip_last_error()) return 2;
  hmvm_hip_bench(&mat, &mat2, b);
  if (hmvm_hip_last_error()) return 3;
  /* the inner kernel accumulates, as hmvm_seq_2 does */
  double *v = (double *)calloc(nd, sizeof(double));
  hmvm_hip_2(v, &mat2, b);
  hmvm_hip_2(v, &mat2, b);
  FILE *F = fopen("result_hip_2x2_d.txt", "w");
  for (int i = 0; i < nd; i++) fprintf(F, "%.17g\n", v[i]);
  fclose(F);
  hmvm_hip_release(&mat);
  hmvm_hip_release(&mat2);
  freeHmatrix(&mat, &mat2);
  free(b);
  free(v);
  return hmvm_hip_last_error() ? 4 : 0;
}
