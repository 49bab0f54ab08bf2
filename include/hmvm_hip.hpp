// hmvm_hip.hpp -- C++ drop-in of the reference GPU entry points.
//
//   reference (src_cpp/hmvm_cuda.h:4-7)              here
//   hmvm_cuda1<T>(mat2, b, kernel, dump, nbench)     hmvm_hip<T>(mat2, b, kernel, dump, nbench)
//
// T = double: FP64 storage.  T = float: the FP32 matrix made by convertD2F,
// kept as FP32 in HBM with FP32 accumulation inside a tile and FP64 for the
// cross-tile sums of y.
// kernel: -1 = the default (tuned) launch; 0..7 = a tile-kernel variant
//   (bit0 batched vector entries, bit1 software pipelining, bit2 non-temporal).
// dump_result != 0: computes v = A b once and writes
//   result_hip_k<kernel>_<d|f>.txt, one "%.3E" per line (src_cpp/hmvm_cuda1.cu:273-279).
// nbench > 0: 5 warm-up + nbench timed HMVMs (kernels only, b already on the
//   GPU) printed as "TIME <nbench> hmvm_hip_k<kernel>_<d|f> min <s> max <s> avg <s>"
//   (src_cpp/hmvm_cuda1.cu:280-300).
// The device copy is cached per matrix2 address; hmvm_hip_release_cpp drops it.
#ifndef HACAPK_AMD_HMVM_HIP_HPP
#define HACAPK_AMD_HMVM_HIP_HPP

#include "hacapk.hpp"

template <class T>
void hmvm_hip(matrix2<T> *mat2, T *b, int kernel, int dump_result, int nbench);

extern template void hmvm_hip<double>(matrix2<double> *mat2, double *b, int kernel, int dump_result, int nbench);
extern template void hmvm_hip<float>(matrix2<float> *mat2, float *b, int kernel, int dump_result, int nbench);

template <class T>
void hmvm_hip_release_cpp(const matrix2<T> *mat2);

extern template void hmvm_hip_release_cpp<double>(const matrix2<double> *mat2);
extern template void hmvm_hip_release_cpp<float>(const matrix2<float> *mat2);

#endif  // HACAPK_AMD_HMVM_HIP_HPP
