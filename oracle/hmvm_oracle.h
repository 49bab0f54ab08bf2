/*
 * hmvm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference HMVM used as the correctness oracle and as
 * the timed CPU baseline.  Only tests, the smoke check and the CPU-baseline leg
 * of the benchmark may load it; the HIP library never calls into it.
 *
 * Each function names the reference routine it restates.  The sequential
 * kernels keep the reference loop and accumulation order exactly, so their
 * FP64 output is bit-identical to the reference when both are built with the
 * same flags (-O3 -ffp-contract=off, no FMA).
 *
 * Pinning (what checks each function, beyond retyping the formula):
 *   oracle_hmvm_seq_1/1t/2/2t  bit-exact vs the reference sources compiled
 *       into oracle/_ref (in this container) and vs committed golden vectors;
 *       closed form on all-dense matrices (== dense GEMV); brute force on the
 *       assembled dense matrix; linearity; 1==1t==2==2t.
 *   oracle_hmvm_omp_1/1t/2/2t  equal to the sequential oracle to rounding.
 *   oracle_hmvm_seq_2_f32      bit-exact vs the C++ reference template
 *       hmvm_seq_2<float,0> compiled into oracle/_ref; FP32-rounded inputs.
 */
#ifndef HMVM_ORACLE_H
#define HMVM_ORACLE_H

#include "../include/hacapk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* v += A*b over the matrix form (reference src_c/hmvm_seq.c:9-58). */
void oracle_hmvm_seq_1(double *v, const matrix *mat, const double *b);
/* same, low-rank second stage as row dot over a2t (hmvm_seq.c:61-110). */
void oracle_hmvm_seq_1t(double *v, const matrix *mat, const double *b);
/* v += A*b over matrix2.rowmat (hmvm_seq.c:113-166): the oracle proper. */
void oracle_hmvm_seq_2(double *v, const matrix2 *mat, const double *b);
/* same over matrix2.rowmat_t (hmvm_seq.c:169-222). */
void oracle_hmvm_seq_2t(double *v, const matrix2 *mat, const double *b);

/* zau = A*zu with OpenMP: static leaf partition, per-thread private y of
 * length nd, atomic merge of the touched range (src_c/hmvm_omp.c:9-290).
 * These zero zau themselves, as the reference does. */
void oracle_hmvm_omp_1(double *zau, const matrix *mat, const double *zu);
void oracle_hmvm_omp_1t(double *zau, const matrix *mat, const double *zu);
void oracle_hmvm_omp_2(double *zau, const matrix2 *mat, const double *zu);
void oracle_hmvm_omp_2t(double *zau, const matrix2 *mat, const double *zu);

/* FP32 sequential HMVM with the loop order of the C++ reference template
 * hmvm_seq_2<float,0> (src_cpp/hmvm_seq.cpp:154-229): dense rows use a row
 * temporary.  Structure arrays come from matrix2, values from rowmat_f. */
void oracle_hmvm_seq_2_f32(float *v, const matrix2 *mat, const float *rowmat_f, const float *b);

/* Round-to-nearest FP64 -> FP32 copy (src_cpp/loadmatrix.cpp:252-342). */
void oracle_convert_d2f(const double *src, float *dst, long long n);

/* oracle_hmvm_seq_2 and oracle_hmvm_omp_2 over the int64 form (matrix2_64);
 * pinned by equality with the int32 functions on the same matrix. */
void oracle_hmvm_seq_2_64(double *v, const matrix2_64 *mat, const double *b);
void oracle_hmvm_omp_2_64(double *zau, const matrix2_64 *mat, const double *zu);

/* Number of OpenMP threads the omp kernels will use. */
int oracle_omp_threads(void);

#ifdef __cplusplus
}
#endif

#endif
