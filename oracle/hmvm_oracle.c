/*
 * hmvm_oracle.c -- TEST INFRASTRUCTURE ONLY (see hmvm_oracle.h).
 *
 * CPU restatement of the reference HMVM.  The loop nests and the order of the
 * floating-point accumulations are the reference's, because bit-identical FP64
 * output is what lets this file stand in for the reference on machines
 * where the reference sources are not available.
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "hmvm_oracle.h"

/* Stage 1 of a low-rank leaf: t[k] = sum_it a1[k*ndt+it] * b[c0+it], k-outer,
 * it-inner, accumulated in place (hmvm_seq.c:30-37). */
static void lr_stage1(double *t, const double *a1, int kt, int ndt, const double *bc) {
  for (int k = 0; k < kt; k++) t[k] = 0.0;
  for (int k = 0; k < kt; k++)
    for (int it = 0; it < ndt; it++) t[k] += a1[it + k * ndt] * bc[it];
}

/* Stage 2 over a2 (kt x ndl): v[r0+il] += a2[k*ndl+il]*t[k], k-outer (:38-44). */
static void lr_stage2_a2(double *vr, const double *a2, int kt, int ndl, const double *t) {
  for (int k = 0; k < kt; k++)
    for (int il = 0; il < ndl; il++) vr[il] += a2[il + k * ndl] * t[k];
}

/* Stage 2 over a2t (ndl x kt): row-outer dot, accumulated directly into v
 * (hmvm_seq.c:91-97). */
static void lr_stage2_a2t(double *vr, const double *a2t, int kt, int ndl, const double *t) {
  for (int il = 0; il < ndl; il++)
    for (int k = 0; k < kt; k++) vr[il] += a2t[k + il * kt] * t[k];
}

/* Dense leaf: v[r0+il] += a1[il*ndt+it]*b[c0+it], accumulated directly into v
 * without a row temporary (hmvm_seq.c:46-53). */
static void dense_leaf(double *vr, const double *a1, int ndl, int ndt, const double *bc) {
  for (int il = 0; il < ndl; il++)
    for (int it = 0; it < ndt; it++) vr[il] += a1[it + il * ndt] * bc[it];
}

void oracle_hmvm_seq_1(double *v, const matrix *mat, const double *b) {
  double *t = (double *)malloc(sizeof(double) * (mat->ktmax > 0 ? mat->ktmax : 1));
  for (int i = 0; i < mat->nlf; i++) {
    const submatrix *s = &mat->submat[i];
    double *vr = v + s->nstrtl - 1;
    const double *bc = b + s->nstrtt - 1;
    if (s->ltmtx == 1) {
      lr_stage1(t, s->a1, s->kt, s->ndt, bc);
      lr_stage2_a2(vr, s->a2, s->kt, s->ndl, t);
    } else {
      dense_leaf(vr, s->a1, s->ndl, s->ndt, bc);
    }
  }
  free(t);
}

void oracle_hmvm_seq_1t(double *v, const matrix *mat, const double *b) {
  double *t = (double *)malloc(sizeof(double) * (mat->ktmax > 0 ? mat->ktmax : 1));
  for (int i = 0; i < mat->nlf; i++) {
    const submatrix *s = &mat->submat[i];
    double *vr = v + s->nstrtl - 1;
    const double *bc = b + s->nstrtt - 1;
    if (s->ltmtx == 1) {
      lr_stage1(t, s->a1, s->kt, s->ndt, bc);
      lr_stage2_a2t(vr, s->a2t, s->kt, s->ndl, t);
    } else {
      dense_leaf(vr, s->a1, s->ndl, s->ndt, bc);
    }
  }
  free(t);
}

void oracle_hmvm_seq_2(double *v, const matrix2 *mat, const double *b) {
  double *t = (double *)malloc(sizeof(double) * (mat->ktmax > 0 ? mat->ktmax : 1));
  for (int i = 0; i < mat->nlf; i++) {
    double *vr = v + mat->nstrtl[i] - 1;
    const double *bc = b + mat->nstrtt[i] - 1;
    if (mat->ltmtx[i] == 1) {
      lr_stage1(t, mat->rowmat + mat->a1[i], mat->kt[i], mat->ndt[i], bc);
      lr_stage2_a2(vr, mat->rowmat + mat->a2[i], mat->kt[i], mat->ndl[i], t);
    } else {
      dense_leaf(vr, mat->rowmat + mat->a1[i], mat->ndl[i], mat->ndt[i], bc);
    }
  }
  free(t);
}

void oracle_hmvm_seq_2t(double *v, const matrix2 *mat, const double *b) {
  double *t = (double *)malloc(sizeof(double) * (mat->ktmax > 0 ? mat->ktmax : 1));
  for (int i = 0; i < mat->nlf; i++) {
    double *vr = v + mat->nstrtl[i] - 1;
    const double *bc = b + mat->nstrtt[i] - 1;
    if (mat->ltmtx[i] == 1) {
      lr_stage1(t, mat->rowmat_t + mat->a1[i], mat->kt[i], mat->ndt[i], bc);
      lr_stage2_a2t(vr, mat->rowmat_t + mat->a2[i], mat->kt[i], mat->ndl[i], t);
    } else {
      dense_leaf(vr, mat->rowmat_t + mat->a1[i], mat->ndl[i], mat->ndt[i], bc);
    }
  }
  free(t);
}

/* ---- OpenMP baseline (src_c/hmvm_omp.c) ------------------------------------
 * Every thread: zero a private y of length nd, take a static share of the
 * leaves, track the touched row range [ls,le], then add the private y into
 * the shared one with one atomic per element of that range.  Leaves with an
 * ltmtx other than 1/2 are skipped, as the reference omp path does. */
typedef struct {
  int ltmtx, kt, ndl, ndt, nstrtl, nstrtt;
  const double *a1, *a2;
  int a2_is_t;
} leaf_view;

static void omp_body(double *zau, int nd, int nlf, int ktmax, leaf_view (*get)(const void *, int),
                     const void *m, const double *zu) {
#pragma omp parallel
  {
#pragma omp for
    for (int i = 0; i < nd; i++) zau[i] = 0.0;
    double *zaut = (double *)malloc(sizeof(double) * nd);
    double *zbut = (double *)malloc(sizeof(double) * (ktmax > 0 ? ktmax : 1));
    for (int i = 0; i < nd; i++) zaut[i] = 0.0;
    int ls = nd, le = 1;
#pragma omp for
    for (int ip = 0; ip < nlf; ip++) {
      leaf_view s = get(m, ip);
      if (s.nstrtl < ls) ls = s.nstrtl;
      if (s.nstrtl + s.ndl - 1 > le) le = s.nstrtl + s.ndl - 1;
      double *vr = zaut + s.nstrtl - 1;
      const double *bc = zu + s.nstrtt - 1;
      if (s.ltmtx == 1) {
        for (int k = 0; k < s.kt; k++) {
          zbut[k] = 0.0;
          for (int it = 0; it < s.ndt; it++) zbut[k] += s.a1[it + k * s.ndt] * bc[it];
        }
        if (s.a2_is_t)
          lr_stage2_a2t(vr, s.a2, s.kt, s.ndl, zbut);
        else
          lr_stage2_a2(vr, s.a2, s.kt, s.ndl, zbut);
      } else if (s.ltmtx == 2) {
        dense_leaf(vr, s.a1, s.ndl, s.ndt, bc);
      }
    }
    for (int il = ls - 1; il <= le - 1; il++) {
#pragma omp atomic
      zau[il] += zaut[il];
    }
    free(zaut);
    free(zbut);
  }
}

static leaf_view view_1(const void *m, int i) {
  const submatrix *s = &((const matrix *)m)->submat[i];
  leaf_view l = {s->ltmtx, s->kt, s->ndl, s->ndt, s->nstrtl, s->nstrtt, s->a1, s->a2, 0};
  return l;
}
static leaf_view view_1t(const void *m, int i) {
  const submatrix *s = &((const matrix *)m)->submat[i];
  leaf_view l = {s->ltmtx, s->kt, s->ndl, s->ndt, s->nstrtl, s->nstrtt, s->a1, s->a2t, 1};
  return l;
}
static leaf_view view_2(const void *m, int i) {
  const matrix2 *s = (const matrix2 *)m;
  leaf_view l = {s->ltmtx[i], s->kt[i], s->ndl[i], s->ndt[i], s->nstrtl[i], s->nstrtt[i],
                 s->rowmat + s->a1[i], s->rowmat + s->a2[i], 0};
  return l;
}
static leaf_view view_2t(const void *m, int i) {
  const matrix2 *s = (const matrix2 *)m;
  leaf_view l = {s->ltmtx[i], s->kt[i], s->ndl[i], s->ndt[i], s->nstrtl[i], s->nstrtt[i],
                 s->rowmat_t + s->a1[i], s->rowmat_t + s->a2[i], 1};
  return l;
}

void oracle_hmvm_omp_1(double *zau, const matrix *m, const double *zu) {
  omp_body(zau, m->nd, m->nlf, m->ktmax, view_1, m, zu);
}
void oracle_hmvm_omp_1t(double *zau, const matrix *m, const double *zu) {
  omp_body(zau, m->nd, m->nlf, m->ktmax, view_1t, m, zu);
}
void oracle_hmvm_omp_2(double *zau, const matrix2 *m, const double *zu) {
  omp_body(zau, m->nd, m->nlf, m->ktmax, view_2, m, zu);
}
void oracle_hmvm_omp_2t(double *zau, const matrix2 *m, const double *zu) {
  omp_body(zau, m->nd, m->nlf, m->ktmax, view_2t, m, zu);
}

/* FP32, loop order of src_cpp/hmvm_seq.cpp:154-229 with a2i=0: stage 1 and
 * stage 2 as the C version; dense rows summed into a row temporary first. */
void oracle_hmvm_seq_2_f32(float *v, const matrix2 *mat, const float *rowmat_f, const float *b) {
  float *t = (float *)malloc(sizeof(float) * (mat->ktmax > 0 ? mat->ktmax : 1));
  for (int i = 0; i < mat->nlf; i++) {
    const int kt = mat->kt[i], ndl = mat->ndl[i], ndt = mat->ndt[i];
    float *vr = v + mat->nstrtl[i] - 1;
    const float *bc = b + mat->nstrtt[i] - 1;
    if (mat->ltmtx[i] == 1) {
      const float *a1 = rowmat_f + mat->a1[i], *a2 = rowmat_f + mat->a2[i];
      for (int k = 0; k < kt; k++) t[k] = 0.0f;
      for (int k = 0; k < kt; k++)
        for (int it = 0; it < ndt; it++) t[k] += a1[it + k * ndt] * bc[it];
      for (int k = 0; k < kt; k++)
        for (int il = 0; il < ndl; il++) vr[il] += a2[il + k * ndl] * t[k];
    } else {
      const float *a1 = rowmat_f + mat->a1[i];
      for (int il = 0; il < ndl; il++) {
        float row = 0.0f;
        for (int it = 0; it < ndt; it++) row += a1[it + il * ndt] * bc[it];
        vr[il] += row;
      }
    }
  }
  free(t);
}

void oracle_convert_d2f(const double *src, float *dst, long long n) {
  for (long long i = 0; i < n; i++) dst[i] = (float)src[i];
}

int oracle_omp_threads(void) { return omp_get_max_threads(); }

/* The same two kernels over the int64 form (matrix2_64): only the offset type
 * differs from oracle_hmvm_seq_2 / oracle_hmvm_omp_2. */
void oracle_hmvm_seq_2_64(double *v, const matrix2_64 *mat, const double *b) {
  double *t = (double *)malloc(sizeof(double) * (mat->ktmax > 0 ? mat->ktmax : 1));
  for (int i = 0; i < mat->nlf; i++) {
    double *vr = v + mat->nstrtl[i] - 1;
    const double *bc = b + mat->nstrtt[i] - 1;
    if (mat->ltmtx[i] == 1) {
      lr_stage1(t, mat->rowmat + mat->a1[i], mat->kt[i], mat->ndt[i], bc);
      lr_stage2_a2(vr, mat->rowmat + mat->a2[i], mat->kt[i], mat->ndl[i], t);
    } else {
      dense_leaf(vr, mat->rowmat + mat->a1[i], mat->ndl[i], mat->ndt[i], bc);
    }
  }
  free(t);
}

static leaf_view view_2_64(const void *m, int i) {
  const matrix2_64 *s = (const matrix2_64 *)m;
  leaf_view l = {s->ltmtx[i], s->kt[i], s->ndl[i], s->ndt[i], s->nstrtl[i], s->nstrtt[i],
                 s->rowmat + s->a1[i], s->rowmat + s->a2[i], 0};
  return l;
}

void oracle_hmvm_omp_2_64(double *zau, const matrix2_64 *m, const double *zu) {
  omp_body(zau, m->nd, m->nlf, m->ktmax, view_2_64, m, zu);
}
