// hmvm_cpp_api.cpp -- the C++ template entry points (hacapk.hpp, hmvm_hip.hpp)
// on top of the C ABI.  The C structs of hacapk.h share their names with the
// templates, so this file talks to the C ABI through untyped pointers: the C
// matrix2 is the leading part of matrix2<double> (same fields, same order).
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <typeinfo>
#include <vector>

#include "hacapk.hpp"
#include "hmvm_hip.hpp"

extern "C" {
// C ABI (hacapk.h / hmvm_hip.h) seen with untyped struct pointers.
int loadHmatrix(const char *fname, void *mat, void *mat2);
int loadHmatrix2(const char *fname, void *mat2);
void freeHmatrix(void *mat, void *mat2);
struct hmvm_plan;
int hmvm_plan_create(const void *mat2, int a2_source, const void *opt, hmvm_plan **out);
void hmvm_plan_destroy(hmvm_plan *plan);
int hmvm_plan_apply_host(hmvm_plan *plan, const double *x, double *y, int accumulate);
int hmvm_plan_set_launch(hmvm_plan *plan, int kernel, int block_threads);
int hmvm_plan_reset_launch(hmvm_plan *plan);
void hmvm_options_default(void *opt);
int hmvm_plan_bench_apply(hmvm_plan *plan, const double *x, int iters, double *seconds);
int hmvm_hip_last_error(void);
}

namespace {

// Layout of the C matrix2 (hacapk.h); must match the leading fields of matrix2<T>.
struct CMatrix2 {
  int nd, nlf, ktmax, len;
  int *nleavs;
  int *ltmtx, *kt, *ndl, *ndt, *nstrtl, *nstrtt;
  int *a1, *a2;
  double *rowmat, *rowmat_t;
};
struct COptions {
  int storage, w1, w2, block_threads, kernel, order, mode, t_mode, max_rows, max_rows_b;
};
static_assert(sizeof(CMatrix2) == 104, "C matrix2 layout");
static_assert(sizeof(COptions) == 10 * sizeof(int), "hmvm_options layout");
static_assert(offsetof(matrix2<double>, rowmat_t) == offsetof(CMatrix2, rowmat_t), "matrix2<double> prefix");

template <class T>
void clear_extra(matrix2<T> *m) {
  m->napprox = m->ndense = 0;
  m->approx = m->dense = nullptr;
  m->nleavs = nullptr;
}

struct Entry {
  hmvm_plan *plan = nullptr;
  const void *rowmat = nullptr;
  int len = 0, nlf = 0;
};
std::mutex g_mu;
std::map<std::pair<const void *, int>, Entry> g_cache;  // (mat2, is_float)

template <class T>
hmvm_plan *plan_for(matrix2<T> *m) {
  const int is_float = sizeof(T) == 4;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair((const void *)m, is_float);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    if (it->second.rowmat == m->rowmat && it->second.len == m->len && it->second.nlf == m->nlf) return it->second.plan;
    hmvm_plan_destroy(it->second.plan);
    g_cache.erase(it);
  }
  CMatrix2 c;
  std::memcpy(&c, m, sizeof(int) * 4);
  c.nleavs = nullptr;
  c.ltmtx = m->ltmtx; c.kt = m->kt; c.ndl = m->ndl; c.ndt = m->ndt;
  c.nstrtl = m->nstrtl; c.nstrtt = m->nstrtt; c.a1 = m->a1; c.a2 = m->a2;
  std::vector<double> widened;
  if (is_float) {  // FP32 values are exact in FP64; the plan stores them as FP32 again
    widened.resize((size_t)std::max(m->len, 1));
    for (int i = 0; i < m->len; i++) widened[i] = (double)m->rowmat[i];
    c.rowmat = widened.data();
  } else {
    c.rowmat = (double *)m->rowmat;
  }
  c.rowmat_t = nullptr;
  COptions o;
  hmvm_options_default(&o);
  o.storage = is_float ? 1 : 0;
  hmvm_plan *p = nullptr;
  if (hmvm_plan_create(&c, 0, &o, &p)) return nullptr;
  Entry e;
  e.plan = p;
  e.rowmat = m->rowmat;
  e.len = m->len;
  e.nlf = m->nlf;
  g_cache[key] = e;
  return p;
}

}  // namespace

int loadHmatrix(const char *fname, matrix<double> *mat, matrix2<double> *mat2) {
  const int rc = ::loadHmatrix(fname, (void *)mat, (void *)mat2);
  clear_extra(mat2);
  return rc;
}

int loadHmatrix2(const char *fname, matrix2<double> *mat2) {
  const int rc = ::loadHmatrix2(fname, (void *)mat2);
  clear_extra(mat2);
  return rc;
}

// Element-wise round-to-nearest cast, offsets unchanged (src_cpp/loadmatrix.cpp:252-342).
int convertD2F(matrix<float> *matF, matrix2<float> *matF2, const matrix<double> *matD, const matrix2<double> *matD2) {
  if (matD && matF) {
    matF->nd = matD->nd;
    matF->nlf = matD->nlf;
    matF->ktmax = matD->ktmax;
    matF->nleavs = nullptr;
    matF->b1 = matF->b2 = matF->b3 = matF->e1 = matF->e2 = matF->e3 = 0;
    matF->submat = (submatrix<float> *)std::calloc((size_t)matD->nlf, sizeof(submatrix<float>));
    if (!matF->submat) return -7;
    for (int i = 0; i < matD->nlf; i++) {
      const submatrix<double> &d = matD->submat[i];
      submatrix<float> &f = matF->submat[i];
      f.ltmtx = d.ltmtx; f.kt = d.kt; f.ndl = d.ndl; f.ndt = d.ndt; f.nstrtl = d.nstrtl; f.nstrtt = d.nstrtt;
      auto cvt = [](const double *s, size_t n) {
        float *o = (float *)std::malloc(sizeof(float) * std::max<size_t>(n, 1));
        if (o)
          for (size_t j = 0; j < n; j++) o[j] = (float)s[j];
        return o;
      };
      if (d.ltmtx == 1) {
        f.a1 = cvt(d.a1, (size_t)d.kt * d.ndt);
        f.a2 = cvt(d.a2, (size_t)d.kt * d.ndl);
        f.a2t = cvt(d.a2t, (size_t)d.kt * d.ndl);
      } else {
        f.a1 = cvt(d.a1, (size_t)d.ndl * d.ndt);
        f.a2 = f.a2t = nullptr;
      }
    }
  }
  if (matD2 && matF2) {
    const int nlf = matD2->nlf;
    std::memcpy(matF2, matD2, sizeof(int) * 4);
    clear_extra(matF2);
    int **dst[8] = {&matF2->ltmtx, &matF2->kt, &matF2->ndl, &matF2->ndt, &matF2->nstrtl, &matF2->nstrtt,
                    &matF2->a1, &matF2->a2};
    int *const src[8] = {matD2->ltmtx, matD2->kt, matD2->ndl, matD2->ndt, matD2->nstrtl, matD2->nstrtt,
                         matD2->a1, matD2->a2};
    for (int k = 0; k < 8; k++) {
      *dst[k] = (int *)std::malloc(sizeof(int) * (size_t)std::max(nlf, 1));
      if (!*dst[k]) return -7;
      std::memcpy(*dst[k], src[k], sizeof(int) * (size_t)nlf);
    }
    matF2->rowmat = (float *)std::malloc(sizeof(float) * (size_t)std::max(matD2->len, 1));
    matF2->rowmat_t = (float *)std::malloc(sizeof(float) * (size_t)std::max(matD2->len, 1));
    if (!matF2->rowmat || !matF2->rowmat_t) return -7;
    for (int j = 0; j < matD2->len; j++) {
      matF2->rowmat[j] = (float)matD2->rowmat[j];
      matF2->rowmat_t[j] = (float)matD2->rowmat_t[j];
    }
  }
  return 0;
}

template <class T>
static void free_any(matrix<T> *mat, matrix2<T> *mat2) {
  if (mat && mat->submat) {
    for (int i = 0; i < mat->nlf; i++) {
      std::free(mat->submat[i].a1);
      std::free(mat->submat[i].a2);
      std::free(mat->submat[i].a2t);
    }
    std::free(mat->submat);
    mat->submat = nullptr;
  }
  if (mat2) {
    int *arrs[10] = {mat2->ltmtx, mat2->kt, mat2->ndl, mat2->ndt, mat2->nstrtl, mat2->nstrtt, mat2->a1, mat2->a2,
                     mat2->approx, mat2->dense};
    for (int *a : arrs) std::free(a);
    std::free(mat2->rowmat);
    std::free(mat2->rowmat_t);
    std::memset(mat2, 0, sizeof(*mat2));
  }
}
void freeHmatrix(matrix<double> *mat, matrix2<double> *mat2) { free_any(mat, mat2); }
void freeHmatrix(matrix<float> *mat, matrix2<float> *mat2) { free_any(mat, mat2); }

template <class T>
void hmvm_hip(matrix2<T> *mat2, T *b, int kernel, int dump_result, int nbench) {
  if (!mat2 || !b) return;
  hmvm_plan *p = plan_for(mat2);
  if (!p) return;
  // kernel < 0: the measured per-storage launch defaults; 0..7: that variant
  if (kernel < 0 ? hmvm_plan_reset_launch(p) : hmvm_plan_set_launch(p, kernel, 256)) return;
  const int nd = mat2->nd;
  const char tc = sizeof(T) == 4 ? 'f' : 'd';
  std::vector<double> x((size_t)nd), y((size_t)nd, 0.0);
  for (int i = 0; i < nd; i++) x[i] = (double)b[i];
  if (dump_result) {
    if (hmvm_plan_apply_host(p, x.data(), y.data(), 0)) return;
    char fname[96];
    std::snprintf(fname, sizeof fname, "result_hip_k%d_%c.txt", kernel, tc);
    FILE *F = std::fopen(fname, "w");
    if (!F) return;
    for (int i = 0; i < nd; i++) std::fprintf(F, "%.3E\n", (double)(T)y[i]);
    std::fclose(F);
  }
  if (nbench > 0) {
    std::vector<double> dt((size_t)(5 + nbench));
    for (int l = 0; l < 5 + nbench; l++)
      if (hmvm_plan_bench_apply(p, x.data(), 1, &dt[l])) return;
    double mn = 9999.99, mx = 0.0, avg = 0.0;
    for (int l = 5; l < 5 + nbench; l++) {
      avg += dt[l];
      mn = std::min(mn, dt[l]);
      mx = std::max(mx, dt[l]);
    }
    avg /= nbench;
    std::printf("TIME %d hmvm_hip_k%d_%c min %e max %e avg %e\n", nbench, kernel, tc, mn, mx, avg);
    std::fflush(stdout);
  }
}

template <class T>
void hmvm_hip_release_cpp(const matrix2<T> *mat2) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(std::make_pair((const void *)mat2, (int)(sizeof(T) == 4)));
  if (it != g_cache.end()) {
    hmvm_plan_destroy(it->second.plan);
    g_cache.erase(it);
  }
}

template void hmvm_hip<double>(matrix2<double> *, double *, int, int, int);
template void hmvm_hip<float>(matrix2<float> *, float *, int, int, int);
template void hmvm_hip_release_cpp<double>(const matrix2<double> *);
template void hmvm_hip_release_cpp<float>(const matrix2<float> *);
