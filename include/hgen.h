/*
 * hgen.h -- H-matrix generator producing the inputs of the HMVM path.
 *
 * Restates the reference HACApK 1.0 construction (src_hdump/m_HACApK_base.f90)
 * for a point kernel instead of the BEM face integral:
 *   cluster tree    bisection of the largest bounding-box extent at its
 *                   midpoint, leaves of at most `leaf` points (:994-1054)
 *   bounding boxes  per cluster, width = box diagonal (:953-991)
 *   block tree      admissible (low-rank) iff min(width_l, width_t) <= eta*dist
 *                   and both sides have more than `leaf` points; otherwise dense
 *                   if either side is a leaf or has at most leaf+1 points;
 *                   otherwise recurse into the 2x2 sons (:781-891)
 *   leaf order      sorted by nstrtl, then nstrtt (:245-255)
 *   low-rank fill   ACA with partial pivoting, stopping rule and pivot masks of
 *                   HACApK_aca (:494-577)
 * Kernel: a_ij = 1/(4*pi*|p_i - p_j|), diagonal 1/(4*pi*h), h = mean spacing.
 * The result is a matrix2 exactly as loadHmatrix would build it from the
 * files the reference dump writes (m_HACApK_hdump.f90:78-180).
 */
#ifndef HACAPK_AMD_HGEN_H
#define HACAPK_AMD_HGEN_H

#include "hacapk.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { HGEN_SPHERE = 0, HGEN_CUBE = 1 };

typedef struct hgen_params {
  int n;                   /* matrix order (number of points)                      */
  int geometry;            /* HGEN_SPHERE: Fibonacci lattice on the unit sphere;    */
                           /* HGEN_CUBE: jittered grid on the unit-cube surface     */
  int leaf;                /* max points per leaf cluster (HACApK param(21))        */
  int kmax;                /* ACA rank cap (param(63))                              */
  double eta;              /* admissibility parameter (param(51))                   */
  double eps;              /* ACA tolerance (param(71))                             */
  double aca_eps_factor;   /* absolute pivot floor = factor*eps (param(72))         */
  double jitter;           /* cube: jitter half-width as a fraction of the spacing  */
  unsigned long long seed; /* cube jitter seed                                      */
} hgen_params;

/* Defaults: sphere, leaf 32, kmax 1000, eta 2, eps 1e-4, factor 1e-3, jitter 0.25. */
void hgen_default_params(hgen_params *p, int n);

/* Build the H-matrix in memory (matrix2 with rowmat and rowmat_t). */
int hgen_generate(const hgen_params *p, matrix2 *mat2);

/* Same construction into the int64 form (for matrices beyond 2^31-1
 * elements); rowmat_t is built only with HACAPK_WITH_ROWMAT_T in flags. */
int hgen_generate64(const hgen_params *p, matrix2_64 *mat2, int flags);

/* Build the array-of-structs form (with a2t) from a packed matrix2, as the
 * format-1 loader does; the per-leaf arrays are copies. */
int hgen_matrix_from_matrix2(const matrix2 *mat2, matrix *mat);

/* Points of the generated problem in cluster (= matrix index) order, 3*n
 * doubles x0,y0,z0,x1,... ; and the diagonal value used. */
int hgen_points(const hgen_params *p, double *xyz_cluster_order, double *diag);

/* Kernel entries A[i0+r][j0+c] (0-based, cluster order) into out[r*nc+c]. */
int hgen_kernel_block(const hgen_params *p, int i0, int nr, int j0, int nc, double *out);

/* ||A - H||_F / ||A||_F over the leaves (every stride-th leaf when stride>1),
 * with the entries of A recomputed (HACApK_accuracy_leafmtx, :407-466). */
double hgen_accuracy(const hgen_params *p, const matrix2 *mat2, int stride);
double hgen_accuracy64(const hgen_params *p, const matrix2_64 *mat2, int stride);

#ifdef __cplusplus
}
#endif

#endif /* HACAPK_AMD_HGEN_H */
