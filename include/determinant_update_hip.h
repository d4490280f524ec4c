/*
 * determinant_update_hip.h -- batched Slater-determinant kernels on AMD
 * Instinct MI355X (gfx950): the determinant ratio of a row replacement, the
 * Sherman-Morrison update of the inverse, and the orbital row fill fused
 * with the ratio.  They consume the rows that the multi-spline evaluation
 * produces, one matrix per walker.
 *
 * Replaces the reference's prototypes:
 *   src/determinant_update.h:8-104   update_inverse_cuda1 + update_inverse_cuda2
 *                                    (row k of A replaced by u; A updated too)
 *   src/determinant_update.cu:10-98  the same pair where u is the change of row k
 *   src/cuda_walker.cu:150-185       calc_ratios (new row . column k of A^-1)
 *   src/cuda_walker.cu:69-95         calc_new_row (the spline row fill), here
 *                                    fused with calc_ratios
 *
 * Layout, as in the reference: matrices are row-major with a row stride,
 * M[i*rowstride + j].  Row i of A holds electron i's orbital values (orbital
 * j in column j), so A^-1 has orbitals on its rows and electrons on its
 * columns, and the ratio for a new row u of electron k is
 *   det(A') / det(A) = sum_i u_i * Ainv[i*rowstride + k].
 * Every *_list is a device array of num_mats device pointers (one matrix or
 * vector per walker), like the reference's float *X_g[] arguments.
 *
 * All functions enqueue on `stream` (a hipStream_t; NULL = the null stream),
 * do not synchronise, and return 0 or a negative EINSPLINE_E* code
 * (multi_bspline_hip.h); einspline_hip_last_error() describes a failure.
 * num_mats == 0 is a no-op.  Calls that touch the same matrices must be
 * ordered by the caller (same stream or events).
 */
#ifndef EINSPLINE_DETERMINANT_UPDATE_HIP_H
#define EINSPLINE_DETERMINANT_UPDATE_HIP_H

#include "multi_bspline_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sherman-Morrison update of Ainv (N x N) after row k of A changes.
 *
 * With A_list non-NULL (src/determinant_update.h:8-104): u_list holds the new
 * rows; row k of each A is overwritten with u and delta = u - (old row k).
 * With A_list NULL (src/determinant_update.cu:10-98): u_list holds delta, the
 * change of row k, itself.
 * Then, with d = delta^T Ainv and c = column k of the old Ainv,
 *   Ainv += (-1 / (1 + d_k)) c d^T.
 * Ainv_colk_list and Ainv_delta_list are work vectors of N reals per matrix
 * (the reference's own arguments).  On return Ainv_colk holds column k of the
 * old inverse, as in the reference; Ainv_delta holds delta in the A_list
 * form (the reference leaves d there) and is untouched in the delta form.
 * Requires 1 <= N, N <= rowstride, 0 <= k < N. */
int einspline_update_inverse_s(float **A_list, float **Ainv_list, float **u_list,
                               float **Ainv_delta_list, float **Ainv_colk_list,
                               int N, int rowstride, int k, int num_mats, void *stream);
int einspline_update_inverse_d(double **A_list, double **Ainv_list, double **u_list,
                               double **Ainv_delta_list, double **Ainv_colk_list,
                               int N, int rowstride, int k, int num_mats, void *stream);

/* Determinant ratios (src/cuda_walker.cu:150-185 calc_ratios):
 *   ratio_d[m] = sum_{i<N} Ainv_m[i*rowstride + elec] * new_row_m[i].
 * Requires 1 <= N <= rowstride and 0 <= elec < N; unlike the reference
 * kernel, N is not limited to 128. */
int einspline_calc_ratios_s(float **Ainv_list, float **new_row_list, float *ratio_d,
                            int N, int rowstride, int elec, int num_mats, void *stream);
int einspline_calc_ratios_d(double **Ainv_list, double **new_row_list, double *ratio_d,
                            int N, int rowstride, int elec, int num_mats, void *stream);

/* Row fill fused with the ratio (src/cuda_walker.cu:69-95 calc_new_row, then
 * calc_ratios): for walker m at pos_d[3m..3m+2], writes the values of every
 * orbital of `spline` to new_row_list[m] (num_splines reals, exactly the
 * values eval_multi_multi_UBspline_3d_{s,d}_hip writes) and
 *   ratio_d[m] = sum_{j<N} new_row_m[j] * Ainv_m[j*rowstride + elec],
 * without reading the row back from memory.  The _s form takes float tables
 * (any storage) and float positions, the _d form double tables and double
 * positions.  Requires 1 <= N <= num_splines, N <= rowstride, 0 <= elec < N. */
int einspline_fill_ratios_s(const multi_UBspline_3d_s_hip *spline, const float *pos_d,
                            float **new_row_list, float **Ainv_list, float *ratio_d,
                            int N, int rowstride, int elec, int num, void *stream);
int einspline_fill_ratios_d(const multi_UBspline_3d_d_hip *spline, const double *pos_d,
                            double **new_row_list, double **Ainv_list, double *ratio_d,
                            int N, int rowstride, int elec, int num, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EINSPLINE_DETERMINANT_UPDATE_HIP_H */
