/*
 * ha_searchlight.h -- searchlight weights, the two searchlight Procrustes
 * entry points and searchlight ridge, served by the HIP backend.
 *
 * Signatures are those of the reference's csrc/ha_searchlight.h:14-80. The
 * only backend this library executes is kHaBackendHIP; any other value
 * returns kHaErrorArg with a message on stderr (there is no silent CPU
 * fallback: the reference's CPU paths stay in the reference library).
 */
#ifndef HA_SEARCHLIGHT_H
#define HA_SEARCHLIGHT_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ha_searchlight.h:14 / ha_searchlight.c:20-26. Sets the number of
 * host threads used for the host-side parts (layout conversion of inputs and
 * the dense write-back). */
void ha_set_num_threads(int n);

/* Replaces ha_searchlight.h:20 / ha_searchlight.c:28-34. */
int ha_get_num_threads(void);

/* Replaces ha_searchlight.h:31 / ha_searchlight.c:36-92. Weights are computed
 * on the GPU, bit-identical to the reference's FP64 loop (same per-vertex
 * summation order). weights_out[s] is malloc'ed here; the caller frees each. */
int ha_searchlight_weights(const TSearchlights *sls, double **weights_out);

/* Replaces ha_searchlight.h:46-52 / ha_searchlight.c:147-413 (the CSC API).
 * X, Y: nt x nv row-major. mat: pattern from ha_sparse_init, accumulated
 * in place (data += contributions, in searchlight order). weights: NULL or
 * per-searchlight row weights. sls_Y: NULL or the same sizes as sls_X.
 * backend must be kHaBackendHIP. */
int ha_searchlight_procrustes(const TMat *X, const TMat *Y,
                              const TSearchlights *sls_X,
                              const TSearchlights *sls_Y,
                              TSparseCSC *mat,
                              double **weights,
                              bool isReflection, bool isScaling,
                              THaBackend backend);

/* Replaces ha_searchlight.h:69-80 / ha_searchlight.c:950-1376 (the flat-array
 * dense API used by hyperalignment_c.py:198-213). T_out (nv x nv row-major)
 * is accumulated: T_out[sl_i*nv + sl_j] += T_s[i,j] * w_s[i], in searchlight
 * order, exactly the reference CPU semantics (ha_searchlight.c:1072-1085). */
int ha_searchlight_procrustes_dense(
    const double *X_data, const double *Y_data,
    int32_t nt, int32_t nv,
    const int32_t *sl_indices,
    const int32_t *sl_offsets,
    const double *sl_dists,
    int32_t count,
    double radius,
    double *T_out,
    bool isReflection, bool isScaling,
    THaBackend backend,
    bool col_major);

/* Replaces ha_searchlight.h:83-91 / ha_searchlight.c:1378-1429 (searchlight
 * ridge). Per searchlight the ridge betas of ha_ridge (ha_ridge.c:13-54), in
 * FP64 on the GPU, scattered like the Procrustes transforms: mat->data +=
 * w_s[i] * betas_s[i, j], in searchlight order. alpha must be >= 0; a
 * searchlight whose X_sl^T X_sl + alpha I is not positive definite (alpha = 0
 * with fewer time points than vertices, or a zero column) returns kHaErrorSvd.
 * sls_Y: NULL or the same sizes as sls_X. The reference's function has no
 * backend argument; this one runs on the HIP backend. */
int ha_searchlight_ridge(const TMat *X, const TMat *Y,
                         const TSearchlights *sls_X,
                         const TSearchlights *sls_Y,
                         TSparseCSC *mat,
                         double **weights, double alpha);

#ifdef __cplusplus
}
#endif

#endif /* HA_SEARCHLIGHT_H */
