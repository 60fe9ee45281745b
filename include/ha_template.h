/*
 * ha_template.h -- Procrustes templates on the GPU.
 *
 * ha_template_procrustes keeps the signature of the reference's
 * csrc/ha_template.h:18-22. ha_searchlight_template has no C counterpart in
 * the reference: it is the searchlight combination of
 * src/hyperalignment/searchlight.py:215-236 (local templates added into the
 * template with the searchlight weights, tpl[:, sl] += local * w) with the
 * Procrustes template kind of ha_template.c:14-114 as the local template.
 */
#ifndef HA_TEMPLATE_H
#define HA_TEMPLATE_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ha_template.h:18-22 / ha_template.c:14-114. dms: n_subj matrices of
 * the same nt x nv shape (row-major); tpl_out: nt x nv, overwritten. Sequential
 * alignment to the running common space, level2_iter leave-one-out passes,
 * then the sum of the aligned subjects (z-scored per column with the one-pass
 * formula of ha_linalg.c:263-291 when isZscoreCommon, else averaged). Runs as
 * one searchlight holding every column, so nv is meant to be searchlight-sized
 * (up to a few thousand). isReflection = false gives every Procrustes solve of
 * the chain det(T) = +1 (ha_template.c:40,80 pass it to ha_procrustes). */
int ha_template_procrustes(const TMat **dms, int32_t n_subj, bool isReflection, bool isScaling,
                           bool isZscoreCommon, int32_t level2_iter, TMat *tpl_out);

/* Searchlight template: for every searchlight, the local Procrustes template
 * of the subjects' columns in it (as ha_template_procrustes), added into
 * tpl_out[:, sl] with the searchlight weights of ha_searchlight_weights
 * (distance weights when sls->dists is set, else 1/count). tpl_out: nt x nv
 * row-major, overwritten. backend must be kHaBackendHIP. */
int ha_searchlight_template(const TMat **dms, int32_t n_subj, const TSearchlights *sls, bool isReflection,
                            bool isScaling, bool isZscoreCommon, int32_t level2_iter, TMat *tpl_out,
                            THaBackend backend);

/* Flat-array form of ha_searchlight_template (the layout of
 * ha_searchlight_procrustes_dense): dms[s] nt x nv FP64, column-major when
 * col_major; searchlights as concatenated indices, offsets[count + 1] and
 * optional distances; tpl_out in the same order as the inputs. */
int ha_searchlight_template_flat(const double *const *dms, int32_t n_subj, int32_t nt, int32_t nv,
                                 const int32_t *sl_indices, const int32_t *sl_offsets, const double *sl_dists,
                                 int32_t count, double radius, bool isReflection, bool isScaling,
                                 bool isZscoreCommon, int32_t level2_iter, double *tpl_out, bool col_major);

#ifdef __cplusplus
}
#endif

#endif /* HA_TEMPLATE_H */
