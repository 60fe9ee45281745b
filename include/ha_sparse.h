/*
 * ha_sparse.h -- CSC pattern construction and scatter-add, on the GPU.
 *
 * Signatures are those of the reference's csrc/ha_sparse.h:20,35.
 */
#ifndef HA_SPARSE_H
#define HA_SPARSE_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ha_sparse.h:20 / ha_sparse.c:31-112. Builds the union of all
 * (sl[i], sl[j]) pairs as a CSC pattern with zero data. Rows within a column
 * are sorted and unique, so indptr/indices are bit-identical to the
 * reference's qsort-and-deduplicate result. The GPU builds each column as a
 * bitmap union instead of materialising the sum-of-squares pair list.
 * nv <= 0 auto-detects nv = max index + 1. Returns NULL on failure. */
TSparseCSC *ha_sparse_init(const TSearchlights *sls, int32_t nv);

/* Replaces ha_sparse.h:35 / ha_sparse.c:128-142. Host routine (one local
 * transform): mat[sl_rows[i], sl_cols[j]] += local_T[i*n_cols + j] * w[i],
 * entries outside the pattern are skipped. */
void ha_sparse_scatter_add(TSparseCSC *mat,
                           const int32_t *sl_rows, const int32_t *sl_cols,
                           int32_t sl_n_rows, int32_t sl_n_cols,
                           const double *local_T, const double *weights);

#ifdef __cplusplus
}
#endif

#endif /* HA_SPARSE_H */
