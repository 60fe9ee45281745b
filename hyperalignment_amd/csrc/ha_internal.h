/*
 * ha_internal.h -- declarations shared by the host C files of the library.
 */
#ifndef HA_INTERNAL_H
#define HA_INTERNAL_H

#include <stdint.h>
#include <stdbool.h>
#include "ha_common.h"
#include "ha_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fills out[max_steps][3] with the polar coefficient schedule for an initial
 * lower bound ell0 on sigma_min / scale; growth_len (may be NULL) receives the
 * number of minimax steps before the Newton-Schulz tail. Returns the step
 * count or -1. */
int ha_polar_schedule(double ell0, int max_steps, float *out, int *growth_len);

/* Schedule table geometry shared by host and kernels. */
#define HA_SCHED_BUCKETS 37           /* ell_b = 10^(-b/4), b = 0..36 (down to 1e-9) */
#define HA_SCHED_PER_DECADE 4.0f
#define HA_SCHED_STEPS 28
/* kp buckets of the iteration tile choice: kp <= 64, 128, 256, 384, 512, larger */
#define HA_TILE_BUCKETS 6
#define HA_TILE_LISTS 4  /* iteration launches: 64, 128 (teams), rank-deficient, 128 alone */

/* Plan options beyond the public ones. */
typedef struct {
	bool build_pattern;       /* build the union CSC pattern on the GPU */
	bool build_weights_index; /* vertex->searchlight index from the X lists (weights) */
	bool index_only;          /* only the index / weights / pattern (no data, no transforms) */
	int32_t nv_x, nv_y;       /* vertices of X and Y */
} HaPlanSetup;

int ha_plan_create_ex(int32_t nt, const HaPlanSetup *setup,
                      const int32_t *slx_idx, const int32_t *sly_idx, const int32_t *sl_off,
                      const double *sl_dists, int32_t count, double radius, HaHipPlan **out);
/* Host check of a caller's CSC pattern (indptr from 0, non-decreasing, within
 * nnz; rows strictly increasing per column and below nrows): kHaErrorArg with a
 * message when it is malformed. */
int ha_check_csc(int32_t ncols, int32_t nrows, const int32_t *indptr, const int32_t *indices, int64_t nnz);
/* Uploads an external CSC pattern (and its current data) into the plan. */
int ha_plan_set_pattern(HaHipPlan *p, int32_t ncols, const int32_t *indptr, const int32_t *indices,
                        int64_t nnz, const double *data);
/* Uploads caller weights (flat, one per searchlight position) or NULL for none. */
int ha_plan_set_weights(HaHipPlan *p, const double *w_flat, bool use_computed);
int ha_plan_load_xy(HaHipPlan *p, const double *X, const double *Y, bool col_major);
/* One pass. zero_data: start from zero (else accumulate onto current data).
 * method < 0: use the global options. */
int ha_plan_run(HaHipPlan *p, bool isReflection, bool isScaling, bool zero_data, int method);
int ha_plan_download_data(const HaHipPlan *p, double *data);
/* neg[s] = 1 when the local transform in the plan's store has det < 0 (the LU
 * sign of k_detfix), 0 otherwise; the store is left as it is. */
int ha_plan_det_negative(HaHipPlan *p, int32_t *neg);
/* FP64 local transforms (Jacobi-64 path) of every searchlight, row-major k x k each. */
int ha_plan_local64(HaHipPlan *p, bool isReflection, bool isScaling, double *out);
int ha_plan_ridge(HaHipPlan *p, double alpha, bool zero_data, double *betas64);
/* Weights computed on the GPU for the plan's X lists (flat). */
int ha_plan_weights(HaHipPlan *p, double *w_flat);
int ha_plan_pattern(HaHipPlan *p, int32_t **indptr, int32_t **indices, int64_t *nnz);
int ha_plan_run_features(HaHipPlan *p, float *features, float *sigma_minmax);
/* Searchlight Procrustes template of ns subjects (nt x nv FP64 each) into tpl_out. */
int ha_plan_template(HaHipPlan *p, const double *const *dms, int32_t ns, bool col_major, bool isReflection,
                     bool isScaling, bool isZscore, int32_t level2, double *tpl_out);

void *ha_hip_stream(void);
int ha_hip_ensure(void);
int ha_host_threads(void);
int ha_check_finite(const double *A, int64_t n, const char *what);

#ifdef __cplusplus
}
#endif

#endif
