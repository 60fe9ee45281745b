/*
 * ha_api.c -- the reference's public entry points for the searchlight
 * Procrustes path, served by the HIP backend.
 *
 * Entry point                       reference
 * ha_searchlight_procrustes_dense   ha_searchlight.h:69-80, ha_searchlight.c:950-1376
 * ha_searchlight_procrustes         ha_searchlight.h:46-52, ha_searchlight.c:147-413
 * ha_searchlight_weights            ha_searchlight.h:31,    ha_searchlight.c:36-92
 * ha_sparse_init                    ha_sparse.h:20,         ha_sparse.c:31-112
 * ha_sparse_scatter_add             ha_sparse.h:35,         ha_sparse.c:128-142
 * ha_sparse_free                    ha_common.h:187-194
 * ha_procrustes / ha_procrustes_f32 ha_procrustes.h:20,27,  ha_procrustes.c:17-222
 * ha_template_procrustes            ha_template.h:18-22,    ha_template.c:14-114
 * ha_searchlight_template[_flat]    searchlight.py:215-236 with ha_template.c:14-114
 *
 * Only kHaBackendHIP is executed; other backend values are rejected with
 * kHaErrorArg (the reference's CPU backends are not part of this library).
 */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ha_internal.h"
#include "ha_searchlight.h"
#include "ha_sparse.h"
#include "ha_procrustes.h"
#include "ha_template.h"
#include "ha_ridge.h"
#include "hak.h"

#define ERR(...) fprintf(stderr, "ha_hip: " __VA_ARGS__)

void ha_sparse_free(TSparseCSC *m) {
	if (!m) return;
	free(m->data);
	free(m->indices);
	free(m->indptr);
	free(m);
}

/* flat arrays from a TSearchlights (hyperalignment_c.py:161-169 layout) */
typedef struct {
	int32_t *idx, *off;
	double *dist;
	int64_t total;
} Flat;

static void flat_free(Flat *f) { free(f->idx); free(f->off); free(f->dist); memset(f, 0, sizeof(*f)); }

static int flatten(const TSearchlights *sls, bool with_dists, Flat *f) {
	memset(f, 0, sizeof(*f));
	if (!sls || sls->count <= 0 || !sls->indices || !sls->sizes) return kHaErrorArg;
	int64_t total = 0;
	for (int32_t s = 0; s < sls->count; ++s) {
		if (sls->sizes[s] <= 0 || !sls->indices[s]) return kHaErrorArg;
		total += sls->sizes[s];
	}
	if (total >= ((int64_t)1 << 31)) { ERR("more than 2^31 searchlight entries\n"); return kHaErrorArg; }
	f->total = total;
	f->idx = (int32_t *)malloc((size_t)total * sizeof(int32_t));
	f->off = (int32_t *)malloc((size_t)(sls->count + 1) * sizeof(int32_t));
	if (with_dists && sls->dists) f->dist = (double *)malloc((size_t)total * sizeof(double));
	if (!f->idx || !f->off || (with_dists && sls->dists && !f->dist)) { flat_free(f); return kHaErrorAlloc; }
	int64_t o = 0;
	f->off[0] = 0;
	for (int32_t s = 0; s < sls->count; ++s) {
		memcpy(f->idx + o, sls->indices[s], (size_t)sls->sizes[s] * sizeof(int32_t));
		if (f->dist) {
			if (!sls->dists[s]) { flat_free(f); return kHaErrorArg; }
			memcpy(f->dist + o, sls->dists[s], (size_t)sls->sizes[s] * sizeof(double));
		}
		o += sls->sizes[s];
		f->off[s + 1] = (int32_t)o;
	}
	return kHaSuccess;
}

static int32_t max_index_plus_one(const int32_t *idx, int64_t n) {
	int32_t m = -1;
	for (int64_t k = 0; k < n; ++k) if (idx[k] > m) m = idx[k];
	return m + 1;
}

int ha_searchlight_weights(const TSearchlights *sls, double **weights_out) {
	if (!weights_out) return kHaErrorArg;
	Flat f;
	int rc = flatten(sls, true, &f);
	if (rc) return rc;
	const int32_t nv = max_index_plus_one(f.idx, f.total);
	HaPlanSetup su = {.build_pattern = false, .build_weights_index = true, .index_only = true,
	                  .nv_x = nv, .nv_y = nv};
	HaHipPlan *p = NULL;
	double *w = (double *)malloc((size_t)f.total * sizeof(double));
	if (!w) { flat_free(&f); return kHaErrorAlloc; }
	rc = ha_plan_create_ex(1, &su, f.idx, NULL, f.off, f.dist, sls->count, sls->radius, &p);
	if (!rc) rc = ha_plan_weights(p, w);
	ha_hip_plan_destroy(p);
	if (!rc) {
		for (int32_t s = 0; s < sls->count; ++s) {
			weights_out[s] = (double *)malloc((size_t)sls->sizes[s] * sizeof(double));
			if (!weights_out[s]) {
				for (int32_t t = 0; t < s; ++t) free(weights_out[t]);
				rc = kHaErrorAlloc;
				break;
			}
			memcpy(weights_out[s], w + f.off[s], (size_t)sls->sizes[s] * sizeof(double));
		}
	}
	free(w);
	flat_free(&f);
	return rc;
}

TSparseCSC *ha_sparse_init(const TSearchlights *sls, int32_t nv) {
	Flat f;
	if (flatten(sls, false, &f)) { ERR("ha_sparse_init: invalid searchlights\n"); return NULL; }
	int32_t nmax = max_index_plus_one(f.idx, f.total);
	if (nv <= 0) nv = nmax;
	if (nmax > nv) { ERR("ha_sparse_init: vertex index beyond nv\n"); flat_free(&f); return NULL; }
	HaPlanSetup su = {.build_pattern = true, .build_weights_index = false, .index_only = true,
	                  .nv_x = nv, .nv_y = nv};
	HaHipPlan *p = NULL;
	TSparseCSC *m = NULL;
	int rc = ha_plan_create_ex(1, &su, f.idx, NULL, f.off, NULL, sls->count, 0.0, &p);
	flat_free(&f);
	if (rc) return NULL;
	m = (TSparseCSC *)calloc(1, sizeof(TSparseCSC));
	if (!m) { ha_hip_plan_destroy(p); return NULL; }
	m->rows = nv;
	m->cols = nv;
	rc = ha_plan_pattern(p, &m->indptr, &m->indices, &m->nnz);
	ha_hip_plan_destroy(p);
	if (rc) { free(m); return NULL; }
	m->data = (double *)calloc((size_t)(m->nnz > 0 ? m->nnz : 1), sizeof(double));
	if (!m->data) { ha_sparse_free(m); return NULL; }
	return m;
}

void ha_sparse_scatter_add(TSparseCSC *mat, const int32_t *sl_rows, const int32_t *sl_cols, int32_t nr,
                           int32_t nc, const double *local_T, const double *weights) {
	if (!mat || !sl_rows || !sl_cols || !local_T || nr <= 0 || nc <= 0) return;
	if (ha_hip_ensure()) { ERR("ha_sparse_scatter_add: no GPU; matrix left unchanged\n"); return; }
	for (int32_t j = 0; j < nc; ++j)
		if (sl_cols[j] < 0 || sl_cols[j] >= mat->cols) { ERR("ha_sparse_scatter_add: column out of range\n"); return; }
	hipStream_t st = (hipStream_t)ha_hip_stream();
	int32_t *d_ptr = NULL, *d_ind = NULL, *d_r = NULL, *d_c = NULL;
	double *d_data = NULL, *d_T = NULL, *d_w = NULL;
	const size_t nnz = (size_t)mat->nnz;
	bool ok = hipMalloc((void **)&d_ptr, (size_t)(mat->cols + 1) * 4) == hipSuccess &&
	          hipMalloc((void **)&d_ind, (nnz ? nnz : 1) * 4) == hipSuccess &&
	          hipMalloc((void **)&d_data, (nnz ? nnz : 1) * 8) == hipSuccess &&
	          hipMalloc((void **)&d_r, (size_t)nr * 4) == hipSuccess &&
	          hipMalloc((void **)&d_c, (size_t)nc * 4) == hipSuccess &&
	          hipMalloc((void **)&d_T, (size_t)nr * nc * 8) == hipSuccess &&
	          (!weights || hipMalloc((void **)&d_w, (size_t)nr * 8) == hipSuccess);
	if (ok) {
		ok = hipMemcpyAsync(d_ptr, mat->indptr, (size_t)(mat->cols + 1) * 4, hipMemcpyHostToDevice, st) == hipSuccess &&
		     (!nnz || hipMemcpyAsync(d_ind, mat->indices, nnz * 4, hipMemcpyHostToDevice, st) == hipSuccess) &&
		     (!nnz || hipMemcpyAsync(d_data, mat->data, nnz * 8, hipMemcpyHostToDevice, st) == hipSuccess) &&
		     hipMemcpyAsync(d_r, sl_rows, (size_t)nr * 4, hipMemcpyHostToDevice, st) == hipSuccess &&
		     hipMemcpyAsync(d_c, sl_cols, (size_t)nc * 4, hipMemcpyHostToDevice, st) == hipSuccess &&
		     hipMemcpyAsync(d_T, local_T, (size_t)nr * nc * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
		     (!weights || hipMemcpyAsync(d_w, weights, (size_t)nr * 8, hipMemcpyHostToDevice, st) == hipSuccess);
	}
	if (ok) ok = hak_scatter_one(d_ptr, d_ind, d_data, d_r, d_c, nr, nc, d_T, d_w, st) == 0;
	if (ok && nnz) ok = hipMemcpyAsync(mat->data, d_data, nnz * 8, hipMemcpyDeviceToHost, st) == hipSuccess;
	if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
	if (!ok) ERR("ha_sparse_scatter_add: GPU call failed\n");
	hipStreamSynchronize(st);
	hipFree(d_ptr); hipFree(d_ind); hipFree(d_data); hipFree(d_r); hipFree(d_c); hipFree(d_T);
	if (d_w) hipFree(d_w);
}

/* ---- dense entry ---- */

int ha_searchlight_procrustes_hip(const double *X_data, const double *Y_data, int32_t nt, int32_t nv,
                                  const int32_t *sl_indices, const int32_t *sl_offsets, const double *sl_dists,
                                  int32_t count, double radius, double *T_out, bool isReflection, bool isScaling,
                                  bool col_major) {
	if (!X_data || !Y_data || !sl_indices || !sl_offsets || !T_out) return kHaErrorArg;
	if (nt <= 0 || nv <= 0 || count <= 0) return kHaErrorArg;
	if (ha_check_finite(X_data, (int64_t)nt * nv, "X") || ha_check_finite(Y_data, (int64_t)nt * nv, "Y"))
		return kHaErrorArg;
	HaHipPlan *p = NULL;
	int rc = ha_hip_plan_create(nt, nv, sl_indices, sl_offsets, sl_dists, count, radius, &p);
	if (rc) return rc;
	int32_t *indptr = NULL, *indices = NULL;
	double *data = NULL;
	int64_t nnz = 0;
	rc = ha_plan_pattern(p, &indptr, &indices, &nnz);
	if (rc) goto done;
	data = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
	if (!data) { rc = kHaErrorAlloc; goto done; }
	/* start from the caller's T_out values on the pattern (the CPU path accumulates) */
#pragma omp parallel for schedule(dynamic, 64) num_threads(ha_host_threads())
	for (int32_t c = 0; c < nv; ++c)
		for (int32_t q = indptr[c]; q < indptr[c + 1]; ++q) data[q] = T_out[(int64_t)indices[q] * nv + c];
	rc = ha_plan_set_pattern(p, nv, indptr, indices, nnz, data);
	if (rc) goto done;
	rc = ha_plan_load_xy(p, X_data, Y_data, col_major);
	if (rc) goto done;
	rc = ha_plan_run(p, isReflection, isScaling, false, -1);
	if (rc) goto done;
	rc = ha_plan_download_data(p, data);
	if (rc) goto done;
#pragma omp parallel for schedule(dynamic, 64) num_threads(ha_host_threads())
	for (int32_t c = 0; c < nv; ++c)
		for (int32_t q = indptr[c]; q < indptr[c + 1]; ++q) T_out[(int64_t)indices[q] * nv + c] = data[q];
done:
	free(indptr);
	free(indices);
	free(data);
	ha_hip_plan_destroy(p);
	return rc;
}

int ha_searchlight_procrustes_dense(const double *X_data, const double *Y_data, int32_t nt, int32_t nv,
                                    const int32_t *sl_indices, const int32_t *sl_offsets, const double *sl_dists,
                                    int32_t count, double radius, double *T_out, bool isReflection,
                                    bool isScaling, THaBackend backend, bool col_major) {
	if (!X_data || !Y_data || !sl_indices || !sl_offsets || !T_out) return kHaErrorArg;
	if (nt <= 0 || nv <= 0 || count <= 0) return kHaErrorArg;
	if (backend != kHaBackendHIP) {
		ERR("backend %d is not provided by this library (use kHaBackendHIP = 4)\n", (int)backend);
		return kHaErrorArg;
	}
	return ha_searchlight_procrustes_hip(X_data, Y_data, nt, nv, sl_indices, sl_offsets, sl_dists, count,
	                                     radius, T_out, isReflection, isScaling, col_major);
}

/* ---- CSC entry ---- */

int ha_searchlight_procrustes_hip_csc(const TMat *X, const TMat *Y, const TSearchlights *sls_X,
                                      const TSearchlights *sls_Y, TSparseCSC *mat, double **weights,
                                      bool isReflection, bool isScaling) {
	if (!X || !Y || !sls_X || !mat || !X->data || !Y->data) return kHaErrorArg;
	if (X->rows != Y->rows || X->rows <= 0) return kHaErrorArg;
	const TSearchlights *sy = sls_Y ? sls_Y : sls_X;
	if (sy->count != sls_X->count) return kHaErrorArg;
	if (mat->cols > Y->cols) { ERR("CSC pattern has more columns than Y vertices\n"); return kHaErrorArg; }
	if (ha_check_csc(mat->cols, X->cols, mat->indptr, mat->indices, mat->nnz)) return kHaErrorArg;
	if (ha_check_finite(X->data, (int64_t)X->rows * X->cols, "X") ||
	    ha_check_finite(Y->data, (int64_t)Y->rows * Y->cols, "Y"))
		return kHaErrorArg;
	for (int32_t s = 0; s < sls_X->count; ++s)
		if (sy->sizes[s] != sls_X->sizes[s]) {
			ERR("searchlight %d has %d X and %d Y vertices; this backend needs equal sizes\n", s,
			    sls_X->sizes[s], sy->sizes[s]);
			return kHaErrorArg;
		}
	Flat fx, fy;
	int rc = flatten(sls_X, false, &fx);
	if (rc) return rc;
	memset(&fy, 0, sizeof(fy));
	if (sls_Y && sls_Y != sls_X) {
		rc = flatten(sls_Y, false, &fy);
		if (rc) { flat_free(&fx); return rc; }
	}
	double *wflat = NULL;
	HaHipPlan *p = NULL;
	if (weights) {
		wflat = (double *)malloc((size_t)fx.total * sizeof(double));
		if (!wflat) { rc = kHaErrorAlloc; goto done; }
		for (int32_t s = 0; s < sls_X->count; ++s)
			memcpy(wflat + fx.off[s], weights[s], (size_t)sls_X->sizes[s] * sizeof(double));
	}
	HaPlanSetup su = {.build_pattern = false, .build_weights_index = false, .index_only = false,
	                  .nv_x = X->cols, .nv_y = Y->cols};
	rc = ha_plan_create_ex(X->rows, &su, fx.idx, fy.idx ? fy.idx : NULL, fx.off, NULL, sls_X->count, 0.0, &p);
	if (rc) goto done;
	rc = ha_plan_set_pattern(p, mat->cols, mat->indptr, mat->indices, mat->nnz, mat->data);
	if (rc) goto done;
	rc = ha_plan_set_weights(p, wflat, false);
	if (rc) goto done;
	rc = ha_plan_load_xy(p, X->data, Y->data, false);
	if (rc) goto done;
	rc = ha_plan_run(p, isReflection, isScaling, false, -1);
	if (rc) goto done;
	rc = ha_plan_download_data(p, mat->data);
done:
	ha_hip_plan_destroy(p);
	free(wflat);
	flat_free(&fx);
	flat_free(&fy);
	return rc;
}

int ha_searchlight_procrustes(const TMat *X, const TMat *Y, const TSearchlights *sls_X,
                              const TSearchlights *sls_Y, TSparseCSC *mat, double **weights, bool isReflection,
                              bool isScaling, THaBackend backend) {
	if (backend != kHaBackendHIP) {
		ERR("backend %d is not provided by this library (use kHaBackendHIP = 4)\n", (int)backend);
		return kHaErrorArg;
	}
	return ha_searchlight_procrustes_hip_csc(X, Y, sls_X, sls_Y, mat, weights, isReflection, isScaling);
}

/* ---- searchlight ridge ---- */

int ha_searchlight_ridge(const TMat *X, const TMat *Y, const TSearchlights *sls_X, const TSearchlights *sls_Y,
                         TSparseCSC *mat, double **weights, double alpha) {
	if (!X || !Y || !sls_X || !mat || !X->data || !Y->data) return kHaErrorArg;
	if (X->rows != Y->rows || X->rows <= 0) return kHaErrorArg;
	const TSearchlights *sy = sls_Y ? sls_Y : sls_X;
	if (sy->count != sls_X->count) return kHaErrorArg;
	if (mat->cols > Y->cols) { ERR("CSC pattern has more columns than Y vertices\n"); return kHaErrorArg; }
	if (ha_check_csc(mat->cols, X->cols, mat->indptr, mat->indices, mat->nnz)) return kHaErrorArg;
	if (ha_check_finite(X->data, (int64_t)X->rows * X->cols, "X") ||
	    ha_check_finite(Y->data, (int64_t)Y->rows * Y->cols, "Y"))
		return kHaErrorArg;
	for (int32_t s = 0; s < sls_X->count; ++s)
		if (sy->sizes[s] != sls_X->sizes[s]) {
			ERR("searchlight %d has %d X and %d Y vertices; this backend needs equal sizes\n", s,
			    sls_X->sizes[s], sy->sizes[s]);
			return kHaErrorArg;
		}
	Flat fx, fy;
	int rc = flatten(sls_X, false, &fx);
	if (rc) return rc;
	memset(&fy, 0, sizeof(fy));
	if (sls_Y && sls_Y != sls_X) {
		rc = flatten(sls_Y, false, &fy);
		if (rc) { flat_free(&fx); return rc; }
	}
	double *wflat = NULL;
	HaHipPlan *p = NULL;
	if (weights) {
		wflat = (double *)malloc((size_t)fx.total * sizeof(double));
		if (!wflat) { rc = kHaErrorAlloc; goto done; }
		for (int32_t s = 0; s < sls_X->count; ++s)
			memcpy(wflat + fx.off[s], weights[s], (size_t)sls_X->sizes[s] * sizeof(double));
	}
	HaPlanSetup su = {.build_pattern = false, .build_weights_index = false, .index_only = false,
	                  .nv_x = X->cols, .nv_y = Y->cols};
	rc = ha_plan_create_ex(X->rows, &su, fx.idx, fy.idx ? fy.idx : NULL, fx.off, NULL, sls_X->count, 0.0, &p);
	if (rc) goto done;
	rc = ha_plan_set_pattern(p, mat->cols, mat->indptr, mat->indices, mat->nnz, mat->data);
	if (rc) goto done;
	rc = ha_plan_set_weights(p, wflat, false);
	if (rc) goto done;
	rc = ha_plan_load_xy(p, X->data, Y->data, false);
	if (rc) goto done;
	rc = ha_plan_ridge(p, alpha, false, NULL);
	if (rc) goto done;
	rc = ha_plan_download_data(p, mat->data);
done:
	ha_hip_plan_destroy(p);
	free(wflat);
	flat_free(&fx);
	flat_free(&fy);
	return rc;
}

/* ---- ridge (one problem) ---- */

/* X (M x N) and Y (M x P) are padded with zero columns to K = max(N, P) and
 * solved as one K-vertex searchlight: for alpha > 0 the padded system is block
 * diagonal, diag(X^T X + alpha I, alpha I), so betas[:N, :P] are unchanged. */
/* One M x N by M x P problem as ceil(P / N) searchlights of N vertices: each
 * takes every column of X and the next N columns of Y (the last ones padded
 * with zero columns, whose betas are zero), so the systems stay N x N (or M x M
 * in the dual form when M < N) and alpha = 0 needs only X of full rank, as the
 * reference's SVD form (ha_ridge.c:13-54). */
int ha_ridge(const TMat *X, const TMat *Y, double alpha, TMat *betas) {
	if (!X || !Y || !betas || !X->data || !Y->data || !betas->data) return kHaErrorArg;
	const int32_t M = X->rows, N = X->cols, P = Y->cols;
	if (M <= 0 || N <= 0 || P <= 0 || Y->rows != M || betas->rows != N || betas->cols != P) return kHaErrorArg;
	if (ha_check_finite(X->data, (int64_t)M * N, "X") || ha_check_finite(Y->data, (int64_t)M * P, "Y"))
		return kHaErrorArg;
	const int32_t nch = (P + N - 1) / N;
	if ((int64_t)nch * N > INT32_MAX) return kHaErrorArg;
	const int32_t PY = nch * N;
	double *yp = (double *)calloc((size_t)M * PY, sizeof(double));
	double *bk = (double *)malloc((size_t)nch * N * N * sizeof(double));
	int32_t *ix = (int32_t *)malloc((size_t)PY * sizeof(int32_t));
	int32_t *iy = (int32_t *)malloc((size_t)PY * sizeof(int32_t));
	int32_t *off = (int32_t *)malloc((size_t)(nch + 1) * sizeof(int32_t));
	HaHipPlan *p = NULL;
	int rc = kHaErrorAlloc;
	if (!yp || !bk || !ix || !iy || !off) goto done;
	for (int32_t t = 0; t < M; ++t)
		memcpy(yp + (int64_t)t * PY, Y->data + (int64_t)t * P, (size_t)P * sizeof(double));
	for (int32_t c = 0; c < nch; ++c) {
		off[c] = c * N;
		for (int32_t i = 0; i < N; ++i) {
			ix[c * N + i] = i;
			iy[c * N + i] = c * N + i;
		}
	}
	off[nch] = PY;
	HaPlanSetup su = {.build_pattern = false, .build_weights_index = false, .index_only = false,
	                  .nv_x = N, .nv_y = PY};
	rc = ha_plan_create_ex(M, &su, ix, iy, off, NULL, nch, 0.0, &p);
	if (!rc) rc = ha_plan_load_xy(p, X->data, yp, false);
	if (!rc) rc = ha_plan_ridge(p, alpha, false, bk);
	if (!rc)
		for (int32_t c = 0; c < nch; ++c)
			for (int32_t r = 0; r < N; ++r)
				for (int32_t j = 0; j < N && c * N + j < P; ++j)
					betas->data[(int64_t)r * P + c * N + j] = bk[((int64_t)c * N + r) * N + j];
done:
	ha_hip_plan_destroy(p);
	free(yp); free(bk); free(ix); free(iy); free(off);
	return rc;
}

int ha_ridge_vec(const TMat *X, const double *y, double alpha, double *betas_out) {
	if (!X || !y || !betas_out) return kHaErrorArg;
	TMat Ym = {(double *)y, X->rows, 1}, Bm = {betas_out, X->cols, 1};
	return ha_ridge(X, &Ym, alpha, &Bm);
}

/* ---- local Procrustes (one problem) on the FP64 GPU path ---- */

int ha_procrustes(const TMat *X, const TMat *Y, TMat *T, bool isReflection, bool isScaling) {
	if (!X || !Y || !T || !X->data || !Y->data || !T->data) return kHaErrorArg;
	if (X->rows != Y->rows || X->cols != Y->cols || T->rows != X->cols || T->cols != X->cols) return kHaErrorArg;
	if (X->rows <= 0 || X->cols <= 0) return kHaErrorArg;
	if (ha_check_finite(X->data, (int64_t)X->rows * X->cols, "X") ||
	    ha_check_finite(Y->data, (int64_t)Y->rows * Y->cols, "Y"))
		return kHaErrorArg;
	const int32_t N = X->cols;
	int32_t *idx = (int32_t *)malloc((size_t)N * sizeof(int32_t));
	int32_t off[2] = {0, N};
	if (!idx) return kHaErrorAlloc;
	for (int32_t i = 0; i < N; ++i) idx[i] = i;
	HaPlanSetup su = {.build_pattern = false, .build_weights_index = false, .index_only = false,
	                  .nv_x = N, .nv_y = N};
	HaHipPlan *p = NULL;
	int rc = ha_plan_create_ex(X->rows, &su, idx, NULL, off, NULL, 1, 0.0, &p);
	if (!rc) rc = ha_plan_load_xy(p, X->data, Y->data, false);
	if (!rc) rc = ha_plan_local64(p, isReflection, isScaling, T->data);
	ha_hip_plan_destroy(p);
	free(idx);
	return rc;
}

int ha_procrustes_f32(const TMatF *X, const TMatF *Y, TMatF *T, bool isReflection, bool isScaling) {
	if (!X || !Y || !T || !X->data || !Y->data || !T->data) return kHaErrorArg;
	const size_t nx = (size_t)X->rows * X->cols, nn = (size_t)X->cols * X->cols;
	double *xd = (double *)malloc(nx * sizeof(double)), *yd = (double *)malloc(nx * sizeof(double));
	double *td = (double *)malloc(nn * sizeof(double));
	int rc = kHaErrorAlloc;
	if (xd && yd && td && (size_t)Y->rows * Y->cols == nx) {
		for (size_t i = 0; i < nx; ++i) { xd[i] = X->data[i]; yd[i] = Y->data[i]; }
		TMat Xm = {xd, X->rows, X->cols}, Ym = {yd, Y->rows, Y->cols}, Tm = {td, T->rows, T->cols};
		rc = ha_procrustes(&Xm, &Ym, &Tm, isReflection, isScaling);
		if (!rc)
			for (size_t i = 0; i < nn; ++i) T->data[i] = (float)td[i];
	} else if (xd && yd && td) {
		rc = kHaErrorArg;
	}
	free(xd); free(yd); free(td);
	return rc;
}

/* ---- polar factor of a given matrix (ha_procrustes.c:271-366) ---- */

/* The reference runs the unscaled Newton iteration X <- (X + X^-T)/2 on A in
 * FP32 (at most 20 steps, relative change below 1e-6); here A goes through the
 * GPU polar path as a one-searchlight pass with X = I (so the cross-covariance
 * is A itself, formed with FP32 products) and Y = A: the FP32-product iteration
 * of the nearly rank-deficient searchlights, from the smallest bucket's
 * schedule, with the FP64 Jacobi path behind it. Both converge to the
 * orthogonal polar factor of A. The reference's reflection and scaling rules
 * follow: det(T) < 0 (LU sign of T) negates T's last column; with isScaling and
 * X_data, T is multiplied by trace(T^T A) / (sum_j var(X[:, j]) * M). */
int ha_polar_newton(const float *A, float *T, int32_t N, bool isReflection, bool isScaling, const float *X_data,
                    int32_t M) {
	if (!A || !T || N <= 0) return kHaErrorArg;
	if (isScaling && X_data && M <= 0) return kHaErrorArg;
	const size_t nn = (size_t)N * N;
	double *xi = (double *)calloc(nn, sizeof(double)), *ya = (double *)malloc(nn * sizeof(double));
	int32_t *idx = (int32_t *)malloc((size_t)N * sizeof(int32_t));
	int32_t off[2] = {0, N}, neg = 0, which = 0;
	HaHipPlan *p = NULL;
	int rc = kHaErrorAlloc;
	if (!xi || !ya || !idx) goto done;
	for (int32_t i = 0; i < N; ++i) {
		xi[(size_t)i * N + i] = 1.0;
		idx[i] = i;
	}
	for (size_t e = 0; e < nn; ++e) ya[e] = A[e];
	rc = ha_hip_plan_create(N, N, idx, off, NULL, 1, 0.0, &p);
	if (!rc) rc = ha_plan_load_xy(p, xi, ya, false);
	if (!rc) rc = ha_plan_run(p, true, false, true, kHaHipMethodPolar);
	if (!rc) rc = ha_hip_plan_download_local(p, &which, 1, T);
	if (!rc && !isReflection) rc = ha_plan_det_negative(p, &neg);
	if (rc) goto done;
	if (neg)
		for (int32_t i = 0; i < N; ++i) T[(size_t)i * N + (N - 1)] = -T[(size_t)i * N + (N - 1)];
	if (isScaling && X_data) {
		double tr = 0.0, var_sum = 0.0;
		for (size_t e = 0; e < nn; ++e) tr += (double)T[e] * (double)A[e];  /* trace(T^T A) */
		for (int32_t j = 0; j < N; ++j) {
			double mean = 0.0, var = 0.0;
			for (int32_t t = 0; t < M; ++t) mean += X_data[(size_t)t * N + j];
			mean /= M;
			for (int32_t t = 0; t < M; ++t) {
				const double d = X_data[(size_t)t * N + j] - mean;
				var += d * d;
			}
			var_sum += var / M;
		}
		const double scale = tr / (var_sum * M);
		for (size_t e = 0; e < nn; ++e) T[e] = (float)(T[e] * scale);
	}
done:
	ha_hip_plan_destroy(p);
	free(xi); free(ya); free(idx);
	return rc;
}

/* The statistics of the last ha_searchlight_template_flat call (its plan is
 * gone when the call returns): kernel time, step counts and FP64 flops. */
static HaHipStats g_last_template_stats;

int ha_hip_last_template_stats(HaHipStats *stats) {
	if (!stats) return kHaErrorArg;
	*stats = g_last_template_stats;
	return kHaSuccess;
}

/* ---- templates ---- */

int ha_searchlight_template_flat(const double *const *dms, int32_t n_subj, int32_t nt, int32_t nv,
                                 const int32_t *sl_indices, const int32_t *sl_offsets, const double *sl_dists,
                                 int32_t count, double radius, bool isReflection, bool isScaling,
                                 bool isZscoreCommon, int32_t level2_iter, double *tpl_out, bool col_major) {
	if (!dms || !sl_indices || !sl_offsets || !tpl_out || n_subj < 2 || nt <= 0 || nv <= 0 || count <= 0 ||
	    level2_iter < 0)
		return kHaErrorArg;
	for (int32_t s = 0; s < n_subj; ++s)
		if (!dms[s] || ha_check_finite(dms[s], (int64_t)nt * nv, "a subject")) return kHaErrorArg;
	HaPlanSetup su = {.build_pattern = false, .build_weights_index = true, .index_only = true,
	                  .nv_x = nv, .nv_y = nv};
	HaHipPlan *p = NULL;
	int rc = ha_plan_create_ex(nt, &su, sl_indices, NULL, sl_offsets, sl_dists, count, radius, &p);
	if (rc) return rc;
	rc = ha_plan_template(p, dms, n_subj, col_major, isReflection, isScaling, isZscoreCommon, level2_iter, tpl_out);
	ha_hip_plan_stats(p, &g_last_template_stats);
	ha_hip_plan_destroy(p);
	return rc;
}

static int template_tmats(const TMat **dms, int32_t n_subj, const TMat *tpl_out, const double ***flat) {
	if (!dms || !tpl_out || !tpl_out->data || n_subj < 2) return kHaErrorArg;
	const double **f = (const double **)malloc((size_t)n_subj * sizeof(double *));
	if (!f) return kHaErrorAlloc;
	for (int32_t s = 0; s < n_subj; ++s) {
		if (!dms[s] || !dms[s]->data || dms[s]->rows != tpl_out->rows || dms[s]->cols != tpl_out->cols) {
			ERR("template subjects must all be %d x %d\n", tpl_out->rows, tpl_out->cols);
			free(f);
			return kHaErrorArg;
		}
		f[s] = dms[s]->data;
	}
	*flat = f;
	return kHaSuccess;
}

int ha_searchlight_template(const TMat **dms, int32_t n_subj, const TSearchlights *sls, bool isReflection,
                            bool isScaling, bool isZscoreCommon, int32_t level2_iter, TMat *tpl_out,
                            THaBackend backend) {
	if (backend != kHaBackendHIP) {
		ERR("backend %d is not provided by this library (use kHaBackendHIP = 4)\n", (int)backend);
		return kHaErrorArg;
	}
	const double **flat = NULL;
	int rc = template_tmats(dms, n_subj, tpl_out, &flat);
	if (rc) return rc;
	Flat f;
	rc = flatten(sls, true, &f);
	if (rc) { free(flat); return rc; }
	rc = ha_searchlight_template_flat(flat, n_subj, tpl_out->rows, tpl_out->cols, f.idx, f.off, f.dist,
	                                  sls->count, sls->radius, isReflection, isScaling, isZscoreCommon,
	                                  level2_iter, tpl_out->data, false);
	flat_free(&f);
	free(flat);
	return rc;
}

int ha_template_procrustes(const TMat **dms, int32_t n_subj, bool isReflection, bool isScaling,
                           bool isZscoreCommon, int32_t level2_iter, TMat *tpl_out) {
	const double **flat = NULL;
	int rc = template_tmats(dms, n_subj, tpl_out, &flat);
	if (rc) return rc;
	const int32_t nv = tpl_out->cols;
	int32_t *idx = (int32_t *)malloc((size_t)nv * sizeof(int32_t));
	int32_t off[2] = {0, nv};
	if (!idx) { free(flat); return kHaErrorAlloc; }
	for (int32_t v = 0; v < nv; ++v) idx[v] = v;
	/* one searchlight over every column: each vertex has weight 1 */
	rc = ha_searchlight_template_flat(flat, n_subj, tpl_out->rows, nv, idx, off, NULL, 1, 0.0, isReflection,
	                                  isScaling, isZscoreCommon, level2_iter, tpl_out->data, false);
	free(idx);
	free(flat);
	return rc;
}
