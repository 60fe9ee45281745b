/*
 * ha_hip.c -- HIP backend of searchlight Procrustes: device state, plans and
 * the pass orchestration. Plain C on the HIP runtime API; all arithmetic runs
 * in the gfx950 kernels behind hak.h.
 *
 * Reference interfaces replaced: ha_cuda.h:38-66 / ha_cuda.cu:47-70, 355-468
 * (init / cleanup / available / dense entry) and the CPU loop of
 * ha_searchlight.c:147-202 (CSC entry, which has no GPU branch there).
 */
#define _GNU_SOURCE
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "ha_internal.h"
#include "hak.h"

/* ------------------------------------------------------------------------ */
/* global state                                                              */
/* ------------------------------------------------------------------------ */

static struct {
	bool ready;
	hipStream_t stream;
	int n_cu;
	HaHipOptions opt;
	float *d_sched;
	int32_t *d_sched_len;
	float *d_mlp;
	float *d_mlp_norm;
	bool has_model;
	int host_threads;
} g = {
	.ready = false,
	.opt = {
		.method = kHaHipMethodAuto,
		.fp64_fallback = 1,
		.max_polar_steps = HA_SCHED_STEPS,
		.polar_tol = 1e-2,
		.plain_ell = 1e-7,
		.tile = 0,
		.use_graph = 1,
		.chunk = 0,
		.verbose = 0,
		.learned_margin_scale = 0.15f,
		.learned_margin_ell = 1.0f,
		.team = 1,
		.team_spin = 0,
		.orth_tol = 1e-3,
		.template_ell = 3e-4,
	},
	.host_threads = 0,
};

#define ERR(...) fprintf(stderr, "ha_hip: " __VA_ARGS__)

/* hip call check: on failure print and jump to `fail` with rc set */
#define HCK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
	ERR("%s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
	rc = kHaErrorInternal; goto fail; } } while (0)
/* kernel launcher check */
#define KCK(x) do { int e_ = (x); if (e_ != 0) { \
	ERR("%s failed: %s (%s:%d)\n", #x, hipGetErrorString((hipError_t)e_), __FILE__, __LINE__); \
	rc = kHaErrorInternal; goto fail; } } while (0)
#define ACK(p, bytes) do { size_t b_ = (size_t)(bytes); if (b_ == 0) b_ = 16; \
	if (hipMalloc((void **)&(p), b_) != hipSuccess) { \
	ERR("device allocation of %zu bytes failed (%s:%d)\n", b_, __FILE__, __LINE__); \
	(void)hipGetLastError(); rc = kHaErrorAlloc; goto fail; } } while (0)

void *ha_hip_stream(void) { return (void *)g.stream; }

int ha_host_threads(void) {
#ifdef _OPENMP
	return g.host_threads > 0 ? g.host_threads : omp_get_max_threads();
#else
	return 1;
#endif
}

void ha_set_num_threads(int n) { g.host_threads = n > 0 ? n : 0; }

/* kHaErrorArg unless every one of the n values is finite (a host scan before
 * any upload: NaN or Inf in X or Y would otherwise come back as NaN transforms,
 * or as the FP64 path's best effort, with no error). */
int ha_check_finite(const double *A, int64_t n, const char *what) {
	int bad = 0;
#pragma omp parallel for schedule(static) reduction(|| : bad) num_threads(ha_host_threads())
	for (int64_t i = 0; i < n; ++i) bad = bad || !isfinite(A[i]);
	if (bad) {
		ERR("%s holds a value that is not finite\n", what);
		return kHaErrorArg;
	}
	return kHaSuccess;
}
int ha_get_num_threads(void) { return ha_host_threads(); }

static int build_schedules(void) {
	int rc = kHaSuccess;
	size_t n = (size_t)HA_SCHED_BUCKETS * HA_SCHED_STEPS * 3;
	float *h = (float *)malloc(n * sizeof(float));
	int32_t len[HA_SCHED_BUCKETS];
	if (!h) return kHaErrorAlloc;
	for (int b = 0; b < HA_SCHED_BUCKETS; ++b) {
		double ell = pow(10.0, -(double)b / HA_SCHED_PER_DECADE);
		if (b == 0) ell = 0.96; /* already converged: quintic Newton-Schulz */
		int gl = 0;
		if (ha_polar_schedule(ell, HA_SCHED_STEPS, h + (size_t)b * HA_SCHED_STEPS * 3, &gl) < 0) {
			ERR("polar schedule for ell=%g failed\n", ell);
			free(h);
			return kHaErrorInternal;
		}
		len[b] = gl;
	}
	if (!g.d_sched) ACK(g.d_sched, n * sizeof(float));
	if (!g.d_sched_len) ACK(g.d_sched_len, sizeof(len));
	HCK(hipMemcpy(g.d_sched, h, n * sizeof(float), hipMemcpyHostToDevice));
	HCK(hipMemcpy(g.d_sched_len, len, sizeof(len), hipMemcpyHostToDevice));
fail:
	free(h);
	return rc;
}

int ha_hip_init(void) {
	int rc = kHaSuccess;
	if (g.ready) return kHaSuccess;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
		(void)hipGetLastError();
		ERR("no HIP device available\n");
		return kHaErrorInternal;
	}
	HCK(hipSetDevice(0));
	hipDeviceProp_t prop;
	HCK(hipGetDeviceProperties(&prop, 0));
	g.n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
	HCK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
	/* device address of the no-allocation entry point's scratch array (device 0) */
	KCK(hak_noalloc_attach());
	g.ready = true;
	rc = build_schedules();
	if (rc != kHaSuccess) { ha_hip_cleanup(); return rc; }
	return kHaSuccess;
fail:
	return rc;
}

void ha_hip_cleanup(void) {
	if (!g.ready) return;
	hipStreamSynchronize(g.stream);
	if (g.d_sched) hipFree(g.d_sched);
	if (g.d_sched_len) hipFree(g.d_sched_len);
	g.d_sched_len = NULL;
	if (g.d_mlp) hipFree(g.d_mlp);
	if (g.d_mlp_norm) hipFree(g.d_mlp_norm);
	g.d_sched = g.d_mlp = g.d_mlp_norm = NULL;
	g.has_model = false;
	hipStreamDestroy(g.stream);
	g.ready = false;
}

bool ha_hip_available(void) { return g.ready; }

int ha_hip_ensure(void) {
	if (g.ready) return kHaSuccess;
	return ha_hip_init();
}

void ha_hip_get_options(HaHipOptions *opt) { if (opt) *opt = g.opt; }

int ha_hip_set_options(const HaHipOptions *opt) {
	if (!opt) return kHaErrorArg;
	if (opt->method < kHaHipMethodAuto || opt->method > kHaHipMethodJacobi64) {
		ERR("unknown method %d\n", opt->method);
		return kHaErrorArg;
	}
	if (opt->tile != 0 && opt->tile != 64 && opt->tile != 128) {
		ERR("tile must be 0, 64 or 128\n");
		return kHaErrorArg;
	}
	if (opt->max_polar_steps < 1 || opt->max_polar_steps > 64 || !(opt->polar_tol > 0) ||
	    !(opt->plain_ell > 0 && opt->plain_ell < 1) || !(opt->template_ell > 0 && opt->template_ell < 1)) {
		ERR("invalid polar options\n");
		return kHaErrorArg;
	}
	g.opt = *opt;
	return kHaSuccess;
}

int ha_hip_set_model(const float *params, const float *feat_norm) {
	int rc = ha_hip_ensure();
	if (rc) return rc;
	if (!params) { g.has_model = false; return kHaSuccess; }
	if (!feat_norm) return kHaErrorArg;
	if (!g.d_mlp) ACK(g.d_mlp, HA_HIP_MLP_NPARAM * sizeof(float));
	if (!g.d_mlp_norm) ACK(g.d_mlp_norm, 2 * HA_HIP_MLP_IN * sizeof(float));
	HCK(hipMemcpy(g.d_mlp, params, HA_HIP_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
	HCK(hipMemcpy(g.d_mlp_norm, feat_norm, 2 * HA_HIP_MLP_IN * sizeof(float), hipMemcpyHostToDevice));
	g.has_model = true;
	return kHaSuccess;
fail:
	return rc;
}

/* ------------------------------------------------------------------------ */
/* plans                                                                     */
/* ------------------------------------------------------------------------ */

struct HaHipPlan {
	int32_t nt, nv_x, nv_y, count, ld;
	int64_t total;
	double radius;
	bool has_dists, own_pattern;
	int32_t kmax, kpmax;
	/* searchlights */
	int32_t *slx, *sly, *off, *sl_of;
	double *dist;
	int32_t *h_k; /* host copy of sizes */
	/* inverse index over the Y lists (scatter) and, when they differ, X lists (weights) */
	int64_t *occ_ptr; int32_t *occ;
	int64_t *occx_ptr; int32_t *occx;
	/* weights */
	double *wsum, *w;
	bool weights_computed; /* w comes from the GPU weight kernel */
	bool weights_none;
	/* CSC */
	int32_t ncols;
	int32_t max_col;   /* longest CSC column (entries) */
	int64_t nnz;
	int64_t *indptr64; int32_t *indptr; int32_t *indices; double *data;
	/* local transforms */
	int32_t *qld; int64_t *qoff; float *q; int64_t q_elems;
	double *q64;
	float *X, *Y; double *X64, *Y64; double *stage;
	uint16_t *Xs, *Ys;   /* bf16 hi plane then lo plane of X and Y (ld x nv each) */
	float *work; int64_t slot_stride; int32_t n_slots, n_slots_polar;
	float *sl_par;
	double *work64; int64_t stride64; int32_t n_slots64;
	int32_t *iota;   /* 0, 1, ..., kpmax - 1 (the reflection fix's A^T A product) */
	int32_t *det_neg;  /* reflection fix: per searchlight det Q < 0, from k_lu_sign */
	int64_t work_floats;  /* size of the work area */
	double *work_ridge; int32_t n_wg_ridge;  /* ridge: 2 x stride64 doubles per workgroup */
	void *grp; int32_t n_teams;               /* polar teams (128 x 128 tiles): state per team */
	int32_t *flag, *steps, *queues, *order;
	double *sigma_sum, *var_sum, *scale;
	double *orth;        /* per searchlight FP64 orthogonality of Q (the pass's check) */
	bool orth_ran;       /* the last pass ran the check */
	float *features, *pred, *trace, *sig;
	int32_t *d_which;
	HaHipStats stats;
	hipEvent_t ev[8];
	hipEvent_t ev_graph[2];  /* around a graph launch */
	bool events;
	/* polar iteration tiles: per kp bucket, and the k-sorted searchlights of each tile */
	int32_t bucket_tile[HA_TILE_BUCKETS];
	bool tuned;
	int32_t tuned_opt;  /* the tile option the choice was made under */
	int32_t bucket_team[HA_TILE_BUCKETS];  /* 128-tile buckets: 1 = teams of HAK_GRP workgroups */
	int32_t tune_sample[HA_TILE_BUCKETS];  /* searchlights timed per bucket (0: not timed) */
	float tune_ms[HA_TILE_BUCKETS][3];     /* their iteration time: 64, 128 alone, 128 in teams */
	/* k-sorted searchlights per launch: 64, 128 (teams first), rank-deficient
	 * (FP32 products), 128 one workgroup each */
	int32_t *tile_order[HA_TILE_LISTS]; int32_t tile_count[HA_TILE_LISTS];
	/* captured pass (use_graph) and the arguments it was captured with */
	hipGraphExec_t gexec;
	bool gevents;       /* the captured pass records the stage events */
	int64_t n_passes;   /* polar passes run on this plan */
	struct { HakLocalArgs a; HaHipOptions opt; int method; bool refl, scal, zero; } gkey;
	int32_t tuned_team;
	int32_t tuned_fb;   /* the fp64_fallback option the lists were built under */
	bool loaded;
};

static void plan_free(HaHipPlan *p) {
	if (!p) return;
	void *ptrs[] = {p->slx, p->sly == p->slx ? NULL : p->sly, p->off, p->sl_of, p->dist, p->occ_ptr, p->occ,
	                p->occx_ptr, p->occx, p->wsum, p->w, p->indptr64, p->indptr, p->indices, p->data,
	                p->qld, p->qoff, p->q, p->q64, p->X, p->Y, p->X64, p->Y64, p->Xs, p->Ys, p->stage, p->work, p->work64,
	                p->flag, p->steps, p->queues, p->order, p->sigma_sum, p->var_sum, p->scale,
	                p->features, p->pred, p->trace, p->sig, p->d_which, p->sl_par, p->tile_order[0],
	                p->tile_order[1], p->tile_order[2], p->tile_order[3], p->work_ridge, p->grp, p->iota, p->orth, p->det_neg};
	for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
		if (ptrs[i]) hipFree(ptrs[i]);
	if (p->gexec) hipGraphExecDestroy(p->gexec);
	if (p->events) {
		for (int i = 0; i < 8; ++i) hipEventDestroy(p->ev[i]);
		for (int i = 0; i < 2; ++i) hipEventDestroy(p->ev_graph[i]);
	}
	free(p->h_k);
	free(p);
}

void ha_hip_plan_destroy(HaHipPlan *p) {
	if (p && g.ready) hipStreamSynchronize(g.stream);
	plan_free(p);
}

static int cmp_k_desc(const void *a, const void *b, void *ctx) {
	const int32_t *k = (const int32_t *)ctx;
	int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
	if (k[x] != k[y]) return k[x] > k[y] ? -1 : 1;
	return x < y ? -1 : (x > y);
}

/* vertex -> occurrence index (sorted by flat position) over list idx */
static int build_occ(HaHipPlan *p, const int32_t *idx, int32_t nv, int64_t **ptr_out, int32_t **occ_out) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	int32_t *cnt = NULL, *err = NULL;
	int64_t *ptr = NULL;
	int32_t *occ = NULL;
	ACK(cnt, (size_t)nv * sizeof(int32_t));
	ACK(err, sizeof(int32_t));
	ACK(ptr, (size_t)(nv + 1) * sizeof(int64_t));
	ACK(occ, (size_t)p->total * sizeof(int32_t));
	KCK(hak_fill_i32(cnt, nv, 0, st));
	KCK(hak_fill_i32(err, 1, 0, st));
	KCK(hak_count_occ(idx, p->total, cnt, st));
	KCK(hak_scan_i32_i64(cnt, nv, ptr, st));
	KCK(hak_fill_i32(cnt, nv, 0, st));
	KCK(hak_fill_occ(idx, p->total, ptr, cnt, occ, st));
	KCK(hak_sort_occ(ptr, nv, p->sl_of, occ, err, st));
	int32_t h_err = 0;
	HCK(hipMemcpyAsync(&h_err, err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
	if (h_err == 1) {
		ERR("a vertex belongs to more than 8192 searchlights (unsupported)\n");
		rc = kHaErrorArg;
		goto fail;
	}
	if (h_err == 2) {
		ERR("a searchlight lists the same vertex twice\n");
		rc = kHaErrorArg;
		goto fail;
	}
	hipFree(cnt);
	hipFree(err);
	*ptr_out = ptr;
	*occ_out = occ;
	return kHaSuccess;
fail:
	if (cnt) hipFree(cnt);
	if (err) hipFree(err);
	if (ptr) hipFree(ptr);
	if (occ) hipFree(occ);
	return rc;
}

static int pick_tile(int kpmax) {
	if (g.opt.tile) return g.opt.tile;
	(void)kpmax;
	return 64; /* measured: 64x64 tiles at two workgroups per CU beat 128x128 at one */
}

int ha_plan_create_ex(int32_t nt, const HaPlanSetup *su, const int32_t *slx_idx, const int32_t *sly_idx,
                      const int32_t *sl_off, const double *sl_dists, int32_t count, double radius,
                      HaHipPlan **out) {
	/* the arguments first (host only), so that they are reported as such with or without a device */
	if (!out || !slx_idx || !sl_off || count <= 0 || nt <= 0 || su->nv_x <= 0 || su->nv_y <= 0) {
		ERR("invalid plan arguments\n");
		return kHaErrorArg;
	}
	if (sl_off[0] != 0) { ERR("sl_offsets[0] must be 0\n"); return kHaErrorArg; }
	for (int32_t s = 0; s < count; ++s)
		if (sl_off[s + 1] <= sl_off[s]) { ERR("searchlight %d is empty or offsets decrease\n", s); return kHaErrorArg; }
	const int64_t total = sl_off[count];
	for (int64_t k = 0; k < total; ++k) {
		if (slx_idx[k] < 0 || slx_idx[k] >= su->nv_x) { ERR("vertex index %d out of range\n", slx_idx[k]); return kHaErrorArg; }
		if (sly_idx && (sly_idx[k] < 0 || sly_idx[k] >= su->nv_y)) { ERR("vertex index %d out of range\n", sly_idx[k]); return kHaErrorArg; }
	}
	int rc = ha_hip_ensure();
	if (rc) return rc;
	HaHipPlan *p = (HaHipPlan *)calloc(1, sizeof(HaHipPlan));
	if (!p) return kHaErrorAlloc;
	hipStream_t st = g.stream;
	p->nt = nt; p->nv_x = su->nv_x; p->nv_y = su->nv_y; p->count = count; p->total = total;
	p->radius = radius; p->has_dists = sl_dists != NULL;
	p->ld = (nt + 31) / 32 * 32;
	p->h_k = (int32_t *)malloc((size_t)count * sizeof(int32_t));
	int32_t *h_qld = (int32_t *)malloc((size_t)count * sizeof(int32_t));
	int64_t *h_qoff = (int64_t *)malloc((size_t)count * sizeof(int64_t));
	int32_t *h_order = (int32_t *)malloc((size_t)count * sizeof(int32_t));
	if (!p->h_k || !h_qld || !h_qoff || !h_order) { rc = kHaErrorAlloc; goto fail; }
	int64_t qe = 0;
	for (int32_t s = 0; s < count; ++s) {
		int32_t k = sl_off[s + 1] - sl_off[s];
		p->h_k[s] = k;
		if (k > p->kmax) p->kmax = k;
		h_qld[s] = (k + 31) / 32 * 32;
		h_qoff[s] = qe;
		qe += (int64_t)h_qld[s] * h_qld[s];
		h_order[s] = s;
	}
	p->kpmax = (p->kmax + 31) / 32 * 32;
	p->q_elems = qe;
	qsort_r(h_order, (size_t)count, sizeof(int32_t), cmp_k_desc, p->h_k);

	for (int i = 0; i < 8; ++i) HCK(hipEventCreate(&p->ev[i]));
	for (int i = 0; i < 2; ++i) HCK(hipEventCreate(&p->ev_graph[i]));
	p->events = true;

	ACK(p->slx, total * sizeof(int32_t));
	HCK(hipMemcpyAsync(p->slx, slx_idx, total * sizeof(int32_t), hipMemcpyHostToDevice, st));
	if (sly_idx && sly_idx != slx_idx) {
		ACK(p->sly, total * sizeof(int32_t));
		HCK(hipMemcpyAsync(p->sly, sly_idx, total * sizeof(int32_t), hipMemcpyHostToDevice, st));
	} else {
		p->sly = p->slx;
	}
	ACK(p->off, (size_t)(count + 1) * sizeof(int32_t));
	HCK(hipMemcpyAsync(p->off, sl_off, (size_t)(count + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
	if (sl_dists) {
		ACK(p->dist, total * sizeof(double));
		HCK(hipMemcpyAsync(p->dist, sl_dists, total * sizeof(double), hipMemcpyHostToDevice, st));
	}
	ACK(p->sl_of, total * sizeof(int32_t));
	KCK(hak_sl_of(p->off, count, p->sl_of, st));

	/* scatter index over the Y lists; weights index over the X lists */
	rc = build_occ(p, p->sly, p->nv_y, &p->occ_ptr, &p->occ);
	if (rc) goto fail;
	if (su->build_weights_index) {
		if (p->sly == p->slx && p->nv_x == p->nv_y) {
			p->occx_ptr = NULL; p->occx = NULL;
		} else {
			rc = build_occ(p, p->slx, p->nv_x, &p->occx_ptr, &p->occx);
			if (rc) goto fail;
		}
		ACK(p->wsum, (size_t)p->nv_x * sizeof(double));
		ACK(p->w, total * sizeof(double));
		p->weights_computed = true;
	}
	p->weights_none = !su->build_weights_index;

	if (su->build_pattern) {
		if (p->sly != p->slx || p->nv_x != p->nv_y) {
			ERR("the union pattern needs identical X and Y searchlights\n");
			rc = kHaErrorArg;
			goto fail;
		}
		hipEventRecord(p->ev[0], st);
		int32_t *slmin = NULL, *slmax = NULL, *colcnt = NULL;
		ACK(slmin, (size_t)count * sizeof(int32_t));
		ACK(slmax, (size_t)count * sizeof(int32_t));
		ACK(colcnt, (size_t)p->nv_x * sizeof(int32_t));
		ACK(p->indptr64, (size_t)(p->nv_x + 1) * sizeof(int64_t));
		KCK(hak_sl_minmax(p->slx, p->off, count, slmin, slmax, st));
		KCK(hak_pattern_count(p->slx, p->off, p->sl_of, p->occ_ptr, p->occ, slmin, slmax, p->nv_x, colcnt, st));
		KCK(hak_scan_i32_i64(colcnt, p->nv_x, p->indptr64, st));
		int64_t nnz = 0;
		HCK(hipMemcpyAsync(&nnz, p->indptr64 + p->nv_x, sizeof(int64_t), hipMemcpyDeviceToHost, st));
		HCK(hipStreamSynchronize(st));
		if (nnz >= ((int64_t)1 << 31)) {
			ERR("pattern has %lld nonzeros; int32 indptr cannot hold it\n", (long long)nnz);
			hipFree(slmin); hipFree(slmax); hipFree(colcnt);
			rc = kHaErrorArg;
			goto fail;
		}
		p->nnz = nnz;
		p->ncols = p->nv_x;
		{
			int32_t *hc = (int32_t *)malloc((size_t)p->nv_x * sizeof(int32_t));
			if (!hc) { hipFree(slmin); hipFree(slmax); hipFree(colcnt); rc = kHaErrorAlloc; goto fail; }
			if (hipMemcpy(hc, colcnt, (size_t)p->nv_x * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
				free(hc); hipFree(slmin); hipFree(slmax); hipFree(colcnt);
				rc = kHaErrorInternal;
				goto fail;
			}
			p->max_col = 0;
			for (int32_t c = 0; c < p->nv_x; ++c)
				if (hc[c] > p->max_col) p->max_col = hc[c];
			free(hc);
		}
		ACK(p->indptr, (size_t)(p->nv_x + 1) * sizeof(int32_t));
		ACK(p->indices, (size_t)nnz * sizeof(int32_t));
		ACK(p->data, (size_t)nnz * sizeof(double));
		KCK(hak_i64_to_i32(p->indptr64, p->nv_x + 1, p->indptr, st));
		KCK(hak_pattern_fill(p->slx, p->off, p->sl_of, p->occ_ptr, p->occ, slmin, slmax, p->nv_x, p->indptr64,
		                     p->indices, st));
		hipEventRecord(p->ev[1], st);
		HCK(hipStreamSynchronize(st));
		float ms = 0;
		hipEventElapsedTime(&ms, p->ev[0], p->ev[1]);
		p->stats.ms_pattern = ms;
		hipFree(slmin); hipFree(slmax); hipFree(colcnt);
		p->own_pattern = true;
	}

	if (su->index_only) {
		HCK(hipStreamSynchronize(st));
		free(h_qld); free(h_qoff); free(h_order);
		*out = p;
		return kHaSuccess;
	}
	/* local transform storage */
	ACK(p->qld, (size_t)count * sizeof(int32_t));
	ACK(p->qoff, (size_t)count * sizeof(int64_t));
	ACK(p->order, (size_t)count * sizeof(int32_t));
	HCK(hipMemcpyAsync(p->qld, h_qld, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(p->qoff, h_qoff, (size_t)count * sizeof(int64_t), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(p->order, h_order, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice, st));
	ACK(p->q, (size_t)p->q_elems * sizeof(float));
	ACK(p->flag, (size_t)count * sizeof(int32_t));
	ACK(p->steps, (size_t)count * sizeof(int32_t));
	ACK(p->queues, 16 * sizeof(int32_t));
	ACK(p->sigma_sum, (size_t)count * sizeof(double));
	ACK(p->var_sum, (size_t)count * sizeof(double));
	ACK(p->scale, (size_t)count * sizeof(double));
	p->slot_stride = (int64_t)p->kpmax * p->kpmax;
	/* Jacobi-32 and the cross-covariance stage: 2 workgroups per CU, 5 buffers;
	 * the polar iteration: HAK_POLAR_OCC per CU, 4 buffers (the work area serves both) */
	p->n_slots = 2 * g.n_cu;
	if (p->n_slots > count) p->n_slots = count;
	p->n_slots_polar = HAK_POLAR_OCC * g.n_cu;
	if (p->n_slots_polar > count) p->n_slots_polar = count;
	{
		const int64_t nbuf = (int64_t)p->n_slots * 5 > (int64_t)p->n_slots_polar * 4
		                         ? (int64_t)p->n_slots * 5 : (int64_t)p->n_slots_polar * 4;
		ACK(p->work, (size_t)nbuf * p->slot_stride * sizeof(float));
		p->work_floats = nbuf * p->slot_stride;
	}
	ACK(p->sl_par, (size_t)count * 3 * sizeof(float));
	p->stride64 = (int64_t)p->kpmax * p->kpmax;
	p->n_slots64 = g.n_cu;
	if (p->n_slots64 > count) p->n_slots64 = count;
	if (p->n_slots64 > 128) p->n_slots64 = 128;
	ACK(p->work64, (size_t)p->n_slots64 * 4 * p->stride64 * sizeof(double));
	ACK(p->iota, (size_t)p->kpmax * sizeof(int32_t));
	{
		int32_t *h_iota = (int32_t *)malloc((size_t)p->kpmax * sizeof(int32_t));
		if (!h_iota) { rc = kHaErrorAlloc; goto fail; }
		for (int32_t i = 0; i < p->kpmax; ++i) h_iota[i] = i;
		const hipError_t e_ = hipMemcpy(p->iota, h_iota, (size_t)p->kpmax * sizeof(int32_t), hipMemcpyHostToDevice);
		free(h_iota);
		if (e_ != hipSuccess) { rc = kHaErrorInternal; goto fail; }
	}
	ACK(p->X, (size_t)p->ld * p->nv_x * sizeof(float));
	ACK(p->Y, (size_t)p->ld * p->nv_y * sizeof(float));
	ACK(p->Xs, (size_t)2 * p->ld * p->nv_x * sizeof(uint16_t));
	ACK(p->Ys, (size_t)2 * p->ld * p->nv_y * sizeof(uint16_t));
	ACK(p->X64, (size_t)nt * p->nv_x * sizeof(double));
	ACK(p->Y64, (size_t)nt * p->nv_y * sizeof(double));
	HCK(hipStreamSynchronize(st));
	free(h_qld); free(h_qoff); free(h_order);
	*out = p;
	return kHaSuccess;
fail:
	hipStreamSynchronize(st);
	free(h_qld); free(h_qoff); free(h_order);
	plan_free(p);
	return rc;
}

int ha_hip_plan_create(int32_t nt, int32_t nv, const int32_t *sl_indices, const int32_t *sl_offsets,
                       const double *sl_dists, int32_t count, double radius, HaHipPlan **plan_out) {
	HaPlanSetup su = {.build_pattern = true, .build_weights_index = true, .nv_x = nv, .nv_y = nv};
	return ha_plan_create_ex(nt, &su, sl_indices, NULL, sl_offsets, sl_dists, count, radius, plan_out);
}

/* A caller's CSC pattern must be well formed before the scatter trusts it
 * (data + indptr[c] + position is written on the GPU): indptr starts at 0, never
 * decreases and ends within nnz; each column's rows are strictly increasing and
 * below nrows. O(ncols + nnz) on the host. */
int ha_check_csc(int32_t ncols, int32_t nrows, const int32_t *indptr, const int32_t *indices, int64_t nnz) {
	if (ncols < 0 || nnz < 0 || !indptr || (nnz > 0 && !indices)) {
		ERR("CSC pattern: missing arrays or negative sizes\n");
		return kHaErrorArg;
	}
	if (indptr[0] != 0) { ERR("CSC pattern: indptr[0] = %d, not 0\n", indptr[0]); return kHaErrorArg; }
	for (int32_t c = 0; c < ncols; ++c) {
		const int32_t b = indptr[c], e = indptr[c + 1];
		if (e < b || (int64_t)e > nnz) {
			ERR("CSC pattern: column %d spans [%d, %d) outside [0, nnz = %lld)\n", c, b, e, (long long)nnz);
			return kHaErrorArg;
		}
		for (int32_t q = b; q < e; ++q) {
			const int32_t r = indices[q];
			if (r < 0 || r >= nrows || (q > b && r <= indices[q - 1])) {
				ERR("CSC pattern: column %d has row %d at entry %d (rows must be increasing and below %d)\n",
				    c, r, q, nrows);
				return kHaErrorArg;
			}
		}
	}
	return kHaSuccess;
}

/* Drops the captured pass: called wherever a buffer or a host-side setting the
 * pass was captured with is replaced, so that no replay reads freed memory. */
static void graph_drop(HaHipPlan *p) {
	if (p->gexec) { hipGraphExecDestroy(p->gexec); p->gexec = NULL; }
}

int ha_plan_set_pattern(HaHipPlan *p, int32_t ncols, const int32_t *indptr, const int32_t *indices, int64_t nnz,
                        const double *data) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	graph_drop(p);
	if (ncols > p->nv_y) { ERR("pattern has more columns than Y vertices\n"); return kHaErrorArg; }
	rc = ha_check_csc(ncols, p->nv_x, indptr, indices, nnz);
	if (rc) return rc;
	if (p->indptr) { hipFree(p->indptr); p->indptr = NULL; }
	if (p->indptr64) { hipFree(p->indptr64); p->indptr64 = NULL; }
	if (p->indices) { hipFree(p->indices); p->indices = NULL; }
	if (p->data) { hipFree(p->data); p->data = NULL; }
	p->ncols = ncols;
	p->nnz = nnz;
	int64_t *h64 = (int64_t *)malloc((size_t)(ncols + 1) * sizeof(int64_t));
	if (!h64) return kHaErrorAlloc;
	p->max_col = 0;
	for (int32_t c = 0; c <= ncols; ++c) {
		h64[c] = indptr[c];
		if (c < ncols && indptr[c + 1] - indptr[c] > p->max_col) p->max_col = indptr[c + 1] - indptr[c];
	}
	ACK(p->indptr, (size_t)(ncols + 1) * sizeof(int32_t));
	ACK(p->indptr64, (size_t)(ncols + 1) * sizeof(int64_t));
	ACK(p->indices, (size_t)nnz * sizeof(int32_t));
	ACK(p->data, (size_t)nnz * sizeof(double));
	HCK(hipMemcpyAsync(p->indptr, indptr, (size_t)(ncols + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(p->indptr64, h64, (size_t)(ncols + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
	if (nnz > 0) {
		HCK(hipMemcpyAsync(p->indices, indices, (size_t)nnz * sizeof(int32_t), hipMemcpyHostToDevice, st));
		if (data)
			HCK(hipMemcpyAsync(p->data, data, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, st));
		else
			KCK(hak_fill_f64(p->data, nnz, 0.0, st));
	}
	HCK(hipStreamSynchronize(st));
	free(h64);
	return kHaSuccess;
fail:
	free(h64);
	return rc;
}

int ha_plan_set_weights(HaHipPlan *p, const double *w_flat, bool use_computed) {
	int rc = kHaSuccess;
	graph_drop(p);
	if (use_computed) {
		if (!p->weights_computed) { ERR("plan has no weight index\n"); return kHaErrorArg; }
		p->weights_none = false;
		return kHaSuccess;
	}
	p->weights_computed = false;
	if (!w_flat) { p->weights_none = true; return kHaSuccess; }
	p->weights_none = false;
	if (!p->w) ACK(p->w, (size_t)p->total * sizeof(double));
	HCK(hipMemcpy(p->w, w_flat, (size_t)p->total * sizeof(double), hipMemcpyHostToDevice));
	return kHaSuccess;
fail:
	return rc;
}

static int upload_matrix(HaHipPlan *p, const double *M, int32_t nv, bool col_major, float *d32, double *d64,
                         uint16_t *split) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	size_t bytes = (size_t)p->nt * nv * sizeof(double);
	if (!p->stage) {
		int32_t nvm = p->nv_x > p->nv_y ? p->nv_x : p->nv_y;
		ACK(p->stage, (size_t)p->nt * nvm * sizeof(double));
	}
	HCK(hipMemcpyAsync(p->stage, M, bytes, hipMemcpyHostToDevice, st));
	KCK(hak_convert_data(p->stage, p->nt, nv, col_major ? 1 : 0, d32, p->ld, d64, split,
	                     split + (size_t)p->ld * nv, st));
	HCK(hipStreamSynchronize(st));
	return kHaSuccess;
fail:
	return rc;
}

int ha_plan_load_xy(HaHipPlan *p, const double *X, const double *Y, bool col_major) {
	if (!p || !X || !Y) return kHaErrorArg;
	int rc = upload_matrix(p, X, p->nv_x, col_major, p->X, p->X64, p->Xs);
	if (rc) return rc;
	rc = upload_matrix(p, Y, p->nv_y, col_major, p->Y, p->Y64, p->Ys);
	if (rc) return rc;
	p->loaded = true;
	return kHaSuccess;
}

int ha_hip_plan_load(HaHipPlan *plan, const double *X, const double *Y, bool col_major) {
	/* the entry points check their data before they make a plan; a plan's own load checks here */
	if (!plan || !X || !Y) return kHaErrorArg;
	if (ha_check_finite(X, (int64_t)plan->nt * plan->nv_x, "X") ||
	    ha_check_finite(Y, (int64_t)plan->nt * plan->nv_y, "Y"))
		return kHaErrorArg;
	return ha_plan_load_xy(plan, X, Y, col_major);
}

static HakLocalArgs local_args(HaHipPlan *p) {
	HakLocalArgs a;
	memset(&a, 0, sizeof(a));
	a.X = p->X; a.Y = p->Y; a.X64 = p->X64; a.Y64 = p->Y64;
	a.Xh = p->Xs; a.Xl = p->Xs + (size_t)p->ld * p->nv_x;
	a.Yh = p->Ys; a.Yl = p->Ys + (size_t)p->ld * p->nv_y;
	a.nt = p->nt; a.ld = p->ld;
	a.slx_idx = p->slx; a.sly_idx = p->sly; a.sl_off = p->off;
	a.order = p->order; a.n_order = p->count;
	a.q = p->q; a.qoff = p->qoff; a.qld = p->qld;
	a.flag = p->flag; a.steps = p->steps;
	a.work = p->work; a.slot_stride = p->slot_stride; a.n_slots = p->n_slots;
	a.sl_par = p->sl_par;
	a.sched = g.d_sched;
	a.sched_len = g.d_sched_len;
	a.n_buckets = HA_SCHED_BUCKETS; a.max_sched = HA_SCHED_STEPS;
	a.bucket_per_decade = HA_SCHED_PER_DECADE;
	a.max_steps = g.opt.max_polar_steps;
	a.plain_ell = (float)g.opt.plain_ell;
	a.tol = (float)g.opt.polar_tol;
	a.tile = pick_tile(p->kpmax);
	a.fallback_enabled = g.opt.fp64_fallback;
	a.rank_check_nt = g.opt.fp64_fallback;
	a.mlp_margin_scale = g.opt.learned_margin_scale;
	a.mlp_margin_ell = g.opt.learned_margin_ell;
	a.team_count = p->queues + 12;  /* zeroed with the queues at the start of a pass */
	return a;
}

static int resolve_method(const HaHipPlan *p, int method) {
	if (method < 0) method = g.opt.method;
	if (method == kHaHipMethodAuto) {
		if (p->kmax <= 32) return kHaHipMethodJacobi64;
		return g.has_model ? kHaHipMethodPolarLearned : kHaHipMethodPolar;
	}
	if (method == kHaHipMethodPolarLearned && !g.has_model) {
		ERR("learned polar scaling requested but no model is loaded\n");
		return -1;
	}
	return method;
}

static void stats_update(HaHipPlan *p, int method) {
	int32_t *st = (int32_t *)malloc((size_t)p->count * sizeof(int32_t));
	int32_t *fl = (int32_t *)malloc((size_t)p->count * sizeof(int32_t));
	if (!st || !fl) { free(st); free(fl); return; }
	hipMemcpy(st, p->steps, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost);
	hipMemcpy(fl, p->flag, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost);
	HaHipStats *S = &p->stats;
	memset(S->steps_hist, 0, sizeof(S->steps_hist));
	S->n_fallback = 0; S->n_unconverged = 0; S->steps_total = 0;
	double flops = 0;
	for (int32_t s = 0; s < p->count; ++s) {
		double k = p->h_k[s];
		int32_t v = st[s];
		if (fl[s]) S->n_fallback++;
		if (method == kHaHipMethodPolar || method == kHaHipMethodPolarLearned) {
			if (v >= 1000) v = 0;      /* solved by the FP64 path */
			if (v < 0) { S->n_unconverged++; v = -v; }
			S->steps_total += v;
			S->steps_hist[v < 31 ? v : 31]++;
			/* per step: X^T X and G^2 are symmetric (k^3 each, only tiles I <= J
			 * are computed), X P is a full product (2 k^3) */
			flops += v * 4.0 * k * k * k;
		} else {
			if (v >= 1000) v -= 1000;
			if (v < 0) { S->n_unconverged++; v = -v; }
			S->steps_total += v;
			S->steps_hist[v < 31 ? v : 31]++;
		}
	}
	S->gemm_flops = flops;
	S->nnz = p->nnz;
	S->n_orth_checked = 0; S->n_orth_flagged = 0; S->orth_max = 0;
	if (p->orth_ran) {
		double *o = (double *)malloc((size_t)p->count * sizeof(double));
		if (o && hipMemcpy(o, p->orth, (size_t)p->count * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess) {
			for (int32_t s = 0; s < p->count; ++s) {
				if (o[s] < 0) continue;  /* already on the FP64 path when the check ran */
				S->n_orth_checked++;
				if (!(o[s] <= g.opt.orth_tol)) S->n_orth_flagged++;
				if (o[s] > S->orth_max || !isfinite(o[s])) S->orth_max = o[s];
			}
		}
		free(o);
	}
	int32_t tc[2] = {0, 0};
	if (hipMemcpy(tc, p->queues + 12, sizeof(tc), hipMemcpyDeviceToHost) == hipSuccess) {
		S->n_team_giveup = tc[0];
		S->n_team_handback = tc[1];
	}
	free(st);
	free(fl);
}

/* ---- polar tiles per kp bucket: selection and autotuning ---- */

/* iteration workgroups with 128 x 128 tiles (each needs 4 slot buffers, as with 64) */
static int32_t slots128(const HaHipPlan *p) {
	int32_t n = HAK_POLAR_OCC128 * g.n_cu;
	return n < p->n_slots_polar ? n : p->n_slots_polar;
}

/* Teams of HAK_GRP workgroups for the 128-tile iteration: as many as fit
 * HAK_POLAR_OCC128 workgroups per CU and the work area (4 buffers per team), a
 * multiple of 8 (one team per XCD per group of 8); 0 when fewer than 8. */
static int32_t team_count(const HaHipPlan *p) {
	const int64_t nbuf = (int64_t)p->n_slots * 5 > (int64_t)p->n_slots_polar * 4
	                         ? (int64_t)p->n_slots * 5 : (int64_t)p->n_slots_polar * 4;
	int64_t n = (int64_t)HAK_POLAR_OCC128 * g.n_cu / HAK_GRP;
	if (n > nbuf / 4) n = nbuf / 4;
	n -= n % 8;
	return n >= 8 ? (int32_t)n : 0;
}

/* The 128-tile iteration over b's searchlights (b.order, b.n_order): the team
 * kernel first when teams are on, then the one-workgroup kernel over what the
 * teams left (nothing, unless a team barrier gave up). Needs steps[] zeroed. */
/* Allocates the team state (outside any graph capture). */
static int ensure_teams(HaHipPlan *p) {
	int rc = kHaSuccess;
	if (!g.opt.team || p->grp || p->n_teams < 0) return kHaSuccess;
	graph_drop(p);
	p->n_teams = team_count(p);
	if (p->n_teams == 0) { p->n_teams = -1; return kHaSuccess; }
	ACK(p->grp, (size_t)p->n_teams * HAK_GRP_STATE_BYTES);
fail:
	return rc;
}

static int launch_polar128(HaHipPlan *p, HakLocalArgs b, bool teams, hipStream_t st) {
	int rc = kHaSuccess;
	b.tile = 128;
	b.skip_done = 0;
	if (teams && g.opt.team && b.n_order > 0) {
		if (p->grp) {
			KCK(hak_fill_i32((int32_t *)p->grp, (int64_t)p->n_teams * (HAK_GRP_STATE_BYTES / 4), 0, st));
			HakLocalArgs t = b;
			t.grp = p->grp;
			t.grp_spin = g.opt.team_spin;
			t.queue = p->queues + 6;
			KCK(hak_polar_grp(&t, p->n_teams, st));
			b.skip_done = 1;
		}
	}
	b.n_slots = slots128(p);
	KCK(hak_polar(&b, st));
fail:
	return rc;
}

static int kp_bucket(int32_t kp) {
	static const int32_t edges[HA_TILE_BUCKETS - 1] = {64, 128, 256, 384, 512};
	int b = 0;
	while (b < HA_TILE_BUCKETS - 1 && kp > edges[b]) ++b;
	return b;
}

/* Rebuilds the per-tile searchlight lists (k-descending order kept) from bucket_tile. */
static int build_tile_orders(HaHipPlan *p) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	int32_t *h = (int32_t *)malloc((size_t)p->count * HA_TILE_LISTS * sizeof(int32_t));
	int32_t *ord = (int32_t *)malloc((size_t)p->count * sizeof(int32_t));
	if (!h || !ord) { free(h); free(ord); return kHaErrorAlloc; }
	HCK(hipMemcpy(ord, p->order, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost));
	int32_t n[HA_TILE_LISTS] = {0, 0, 0, 0};
	for (int32_t i = 0; i < p->count; ++i) {
		const int32_t s = ord[i];
		const int b = kp_bucket((p->h_k[s] + 31) / 32 * 32);
		/* rank deficient (stage 1 starts them from the smallest bucket): FP32 products */
		const int t = (g.opt.fp64_fallback && HAK_NEAR_DEFICIENT(p->h_k[s], p->nt))
		                  ? 2
		                  : (p->bucket_tile[b] == 128 ? (p->bucket_team[b] ? 1 : 3) : 0);
		h[(size_t)t * p->count + n[t]++] = s;
	}
	p->tuned_fb = g.opt.fp64_fallback;
	for (int t = 0; t < HA_TILE_LISTS; ++t) {
		if (!p->tile_order[t]) ACK(p->tile_order[t], (size_t)p->count * sizeof(int32_t));
		if (n[t]) HCK(hipMemcpyAsync(p->tile_order[t], h + (size_t)t * p->count, (size_t)n[t] * sizeof(int32_t),
		                             hipMemcpyHostToDevice, st));
		p->tile_count[t] = n[t];
	}
	HCK(hipStreamSynchronize(st));
fail:
	free(h);
	free(ord);
	return rc;
}

/* Time of the iteration stage for the searchlights `sel` (device list, n of
 * them) with the given tile (128: in teams or one workgroup each); the
 * cross-covariance stage is rerun first because the iteration overwrites A
 * with Q. */
static int time_tile(HaHipPlan *p, const int32_t *sel, int32_t n, int tile, bool teams, float *ms) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	HakLocalArgs a = local_args(p);
	a.order = sel;
	a.n_order = n;
	a.tile = tile;
	KCK(hak_fill_i32(p->queues, 16, 0, st));
	KCK(hak_fill_i32(p->flag, p->count, 0, st));
	KCK(hak_fill_i32(p->steps, p->count, 0, st));
	a.queue = p->queues + 3;
	KCK(hak_crosscov(&a, p->n_slots, st));
	a.queue = p->queues + 5;
	HCK(hipEventRecord(p->ev[6], st));
	if (tile == 64) {
		a.n_slots = p->n_slots_polar;
		KCK(hak_polar(&a, st));
	} else {
		rc = launch_polar128(p, a, teams, st);
		if (rc) goto fail;
	}
	HCK(hipEventRecord(p->ev[7], st));
	HCK(hipStreamSynchronize(st));
	HCK(hipEventElapsedTime(ms, p->ev[6], p->ev[7]));
fail:
	return rc;
}

/* Picks the iteration launch of every kp bucket: 64 x 64 tiles, 128 x 128
 * tiles with one workgroup per searchlight, or 128 x 128 tiles in teams of
 * HAK_GRP workgroups. With the tile option set, that tile everywhere (128 in
 * teams when the team option is on); otherwise buckets up to kp 64 and buckets
 * with fewer searchlights than the GPU has iteration slots take 64, and the
 * others are timed with each candidate on an evenly spaced sample of up to two
 * slot-fuls of their searchlights (plain schedule), keeping the fastest. The
 * candidates' times stay in the plan (ha_hip_plan_tuning). Needs loaded data. */
static int tune_tiles(HaHipPlan *p) {
	int rc = kHaSuccess;
	int32_t *ord = NULL, *sel = NULL, *d_sel = NULL;
	rc = ensure_teams(p);
	if (rc) return rc;
	graph_drop(p);
	const bool teams = g.opt.team && p->grp;
	for (int b = 0; b < HA_TILE_BUCKETS; ++b) {
		p->bucket_tile[b] = g.opt.tile ? g.opt.tile : 64;
		p->bucket_team[b] = p->bucket_tile[b] == 128 && teams;
		p->tune_sample[b] = 0;
		for (int c = 0; c < 3; ++c) p->tune_ms[b][c] = 0.f;
	}
	if (!g.opt.tile) {
		ord = (int32_t *)malloc((size_t)p->count * sizeof(int32_t));
		sel = (int32_t *)malloc((size_t)p->count * sizeof(int32_t));
		if (!ord || !sel) { rc = kHaErrorAlloc; goto fail; }
		HCK(hipMemcpy(ord, p->order, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost));
		ACK(d_sel, (size_t)p->count * sizeof(int32_t));
		for (int b = 1; b < HA_TILE_BUCKETS; ++b) {
			int32_t nb = 0;
			for (int32_t i = 0; i < p->count; ++i)
				if (kp_bucket((p->h_k[ord[i]] + 31) / 32 * 32) == b) sel[nb++] = ord[i];
			if (nb < p->n_slots_polar) continue;
			const int32_t want = 2 * p->n_slots_polar;
			int32_t n = 0;
			for (int32_t j = 0; j < want && j < nb; ++j) sel[n++] = sel[(int64_t)j * nb / (want < nb ? want : nb)];
			HCK(hipMemcpy(d_sel, sel, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
			float w = 0, t[3] = {0, 0, 0};
			rc = time_tile(p, d_sel, n, 64, false, &w);      /* warm-up */
			if (!rc) rc = time_tile(p, d_sel, n, 64, false, &t[0]);
			if (!rc) rc = time_tile(p, d_sel, n, 128, false, &t[1]);
			if (!rc && teams) rc = time_tile(p, d_sel, n, 128, true, &t[2]);
			if (rc) goto fail;
			int best = 0;
			for (int c = 1; c < (teams ? 3 : 2); ++c)
				if (t[c] < t[best]) best = c;
			p->bucket_tile[b] = best ? 128 : 64;
			p->bucket_team[b] = best == 2;
			p->tune_sample[b] = n;
			for (int c = 0; c < 3; ++c) p->tune_ms[b][c] = t[c];
			if (g.opt.verbose)
				fprintf(stderr, "ha_hip: tile bucket %d (%d searchlights, sample %d): 64 -> %.2f ms, "
				        "128 -> %.2f ms, 128 in teams -> %.2f ms\n", b, nb, n, t[0], t[1], t[2]);
		}
	}
	rc = build_tile_orders(p);
	if (!rc) {
		p->tuned = true;
		p->tuned_opt = g.opt.tile;
		p->tuned_team = g.opt.team;
	}
fail:
	free(ord);
	free(sel);
	if (d_sel) hipFree(d_sel);
	return rc;
}

int ha_hip_plan_tiles(HaHipPlan *p, int32_t bucket_tile[6]) {
	if (!p || !bucket_tile) return kHaErrorArg;
	if (!p->loaded) { ERR("plan has no data loaded\n"); return kHaErrorArg; }
	if (!p->tuned || p->tuned_opt != g.opt.tile || p->tuned_team != g.opt.team ||
	    p->tuned_fb != g.opt.fp64_fallback) {
		int rc = tune_tiles(p);
		if (rc) return rc;
	}
	memcpy(bucket_tile, p->bucket_tile, sizeof(p->bucket_tile));
	return kHaSuccess;
}

int ha_hip_plan_tuning(HaHipPlan *p, HaHipTuning *out) {
	if (!p || !out) return kHaErrorArg;
	int32_t bt[HA_TILE_BUCKETS];
	const int rc = ha_hip_plan_tiles(p, bt);
	if (rc) return rc;
	memset(out, 0, sizeof(*out));
	for (int b = 0; b < HA_TILE_BUCKETS; ++b) {
		out->tile[b] = p->bucket_tile[b];
		out->team[b] = p->bucket_team[b];
		out->sample[b] = p->tune_sample[b];
		for (int c = 0; c < 3; ++c) out->ms[b][c] = p->tune_ms[b][c];
	}
	for (int32_t i = 0; i < p->count; ++i) ++out->searchlights[kp_bucket((p->h_k[i] + 31) / 32 * 32)];
	return kHaSuccess;
}

/* ---- one pass ---- */

/* The pass's orthogonality check: on for polar passes with orth_tol > 0 when
 * the plan's searchlights fit the probe kernel's LDS. */
static bool orth_check_on(const HaHipPlan *p) {
	return g.opt.orth_tol > 0 && p->orth && p->kpmax <= HAK_ORTH_PROBE_MAX_KP;
}
static int32_t orth_workgroups(const HaHipPlan *p) {
	const int32_t n = 8 * g.n_cu;
	return p->count < n ? p->count : n;
}

/* The LU-sign kernel's slots in the work area (free once the iteration is
 * done): a copy of Q and the 32 x kpmax U12 block per workgroup. */
static int64_t lu_slot_floats(const HaHipPlan *p) {
	return ((int64_t)p->kpmax * p->kpmax + 32 * (int64_t)p->kpmax + 63) / 64 * 64;
}
static int32_t lu_workgroups(const HaHipPlan *p) {
	int64_t n = 2 * (int64_t)g.n_cu;
	if (n > p->count) n = p->count;
	if (n > p->work_floats / lu_slot_floats(p)) n = p->work_floats / lu_slot_floats(p);
	return (int32_t)n;
}
static bool lu_kernel_on(const HaHipPlan *p) {
	return p->det_neg && p->kpmax <= HAK_LU_KPMAX && lu_workgroups(p) > 0;
}

/* Enqueues the whole pass on st: no allocation, no host synchronisation, so
 * it can be captured into a graph. timed: records the stage events. */
static int enqueue_pass(HaHipPlan *p, HakLocalArgs a, int method, bool isReflection, bool isScaling,
                        bool zero_data, hipStream_t st, bool timed) {
	int rc = kHaSuccess;
	if (timed) HCK(hipEventRecord(p->ev[0], st));
	KCK(hak_fill_i32(p->queues, 16, 0, st));
	KCK(hak_fill_i32(p->flag, p->count, 0, st));
	KCK(hak_fill_i32(p->steps, p->count, 0, st));
	if (zero_data) KCK(hak_fill_f64(p->data, p->nnz, 0.0, st));
	/* weights (ha_searchlight.c:970-1001) */
	if (p->weights_computed && !p->weights_none) {
		const int64_t *optr = p->occx_ptr ? p->occx_ptr : p->occ_ptr;
		const int32_t *occ = p->occx ? p->occx : p->occ;
		KCK(hak_weights(p->slx, p->dist, p->radius, p->total, optr, occ, p->nv_x, p->wsum, p->w, st));
	}
	if (timed) HCK(hipEventRecord(p->ev[1], st));
	/* local transforms; without reflections the polar path fixes det(Q) < 0
	 * afterwards (k_detfix), the FP32 Jacobi path leaves it to the FP64 one */
	const bool det_fix = !isReflection;
	const bool polar_m = method == kHaHipMethodPolar || method == kHaHipMethodPolarLearned;
	if (method == kHaHipMethodJacobi64 || (det_fix && !polar_m)) {
		a.flag = NULL;
		a.queue = p->queues + 1;
		KCK(hak_jacobi64(&a, det_fix ? 1 : 0, p->work64, p->stride64, p->n_slots64, p->count, st));
		a.flag = p->flag;
	} else if (method == kHaHipMethodJacobi32) {
		a.queue = p->queues + 0;
		KCK(hak_jacobi32(&a, st));
	} else {
		a.learned_scale = (method == kHaHipMethodPolarLearned) ? 1 : 0;
		a.queue = p->queues + 3;
		{
			HakLocalArgs c = a;
			/* cross-covariance tile: 128 x 128 halves the panel re-reads of large
			 * searchlights (C2: 134 -> 84 ms per pass) */
			c.tile = p->kpmax >= 256 ? 128 : 64;
			KCK(hak_crosscov(&c, p->n_slots, st));
		}
		if (timed) HCK(hipEventRecord(p->ev[6], st));
		/* one persistent launch per tile over its searchlights (k descending), and
		 * one with FP32 products over the rank-deficient ones */
		for (int t = 0; t < HA_TILE_LISTS; ++t) {
			if (!p->tile_count[t]) continue;
			HakLocalArgs b = a;
			b.tile = (t == 1 || t == 3) ? 128 : 64;
			b.prec_f32 = t == 2;
			b.order = p->tile_order[t];
			b.n_order = p->tile_count[t];
			b.queue = p->queues + (t == 3 ? 9 : t == 2 ? 7 : 4 + t);
			if (t == 1 || t == 3) {
				rc = launch_polar128(p, b, t == 1, st);
				if (rc) goto fail;
			} else {
				b.n_slots = p->n_slots_polar;
				b.skip_done = 0;
				if (t == 2 && g.opt.team && p->grp) {
					/* the rank-deficient searchlights are few and long: teams
					 * with FP32 products first, then one workgroup each for
					 * what a team left */
					KCK(hak_fill_i32((int32_t *)p->grp, (int64_t)p->n_teams * (HAK_GRP_STATE_BYTES / 4), 0, st));
					HakLocalArgs tt = b;
					tt.grp = p->grp;
					tt.grp_spin = g.opt.team_spin;
					tt.queue = p->queues + 8;
					KCK(hak_polar_grp(&tt, p->n_teams, st));
					b.skip_done = 1;
				}
				/* one workgroup per CU at most for them, so that no two share
				 * one (which made the launch's time vary by a third) */
				if (t == 2 && b.n_slots > g.n_cu) b.n_slots = g.n_cu;
				KCK(hak_polar(&b, st));
			}
		}
	}
	if (det_fix && polar_m) {
		HakLocalArgs b = a;
		if (lu_kernel_on(p)) {
			/* the signs first, by the LU kernel, in the idle iteration slots */
			HakLocalArgs l = a;
			l.queue = p->queues + 10;
			KCK(hak_lu_sign(&l, p->work, lu_slot_floats(p), p->kpmax, p->det_neg, lu_workgroups(p), st));
			b.det_neg_in = p->det_neg;
		}
		b.queue = p->queues + 0;
		KCK(hak_detfix(&b, p->work64, p->stride64, p->iota, p->n_slots64, st));
	}
	/* FP64 orthogonality of every local transform the iteration wrote; those
	 * above orth_tol are flagged for the FP64 path below */
	if (timed) HCK(hipEventRecord(p->ev[7], st));
	if (polar_m && orth_check_on(p))
		KCK(hak_orth(p->q, p->qoff, p->qld, p->off, NULL, p->count, p->kpmax, 0, p->orth, p->flag, g.opt.orth_tol,
		             orth_workgroups(p), st));
	if (timed) HCK(hipEventRecord(p->ev[2], st));
	if (method != kHaHipMethodJacobi64 && (polar_m || !det_fix) && g.opt.fp64_fallback) {
		HakLocalArgs b = a;
		b.queue = p->queues + 2;
		KCK(hak_jacobi64(&b, det_fix ? 1 : 0, p->work64, p->stride64, p->n_slots64, p->count, st));
	}
	if (timed) HCK(hipEventRecord(p->ev[3], st));
	if (isScaling) {
		KCK(hak_var_sum(p->X64, p->nt, p->slx, p->off, p->count, p->var_sum, st));
		KCK(hak_scale(p->sigma_sum, p->var_sum, p->nt, p->count, p->scale, st));
	}
	if (timed) HCK(hipEventRecord(p->ev[4], st));
	KCK(hak_scatter(p->slx, p->off, p->sl_of, p->occ_ptr, p->occ, p->ncols, p->indptr64, p->indices, p->data,
	                p->q, p->qoff, p->qld, isScaling ? p->scale : NULL, p->weights_none ? NULL : p->w, 0,
	                p->count, p->max_col, st));
	if (timed) HCK(hipEventRecord(p->ev[5], st));
fail:
	return rc;
}

/* Debug views of the polar path (HA_HIP_TRACE=n: residuals of the first n
 * searchlights; HA_HIP_CLOCK: per-phase busy time of the slots). Eager only. */
static void debug_report(HaHipPlan *p, const HakLocalArgs *a, uint64_t *clk, int32_t nclk) {
	hipStream_t st = g.stream;
	if (clk) {
		uint64_t *h = (uint64_t *)calloc((size_t)nclk * 4, sizeof(uint64_t));
		if (h && hipMemcpy(h, clk, (size_t)nclk * 4 * sizeof(uint64_t), hipMemcpyDeviceToHost) == hipSuccess) {
			double sum[4] = {0, 0, 0, 0};
			for (int i = 0; i < nclk; ++i)
				for (int j = 0; j < 4; ++j) sum[j] += (double)h[4 * i + j];
			fprintf(stderr, "ha_hip clock: per slot ms gather+crosscov %.1f, scale+features %.1f "
			        "(%d slots), iteration %.1f (%d slots); searchlights %.0f\n",
			        sum[0] / p->n_slots * 1e-5, sum[1] / p->n_slots * 1e-5, p->n_slots,
			        sum[2] / p->n_slots_polar * 1e-5, p->n_slots_polar, sum[3]);
		}
		free(h);
	}
	if (a->trace) {
		int ns = atoi(getenv("HA_HIP_TRACE"));
		if (ns > p->count) ns = p->count;
		float *h = (float *)malloc((size_t)(ns > 0 ? ns : 1) * a->max_steps * sizeof(float));
		if (h && hipMemcpyAsync(h, p->trace, (size_t)ns * a->max_steps * sizeof(float), hipMemcpyDeviceToHost,
		                        st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess) {
			for (int i = 0; i < ns; ++i) {
				fprintf(stderr, "ha_hip trace s=%d k=%d:", i, p->h_k[i]);
				for (int t = 0; t < a->max_steps; ++t) fprintf(stderr, " %.3g", h[(size_t)i * a->max_steps + t]);
				fprintf(stderr, "\n");
			}
		}
		free(h);
	}
}

/* More FP64 slots (k_jacobi64 fits four workgroups per CU) for passes that run
 * every searchlight through the FP64 path; at most 8 GiB of workspace. Called
 * before any capture: a captured pass holding the old workspace is dropped. An
 * allocation that fails leaves the slots as they were. */
static int ensure_work64(HaHipPlan *p, int32_t want) {
	if (want > p->count) want = p->count;
	const size_t per = (size_t)4 * p->stride64 * sizeof(double);
	const int64_t cap = per ? (int64_t)(((size_t)8 << 30) / per) : want;
	if (want > cap) want = (int32_t)cap;
	if (want <= p->n_slots64) return kHaSuccess;
	double *w = NULL;
	if (hipMalloc((void **)&w, (size_t)want * per) != hipSuccess) {
		(void)hipGetLastError();
		return kHaSuccess;
	}
	if (hipStreamSynchronize(g.stream) != hipSuccess) {
		hipFree(w);
		return kHaErrorInternal;
	}
	graph_drop(p);
	hipFree(p->work64);
	p->work64 = w;
	p->n_slots64 = want;
	return kHaSuccess;
}

/* Stage times of the last eager pass from the events enqueue_pass recorded. */
static void stage_times(HaHipPlan *p, bool is_polar) {
	float t[5] = {0, 0, 0, 0, 0};
	for (int i = 0; i < 5; ++i) hipEventElapsedTime(&t[i], p->ev[i], p->ev[i + 1]);
	float tot = 0, tc = 0, ti = t[1], tk = t[1], to = 0;
	hipEventElapsedTime(&tot, p->ev[0], p->ev[5]);
	if (is_polar) {
		hipEventElapsedTime(&tc, p->ev[1], p->ev[6]);
		hipEventElapsedTime(&ti, p->ev[6], p->ev[2]);
		hipEventElapsedTime(&tk, p->ev[6], p->ev[7]);  /* iteration (+ reflection fix) */
		hipEventElapsedTime(&to, p->ev[7], p->ev[2]);  /* orthogonality check */
	}
	p->stats.ms_weights = t[0];
	p->stats.ms_crosscov = tc;  /* polar path: gather + cross-covariance (+ features) stage */
	p->stats.ms_polar = ti;     /* local solver (the iteration stage on the polar path) */
	p->stats.ms_orth = to;
	p->stats.ms_fallback = t[2];
	p->stats.ms_scatter = t[4] + t[3];
	p->stats.ms_total = tot;
	p->stats.kernel_ms_polar_avg = tk;
	p->stats.polar_launches =
	    is_polar ? (p->tile_count[0] > 0) + (p->tile_count[1] > 0) + (p->tile_count[2] > 0) + (p->tile_count[3] > 0)
	             : 1;
	(void)hipGetLastError();
}

int ha_plan_run(HaHipPlan *p, bool isReflection, bool isScaling, bool zero_data, int method_in) {
	int rc = kHaSuccess;
	uint64_t *clk = NULL;
	hipGraph_t graph = NULL;
	bool capturing = false;
	if (!p || !p->loaded) { ERR("plan has no data loaded\n"); return kHaErrorArg; }
	if (!p->data || !p->indptr64) { ERR("plan has no CSC pattern\n"); return kHaErrorArg; }
	const int method = resolve_method(p, method_in);
	if (method < 0) return kHaErrorArg;
	hipStream_t st = g.stream;
	const bool is_polar = method == kHaHipMethodPolar || method == kHaHipMethodPolarLearned;
	if (is_polar) {
		rc = ensure_teams(p);
		if (rc) return rc;
	}
	if (!isReflection || method == kHaHipMethodJacobi64) {
		rc = ensure_work64(p, 4 * g.n_cu);
		if (rc) return rc;
	}
	if (!isReflection && is_polar && !p->det_neg) {
		graph_drop(p);
		ACK(p->det_neg, (size_t)p->count * sizeof(int32_t));
	}
	if (is_polar && (!p->tuned || p->tuned_opt != g.opt.tile || p->tuned_team != g.opt.team ||
	                 p->tuned_fb != g.opt.fp64_fallback)) {
		rc = tune_tiles(p);
		if (rc) return rc;
	}
	if (is_polar && g.opt.orth_tol > 0 && !p->orth) ACK(p->orth, (size_t)p->count * sizeof(double));
	HakLocalArgs a = local_args(p);
	if (isScaling) a.sigma_sum = p->sigma_sum;
	if (method == kHaHipMethodPolarLearned) {
		a.mlp = g.d_mlp;
		a.mlp_norm = g.d_mlp_norm;
		if (!p->pred) ACK(p->pred, (size_t)p->count * 2 * sizeof(float));
		a.pred_out = p->pred;
	}
	/* graph mode: the first pass of a plan runs eagerly (per-stage debug views
	 * and the stage events as usual), later ones are captured once and replayed
	 * while nothing they depend on changes */
	const bool debug_env = getenv("HA_HIP_TRACE") || getenv("HA_HIP_CLOCK");
	const bool use_graph = g.opt.use_graph != 0 && p->n_passes > 0 && !debug_env;
	const int32_t nclk = p->n_slots > p->n_slots_polar ? p->n_slots : p->n_slots_polar;
	if (!use_graph && is_polar) {
		if (getenv("HA_HIP_TRACE")) {
			if (!p->trace) ACK(p->trace, (size_t)p->count * a.max_steps * sizeof(float));
			KCK(hak_fill_i32((int32_t *)p->trace, (int64_t)p->count * a.max_steps, 0, st));
			a.trace = p->trace;
		}
		if (getenv("HA_HIP_CLOCK")) {
			ACK(clk, (size_t)nclk * 4 * sizeof(uint64_t));
			KCK(hak_fill_i32((int32_t *)clk, (int64_t)nclk * 8, 0, st));
			a.clk = clk;
		}
	}
	if (!use_graph) {
		rc = enqueue_pass(p, a, method, isReflection, isScaling, zero_data, st, true);
		if (rc) goto fail;
		HCK(hipStreamSynchronize(st));
		stage_times(p, is_polar);
		debug_report(p, &a, clk, nclk);
	} else {
		/* replay the captured pass when nothing it depends on changed: its
		 * arguments, the options, the method and the flags */
		const bool same = p->gexec && p->gkey.method == method && p->gkey.refl == isReflection &&
		                  p->gkey.scal == isScaling && p->gkey.zero == zero_data &&
		                  memcmp(&p->gkey.opt, &g.opt, sizeof(g.opt)) == 0 &&
		                  memcmp(&p->gkey.a, &a, sizeof(a)) == 0;
		if (!same) {
			if (p->gexec) { hipGraphExecDestroy(p->gexec); p->gexec = NULL; }
			/* the stage events are recorded inside the graph; a runtime that
			 * cannot capture them gets a graph without them */
			for (int with_events = 1; with_events >= 0 && !p->gexec; --with_events) {
				HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
				capturing = true;
				rc = enqueue_pass(p, a, method, isReflection, isScaling, zero_data, st, with_events != 0);
				capturing = false;
				hipError_t e = hipStreamEndCapture(st, &graph);
				if (e == hipSuccess && rc == kHaSuccess) {
					e = hipGraphInstantiate(&p->gexec, graph, NULL, NULL, 0);
					if (e != hipSuccess) p->gexec = NULL;
				}
				if (graph) { hipGraphDestroy(graph); graph = NULL; }
				(void)hipGetLastError();
				if (!p->gexec && !with_events) {
					ERR("graph capture failed (%s)\n", hipGetErrorString(e));
					if (!rc) rc = kHaErrorInternal;
					goto fail;
				}
				rc = kHaSuccess;
				p->gevents = with_events != 0;
			}
			memset(&p->gkey, 0, sizeof(p->gkey));
			p->gkey.a = a; p->gkey.opt = g.opt; p->gkey.method = method; p->gkey.refl = isReflection;
			p->gkey.scal = isScaling; p->gkey.zero = zero_data;
		}
		HCK(hipEventRecord(p->ev_graph[0], st));
		HCK(hipGraphLaunch(p->gexec, st));
		HCK(hipEventRecord(p->ev_graph[1], st));
		HCK(hipStreamSynchronize(st));
		float tot = 0;
		hipEventElapsedTime(&tot, p->ev_graph[0], p->ev_graph[1]);
		if (p->gevents) {
			stage_times(p, is_polar);
		} else {
			memset(&p->stats.ms_weights, 0, sizeof(double) * 5);  /* ms_weights .. ms_scatter: not split */
			p->stats.ms_orth = 0;
			p->stats.kernel_ms_polar_avg = 0;
			p->stats.polar_launches = 0;
		}
		p->stats.ms_total = tot;
	}
	++p->n_passes;
	p->stats.graph_replay = use_graph ? 1 : 0;
	p->orth_ran = is_polar && orth_check_on(p);
	stats_update(p, method);
	if (!g.opt.fp64_fallback && p->stats.n_fallback > 0 && method != kHaHipMethodJacobi64) {
		/* flagged searchlights hold no valid transform without the FP64 pass,
		 * and the scatter has added them into T */
		ERR("%lld searchlights need the FP64 path (unconverged, rank deficient or failed checks), "
		    "which the fp64_fallback option switches off\n", (long long)p->stats.n_fallback);
		rc = kHaErrorSvd;
	}
	if (g.opt.verbose)
		fprintf(stderr, "ha_hip: pass %.3f ms (weights %.3f, cross-cov %.3f, local %.3f, fp64 %.3f, scatter %.3f)%s; "
		        "fallback %lld\n", p->stats.ms_total, p->stats.ms_weights, p->stats.ms_crosscov, p->stats.ms_polar,
		        p->stats.ms_fallback, p->stats.ms_scatter, use_graph ? " [graph]" : "",
		        (long long)p->stats.n_fallback);
fail:
	if (capturing) {
		hipGraph_t g2 = NULL;
		hipStreamEndCapture(st, &g2);
		if (g2) hipGraphDestroy(g2);
	}
	if (graph) hipGraphDestroy(graph);
	if (clk) hipFree(clk);
	if (rc) hipStreamSynchronize(st);
	return rc;
}

int ha_hip_plan_execute(HaHipPlan *plan, bool isReflection, bool isScaling) {
	return ha_plan_run(plan, isReflection, isScaling, true, -1);
}

int ha_hip_plan_nnz(const HaHipPlan *p, int64_t *nnz) {
	if (!p || !nnz) return kHaErrorArg;
	*nnz = p->nnz;
	return kHaSuccess;
}

int ha_hip_plan_download(const HaHipPlan *p, int32_t *indptr, int32_t *indices, double *data) {
	if (!p) return kHaErrorArg;
	int rc = kHaSuccess;
	if (indptr) HCK(hipMemcpy(indptr, p->indptr, (size_t)(p->ncols + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
	if (indices && p->nnz) HCK(hipMemcpy(indices, p->indices, (size_t)p->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
	if (data && p->nnz) HCK(hipMemcpy(data, p->data, (size_t)p->nnz * sizeof(double), hipMemcpyDeviceToHost));
	return kHaSuccess;
fail:
	return rc;
}

int ha_plan_download_data(const HaHipPlan *p, double *data) { return ha_hip_plan_download(p, NULL, NULL, data); }

int ha_hip_plan_download_local(const HaHipPlan *p, const int32_t *which, int32_t n_which, float *out) {
	if (!p || (!which && n_which) || !out) return kHaErrorArg;
	int rc = kHaSuccess;
	int64_t *qoff = (int64_t *)malloc((size_t)p->count * sizeof(int64_t));
	float *buf = (float *)malloc((size_t)p->kpmax * p->kpmax * sizeof(float));
	if (!qoff || !buf) { free(qoff); free(buf); return kHaErrorAlloc; }
	HCK(hipMemcpy(qoff, p->qoff, (size_t)p->count * sizeof(int64_t), hipMemcpyDeviceToHost));
	int64_t o = 0;
	for (int32_t i = 0; i < n_which; ++i) {
		int32_t s = which[i];
		if (s < 0 || s >= p->count) { rc = kHaErrorArg; goto fail; }
		int32_t k = p->h_k[s], kp = (k + 31) / 32 * 32;
		HCK(hipMemcpy(buf, p->q + qoff[s], (size_t)kp * kp * sizeof(float), hipMemcpyDeviceToHost));
		for (int32_t r = 0; r < k; ++r)
			for (int32_t c = 0; c < k; ++c) out[o + (int64_t)r * k + c] = buf[r + (int64_t)c * kp];
		o += (int64_t)k * k;
	}
fail:
	free(qoff);
	free(buf);
	return rc;
}

int ha_hip_plan_stats(const HaHipPlan *p, HaHipStats *s) {
	if (!p || !s) return kHaErrorArg;
	*s = p->stats;
	return kHaSuccess;
}

int ha_plan_det_negative(HaHipPlan *p, int32_t *neg) { return ha_hip_plan_det_signs(p, 0, neg); }

int ha_hip_plan_det_signs(HaHipPlan *p, int lu_kernel, int32_t *neg) {
	if (!p || !neg) return kHaErrorArg;
	if (!p->loaded || !p->q) { ERR("plan has no local transforms yet\n"); return kHaErrorArg; }
	if (lu_kernel && p->kpmax > HAK_LU_KPMAX) { ERR("searchlights too large for the LU kernel\n"); return kHaErrorArg; }
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	int32_t *d = NULL, *zf = NULL;
	ACK(d, (size_t)p->count * sizeof(int32_t));
	if (lu_kernel) {
		/* every searchlight, flagged or not, as the signs-only mode of k_detfix */
		ACK(zf, (size_t)p->count * sizeof(int32_t));
		KCK(hak_fill_i32(zf, p->count, 0, st));
		HakLocalArgs a = local_args(p);
		a.order = NULL;
		a.n_order = p->count;
		a.queue = p->queues + 10;
		a.flag = zf;
		KCK(hak_fill_i32(p->queues + 10, 1, 0, st));
		KCK(hak_lu_sign(&a, p->work, lu_slot_floats(p), p->kpmax, d, lu_workgroups(p), st));
	} else {
		HakLocalArgs a = local_args(p);
		a.order = NULL;
		a.n_order = p->count;
		a.queue = p->queues + 0;
		a.det_neg = d;
		KCK(hak_fill_i32(p->queues, 1, 0, st));
		KCK(hak_detfix(&a, p->work64, p->stride64, p->iota, p->n_slots64, st));
	}
	HCK(hipMemcpyAsync(neg, d, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
fail:
	if (d) hipFree(d);
	if (zf) hipFree(zf);
	return rc;
}

int ha_hip_plan_orthogonality(HaHipPlan *p, int exact, double *out) {
	if (!p || !out) return kHaErrorArg;
	if (!p->loaded || !p->q) { ERR("plan has no local transforms yet\n"); return kHaErrorArg; }
	if (!exact && p->kpmax > HAK_ORTH_PROBE_MAX_KP) {
		ERR("the probe check holds searchlights up to kp %d (plan: %d); use the exact check\n",
		    HAK_ORTH_PROBE_MAX_KP, p->kpmax);
		return kHaErrorArg;
	}
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	double *d = NULL;
	ACK(d, (size_t)p->count * sizeof(double));
	KCK(hak_orth(p->q, p->qoff, p->qld, p->off, NULL, p->count, p->kpmax, exact ? 1 : 0, d, NULL, 0.0,
	             exact ? (p->count < 65535 ? p->count : 65535) : orth_workgroups(p), st));
	HCK(hipMemcpyAsync(out, d, (size_t)p->count * sizeof(double), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
fail:
	if (d) hipFree(d);
	return rc;
}

/* Searchlight Procrustes template over the plan's searchlights (see
 * k_template.hip): subjects dms[s] are nt x nv FP64 (column- or row-major);
 * tpl_out receives the weighted sum of the local templates in the same order.
 * Without reflections every Procrustes solve of the chain gets the reflection
 * fix (det = +1), as ha_template.c:40,80 passes isReflection to ha_procrustes. */
int ha_plan_template(HaHipPlan *p, const double *const *dms, int32_t ns, bool col_major, bool isReflection,
                     bool isScaling, bool isZscore, int32_t level2, double *tpl_out) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	const int32_t nt = p->nt, nv = p->nv_x, ld = p->ld, count = p->count;
	float *data = NULL, *loc = NULL, *work = NULL;
	double *stage = NULL, *out64 = NULL, *host = NULL;
	int32_t *order = NULL, *status = NULL, *queue = NULL, *h_order = NULL, *h_status = NULL;
	int64_t *loc_off = NULL, *h_loc_off = NULL;
	if (!p->weights_computed || ns < 2 || level2 < 0) { ERR("template needs >= 2 subjects\n"); return kHaErrorArg; }
	for (int32_t s = 0; s < ns; ++s) {
		if (!dms[s]) { ERR("subject %d has no data\n", s); return kHaErrorArg; }
	}
	h_order = (int32_t *)malloc((size_t)count * sizeof(int32_t));
	h_status = (int32_t *)malloc((size_t)count * sizeof(int32_t));
	h_loc_off = (int64_t *)malloc((size_t)(count + 1) * sizeof(int64_t));
	host = (double *)malloc((size_t)nt * nv * sizeof(double));
	if (!h_order || !h_status || !h_loc_off || !host) { rc = kHaErrorAlloc; goto fail; }
	h_loc_off[0] = 0;
	for (int32_t s = 0; s < count; ++s) {
		h_order[s] = s;
		h_loc_off[s + 1] = h_loc_off[s] + (int64_t)nt * p->h_k[s];
	}
	qsort_r(h_order, (size_t)count, sizeof(int32_t), cmp_k_desc, p->h_k);
	{
		const int64_t *optr = p->occx_ptr ? p->occx_ptr : p->occ_ptr;
		const int32_t *occ = p->occx ? p->occx : p->occ;
		KCK(hak_weights(p->slx, p->dist, p->radius, p->total, optr, occ, nv, p->wsum, p->w, st));
	}
	ACK(data, (size_t)ns * nv * ld * sizeof(float));
	ACK(stage, (size_t)nt * nv * sizeof(double));
	for (int32_t s = 0; s < ns; ++s) {
		HCK(hipMemcpyAsync(stage, dms[s], (size_t)nt * nv * sizeof(double), hipMemcpyHostToDevice, st));
		KCK(hak_convert_data(stage, nt, nv, col_major ? 1 : 0, data + (size_t)s * nv * ld, ld, NULL, NULL, NULL, st));
	}
	ACK(order, (size_t)count * sizeof(int32_t));
	ACK(status, (size_t)count * sizeof(int32_t));
	ACK(queue, 8 * sizeof(int32_t));
	ACK(loc_off, (size_t)(count + 1) * sizeof(int64_t));
	ACK(loc, (size_t)h_loc_off[count] * sizeof(float));
	HCK(hipMemcpyAsync(order, h_order, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(loc_off, h_loc_off, (size_t)(count + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
	{
		const int64_t slot = hak_template_slot_floats(ns, ld, p->kpmax);
		int32_t n_slots = 2 * g.n_cu;
		if (n_slots > count) n_slots = count;
		/* keep the work slots within 64 GiB of HBM */
		while (n_slots > 1 && (double)n_slots * slot * sizeof(float) > 64.0 * (1u << 30)) n_slots /= 2;
		ACK(work, (size_t)n_slots * slot * sizeof(float));
		HakTplArgs t;
		memset(&t, 0, sizeof(t));
		t.data = data; t.ns = ns; t.nt = nt; t.ld = ld; t.nv = nv;
		t.sl_idx = p->slx; t.sl_off = p->off; t.order = order; t.count = count; t.kpmax = p->kpmax;
		t.loc = loc; t.loc_off = loc_off;
		t.work = work; t.slot_stride = slot; t.n_slots = n_slots;
		t.queue = queue; t.status = status;
		t.zscore = isZscore ? 1 : 0; t.scaling = isScaling ? 1 : 0; t.level2 = level2;
		t.reflection = isReflection ? 1 : 0;
		HakLocalArgs pa = local_args(p);
		/* the 16-subject order-5 workload's chains (n = 400, 20 mm) sit near
		 * sigma_min / sigma_max = 2.5e-4: a schedule from 3e-4 took 2.76 s
		 * against 3.46 s from 1e-7, at the same error */
		pa.plain_ell = (float)g.opt.template_ell;
		KCK(hak_fill_i32(queue, 8, 0, st));
		HCK(hipEventRecord(p->ev[0], st));
		KCK(hak_template(&t, &pa, st));
		HCK(hipEventRecord(p->ev[1], st));
	}
	ACK(out64, (size_t)nt * nv * sizeof(double));
	KCK(hak_template_gather(loc, loc_off, p->off, p->sl_of, p->occ_ptr, p->occ, p->w, nt, nv, out64, st));
	HCK(hipEventRecord(p->ev[2], st));
	HCK(hipMemcpyAsync(h_status, status, (size_t)count * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	HCK(hipMemcpyAsync(host, out64, (size_t)nt * nv * sizeof(double), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
	{
		float t0 = 0, t1 = 0;
		hipEventElapsedTime(&t0, p->ev[0], p->ev[1]);
		hipEventElapsedTime(&t1, p->ev[1], p->ev[2]);
		memset(&p->stats, 0, sizeof(p->stats));
		p->stats.ms_polar = t0;
		p->stats.kernel_ms_polar_avg = t0;
		p->stats.polar_launches = 1;
		p->stats.ms_scatter = t1;
		p->stats.ms_total = t0 + t1;
		int64_t failed = 0, steps = 0;
		for (int32_t s = 0; s < count; ++s) {
			if (h_status[s] < 0) ++failed;
			else steps += h_status[s];
		}
		p->stats.n_unconverged = failed;
		p->stats.steps_total = steps;
		/* FP64 algorithmic flops: per solve A = D_s^T R (2 n k^2), per polar step
		 * X^T X and G^2 (symmetric, k^3 each) and X P (2 k^3); a Jacobi solve
		 * counts its sweeps as steps */
		{
			const double n_solves = (double)(ns - 1) + (double)level2 * ns;
			double flops = 0;
			for (int32_t s = 0; s < count; ++s) {
				const double k = p->h_k[s];
				flops += n_solves * 2.0 * nt * k * k + (h_status[s] > 0 ? (double)h_status[s] : 0.0) * 4.0 * k * k * k;
			}
			p->stats.gemm_flops = flops;
		}
		if (failed) {
			ERR("template: %lld searchlights had a local Procrustes that did not converge\n", (long long)failed);
			rc = kHaErrorSvd;
		}
	}
	/* device result is nt x nv column-major */
	if (col_major) {
		memcpy(tpl_out, host, (size_t)nt * nv * sizeof(double));
	} else {
		for (int32_t v = 0; v < nv; ++v)
			for (int32_t t = 0; t < nt; ++t) tpl_out[(size_t)t * nv + v] = host[(size_t)v * nt + t];
	}
fail:
	hipStreamSynchronize(st);
	hipFree(data); hipFree(stage); hipFree(order); hipFree(status); hipFree(queue);
	hipFree(loc_off); hipFree(loc); hipFree(work); hipFree(out64);
	free(h_order); free(h_status); free(h_loc_off); free(host);
	return rc;
}

/* Yhat = X T (T: the plan's CSC transform) and, when Y is given, the vertex-wise
 * correlation of Yhat with Y. X, Y: nt x nv_x / nt x ncols FP64. */
int ha_hip_plan_evaluate(HaHipPlan *p, const double *X, const double *Y, int32_t nt, bool col_major,
                         double *Yhat_out, double *r_out) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	double *dx = NULL, *dy = NULL, *dyh = NULL, *dr = NULL, *stage = NULL, *host = NULL;
	if (!p || !X || nt <= 0 || (!Yhat_out && !r_out) || (r_out && !Y)) return kHaErrorArg;
	if (!p->data || !p->indptr64) { ERR("plan has no CSC transform\n"); return kHaErrorArg; }
	const int32_t nvx = p->nv_x, nc = p->ncols;
	ACK(stage, (size_t)nt * (nvx > nc ? nvx : nc) * sizeof(double));
	ACK(dx, (size_t)nt * nvx * sizeof(double));
	ACK(dyh, (size_t)nt * nc * sizeof(double));
	/* column-major FP64 copies (k_convert writes the FP64 copy with leading dim nt) */
	{
		float *f32 = NULL;
		const int32_t ld = (nt + 31) / 32 * 32;
		const int32_t nmax = nvx > nc ? nvx : nc;
		ACK(f32, (size_t)ld * nmax * sizeof(float));
		HCK(hipMemcpyAsync(stage, X, (size_t)nt * nvx * sizeof(double), hipMemcpyHostToDevice, st));
		rc = hak_convert_data(stage, nt, nvx, col_major ? 1 : 0, f32, ld, dx, NULL, NULL, st);
		if (!rc && r_out) {
			ACK(dy, (size_t)nt * nc * sizeof(double));
			HCK(hipMemcpyAsync(stage, Y, (size_t)nt * nc * sizeof(double), hipMemcpyHostToDevice, st));
			rc = hak_convert_data(stage, nt, nc, col_major ? 1 : 0, f32, ld, dy, NULL, NULL, st);
		}
		HCK(hipStreamSynchronize(st));
		hipFree(f32);
		if (rc) { rc = kHaErrorInternal; goto fail; }
	}
	HCK(hipEventRecord(p->ev[0], st));
	KCK(hak_spmm_csc(dx, nt, p->indptr64, p->indices, p->data, nc, dyh, st));
	HCK(hipEventRecord(p->ev[1], st));
	if (r_out) {
		ACK(dr, (size_t)nc * sizeof(double));
		KCK(hak_corr_cols(dyh, dy, nt, nc, dr, st));
		HCK(hipMemcpyAsync(r_out, dr, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, st));
	}
	HCK(hipEventRecord(p->ev[2], st));
	if (Yhat_out) {
		if (col_major) {
			HCK(hipMemcpyAsync(Yhat_out, dyh, (size_t)nt * nc * sizeof(double), hipMemcpyDeviceToHost, st));
			HCK(hipStreamSynchronize(st));
		} else {
			host = (double *)malloc((size_t)nt * nc * sizeof(double));
			if (!host) { rc = kHaErrorAlloc; goto fail; }
			HCK(hipMemcpyAsync(host, dyh, (size_t)nt * nc * sizeof(double), hipMemcpyDeviceToHost, st));
			HCK(hipStreamSynchronize(st));
			for (int32_t c = 0; c < nc; ++c)
				for (int32_t t = 0; t < nt; ++t) Yhat_out[(size_t)t * nc + c] = host[(size_t)c * nt + t];
		}
	}
	HCK(hipStreamSynchronize(st));
	{
		float t0 = 0, t1 = 0;
		hipEventElapsedTime(&t0, p->ev[0], p->ev[1]);
		hipEventElapsedTime(&t1, p->ev[1], p->ev[2]);
		p->stats.ms_apply = t0;
		p->stats.ms_corr = r_out ? t1 : 0.0;
	}
fail:
	hipStreamSynchronize(st);
	hipFree(stage); hipFree(dx); hipFree(dy); hipFree(dyh); hipFree(dr);
	free(host);
	return rc;
}

int ha_plan_weights(HaHipPlan *p, double *w_flat) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	if (!p->weights_computed) return kHaErrorArg;
	const int64_t *optr = p->occx_ptr ? p->occx_ptr : p->occ_ptr;
	const int32_t *occ = p->occx ? p->occx : p->occ;
	KCK(hak_weights(p->slx, p->dist, p->radius, p->total, optr, occ, p->nv_x, p->wsum, p->w, st));
	HCK(hipMemcpyAsync(w_flat, p->w, (size_t)p->total * sizeof(double), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
	return kHaSuccess;
fail:
	return rc;
}

int ha_plan_pattern(HaHipPlan *p, int32_t **indptr, int32_t **indices, int64_t *nnz) {
	int rc = kHaSuccess;
	*indptr = (int32_t *)malloc((size_t)(p->ncols + 1) * sizeof(int32_t));
	*indices = (int32_t *)malloc((size_t)(p->nnz > 0 ? p->nnz : 1) * sizeof(int32_t));
	if (!*indptr || !*indices) { free(*indptr); free(*indices); return kHaErrorAlloc; }
	rc = ha_hip_plan_download(p, *indptr, *indices, NULL);
	*nnz = p->nnz;
	if (rc) { free(*indptr); free(*indices); *indptr = NULL; *indices = NULL; }
	return rc;
}

/* FP64 local transforms for every searchlight (row-major k x k each). */
int ha_plan_local64(HaHipPlan *p, bool isReflection, bool isScaling, double *out) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	double *q64 = NULL;
	double *h = NULL;
	int64_t *qoff = NULL;
	if (!p->loaded) return kHaErrorArg;
	HakLocalArgs a = local_args(p);
	a.flag = NULL;
	a.queue = p->queues + 3;
	a.sigma_sum = isScaling ? p->sigma_sum : NULL;
	ACK(q64, (size_t)p->q_elems * sizeof(double));
	a.q64 = q64;
	KCK(hak_fill_i32(p->queues, 16, 0, st));
	KCK(hak_jacobi64(&a, isReflection ? 0 : 1, p->work64, p->stride64, p->n_slots64, p->count, st));
	if (isScaling) {
		KCK(hak_var_sum(p->X64, p->nt, p->slx, p->off, p->count, p->var_sum, st));
		KCK(hak_scale(p->sigma_sum, p->var_sum, p->nt, p->count, p->scale, st));
	}
	HCK(hipStreamSynchronize(st));
	h = (double *)malloc((size_t)p->q_elems * sizeof(double));
	qoff = (int64_t *)malloc((size_t)p->count * sizeof(int64_t));
	double *scale = (double *)malloc((size_t)p->count * sizeof(double));
	if (!h || !qoff || !scale) { free(scale); rc = kHaErrorAlloc; goto fail; }
	HCK(hipMemcpy(h, q64, (size_t)p->q_elems * sizeof(double), hipMemcpyDeviceToHost));
	HCK(hipMemcpy(qoff, p->qoff, (size_t)p->count * sizeof(int64_t), hipMemcpyDeviceToHost));
	if (isScaling) HCK(hipMemcpy(scale, p->scale, (size_t)p->count * sizeof(double), hipMemcpyDeviceToHost));
	int64_t o = 0;
	for (int32_t s = 0; s < p->count; ++s) {
		int32_t k = p->h_k[s], kp = (k + 31) / 32 * 32;
		double sc = isScaling ? scale[s] : 1.0;
		for (int32_t r = 0; r < k; ++r)
			for (int32_t c = 0; c < k; ++c) {
				double v = h[qoff[s] + r + (int64_t)c * kp];
				out[o + (int64_t)r * k + c] = isScaling ? v * sc : v;
			}
		o += (int64_t)k * k;
	}
	free(scale);
fail:
	if (q64) hipFree(q64);
	free(h);
	free(qoff);
	return rc;
}

/* ---- searchlight ridge ---- */

/* Searchlight ridge (ha_searchlight.c:1378-1429, ha_ridge.c:13-54): the FP64
 * betas (X_sl^T X_sl + alpha I)^-1 X_sl^T Y_sl of every searchlight, then the
 * weighted deterministic scatter into the CSC data (accumulating unless
 * zero_data). With betas64 set, the betas are returned instead (row-major k x k
 * per searchlight, in searchlight order) and nothing is scattered. */
int ha_plan_ridge(HaHipPlan *p, double alpha, bool zero_data, double *betas64) {
	int rc = kHaSuccess;
	hipStream_t st = g.stream;
	double *q64 = NULL, *h = NULL;
	int64_t *qoff = NULL;
	int32_t *fl = NULL;
	if (!p || !p->loaded) { ERR("plan has no data loaded\n"); return kHaErrorArg; }
	if (!betas64 && (!p->data || !p->indptr64)) { ERR("plan has no CSC pattern\n"); return kHaErrorArg; }
	if (!(alpha >= 0.0) || isinf(alpha)) { ERR("ridge alpha must be finite and >= 0 (got %g)\n", alpha); return kHaErrorArg; }
	HakLocalArgs a = local_args(p);
	a.queue = p->queues + 1;
	if (betas64) {
		ACK(q64, (size_t)p->q_elems * sizeof(double));
		a.q64 = q64;
	}
	HCK(hipEventRecord(p->ev[0], st));
	KCK(hak_fill_i32(p->queues, 16, 0, st));
	KCK(hak_fill_i32(p->flag, p->count, 0, st));
	KCK(hak_fill_i32(p->steps, p->count, 0, st));
	if (!betas64) {
		if (zero_data) KCK(hak_fill_f64(p->data, p->nnz, 0.0, st));
		if (p->weights_computed && !p->weights_none) {
			const int64_t *optr = p->occx_ptr ? p->occx_ptr : p->occ_ptr;
			const int32_t *occ = p->occx ? p->occx : p->occ;
			KCK(hak_weights(p->slx, p->dist, p->radius, p->total, optr, occ, p->nv_x, p->wsum, p->w, st));
		}
	}
	HCK(hipEventRecord(p->ev[1], st));
	/* HAK_RIDGE_OCC workgroups per CU, two kp x kp FP64 buffers each (allocated on
	 * the first ridge pass of the plan) */
	if (!p->work_ridge) {
		int32_t n = HAK_RIDGE_OCC * g.n_cu;
		if (n > p->count) n = p->count;
		ACK(p->work_ridge, (size_t)n * 2 * p->stride64 * sizeof(double));
		p->n_wg_ridge = n;
	}
	if (p->stride64 < (int64_t)p->kpmax * p->kpmax) {
		ERR("ridge launch does not fit the FP64 work area\n");
		rc = kHaErrorInternal;
		goto fail;
	}
	{
		/* the order is k-descending: the searchlights with fewer time points
		 * than vertices (the dual form, nt x nt systems) come first */
		int32_t n_dual = 0;
		for (int32_t s = 0; s < p->count; ++s) n_dual += p->h_k[s] > p->nt;
		HakLocalArgs b = a;
		b.queue = p->queues + 2;
		KCK(hak_ridge64(&b, alpha, p->work_ridge, p->stride64, p->n_wg_ridge, 0, n_dual, 1, st));
		KCK(hak_ridge64(&a, alpha, p->work_ridge, p->stride64, p->n_wg_ridge, n_dual, p->count, 0, st));
	}
	HCK(hipEventRecord(p->ev[2], st));
	if (!betas64)
		KCK(hak_scatter(p->slx, p->off, p->sl_of, p->occ_ptr, p->occ, p->ncols, p->indptr64, p->indices, p->data,
		                p->q, p->qoff, p->qld, NULL, p->weights_none ? NULL : p->w, 0, p->count, p->max_col, st));
	HCK(hipEventRecord(p->ev[5], st));
	HCK(hipStreamSynchronize(st));
	{
		float t0 = 0, t1 = 0, t2 = 0;
		hipEventElapsedTime(&t0, p->ev[0], p->ev[1]);
		hipEventElapsedTime(&t1, p->ev[1], p->ev[2]);
		hipEventElapsedTime(&t2, p->ev[2], p->ev[5]);
		memset(&p->stats, 0, sizeof(p->stats));
		p->stats.ms_weights = t0;
		p->stats.ms_polar = t1;   /* the local solver: here the ridge solve */
		p->stats.kernel_ms_polar_avg = t1;
		p->stats.polar_launches = 1;
		p->stats.ms_scatter = t2;
		p->stats.ms_total = t0 + t1 + t2;
		p->stats.nnz = p->nnz;
		/* algorithmic flops 3 n k^2 + 7 k^3 / 3: X^T X (symmetric, n k^2), X^T Y
		 * (2 n k^2), the Cholesky (k^3 / 3) and two triangular solves with k
		 * right-hand sides (2 k^3); dual form (n < k): X X^T (k n^2), the
		 * Cholesky (n^3 / 3), two solves with k right-hand sides (2 k n^2) and
		 * X^T Z (2 n k^2) */
		double flops = 0;
		for (int32_t s = 0; s < p->count; ++s) {
			const double k = p->h_k[s], n = p->nt;
			flops += k > n ? 3.0 * k * n * n + n * n * n / 3.0 + 2.0 * n * k * k
			               : 3.0 * n * k * k + 7.0 * k * k * k / 3.0;
		}
		p->stats.gemm_flops = flops;
	}
	fl = (int32_t *)malloc((size_t)p->count * sizeof(int32_t));
	if (!fl) { rc = kHaErrorAlloc; goto fail; }
	HCK(hipMemcpy(fl, p->flag, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost));
	{
		int32_t bad = 0, first = -1;
		for (int32_t s = 0; s < p->count; ++s)
			if (fl[s]) { if (first < 0) first = s; ++bad; }
		p->stats.n_fallback = bad;
		if (bad) {
			ERR("ridge: X_sl^T X_sl + alpha I (X_sl X_sl^T + alpha I where a searchlight has fewer time points "
			    "than vertices) is not positive definite for %d searchlight(s) (first %d, alpha %g); alpha = 0 "
			    "needs X_sl of full rank\n", bad, first, alpha);
			rc = kHaErrorSvd;
			goto fail;
		}
	}
	if (betas64) {
		h = (double *)malloc((size_t)p->q_elems * sizeof(double));
		qoff = (int64_t *)malloc((size_t)p->count * sizeof(int64_t));
		if (!h || !qoff) { rc = kHaErrorAlloc; goto fail; }
		HCK(hipMemcpy(h, q64, (size_t)p->q_elems * sizeof(double), hipMemcpyDeviceToHost));
		HCK(hipMemcpy(qoff, p->qoff, (size_t)p->count * sizeof(int64_t), hipMemcpyDeviceToHost));
		int64_t o = 0;
		for (int32_t s = 0; s < p->count; ++s) {
			const int32_t k = p->h_k[s], kp = (k + 31) / 32 * 32;
			for (int32_t r = 0; r < k; ++r)
				for (int32_t c = 0; c < k; ++c) betas64[o + (int64_t)r * k + c] = h[qoff[s] + r + (int64_t)c * kp];
			o += (int64_t)k * k;
		}
	}
fail:
	if (q64) hipFree(q64);
	free(h);
	free(qoff);
	free(fl);
	if (rc) hipStreamSynchronize(st);
	return rc;
}

int ha_hip_plan_ridge(HaHipPlan *plan, double alpha) { return ha_plan_ridge(plan, alpha, true, NULL); }

/* ------------------------------------------------------------------------ */
/* learned-scaling support: spectra, per-searchlight steps and predictions   */
/* ------------------------------------------------------------------------ */

/* FP32 Jacobi on the searchlights `which`: HAK_NFEAT features and
 * (sigma_max, sigma_min) of A = X_sl^T Y_sl for each, in the order given. */
int ha_hip_plan_spectra(HaHipPlan *p, const int32_t *which, int32_t n, float *features, float *sig) {
	int rc = kHaSuccess;
	if (!p || !p->loaded || !which || n <= 0 || n > p->count) return kHaErrorArg;
	for (int32_t i = 0; i < n; ++i)
		if (which[i] < 0 || which[i] >= p->count) return kHaErrorArg;
	hipStream_t st = g.stream;
	float *hf = NULL, *hs = NULL;
	if (!p->features) ACK(p->features, (size_t)p->count * HAK_NFEAT * sizeof(float));
	if (!p->sig) ACK(p->sig, (size_t)p->count * 2 * sizeof(float));
	if (p->d_which) { hipFree(p->d_which); p->d_which = NULL; }
	ACK(p->d_which, (size_t)n * sizeof(int32_t));
	HCK(hipMemcpyAsync(p->d_which, which, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, st));
	HakLocalArgs a = local_args(p);
	a.order = p->d_which;
	a.n_order = n;
	a.features = p->features;
	a.sig_out = p->sig;
	a.fallback_enabled = 0;
	a.rank_check_nt = 0;
	a.queue = p->queues + 4;
	KCK(hak_fill_i32(p->queues, 16, 0, st));
	KCK(hak_jacobi32(&a, st));
	hf = (float *)malloc((size_t)p->count * HAK_NFEAT * sizeof(float));
	hs = (float *)malloc((size_t)p->count * 2 * sizeof(float));
	if (!hf || !hs) { rc = kHaErrorAlloc; goto fail; }
	HCK(hipMemcpyAsync(hf, p->features, (size_t)p->count * HAK_NFEAT * sizeof(float), hipMemcpyDeviceToHost, st));
	HCK(hipMemcpyAsync(hs, p->sig, (size_t)p->count * 2 * sizeof(float), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
	for (int32_t i = 0; i < n; ++i) {
		if (features) memcpy(features + (size_t)i * HAK_NFEAT, hf + (size_t)which[i] * HAK_NFEAT, HAK_NFEAT * sizeof(float));
		if (sig) { sig[2 * i] = hs[2 * which[i]]; sig[2 * i + 1] = hs[2 * which[i] + 1]; }
	}
fail:
	hipStreamSynchronize(st);
	free(hf);
	free(hs);
	return rc;
}

/* Per-searchlight outcome of the last pass: steps (polar steps, Jacobi sweeps;
 * negative = budget exhausted, >= 1000 = FP64 path) and fallback flags. */
int ha_hip_plan_steps(const HaHipPlan *p, int32_t *steps, int32_t *flags) {
	int rc = kHaSuccess;
	if (!p) return kHaErrorArg;
	if (steps) HCK(hipMemcpy(steps, p->steps, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost));
	if (flags) HCK(hipMemcpy(flags, p->flag, (size_t)p->count * sizeof(int32_t), hipMemcpyDeviceToHost));
	return kHaSuccess;
fail:
	return rc;
}

/* MLP predictions (y0, y1 per searchlight) of the last learned pass. */
int ha_hip_plan_predictions(const HaHipPlan *p, float *pred) {
	int rc = kHaSuccess;
	if (!p || !pred) return kHaErrorArg;
	if (!p->pred) { ERR("no learned pass has run on this plan\n"); return kHaErrorArg; }
	HCK(hipMemcpy(pred, p->pred, (size_t)p->count * 2 * sizeof(float), hipMemcpyDeviceToHost));
	return kHaSuccess;
fail:
	return rc;
}
