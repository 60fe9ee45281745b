/*
 * hak.h -- thin C ABI between the C host layer (ha_hip.c) and the gfx950
 * kernels (*.hip). Every pointer is a device pointer unless named h_*;
 * `stream` is a hipStream_t passed as void*. Each launcher returns 0 or the
 * hipError_t of the launch.
 */
#ifndef HAK_H
#define HAK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fills (used instead of hipMemsetAsync so passes can be graph-captured) ---- */
int hak_fill_i32(int32_t *p, int64_t n, int32_t v, void *stream);
int hak_fill_f64(double *p, int64_t n, double v, void *stream);

/* ---- searchlight index (k_index.hip) ---- */
/* sl_of[k] = searchlight owning flat position k. */
int hak_sl_of(const int32_t *sl_off, int32_t count, int32_t *sl_of, void *stream);
/* per-vertex occurrence counts (cnt zeroed by caller). */
int hak_count_occ(const int32_t *sl_idx, int64_t total, int32_t *cnt, void *stream);
/* exclusive scan of n int32 counts into n+1 int64 prefixes (out[n] = total). */
int hak_scan_i32_i64(const int32_t *in, int64_t n, int64_t *out, void *stream);
/* int64 prefix -> int32 copy (caller checked the total fits). */
int hak_i64_to_i32(const int64_t *in, int64_t n, int32_t *out, void *stream);
/* fill occ[ptr[v] ..] with the flat positions of vertex v (unordered). */
int hak_fill_occ(const int32_t *sl_idx, int64_t total, const int64_t *occ_ptr,
                 int32_t *cursor, int32_t *occ, void *stream);
/* sort each vertex's occurrence list ascending; *err set to 1 if a list is
 * longer than the in-LDS limit (8192), to 2 if a searchlight lists a vertex twice. */
int hak_sort_occ(const int64_t *occ_ptr, int32_t nv, const int32_t *sl_of, int32_t *occ,
                 int32_t *err, void *stream);
/* weights exactly as ha_searchlight.c:970-1001 (dists may be NULL). */
int hak_weights(const int32_t *sl_idx, const double *dists, double radius, int64_t total,
                const int64_t *occ_ptr, const int32_t *occ, int32_t nv,
                double *wsum, double *w, void *stream);
/* min/max vertex id of each searchlight. */
int hak_sl_minmax(const int32_t *sl_idx, const int32_t *sl_off, int32_t count,
                  int32_t *slmin, int32_t *slmax, void *stream);
/* CSC pattern: counts per column, then sorted row ids. Columns are the
 * vertices; the rows of column c are the union of the searchlights that
 * contain c. */
int hak_pattern_count(const int32_t *sl_idx, const int32_t *sl_off, const int32_t *sl_of,
                      const int64_t *occ_ptr, const int32_t *occ,
                      const int32_t *slmin, const int32_t *slmax, int32_t nv,
                      int32_t *colcnt, void *stream);
int hak_pattern_fill(const int32_t *sl_idx, const int32_t *sl_off, const int32_t *sl_of,
                     const int64_t *occ_ptr, const int32_t *occ,
                     const int32_t *slmin, const int32_t *slmax, int32_t nv,
                     const int64_t *indptr, int32_t *indices, void *stream);

/* Deterministic scatter: for every column c (a Y-vertex) and, in increasing
 * searchlight order, every searchlight s in [s_begin, s_end) that holds c at
 * local position j: data[pos(slx_s[i], c)] += (double)Q_s[i,j] * scale_s * w_s[i].
 * Q_s is k_s x k_s at q + qoff[s], column-major with leading dim qld[s].
 * scale may be NULL (1.0), w may be NULL (1.0). Pairs outside the pattern are
 * skipped. occ lists come from the Y-side searchlight lists. max_col: the
 * longest column (indptr[c + 1] - indptr[c]); it sizes the kernel's LDS. */
int hak_scatter(const int32_t *slx_idx, const int32_t *sl_off, const int32_t *sl_of,
                const int64_t *occ_ptr, const int32_t *occ, int32_t ncols,
                const int64_t *indptr, const int32_t *indices, double *data,
                const float *q, const int64_t *qoff, const int32_t *qld,
                const double *scale, const double *w,
                int32_t s_begin, int32_t s_end, int32_t max_col, void *stream);

/* Single local transform scatter (ha_sparse_scatter_add on the GPU). */
int hak_scatter_one(const int32_t *indptr, const int32_t *indices, double *data,
                    const int32_t *rows, const int32_t *cols, int32_t nr, int32_t nc,
                    const double *T, const double *w, void *stream);

/* ---- data layout (k_gemm.hip) ---- */
/* FP64 nt x nv (col- or row-major) -> FP32 col-major with leading dim ld
 * (rows nt..ld-1 zero) and an FP64 col-major copy with leading dim nt. */
int hak_convert_data(const double *src, int32_t nt, int32_t nv, int col_major,
                     float *dst32, int32_t ld, double *dst64, uint16_t *hi, uint16_t *lo, void *stream);

/* ---- local transforms (k_local.hip) ---- */
#define HAK_NFEAT 12
typedef struct {
	/* inputs */
	const float *X;  const float *Y;     /* FP32 col-major, leading dim ld (rows >= nt are zero) */
	const uint16_t *Xh, *Xl, *Yh, *Yl;   /* the same as bf16 hi/lo planes: hi = bf16(x), lo = bf16(x - hi) */
	const double *X64; const double *Y64;/* FP64 col-major, leading dim nt */
	int32_t nt, ld;
	const int32_t *slx_idx; const int32_t *sly_idx; const int32_t *sl_off;
	const int32_t *order;                /* processing order (NULL = 0..n_order-1) */
	int32_t n_order;
	/* outputs */
	float *q; const int64_t *qoff; const int32_t *qld;  /* Q_s: qld[s] x qld[s], col-major */
	double *q64;                         /* optional FP64 copy of Q (FP64 path only) */
	double *sigma_sum;                   /* per searchlight trace(Q^T A) (may be NULL) */
	int32_t *flag;                       /* per searchlight: 0 ok, 1 needs the FP64 path */
	int32_t *steps;                      /* per searchlight polar steps / Jacobi sweeps (may be NULL) */
	/* work: per slot 5 (Jacobi-32) or 4 (polar) buffers of slot_stride floats */
	float *work; int64_t slot_stride; int32_t n_slots;
	int32_t *queue;                      /* atomic work counter, zeroed before the launch */
	/* polar schedule: sched[n_buckets][max_sched][3]; bucket b assumes
	 * sigma_min(X0) >= 10^(-b / bucket_per_decade) and sigma_max(X0) <= 1 */
	const float *sched;
	const int32_t *sched_len;            /* per bucket: minimax steps before the Newton-Schulz tail */
	int32_t n_buckets, max_sched, max_steps;
	float bucket_per_decade;
	float tol;                           /* finish when ||X^T X - I||_F < tol */
	float plain_ell;                     /* sigma_min / ||A||_F assumed without a model */
	/* polar stage 1 -> stage 2: per searchlight (scale, ell, ||A||_F); scale 0 = FP64 path */
	float *sl_par;
	int32_t learned_scale;               /* sl_par came from the model (retry plain on failure) */
	/* learned scaling (NULL mlp = off) */
	const float *mlp; const float *mlp_norm;
	float mlp_margin_scale;              /* added to the predicted ln(sigma_max / ||A||_F) */
	float mlp_margin_ell;                /* decades below the predicted log10(sigma_min / sigma_max) */
	float *pred_out;                     /* per searchlight 2 predictions (may be NULL) */
	float *features;                     /* per searchlight HAK_NFEAT features (may be NULL) */
	float *trace;                        /* per searchlight max_steps residuals ||G - I||_F (may be NULL) */
	uint64_t *clk;                       /* polar: per slot 4 phase times in 100 MHz ticks (may be NULL) */
	float *sig_out;                      /* Jacobi paths: per searchlight (sigma_max, sigma_min) (may be NULL) */
	int32_t tile;
	int32_t fallback_enabled;
	int32_t rank_check_nt;               /* HAK_NEAR_DEFICIENT(k, nt): started from the smallest bucket,
	                                        FP32 products */
	int32_t prec_f32;                    /* hak_polar: FP32 products (PREC_F32) instead of bf16x3 */
	/* teams of HAK_GRP workgroups per searchlight (k_polar_grp) */
	void *grp;                           /* per team HAK_GRP_STATE_BYTES of state, zeroed before the launch */
	uint32_t grp_spin;                   /* polls of a team barrier before it gives up (0: default) */
	int32_t skip_done;                   /* polar: skip searchlights whose steps[] is already nonzero */
	int32_t *team_count;                 /* [0] teams whose barrier gave up, [1] searchlights the
	                                        one-workgroup kernel took over (skip_done); may be NULL */
	int32_t *det_neg;                    /* hak_detfix: when set, only det_neg[s] = (det Q_s < 0) for
	                                        every searchlight of the order, flagged or not; Q unchanged */
	const int32_t *det_neg_in;           /* hak_detfix: the signs hak_lu_sign recorded (1: det Q_s < 0)
	                                        instead of its own LU (NULL) */
} HakLocalArgs;

/* Searchlights the polar path treats as (nearly) rank deficient: z-scored
 * columns span at most nt - 1 dimensions, so A = X_sl^T Y_sl is singular for
 * k >= nt, and its conditioning degrades as c = k / (nt - 1) approaches 1
 * (for Gaussian columns kappa(X^T X) ~ ((1 + sqrt c) / (1 - sqrt c))^2, about
 * 1.4e3 at c = 0.9). From c >= 0.9 on, A and the iteration run on FP32
 * products: bf16x3 left 1.4e-4 of error at n = 16, k = 14 and 15 (c 0.93,
 * 1.0), FP32 products 3.5e-6. The full-hemisphere workload (c <= 0.884) stays
 * on bf16x3. */
#define HAK_NEAR_DEFICIENT(k, nt) (10 * (int64_t)(k) >= 9 * ((int64_t)(nt) - 1))

/* Workgroups per CU of the polar iteration stage (launch bound and slot count). */
#define HAK_POLAR_OCC 3
#define HAK_POLAR_OCC128 2  /* with 128 x 128 tiles (the autotuner picks them where they win) */
/* Polar path, stage 1: cross-covariance into the Q store, scale and schedule
 * bound per searchlight (learned when a->mlp), n_slots workgroups. */
int hak_crosscov(const HakLocalArgs *a, int n_slots, void *stream);
/* Stage 2: the polar iteration, a->n_slots workgroups, 4 slot buffers each. */
int hak_polar(const HakLocalArgs *a, void *stream);
/* Stage 2 with teams of HAK_GRP workgroups of one XCD per searchlight (128 x 128
 * bf16x3 tiles, or 64 x 64 FP32 tiles when a->prec_f32): n_teams teams (a
 * multiple of 8), 4 slot buffers per team, a->grp holds
 * n_teams states. A searchlight it finishes gets steps[s] != 0; one it leaves
 * (a team barrier gave up) keeps steps[s] == 0 and A in the Q store, for a
 * following hak_polar with skip_done; one cut short while Q was being written is
 * flagged for the FP64 path. */
#define HAK_GRP 4  /* workgroups per team */
#define HAK_GRP_STATE_BYTES 128
int hak_polar_grp(const HakLocalArgs *a, int n_teams, void *stream);
int hak_jacobi32(const HakLocalArgs *a, void *stream);
/* FP64 cross-covariance + one-sided Jacobi + Q = U V^T for every searchlight
 * s < count with flag[s] != 0 (all of them when a->flag is NULL); det-sign
 * correction when want_det_fix. work64: n_slots * 4 * stride64 doubles. */
int hak_jacobi64(const HakLocalArgs *a, int want_det_fix, double *work64, int64_t stride64,
                 int32_t n_slots, int32_t count, void *stream);

/* Sign of det Q_s for every searchlight of a->order with flag 0 (k_detfix.hip):
 * neg[s] = 1 when det Q_s < 0, else 0 (0 for flagged ones). An FP32 LU with
 * partial pivoting of a copy of Q_s in the workgroup's slot of `work`
 * (slot_floats >= kpmax^2 + 32 kpmax floats each, n_wg slots): 32-column
 * panels factored in LDS, row interchanges kept as a permutation, the trailing
 * updates on FP32 MFMA. kpmax <= HAK_LU_KPMAX (the panel lives in LDS). */
#define HAK_LU_KPMAX 1024
int hak_lu_sign(const HakLocalArgs *a, float *work, int64_t slot_floats, int32_t kpmax, int32_t *neg, int n_wg,
                void *stream);

/* Reflection fix after the polar iteration (k_detfix.hip): for every
 * searchlight of a->order with flag 0 whose Q has det < 0, Q <- Q - 2 (Q v) v^T
 * with v the smallest right singular vector of the FP64 A (a->sigma_sum, when
 * set, loses 2 sigma_min); a failed Cholesky sets the flag for the FP64 path.
 * work64: n_wg * 4 * stride64 doubles; iota: 0, 1, ..., kpmax - 1. With
 * a->det_neg set it only records the signs (see HakLocalArgs). */
int hak_detfix(const HakLocalArgs *a, double *work64, int64_t stride64, const int32_t *iota, int n_wg,
               void *stream);

/* ---- FP64 orthogonality check of the local transforms (k_orth.hip) ---- */
/* kp limit of the probe check: kpmax x 4 doubles and kpmax sign bytes of LDS */
#define HAK_ORTH_PROBE_MAX_KP 1920
/* out[s] = ||Q_s^T Q_s - I||_F in FP64 for the searchlights order[0..n)
 * (identity when NULL): every Gram entry when exact, else the four-probe
 * estimate sqrt(mean_p ||Q^T Q z_p - z_p||^2). With flag set, searchlights
 * already flagged get out[s] = -1 and those with out[s] > tol (or not finite)
 * are flagged. n_wg workgroups (0: one per searchlight); kpmax bounds qld. */
int hak_orth(const float *q, const int64_t *qoff, const int32_t *qld, const int32_t *sl_off,
             const int32_t *order, int32_t n, int32_t kpmax, int exact, double *out, int32_t *flag,
             double tol, int32_t n_wg, void *stream);

/* ---- no-allocation ha_procrustes (k_noalloc.hip) ---- */
/* Looks up the device address of the library's static scratch array for the
 * current device (once per device); 0 or a hipError_t. */
int hak_noalloc_attach(void);

/* ---- spectral-scaling MLP (k_mlp.hip) ---- */
/* Y[N][2] = MLP(X[N][12]) with the given parameters and input normalisation. */
int hak_mlp_forward(const float *params, const float *norm, const float *X, int64_t N, float *Y, void *stream);
/* One minibatch step over rows idx[0..nb) (identity when NULL): partial
 * gradients per 64-row chunk (partial: ceil(nb/64) * NPARAM floats, loss_part:
 * ceil(nb/64)), deterministic reduction, then Adam with learning rate lr
 * (lr == 0: gradient only, into grad_out when not NULL). loss_out receives the
 * batch loss sum_j w_j (out_j - y_j)^2 averaged over the batch. */
int hak_mlp_grad_step(float *params, float *adam_m, float *adam_v, const float *norm, const float *X,
                      const float *Y, const int32_t *idx, int64_t nb, float w0, float w1, float lr,
                      int step, float *partial, float *loss_part, float *loss_out, float *grad_out,
                      void *stream);

/* Per-searchlight sum of population variances of X columns (FP64), for scaling. */
int hak_var_sum(const double *X64, int32_t nt, const int32_t *slx_idx, const int32_t *sl_off,
                int32_t count, double *var_sum, void *stream);
/* scale[s] = sigma_sum[s] / (var_sum[s] * nt) (ha_searchlight.c:635-656). */
int hak_scale(const double *sigma_sum, const double *var_sum, int32_t nt, int32_t count,
              double *scale, void *stream);

/* ---- searchlight ridge (k_ridge.hip) ---- */
#define HAK_RIDGE_OCC 4  /* ridge workgroups per CU (33.8 KB LDS, at most 128 VGPRs each) */
/* Per searchlight of a.order[begin .. end-1]: betas = (X^T X + alpha I)^-1 X^T Y
 * in FP64 into the Q store (FP32, and q64 when set); flag[s] = 1 if the system
 * is not positive definite. dual: searchlights with nt < k, solved as
 * X^T (X X^T + alpha I)^-1 Y. n_wg workgroups; work64: n_wg * 2 * stride64
 * doubles (stride64 >= kpmax^2). */
int hak_ridge64(const HakLocalArgs *a, double alpha, double *work64, int64_t stride64, int n_wg, int32_t begin,
                int32_t end, int dual, void *stream);

/* ---- applying and scoring a transform (k_apply.hip) ---- */
/* Y = X T: X nt x nv FP64 column-major (leading dim nt), T CSC with ncols
 * columns (int64 indptr), Y nt x ncols column-major. */
int hak_spmm_csc(const double *X, int32_t nt, const int64_t *indptr, const int32_t *indices,
                 const double *data, int32_t ncols, double *Y, void *stream);
/* r[v] = population-z-score correlation of column v of A and B (nt x ncols, column-major). */
int hak_corr_cols(const double *A, const double *B, int32_t nt, int32_t ncols, double *r, void *stream);

/* ---- searchlight Procrustes template (k_template.hip) ---- */
typedef struct {
	const float *data;       /* ns subjects, each ld x nv FP32 column-major (rows nt..ld-1 zero) */
	int32_t ns, nt, ld, nv;
	const int32_t *sl_idx, *sl_off, *order;
	int32_t count, kpmax;
	float *loc;              /* local templates, searchlight s: nt x k column-major at loc_off[s] */
	const int64_t *loc_off;
	float *work; int64_t slot_stride; int32_t n_slots;
	int32_t *queue;          /* atomic work counter, zeroed before the launch */
	int32_t *status;         /* per searchlight: polar steps over all solves, -1 if one failed */
	int32_t zscore, scaling, level2;
	int32_t reflection;      /* 0: every solve gets the reflection fix (det Q = +1) */
} HakTplArgs;
/* Floats of one template work slot for ns subjects, leading dim ld, kp = kpmax. */
int64_t hak_template_slot_floats(int32_t ns, int32_t ld, int32_t kpmax);
/* Local templates of every searchlight (one workgroup each); pa carries the
 * polar schedule (sched, sched_len, n_buckets, max_sched, max_steps,
 * bucket_per_decade, tol, plain_ell). */
int hak_template(const HakTplArgs *t, const HakLocalArgs *pa, void *stream);
/* out[t + v * nt] = sum over the occurrences f of vertex v (searchlight order)
 * of loc[s(f)](t, f - sl_off[s(f)]) * w[f], FP64. */
int hak_template_gather(const float *loc, const int64_t *loc_off, const int32_t *sl_off, const int32_t *sl_of,
                        const int64_t *occ_ptr, const int32_t *occ, const double *w, int32_t nt, int32_t nv,
                        double *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HAK_H */
