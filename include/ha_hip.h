/*
 * ha_hip.h -- MI355X (gfx950) HIP backend for searchlight Procrustes.
 *
 * This is the plug-in module a reference build would include next to
 * csrc/ha_cuda.h. Its first four entry points have the shape of
 * ha_cuda.h:56-66 (init / cleanup / available / dense searchlight Procrustes);
 * the rest add what the reference's CUDA plug-in lacks: a CSC variant (the
 * reference has no GPU branch in ha_searchlight_procrustes,
 * ha_searchlight.c:204-207), a device-resident plan for repeated passes over
 * data already in HBM, run options, and per-pass statistics.
 *
 * All functions return kHaSuccess (0) or a negative error code from
 * ha_common.h and print a one-line reason to stderr on failure. Nothing falls
 * back to a CPU implementation: without a usable GPU every call fails with
 * kHaErrorInternal. GPU state is process-global and not thread-safe, as in the
 * reference (ha_cuda.cu:39-41): one call at a time per process.
 */
#ifndef HA_HIP_H
#define HA_HIP_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- plug-in lifecycle (ha_cuda.h:38-50) ---- */

/* Selects device 0 and creates the library's stream. Idempotent. */
int ha_hip_init(void);
/* Releases the stream and cached device buffers. Safe without init. */
void ha_hip_cleanup(void);
/* True after a successful ha_hip_init. */
bool ha_hip_available(void);

/* ---- run options (new; the reference has no equivalent) ---- */

enum {
	kHaHipMethodAuto         = 0, /* polar (learned if a model is loaded), Jacobi-64 for k<=32 */
	kHaHipMethodJacobi32     = 1, /* plain FP32 one-sided Jacobi SVD, Q = U V^T */
	kHaHipMethodPolar        = 2, /* bf16x3 MFMA polar iteration, default scaling */
	kHaHipMethodPolarLearned = 3, /* polar iteration with MLP-predicted scaling/schedule */
	kHaHipMethodJacobi64     = 4, /* FP64 one-sided Jacobi SVD for every searchlight */
};

typedef struct {
	int32_t method;          /* kHaHipMethod* */
	int32_t fp64_fallback;   /* 1: rank-deficient (k >= nt) searchlights and those whose
	                            polar iteration does not converge (sigma_min/sigma_max
	                            below about 1e-7) use FP64 Jacobi */
	int32_t max_polar_steps; /* iteration budget per searchlight (default 28) */
	double  polar_tol;       /* stop when ||X^T X - I||_F < polar_tol (default 1e-2),
	                            followed by one Newton-Schulz finishing step */
	double  plain_ell;       /* sigma_min/||A||_F assumed by the plain schedule (default 1e-7) */
	int32_t tile;            /* GEMM tile edge 64 or 128; 0 = per-bucket autotuned */
	int32_t use_graph;       /* 1 (default): from the second pass of a plan on, the pass is
	                            captured in a hipGraph once and replayed while its arguments
	                            and the options stay the same; 0: every pass eager */
	int32_t chunk;           /* searchlights per polar chunk (0 = auto) */
	int32_t verbose;         /* 1: timing lines on stderr */
	float   learned_margin_scale; /* added to the predicted ln(sigma_max/||A||_F) (default 0.15) */
	float   learned_margin_ell;   /* decades subtracted from the predicted log10(sigma_min/sigma_max)
	                                 (default 1.0); a model file carries its calibrated margins */
	int32_t team;            /* 1 (default): with 128 x 128 tiles, 4 workgroups of one XCD
	                            iterate each searchlight together and share panel reads
	                            through L2; 0: one workgroup per searchlight */
	uint32_t team_spin;      /* microseconds a team barrier waits before the team gives its
	                            searchlights back to the one-workgroup kernel (0: default 100000) */
	double  orth_tol;        /* polar paths: every local transform's FP64 orthogonality
	                            ||Q^T Q - I||_F (four-probe estimate) is checked after the
	                            iteration; above orth_tol it goes to the FP64 path (default
	                            1e-3; 0 switches the check off) */
	double  template_ell;    /* sigma_min/||A||_F assumed by the template's FP64 schedule
	                            (default 3e-4; smaller ones are reached by its restarts,
	                            and a solve the iteration cannot finish goes to FP64 Jacobi) */
} HaHipOptions;

void ha_hip_get_options(HaHipOptions *opt);
int  ha_hip_set_options(const HaHipOptions *opt);

/* ---- drop-in entry points ---- */

/* Same contract as ha_searchlight_procrustes_cuda (ha_cuda.h:56-66) with two
 * differences that follow the reference CPU path instead of the CUDA one:
 * T_out is accumulated (+=), not overwritten (ha_cuda.cu:445,458), and
 * isReflection=false / isScaling=true are honoured (ha_cuda.cu:364 ignores
 * them). col_major selects the layout of X_data / Y_data (nt x nv). */
int ha_searchlight_procrustes_hip(
    const double *X_data, const double *Y_data,
    int32_t nt, int32_t nv,
    const int32_t *sl_indices,
    const int32_t *sl_offsets,
    const double *sl_dists,
    int32_t count,
    double radius,
    double *T_out,
    bool isReflection, bool isScaling,
    bool col_major);

/* CSC variant with the arguments of ha_searchlight_procrustes
 * (ha_searchlight.h:46-52) minus the backend. */
int ha_searchlight_procrustes_hip_csc(const TMat *X, const TMat *Y,
                                      const TSearchlights *sls_X,
                                      const TSearchlights *sls_Y,
                                      TSparseCSC *mat,
                                      double **weights,
                                      bool isReflection, bool isScaling);

/* ---- device-resident plan: searchlights uploaded once, passes over data in HBM ---- */

typedef struct HaHipPlan HaHipPlan;

typedef struct {
	double  ms_weights;       /* last pass, HIP events on the plan's stream */
	double  ms_crosscov;
	double  ms_polar;          /* local transforms (polar or Jacobi) */
	double  ms_fallback;       /* FP64 Jacobi fallback */
	double  ms_scatter;
	double  ms_total;
	double  ms_pattern;        /* pattern build at plan creation */
	int64_t n_fallback;        /* searchlights solved by FP64 Jacobi */
	int64_t n_unconverged;     /* polar searchlights that hit the step budget */
	int64_t steps_total;       /* sum of polar steps over searchlights */
	int32_t steps_hist[32];    /* histogram of polar steps (last bin = overflow) */
	int64_t nnz;
	double  gemm_flops;        /* algorithmic flops of the polar iteration stage of the last pass
	                              (4 k^3 per executed step and searchlight: X^T X and G^2
	                              symmetric, k^3 each; X P, 2 k^3) */
	double  kernel_ms_polar_avg; /* average duration of one polar launch */
	int32_t polar_launches;
	double  ms_apply;          /* last ha_hip_plan_evaluate: Yhat = X T */
	double  ms_corr;           /* and the vertex-wise correlation */
	int64_t n_orth_checked;    /* local transforms whose FP64 orthogonality was checked in the pass */
	int64_t n_orth_flagged;    /* of those, sent to the FP64 path (estimate above orth_tol) */
	double  orth_max;          /* largest ||Q^T Q - I||_F estimate among the checked ones */
	int64_t n_team_giveup;     /* teams whose barrier wait ran out (their work went back) */
	int64_t n_team_handback;   /* searchlights the one-workgroup kernel took over from teams */
	double  ms_orth;           /* the orthogonality check (part of ms_polar) */
	int32_t graph_replay;      /* 1: the last pass was a replay of the captured hipGraph (its
	                              stage times come from events recorded inside the graph) */
} HaHipStats;

/* Uploads the flat searchlight arrays (layout of hyperalignment_c.py:161-169)
 * and builds, on the GPU, the vertex->searchlight index and the CSC pattern.
 * sl_dists may be NULL (uniform weights). nt is the number of samples of the
 * data passes that will follow. */
int ha_hip_plan_create(int32_t nt, int32_t nv,
                       const int32_t *sl_indices, const int32_t *sl_offsets,
                       const double *sl_dists, int32_t count, double radius,
                       HaHipPlan **plan_out);
void ha_hip_plan_destroy(HaHipPlan *plan);

/* Host -> HBM upload of X and Y (nt x nv, FP64, col- or row-major). */
int ha_hip_plan_load(HaHipPlan *plan, const double *X, const double *Y, bool col_major);

/* One full pass on data already in HBM: weights, cross-covariance, local
 * transforms, deterministic scatter into the plan's CSC data (zeroed first).
 * Synchronous. */
int ha_hip_plan_execute(HaHipPlan *plan, bool isReflection, bool isScaling);

/* The searchlight ridge pass on data already in HBM (ha_searchlight.c:1378-1429):
 * weights, FP64 ridge betas (X_sl^T X_sl + alpha I)^-1 X_sl^T Y_sl per
 * searchlight, deterministic scatter into the plan's CSC data (zeroed first).
 * kHaErrorSvd if some X_sl^T X_sl + alpha I is not positive definite.
 * Synchronous. */
int ha_hip_plan_ridge(HaHipPlan *plan, double alpha);

/* Iteration tile (64 or 128) chosen for each kp bucket (kp <= 64, 128, 256,
 * 384, 512, larger): the tile option when set, else timed on a sample of each
 * large bucket on the first polar pass (the plan's data must be loaded). */
int ha_hip_plan_tiles(HaHipPlan *plan, int32_t bucket_tile[6]);

/* The autotuner's choice per kp bucket (kp <= 64, 128, 256, 384, 512, larger)
 * and what it measured: for every bucket with at least as many searchlights as
 * the GPU has iteration slots, the iteration time of an evenly spaced sample
 * with 64 x 64 tiles, 128 x 128 tiles one workgroup per searchlight, and
 * 128 x 128 tiles in teams of 4 workgroups of one XCD (0 where not timed). */
typedef struct {
	int32_t tile[6];          /* 64 or 128 */
	int32_t team[6];          /* 1: 128 x 128 tiles in teams */
	int32_t sample[6];        /* searchlights timed (0: not timed) */
	int32_t searchlights[6];  /* searchlights of the plan in the bucket */
	float   ms[6][3];         /* 64, 128 alone, 128 in teams */
} HaHipTuning;
int ha_hip_plan_tuning(HaHipPlan *plan, HaHipTuning *out);

/* Diagnostics of the reflection fix: neg[s] = 1 when the local transform Q_s
 * of the last pass has det Q_s < 0, else 0, for every searchlight, from the
 * LU of the reflection-fix kernel (lu_kernel 0) or from the LU-sign kernel the
 * pass runs before it (lu_kernel 1: panels in LDS, MFMA trailing updates; kp up
 * to 1024). Q is not changed. */
int ha_hip_plan_det_signs(HaHipPlan *plan, int lu_kernel, int32_t *neg);

/* Applies the plan's CSC transform to new data and scores it, the evaluation of
 * benchmark_neuroboros.py:204-211,229-232: Yhat = X T (X: nt x nv samples, the
 * plan's nv source vertices; Yhat: nt x nv target vertices) and, when Y and
 * r_out are given, r[v] = mean_t z(Yhat[:, v]) z(Y[:, v]) with population
 * z-scores (NaN for a constant column). X, Y, Yhat in the same order
 * (col_major or row-major); Yhat_out may be NULL when only r is wanted. FP64. */
int ha_hip_plan_evaluate(HaHipPlan *plan, const double *X, const double *Y, int32_t nt, bool col_major,
                         double *Yhat_out, double *r_out);

int ha_hip_plan_nnz(const HaHipPlan *plan, int64_t *nnz);
/* Downloads the CSC pattern and values; any pointer may be NULL. */
int ha_hip_plan_download(const HaHipPlan *plan, int32_t *indptr, int32_t *indices, double *data);
/* Downloads local transforms of chosen searchlights (k_s x k_s row-major each,
 * concatenated in the order given) for diagnostics and parity tests. */
int ha_hip_plan_download_local(const HaHipPlan *plan, const int32_t *which, int32_t n_which,
                               float *out);
int ha_hip_plan_stats(const HaHipPlan *plan, HaHipStats *stats);
/* No counterpart in the reference: the statistics of the last
 * ha_searchlight_template_flat call (ms_polar: the template kernel's time;
 * steps_total: the polar steps of its solves; gemm_flops: their FP64
 * algorithmic flops), for callers that time the template. */
int ha_hip_last_template_stats(HaHipStats *stats);
/* FP64 orthogonality of every local transform in the plan's store (after a
 * pass): out[s] = ||Q_s^T Q_s - I||_F, from every entry of the Gram matrix
 * when exact, else the four-probe estimate the pass itself uses. */
int ha_hip_plan_orthogonality(HaHipPlan *plan, int exact, double *out);

/* Features and (sigma_max, sigma_min) of the chosen searchlights from the FP32
 * Jacobi SVD path (training labels of the learned scaling). features: n x 12,
 * sig: n x 2, in the order of `which`. Requires loaded data. */
int ha_hip_plan_spectra(HaHipPlan *plan, const int32_t *which, int32_t n, float *features, float *sig);
/* Per-searchlight polar steps / Jacobi sweeps of the last pass (negative: budget
 * exhausted; >= 1000: solved on the FP64 path) and FP64-fallback flags. */
int ha_hip_plan_steps(const HaHipPlan *plan, int32_t *steps, int32_t *flags);
/* MLP predictions (2 per searchlight) of the last learned pass. */
int ha_hip_plan_predictions(const HaHipPlan *plan, float *pred);

/* ---- learned polar scaling ---- */

/* Number of features per searchlight and of MLP parameters (12 -> 64 -> 64 -> 2). */
#define HA_HIP_MLP_IN 12
#define HA_HIP_MLP_HID 64
#define HA_HIP_MLP_OUT 2
#define HA_HIP_MLP_NPARAM (HA_HIP_MLP_IN*HA_HIP_MLP_HID + HA_HIP_MLP_HID + \
                           HA_HIP_MLP_HID*HA_HIP_MLP_HID + HA_HIP_MLP_HID + \
                           HA_HIP_MLP_HID*HA_HIP_MLP_OUT + HA_HIP_MLP_OUT)

/* Installs MLP parameters, FP32, concatenated in the order
 * W1[IN][HID], b1[HID], W2[HID][HID], b2[HID], W3[HID][OUT], b3[OUT]
 * (row index = input unit), plus the feature normalisation: mean[IN] then
 * std[IN]. NULL params unloads the model. */
int ha_hip_set_model(const float *params, const float *feat_norm);

/* Batch inference: Y[N][2] = MLP(X[N][12]) on the GPU. */
int ha_hip_mlp_forward(const float *params, const float *feat_norm, const float *X, int64_t N, float *Y);
/* Full-batch gradient of the loss mean_s sum_j w_j (out_sj - Y_sj)^2 at params
 * (HA_HIP_MLP_NPARAM floats into grad) and the loss itself, on the GPU. */
int ha_hip_mlp_grad(const float *params, const float *feat_norm, const float *X, const float *Y, int64_t N,
                    float w0, float w1, float *grad, float *loss);
/* Adam training on the GPU (beta 0.9 / 0.999, eps 1e-8): `epochs` passes of
 * minibatches of `batch` rows in a splitmix64-seeded shuffle; params are
 * updated in place; loss_hist[epochs] (may be NULL) gets the mean batch loss. */
int ha_hip_mlp_train(float *params, const float *feat_norm, const float *X, const float *Y, int64_t N,
                     int32_t epochs, int32_t batch, float lr, float w0, float w1, uint64_t seed,
                     float *loss_hist);

#ifdef __cplusplus
}
#endif

#endif /* HA_HIP_H */
