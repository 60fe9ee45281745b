/*
 * ha_procrustes.h -- one local orthogonal Procrustes problem on the GPU.
 *
 * Signatures are those of the reference's csrc/ha_procrustes.h:20,27.
 * T minimises ||X T - Y||_F over orthogonal T (A = X^T Y = U S V^T, T = U V^T);
 * isReflection=false forces det(T) > 0 by flipping the smallest singular pair;
 * isScaling multiplies T by sum(S) / (sum_j var(X[:, j]) * M).
 * Both run on the FP64 one-sided Jacobi kernel of the HIP backend.
 */
#ifndef HA_PROCRUSTES_H
#define HA_PROCRUSTES_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ha_procrustes.h:20 / ha_procrustes.c:17-127. X, Y: M x N row-major; T: N x N. */
int ha_procrustes(const TMat *X, const TMat *Y, TMat *T, bool isReflection, bool isScaling);

/* Replaces ha_procrustes.h:27 / ha_procrustes.c:129-222 (computed in FP64, rounded to FP32). */
int ha_procrustes_f32(const TMatF *X, const TMatF *Y, TMatF *T, bool isReflection, bool isScaling);

/* Replaces ha_procrustes.h:44 / ha_procrustes.c:271-366: T (N x N, row-major)
 * = the orthogonal polar factor of A (N x N, row-major), on the GPU polar
 * path (FP32 products) instead of the reference's CPU Newton iteration. With
 * isReflection false and det(T) < 0, T's last column is negated (the
 * reference's rule for this function, not the rank-1 SVD flip of
 * ha_procrustes). With isScaling and X_data (M x N, row-major), T is scaled by
 * trace(T^T A) / (sum_j var(X[:, j]) * M), population variances. Where the
 * reference's FP32 inverse fails on an exactly singular A (kHaErrorInternal),
 * this returns one orthogonal polar factor of it. */
int ha_polar_newton(const float *A, float *T, int32_t N, bool isReflection, bool isScaling, const float *X_data,
                    int32_t M);

/* ---- no-allocation variant (no counterpart in the reference) ----
 *
 * ha_procrustes on device buffers, for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. X, Y (M x N) and T (N x N) are device pointers in the
 * row-major layout of TMat; the work is enqueued on `stream` (NULL: the null
 * stream) and the call returns without waiting, so the caller synchronizes
 * before reading T. No hipMalloc, hipFree or allocating library call is made:
 * the work lives in a static device array of HA_NOALLOC_SCRATCH_BYTES that
 * each GPU holds until the process ends. Sizes beyond N <= HA_NOALLOC_MAX_N and
 * M * N <= HA_NOALLOC_MAX_MN return kHaErrorArg. All calls on one GPU share its
 * array: calls that overlap on one GPU race; calls in order on one stream do
 * not. ha_hip_init looks up the array for device 0; on another device the
 * first call does so and must not itself be under graph capture. Same FP64
 * Jacobi kernel as ha_procrustes, so the same T bit for bit. */
#define HA_NOALLOC_MAX_N 512
#define HA_NOALLOC_MAX_MN (1 << 20)
#define HA_NOALLOC_MAX_DEVICES 16
#define HA_NOALLOC_SCRATCH_BYTES (27 * (1 << 20) + 64 * 1024)
int ha_procrustes_noalloc(const double *X, const double *Y, int32_t M, int32_t N, double *T, bool isReflection,
                          bool isScaling, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HA_PROCRUSTES_H */
