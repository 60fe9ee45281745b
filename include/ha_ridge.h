/*
 * ha_ridge.h -- ridge regression, FP64 on the GPU.
 *
 * Signatures are those of the reference's csrc/ha_ridge.h:16,23. The
 * reference solves through the SVD of X, betas = V diag(s / (alpha + s^2))
 * U^T Y; this library solves the normal equations (X^T X + alpha I) betas =
 * X^T Y by Cholesky, the same betas for alpha > 0 (and for alpha = 0 when X
 * has full column rank). The grid-search and ensemble helpers of
 * ha_ridge.h:33-48 are not provided.
 */
#ifndef HA_RIDGE_H
#define HA_RIDGE_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ha_ridge.h:16 / ha_ridge.c:13-54. X: M x N, Y: M x P, betas:
 * N x P, all row-major FP64. alpha >= 0; kHaErrorSvd if X^T X + alpha I is
 * not positive definite (with alpha = 0 that includes P > N: the problem is
 * solved as one max(N, P)-vertex searchlight, X padded with zero columns). */
int ha_ridge(const TMat *X, const TMat *Y, double alpha, TMat *betas);

/* Replaces ha_ridge.h:23 / ha_ridge.c:56-95: one response vector y (M),
 * betas_out (N). */
int ha_ridge_vec(const TMat *X, const double *y, double alpha, double *betas_out);

#ifdef __cplusplus
}
#endif

#endif /* HA_RIDGE_H */
