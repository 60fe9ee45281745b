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

#ifdef __cplusplus
}
#endif

#endif /* HA_PROCRUSTES_H */
