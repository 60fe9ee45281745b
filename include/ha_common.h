/*
 * ha_common.h -- ABI types shared by every entry point of libhyperalignment_amd.
 *
 * Same names, field order, enum values and error codes as the reference's
 * csrc/ha_common.h:35-76, so a caller written against the reference compiles
 * unchanged. One enum value is added: kHaBackendHIP = 4 (after
 * kHaBackendCUDA = 3, ha_common.h:70-76).
 *
 * Layouts (unchanged from the reference):
 *   TMat / TMatF   row-major, data[i * cols + j]            (ha_common.h:44-49, 82-87)
 *   TSparseCSC     scipy.sparse.csc_matrix layout: int32 indptr[cols+1],
 *                  int32 indices[nnz] (row ids, sorted within a column),
 *                  double data[nnz]                          (ha_common.h:51-59)
 *   TSearchlights  per-searchlight index / distance arrays   (ha_common.h:61-68)
 */
#ifndef HA_COMMON_H
#define HA_COMMON_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (ha_common.h:35-42). */
enum {
	kHaSuccess       =  0,
	kHaErrorAlloc    = -1,
	kHaErrorSvd      = -2,
	kHaErrorArg      = -3,
	kHaErrorInternal = -4,
};

typedef struct {
	double *data;
	int32_t rows;
	int32_t cols;
} TMat;

typedef struct {
	float *data;
	int32_t rows;
	int32_t cols;
} TMatF;

typedef struct {
	double *data;       /* values, length nnz */
	int32_t *indices;   /* row index of each value, length nnz */
	int32_t *indptr;    /* column pointers, length cols + 1 */
	int32_t rows;
	int32_t cols;
	int64_t nnz;
} TSparseCSC;

typedef struct {
	int32_t **indices;  /* vertex ids of each searchlight */
	double **dists;     /* distance of each vertex to the centre; NULL = uniform weights */
	int32_t *sizes;     /* vertices per searchlight */
	int32_t count;      /* number of searchlights */
	double radius;      /* searchlight radius (0 if uniform) */
} TSearchlights;

typedef enum {
	kHaBackendCPU64 = 0,
	kHaBackendCPU32 = 1,
	kHaBackendMetal = 2,
	kHaBackendCUDA  = 3,
	kHaBackendHIP   = 4,   /* MI355X (gfx950) HIP backend: this library */
} THaBackend;

/* Frees a CSC matrix returned by ha_sparse_init (the reference's inline
 * helper at ha_common.h:187-194, exported here so FFI callers can reach it). */
void ha_sparse_free(TSparseCSC *m);

#ifdef __cplusplus
}
#endif

#endif /* HA_COMMON_H */
