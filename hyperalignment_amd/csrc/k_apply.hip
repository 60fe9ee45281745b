// k_apply.hip -- applying a searchlight transform and scoring it (FP64).
//
// Yhat = X T for the CSC transform T (rows: source vertices, columns: target
// vertices), as the reference's evaluation computes X_test @ W
// (benchmark_neuroboros.py:229-232), and the vertex-wise correlation
// r[v] = mean_t z(A[:, v]) z(B[:, v]) with population z-scores
// (benchmark_neuroboros.py:204-211).
//
// k_spmm_csc: one workgroup per target column c; each thread owns up to
// HAK_APPLY_T time points and walks the column's nonzeros in index order, so
// Yhat[t, c] = sum_q X[t, indices[q]] data[q] accumulates in the same order as
// a sequential CSC product. Built with -ffp-contract=off (no fused
// multiply-add), like the reference's scalar loop. X and Yhat are column-major
// with leading dimension nt; each nonzero reads one contiguous column of X.
// HBM-bound: X (nt x nv FP64) is re-read once per nonzero of its row, mostly
// from the Infinity Cache.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"

#define LAUNCH_RET() return (int)hipGetLastError()

namespace {

constexpr int AT = 256;         // threads per workgroup
constexpr int HAK_APPLY_T = 4;  // time points per thread per pass (nt up to 1024 in one pass)

__global__ __launch_bounds__(AT) void k_spmm_csc(const double *__restrict__ X, int32_t nt,
                                                 const int64_t *__restrict__ indptr,
                                                 const int32_t *__restrict__ indices,
                                                 const double *__restrict__ data, int32_t ncols,
                                                 double *__restrict__ Y) {
  for (int c = blockIdx.x; c < ncols; c += gridDim.x) {
    const int64_t q0 = indptr[c], q1 = indptr[c + 1];
    for (int t0 = 0; t0 < nt; t0 += AT * HAK_APPLY_T) {
      double acc[HAK_APPLY_T];
#pragma unroll
      for (int u = 0; u < HAK_APPLY_T; ++u) acc[u] = 0.0;
      for (int64_t q = q0; q < q1; ++q) {
        const double w = data[q];
        const double *x = X + (int64_t)indices[q] * nt;
#pragma unroll
        for (int u = 0; u < HAK_APPLY_T; ++u) {
          const int t = t0 + u * AT + threadIdx.x;
          if (t < nt) acc[u] = acc[u] + x[t] * w;
        }
      }
#pragma unroll
      for (int u = 0; u < HAK_APPLY_T; ++u) {
        const int t = t0 + u * AT + threadIdx.x;
        if (t < nt) Y[t + (int64_t)c * nt] = acc[u];
      }
    }
  }
}

// Two-pass population statistics of one column by one wave.
__device__ __forceinline__ void col_mean_std(const double *c, int nt, int lane, double &mean, double &sd) {
  double s = 0.0;
  for (int t = lane; t < nt; t += 64) s += c[t];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  mean = s / nt;
  double v = 0.0;
  for (int t = lane; t < nt; t += 64) {
    const double d = c[t] - mean;
    v += d * d;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  sd = sqrt(v / nt);
}

// r[v] = mean_t ((A - mean_A) / sd_A) ((B - mean_B) / sd_B); a constant column
// gives NaN (0 / 0), as numpy does.
__global__ __launch_bounds__(AT) void k_corr_cols(const double *__restrict__ A, const double *__restrict__ B,
                                                  int32_t nt, int32_t ncols, double *__restrict__ r) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (AT / 64) + (threadIdx.x >> 6);
  const int nwave = gridDim.x * (AT / 64);
  for (int v = wave; v < ncols; v += nwave) {
    const double *a = A + (int64_t)v * nt, *b = B + (int64_t)v * nt;
    double ma, sa, mb, sb;
    col_mean_std(a, nt, lane, ma, sa);
    col_mean_std(b, nt, lane, mb, sb);
    double s = 0.0;
    for (int t = lane; t < nt; t += 64) s += ((a[t] - ma) / sa) * ((b[t] - mb) / sb);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) r[v] = s / nt;
  }
}

}  // namespace

extern "C" int hak_spmm_csc(const double *X, int32_t nt, const int64_t *indptr, const int32_t *indices,
                            const double *data, int32_t ncols, double *Y, void *stream) {
  if (ncols <= 0 || nt <= 0) return 0;
  const int grid = ncols < 65535 ? ncols : 65535;
  hipLaunchKernelGGL(k_spmm_csc, dim3(grid), dim3(AT), 0, (hipStream_t)stream, X, nt, indptr, indices, data, ncols,
                     Y);
  LAUNCH_RET();
}

extern "C" int hak_corr_cols(const double *A, const double *B, int32_t nt, int32_t ncols, double *r, void *stream) {
  if (ncols <= 0 || nt <= 0) return 0;
  int grid = (ncols + AT / 64 - 1) / (AT / 64);
  if (grid > 65535) grid = 65535;
  hipLaunchKernelGGL(k_corr_cols, dim3(grid), dim3(AT), 0, (hipStream_t)stream, A, B, nt, ncols, r);
  LAUNCH_RET();
}
