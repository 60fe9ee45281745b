// k_ridge_dual.hip -- the searchlight ridge of k_ridge.hip for searchlights
// with fewer time points than vertices (nt < k), on gfx950, FP64:
//   betas = X_sl^T (X_sl X_sl^T + alpha I)^-1 Y_sl,
// an nt x nt system instead of k x k (ha_ridge.c:13-54, ridge.py:7-12: the
// reference's V diag(s / (alpha + s^2)) U^T Y from the thin SVD of X_sl). For
// alpha > 0 these are the same betas as (X^T X + alpha I)^-1 X^T Y; at
// alpha = 0 they are the minimum-norm solution V S^-1 U^T Y when X_sl has full
// row rank nt, which the k x k normal equations cannot give (X^T X is singular
// there). A kernel of its own, so that the k x k kernel keeps its inlined
// Cholesky. The steps and tiles are those of k_chol.h: the nt x nt row Gram
// matrix on FP64 MFMA, blocked Cholesky, two blocked substitutions with the k
// right-hand sides of Y_sl, then X_sl^T Z.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"
#include "k_chol.h"

#define LAUNCH_RET() return (int)hipGetLastError()

namespace {

using namespace kchol;

__device__ __forceinline__ int next_sl(int32_t *queue, int32_t *sh) {
  __syncthreads();
  if (threadIdx.x == 0) *sh = atomicAdd(queue, 1);
  __syncthreads();
  return *sh;
}

// Items begin .. end-1 of a.order. G holds the n x n system (n = roundup(nt,
// 32)), then L, then the betas (k x k, leading dimension kp); C the right-hand
// sides Y_sl (n x kp, leading dimension n), then Z. Both fit the kp x kp
// buffers since n <= kp. A singular X X^T + alpha I (alpha = 0 without full
// row rank) fails the Cholesky and flags the searchlight.
__global__ __launch_bounds__(RT, HAK_RIDGE_OCC) void k_ridge64_dual(HakLocalArgs a, double alpha, double *work64,
                                                                    int64_t stride64, int32_t begin, int32_t end) {
  __shared__ double t0[B][B + 1], t1[B][B + 1], t2[B][B + 1], t3[B][B + 1];
  __shared__ int32_t item_sh, ok_sh;
  double *Gm = work64 + (int64_t)blockIdx.x * 2 * stride64;
  double *Cm = Gm + stride64;
  const int nt = a.nt, n = (nt + B - 1) / B * B;
  for (;;) {
    const int it = begin + next_sl(a.queue, &item_sh);
    if (it >= end) break;
    const int s = a.order ? a.order[it] : it;
    const int o = a.sl_off[s];
    const int k = a.sl_off[s + 1] - o;
    const int kp = a.qld[s];
    gram_rows64(a.X64, a.slx_idx + o, nt, k, n, alpha, 1.0, Gm, t0, t1);
    for (int64_t e = threadIdx.x; e < (int64_t)n * kp; e += RT) {
      const int t = (int)(e % n), c = (int)(e / n);
      Cm[e] = (t < nt && c < k) ? a.Y64[(int64_t)a.sly_idx[o + c] * nt + t] : 0.0;
    }
    __syncthreads();
    const bool ok = cholesky64(Gm, n, t0, t3, t1, t2, &ok_sh);
    trsm_ll(Gm, Cm, n, false, t0, t3, t1, t2, kp, n);
    trsm_ll(Gm, Cm, n, true, t0, t3, t1, t2, kp, n);
    // betas = X_sl^T Z over L's dead buffer, Z(t, c) = Cm[t + c n]
    crossprod64<true>(a.X64, a.slx_idx + o, Cm, nullptr, nt, k, kp, 0.0, 0.0, Gm, false, t0, t1, n);
    float *Q = a.q + a.qoff[s];
    double *Q64 = a.q64 ? a.q64 + a.qoff[s] : nullptr;
    for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += RT) {
      const double v = Gm[e];
      Q[e] = (float)v;
      if (Q64) Q64[e] = v;
    }
    if (threadIdx.x == 0) {
      a.flag[s] = ok ? 0 : 1;
      if (a.steps) a.steps[s] = ok ? 0 : -1;
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" int hak_ridge64_dual(const HakLocalArgs *a, double alpha, double *work64, int64_t stride64, int n_wg,
                                int32_t begin, int32_t end, void *stream) {
  if (end <= begin) return 0;
  hipLaunchKernelGGL(k_ridge64_dual, dim3(n_wg), dim3(RT), 0, (hipStream_t)stream, *a, alpha, work64, stride64,
                     begin, end);
  LAUNCH_RET();
}
