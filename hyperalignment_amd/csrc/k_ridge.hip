// k_ridge.hip -- searchlight ridge regression on gfx950, FP64.
//
// Per searchlight (one 256-thread workgroup, persistent queue) the local ridge
// solution of ha_ridge.c:13-54 / ridge.py:7-12,
//   betas = V diag(s / (alpha + s^2)) U^T Y   for X = U S V^T,
// which for alpha > 0 equals (X^T X + alpha I)^(-1) X^T Y exactly; it is
// computed that way: G = X_sl^T X_sl + alpha I and C = X_sl^T Y_sl (LDS-tiled
// FP64 products over the time points), a blocked Cholesky G = L L^T in place
// (32-column panels: unblocked factor of the diagonal block in LDS, the panel
// below solved 32 rows at a time in LDS, 32 x 32-tiled update of the trailing
// matrix), then C <- L^-T L^-1 C by left-looking blocked substitution. With
// fewer time points than vertices (nt < k) the kernel of k_ridge_dual.hip
// solves the dual form X_sl^T (X_sl X_sl^T + alpha I)^-1 Y_sl on the nt x nt
// row Gram matrix with the same steps. Padding rows and
// columns k..kp-1 carry an identity diagonal, so the padded system stays SPD
// and its solution is zero there. The betas (k x k) go to the searchlight's
// place in the Q store (FP32; FP64 too when a.q64 is set) for the
// deterministic CSC scatter. Each workgroup uses two kp x kp FP64 buffers of
// the work area (G, then L, and C, then the betas).

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


__global__ __launch_bounds__(RT, HAK_RIDGE_OCC) void k_ridge64(HakLocalArgs a, double alpha, double *work64,
                                                               int64_t stride64, int32_t begin, int32_t end) {
  __shared__ double t0[B][B + 1], t1[B][B + 1], t2[B][B + 1], t3[B][B + 1];
  __shared__ int32_t item_sh, ok_sh;
  double *Gm = work64 + (int64_t)blockIdx.x * 2 * stride64;
  double *Cm = Gm + stride64;
  for (;;) {
    const int it = begin + next_sl(a.queue, &item_sh);
    if (it >= end) break;
    const int s = a.order ? a.order[it] : it;
    const int o = a.sl_off[s];
    const int k = a.sl_off[s + 1] - o;
    const int kp = a.qld[s];
    crossprod64(a.X64, a.slx_idx + o, a.X64, a.slx_idx + o, a.nt, k, kp, alpha, 1.0, Gm, true, t0, t1);
    crossprod64(a.X64, a.slx_idx + o, a.Y64, a.sly_idx + o, a.nt, k, kp, 0.0, 0.0, Cm, false, t0, t1);
    const bool ok = cholesky64(Gm, kp, t0, t3, t1, t2, &ok_sh);
    trsm_ll(Gm, Cm, kp, false, t0, t3, t1, t2);
    trsm_ll(Gm, Cm, kp, true, t0, t3, t1, t2);
    float *Q = a.q + a.qoff[s];
    double *Q64 = a.q64 ? a.q64 + a.qoff[s] : nullptr;
    for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += RT) {
      const double v = Cm[e];
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
                                int32_t begin, int32_t end, void *stream);

extern "C" int hak_ridge64(const HakLocalArgs *a, double alpha, double *work64, int64_t stride64, int n_wg,
                           int32_t begin, int32_t end, int dual, void *stream) {
  if (end <= begin) return 0;
  if (dual) return hak_ridge64_dual(a, alpha, work64, stride64, n_wg, begin, end, stream);
  hipLaunchKernelGGL(k_ridge64, dim3(n_wg), dim3(RT), 0, (hipStream_t)stream, *a, alpha, work64, stride64, begin,
                     end);
  LAUNCH_RET();
}
