// k_detfix.hip -- the reflection fix of the polar path on gfx950
// (isReflection = false; ha_searchlight.c:612-632, ha_procrustes.c:56-95):
// Q = U V^T of A = X_sl^T Y_sl becomes U diag(1, ..., 1, det(U V^T)) V^T.
//
// Per searchlight the polar iteration finished (flag 0), one 256-thread
// workgroup from a persistent queue:
//  1. sign(det Q) = sign(det A) (the odd polynomial steps keep the sign of
//     det X) from an LU factorization with partial pivoting of a copy of Q, in
//     FP32: Q is orthogonal, so |det Q| = 1 and the sign is robust;
//  2. when negative, the smallest right singular vector v of A, from the FP64
//     data: A64 = X_sl^T Y_sl, G = A64^T A64 and its Cholesky factor L (FP64,
//     32 x 32 LDS tiles), then block inverse iteration on four vectors
//     (G^-1 = L^-T L^-1) closed by a Rayleigh-Ritz step on
//     V^T G V = (L^T V)^T (L^T V). v does not come from Q = polar(A): Q's
//     rounding would reach v amplified by sigma_max over the gap below it;
//  3. Q <- Q - 2 (Q v) v^T (Q v = u, the matching left vector), and the
//     singular value sum for scaling loses 2 sigma_min.
// A Cholesky that fails (A rank deficient) marks the searchlight for the FP64
// path, which runs afterwards and applies the fix itself.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"
#include "k_chol.h"
#include "k_detfix.h"

#define LAUNCH_RET() return (int)hipGetLastError()

namespace {

using namespace kchol;

using namespace kdet;

__global__ __launch_bounds__(RT, 4) void k_detfix(HakLocalArgs a, double *work64, int64_t stride64,
                                               const int32_t *iota) {
  __shared__ double t0[B][B + 1], t1[B][B + 1], t2[B][B + 1], t3[B][B + 1];
  __shared__ double red[4];
  __shared__ float pmax[4];
  __shared__ int pidx[4], piv_sh, sign_sh, ok_sh, item_sh;
  __shared__ double tsh[NV][NV], ysh[NV], lam_sh;
  double *A64 = work64 + (int64_t)blockIdx.x * 4 * stride64;
  double *G = A64 + stride64;
  float *M = reinterpret_cast<float *>(A64 + 2 * stride64);
  double *V = A64 + 3 * stride64;  // kp x NV, then W = L^T V (kp x NV)
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) item_sh = atomicAdd(a.queue, 1);
    __syncthreads();
    const int it = item_sh;
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    if (!a.det_neg && a.flag[s] != 0) continue;  // the FP64 path (with the fix) takes it
    const int o = a.sl_off[s], k = a.sl_off[s + 1] - o, kp = a.qld[s];
    float *Q = a.q + a.qoff[s];
    for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += RT) M[e] = Q[e];
    __syncthreads();
    const int sgn = lu_sign(M, k, kp, pmax, pidx, &piv_sh, &sign_sh, reinterpret_cast<float (*)[B + 1]>(&t0[0][0]),
                            reinterpret_cast<float (*)[B + 1]>(&t1[0][0]));
    if (a.det_neg) {  // signs only
      if (threadIdx.x == 0) a.det_neg[s] = sgn < 0 ? 1 : 0;
      continue;
    }
    if (sgn >= 0) continue;
    // G = A64^T A64 (lower part; identity on the padding) and its Cholesky factor
    crossprod64(a.X64, a.slx_idx + o, a.Y64, a.sly_idx + o, a.nt, k, kp, 0.0, 1.0, A64, false, t0, t1);
    crossprod64(A64, iota, A64, iota, kp, k, kp, 0.0, 1.0, G, true, t0, t1);
    if (!cholesky64(G, kp, t0, t3, t1, t2, &ok_sh)) {
      if (threadIdx.x == 0) a.flag[s] = 1;
      continue;
    }
    // block inverse iteration; the padding rows stay zero (G is block diagonal)
    for (int64_t e = threadIdx.x; e < (int64_t)kp * NV; e += RT) {
      const int i = (int)(e % kp), q = (int)(e / kp);
      V[e] = i < k ? sin(0.7548776662466927 * (i + 1) * (q + 1) + 0.5698402909980532 * q) : 0.0;
    }
    __syncthreads();
    orthonormalize(V, kp, red);
    for (int step = 0; step < NITER; ++step) {
      solve_lower_nv(G, V, kp, t0, t3, t1);
      solve_lower_t_nv(G, V, kp, t0, t3, t1);
      orthonormalize(V, kp, red);
    }
    // Rayleigh-Ritz: T = (L^T V)^T (L^T V)
    double *W = V + (int64_t)NV * kp;
    for (int64_t e = threadIdx.x; e < (int64_t)kp * NV; e += RT) {
      const int i = (int)(e % kp), q = (int)(e / kp);
      double acc = 0.0;
      for (int r = i; r < kp; ++r) acc += G[r + (int64_t)i * kp] * V[r + (int64_t)q * kp];  // L^T(i, r) = L(r, i)
      W[e] = acc;
    }
    __syncthreads();
    for (int p = 0; p < NV; ++p)
      for (int q = p; q < NV; ++q) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < kp; i += RT) acc += W[i + (int64_t)p * kp] * W[i + (int64_t)q * kp];
        acc = wg_sum(acc, red);
        if (threadIdx.x == 0) tsh[p][q] = tsh[q][p] = acc;
      }
    __syncthreads();
    if (threadIdx.x == 0) lam_sh = small_eig(tsh, ysh);
    __syncthreads();
    // v = V y (into W's first column), u = Q v (into W's second), Q -= 2 u v^T
    double *vv = W, *uu = W + kp;
    for (int i = threadIdx.x; i < kp; i += RT) {
      double acc = 0.0;
      for (int q = 0; q < NV; ++q) acc += V[i + (int64_t)q * kp] * ysh[q];
      vv[i] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kp; i += RT) {
      double acc = 0.0;
      for (int j = 0; j < k; ++j) acc += (double)Q[i + (int64_t)j * kp] * vv[j];
      uu[i] = acc;
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += RT) {
      const int i = (int)(e % kp), j = (int)(e / kp);
      Q[e] = (float)((double)Q[e] - 2.0 * uu[i] * vv[j]);
    }
    if (threadIdx.x == 0 && a.sigma_sum) a.sigma_sum[s] -= 2.0 * sqrt(fmax(lam_sh, 0.0));
    __syncthreads();
  }
}

}  // namespace

extern "C" int hak_detfix(const HakLocalArgs *a, double *work64, int64_t stride64, const int32_t *iota, int n_wg,
                          void *stream) {
  if (a->n_order <= 0) return 0;
  if (n_wg <= 0 || (!a->flag && !a->det_neg) || !iota) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_detfix, dim3(n_wg), dim3(RT), 0, (hipStream_t)stream, *a, work64, stride64, iota);
  LAUNCH_RET();
}
