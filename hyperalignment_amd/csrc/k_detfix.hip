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
    int sgn;
    if (a.det_neg_in) {
      sgn = a.det_neg_in[s] ? -1 : 1;
    } else {
      for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += RT) M[e] = Q[e];
      __syncthreads();
      sgn = lu_sign(M, k, kp, pmax, pidx, &piv_sh, &sign_sh, reinterpret_cast<float (*)[B + 1]>(&t0[0][0]),
                    reinterpret_cast<float (*)[B + 1]>(&t1[0][0]));
    }
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

typedef float f32x16_lu __attribute__((ext_vector_type(16)));
constexpr int PW = 32;  // LU panel width

// LDS of k_lu_sign beyond its static part: the panel (kpmax rows of PW + 1
// floats), the row permutation, and one 32 x 33 staging tile per wave.
__host__ __device__ constexpr size_t lu_lds_bytes(int kpmax) {
  return (size_t)kpmax * (PW + 1) * sizeof(float) + (size_t)kpmax * sizeof(int) + 4 * 32 * 33 * sizeof(float);
}

// sign(det Q_s) of every searchlight of the order with flag 0, one workgroup
// each from the queue. LU with partial pivoting (largest magnitude, the lowest
// row on ties) of a copy M of Q in this workgroup's work slot, in FP32 (Q is
// orthogonal, so |det Q| = 1 and the sign is robust):
//  * rows are never moved in M: logical row i lives in physical row perm[i],
//    and every interchange swaps two perm entries and flips the sign;
//  * each 32-column panel (logical rows j0 .. k-1) is loaded into LDS and
//    factored there, one column at a time (pivot search, interchange of the
//    panel rows, multipliers and the rank-1 update of the panel's remaining
//    columns);
//  * U12 = L11^-1 A12 is solved one column per thread, into U12T (32 values
//    per column, contiguous);
//  * the trailing matrix is updated in place in 32 x 32 tiles, one per wave,
//    A22 -= L21 U12 as 16 FP32 MFMAs (v_mfma_f32_32x32x2f32) with L21 from the
//    LDS panel and U12 staged through LDS.
// det = sign(perm) * prod(pivots). An exactly zero pivot (singular block)
// records 0 (no fix), like lu_sign.
__global__ __launch_bounds__(RT, 2) void k_lu_sign(HakLocalArgs a, float *work, int64_t slot_floats, int32_t kpmax,
                                                int32_t *neg) {
  extern __shared__ float lds_dyn[];
  float (*pan)[PW + 1] = reinterpret_cast<float (*)[PW + 1]>(lds_dyn);
  int *perm = reinterpret_cast<int *>(lds_dyn + (size_t)kpmax * (PW + 1));
  float (*ub)[32][33] = reinterpret_cast<float (*)[32][33]>(perm + kpmax);
  __shared__ float pmax[4];
  __shared__ int pidx[4], piv_sh, sign_sh, item_sh;
  float *M = work + (int64_t)blockIdx.x * slot_floats;
  float *U12T = M + (int64_t)kpmax * kpmax;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) item_sh = atomicAdd(a.queue, 1);
    __syncthreads();
    const int it = item_sh;
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    if (a.flag[s] != 0) {
      if (threadIdx.x == 0) neg[s] = 0;
      continue;
    }
    const int k = a.sl_off[s + 1] - a.sl_off[s], kp = a.qld[s];
    {
      const float4 *src = reinterpret_cast<const float4 *>(a.q + a.qoff[s]);
      float4 *dst = reinterpret_cast<float4 *>(M);
      const int64_t n4 = (int64_t)kp * kp / 4;
      for (int64_t e = threadIdx.x; e < n4; e += RT) dst[e] = src[e];
    }
    for (int i = threadIdx.x; i < k; i += RT) perm[i] = i;
    if (threadIdx.x == 0) sign_sh = 1;
    __syncthreads();
    for (int j0 = 0; j0 < k; j0 += PW) {
      const int nb = k - j0 < PW ? k - j0 : PW, m = k - j0, je = j0 + nb;
      // the panel: logical rows j0 .. k-1 of columns j0 .. je-1
      for (int e = threadIdx.x; e < m * nb; e += RT) {
        const int i = e % m, c = e / m;
        pan[i][c] = M[perm[j0 + i] + (int64_t)(j0 + c) * kp];
      }
      __syncthreads();
      bool singular = false;
      for (int c = 0; c < nb; ++c) {
        float best = -1.f;
        int bi = c;
        for (int i = c + threadIdx.x; i < m; i += RT) {
          const float v = fabsf(pan[i][c]);
          if (v > best) {
            best = v;
            bi = i;
          }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const float ob = __shfl_xor(best, o);
          const int oi = __shfl_xor(bi, o);
          if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
          }
        }
        if (lane == 0) {
          pmax[wave] = best;
          pidx[wave] = bi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          float b0 = pmax[0];
          int i0 = pidx[0];
          for (int q = 1; q < 4; ++q)
            if (pmax[q] > b0 || (pmax[q] == b0 && pidx[q] < i0)) {
              b0 = pmax[q];
              i0 = pidx[q];
            }
          piv_sh = i0;
          if (i0 != c) {
            sign_sh = -sign_sh;
            const int t = perm[j0 + c];
            perm[j0 + c] = perm[j0 + i0];
            perm[j0 + i0] = t;
          }
        }
        __syncthreads();
        const int pr = piv_sh;
        if (pr != c && threadIdx.x < nb) {
          const float t = pan[c][threadIdx.x];
          pan[c][threadIdx.x] = pan[pr][threadIdx.x];
          pan[pr][threadIdx.x] = t;
        }
        __syncthreads();
        const float d = pan[c][c];
        if (d == 0.f) {
          singular = true;
          break;
        }
        if (threadIdx.x == 0 && d < 0.f) sign_sh = -sign_sh;
        for (int i = c + 1 + threadIdx.x; i < m; i += RT) {
          const float l = pan[i][c] / d;
          pan[i][c] = l;
          for (int jj = c + 1; jj < nb; ++jj) pan[i][jj] -= l * pan[c][jj];
        }
        __syncthreads();
      }
      if (singular) {
        if (threadIdx.x == 0) sign_sh = 0;
        break;
      }
      if (je >= k) break;
      const int n2 = k - je;
      // U12 = L11^-1 A12 (unit lower L11 in the panel's first nb rows)
      for (int jc = threadIdx.x; jc < n2; jc += RT) {
        const int64_t col = (int64_t)(je + jc) * kp;
        float u[PW];
#pragma unroll
        for (int i = 0; i < PW; ++i) u[i] = i < nb ? M[perm[j0 + i] + col] : 0.f;
#pragma unroll
        for (int i = 1; i < PW; ++i) {
          if (i < nb) {
            float v = u[i];
#pragma unroll
            for (int q = 0; q < i; ++q) v -= pan[i][q] * u[q];
            u[i] = v;
          }
        }
        float4 *dst = reinterpret_cast<float4 *>(U12T + (int64_t)jc * PW);
#pragma unroll
        for (int q = 0; q < PW / 4; ++q) dst[q] = make_float4(u[4 * q], u[4 * q + 1], u[4 * q + 2], u[4 * q + 3]);
      }
      __syncthreads();
      // A22 -= L21 U12, 32 x 32 tiles, one per wave and round
      const int ntr = (n2 + 31) / 32, ntile = ntr * ntr;
      const int mrow = lane & 31, h = lane >> 5;
      for (int base = 0; base < ntile; base += 4) {
        const int t = base + wave;
        const bool act = t < ntile;
        const int r0 = je + 32 * (act ? t / ntr : 0), c0 = je + 32 * (act ? t % ntr : 0);
        if (act)
          for (int e = lane; e < 32 * 32; e += 64) {
            const int n = e >> 5, kk = e & 31;
            const int col = c0 + n;
            ub[wave][kk][n] = (col < k && kk < nb) ? U12T[(int64_t)(col - je) * PW + kk] : 0.f;
          }
        __syncthreads();
        if (act) {
          const int R = r0 + mrow;
          const bool rv = R < k;
          const int64_t prow = rv ? perm[R] : 0;
          f32x16_lu acc;
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int C = c0 + (r >> 2) * 8 + h * 4 + (r & 3);
            acc[r] = (rv && C < k) ? M[prow + (int64_t)C * kp] : 0.f;
          }
#pragma unroll
          for (int st = 0; st < PW / 2; ++st) {
            const int kk = 2 * st + h;
            const float av = (rv && kk < nb) ? -pan[R - j0][kk] : 0.f;
            const float bv = ub[wave][kk][mrow];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(bv, av, acc, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int C = c0 + (r >> 2) * 8 + h * 4 + (r & 3);
            if (rv && C < k) M[prow + (int64_t)C * kp] = acc[r];
          }
        }
        __syncthreads();
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) neg[s] = sign_sh < 0 ? 1 : 0;
  }
}

}  // namespace

extern "C" int hak_lu_sign(const HakLocalArgs *a, float *work, int64_t slot_floats, int32_t kpmax, int32_t *neg,
                           int n_wg, void *stream) {
  if (a->n_order <= 0) return 0;
  if (n_wg <= 0 || !a->flag || !neg || !work || kpmax <= 0 || kpmax > HAK_LU_KPMAX || kpmax % 32 ||
      slot_floats < (int64_t)kpmax * kpmax + (int64_t)PW * kpmax || slot_floats % 4)
    return (int)hipErrorInvalidValue;
  // the LDS limit is raised once, to the largest panel, on the first (eager) call
  static bool lds_raised = false;
  if (!lds_raised) {
    const hipError_t e = hipFuncSetAttribute((const void *)k_lu_sign, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lu_lds_bytes(HAK_LU_KPMAX));
    if (e != hipSuccess) return (int)e;
    lds_raised = true;
  }
  const size_t lds = lu_lds_bytes(kpmax);
  hipLaunchKernelGGL(k_lu_sign, dim3(n_wg), dim3(RT), lds, (hipStream_t)stream, *a, work, slot_floats, kpmax, neg);
  LAUNCH_RET();
}

extern "C" int hak_detfix(const HakLocalArgs *a, double *work64, int64_t stride64, const int32_t *iota, int n_wg,
                          void *stream) {
  if (a->n_order <= 0) return 0;
  if (n_wg <= 0 || (!a->flag && !a->det_neg) || !iota) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_detfix, dim3(n_wg), dim3(RT), 0, (hipStream_t)stream, *a, work64, stride64, iota);
  LAUNCH_RET();
}
