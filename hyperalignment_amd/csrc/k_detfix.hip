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

#define LAUNCH_RET() return (int)hipGetLastError()

namespace {

using namespace kchol;

constexpr int NV = 4;      // inverse iteration vectors
constexpr int NITER = 6;   // inverse iteration steps

__device__ __forceinline__ double wg_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// sign of det of the k x k block of M (FP32, column-major, leading dim kp),
// destroyed: blocked right-looking LU with partial pivoting (the lowest row
// index on ties), 32-column panels: each panel factored in place with whole
// rows swapped as pivots are chosen, the U block beside it solved one column
// per thread, and the trailing matrix updated once per panel in 32 x 32 LDS
// tiles. 0 for an exactly singular block.
__device__ int lu_sign(float *M, int k, int kp, float *pmax, int *pidx, int *piv_sh, int *sign_sh,
                       float (*ls)[B + 1], float (*us)[B + 1]) {
  if (threadIdx.x == 0) *sign_sh = 1;
  __syncthreads();
  for (int j0 = 0; j0 < k; j0 += B) {
    const int nb = k - j0 < B ? k - j0 : B, je = j0 + nb;
    // panel: columns j0 .. je-1, rows j0 .. k-1
    for (int c = j0; c < je; ++c) {
      float best = -1.f;
      int bi = c;
      for (int r = c + threadIdx.x; r < k; r += RT) {
        const float v = fabsf(M[r + (int64_t)c * kp]);
        if (v > best) { best = v; bi = r; }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) { pmax[wave] = best; pidx[wave] = bi; }
      __syncthreads();
      if (threadIdx.x == 0) {
        float b0 = pmax[0];
        int i0 = pidx[0];
        for (int q = 1; q < 4; ++q)
          if (pmax[q] > b0 || (pmax[q] == b0 && pidx[q] < i0)) { b0 = pmax[q]; i0 = pidx[q]; }
        *piv_sh = i0;
        if (i0 != c) *sign_sh = -*sign_sh;
      }
      __syncthreads();
      const int pr = *piv_sh;
      if (pr != c)
        for (int j = threadIdx.x; j < k; j += RT) {
          const float t = M[c + (int64_t)j * kp];
          M[c + (int64_t)j * kp] = M[pr + (int64_t)j * kp];
          M[pr + (int64_t)j * kp] = t;
        }
      __syncthreads();
      const float d = M[c + (int64_t)c * kp];
      if (d == 0.f) {
        if (threadIdx.x == 0) *sign_sh = 0;
        __syncthreads();
        return 0;
      }
      if (threadIdx.x == 0 && d < 0.f) *sign_sh = -*sign_sh;
      // multipliers below the pivot, then the rest of the panel
      for (int r = c + 1 + threadIdx.x; r < k; r += RT) M[r + (int64_t)c * kp] /= d;
      __syncthreads();
      const int nr = k - c - 1, ncol = je - c - 1;
      for (int64_t e = threadIdx.x; e < (int64_t)nr * ncol; e += RT) {
        const int r = c + 1 + (int)(e % nr), j = c + 1 + (int)(e / nr);
        M[r + (int64_t)j * kp] -= M[r + (int64_t)c * kp] * M[c + (int64_t)j * kp];
      }
      __syncthreads();
    }
    if (je >= k) break;
    // U block: rows j0 .. je-1, columns je .. k-1, unit lower L11 (one column per thread)
    for (int j = je + threadIdx.x; j < k; j += RT)
      for (int i = 1; i < nb; ++i) {
        float v = M[(j0 + i) + (int64_t)j * kp];
        for (int p = 0; p < i; ++p) v -= M[(j0 + i) + (int64_t)(j0 + p) * kp] * M[(j0 + p) + (int64_t)j * kp];
        M[(j0 + i) + (int64_t)j * kp] = v;
      }
    __syncthreads();
    // trailing matrix: M(r, j) -= sum_p L(r, j0 + p) U(j0 + p, j), 32 x 32 tiles
    const int ri = threadIdx.x & 31, nj = threadIdx.x >> 5;
    for (int r0 = je; r0 < k; r0 += B)
      for (int c0 = je; c0 < k; c0 += B) {
        for (int e = threadIdx.x; e < B * B; e += RT) {
          const int a = e & 31, b = e >> 5;  // row / column inside the tile, panel index
          ls[a][b] = (r0 + a < k && b < nb) ? M[(r0 + a) + (int64_t)(j0 + b) * kp] : 0.f;
          us[b][a] = (c0 + a < k && b < nb) ? M[(j0 + b) + (int64_t)(c0 + a) * kp] : 0.f;
        }
        __syncthreads();
        if (r0 + ri < k) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int j = c0 + nj + 8 * q;
            if (j < k) {
              float acc = 0.f;
#pragma unroll 8
              for (int p = 0; p < B; ++p) acc += ls[ri][p] * us[p][nj + 8 * q];
              M[(r0 + ri) + (int64_t)j * kp] -= acc;
            }
          }
        }
        __syncthreads();
      }
  }
  const int sgn = *sign_sh;
  __syncthreads();
  return sgn;
}

// V <- L^-1 V for the NV columns of V (kp x NV, column-major, leading dim kp),
// L lower triangular (kp x kp): 32-row blocks, the rows above a block folded in
// by two threads per (row, column), the block itself solved in LDS.
__device__ void solve_lower_nv(const double *L, double *V, int kp, double (*d)[B + 1], double (*cs)[B + 1],
                               double (*part)[B + 1]) {
  const int i = threadIdx.x & 31, q = (threadIdx.x >> 5) & 3, h = threadIdx.x >> 7;
  for (int j0 = 0; j0 < kp; j0 += B) {
    double s = 0.0;
    for (int c = h; c < j0; c += 2) s += L[(j0 + i) + (int64_t)c * kp] * V[c + (int64_t)q * kp];
    part[h * NV + q][i] = s;
    for (int e = threadIdx.x; e < B * B; e += RT) d[e & 31][e >> 5] = L[(j0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp];
    __syncthreads();
    if (h == 0) cs[i][q] = V[(j0 + i) + (int64_t)q * kp] - (part[q][i] + part[NV + q][i]);
    __syncthreads();
    forward_sub(d, cs);
    if (h == 0) V[(j0 + i) + (int64_t)q * kp] = cs[i][q];
    __syncthreads();
  }
}

// V <- L^-T V: blocks from the bottom, each (row, column) folding in the rows
// below the block down its own column of L.
__device__ void solve_lower_t_nv(const double *L, double *V, int kp, double (*d)[B + 1], double (*cs)[B + 1],
                                 double (*part)[B + 1]) {
  const int i = threadIdx.x & 31, q = (threadIdx.x >> 5) & 3, h = threadIdx.x >> 7;
  for (int j0 = kp - B; j0 >= 0; j0 -= B) {
    double s = 0.0;
    for (int r = j0 + B + h; r < kp; r += 2) s += L[r + (int64_t)(j0 + i) * kp] * V[r + (int64_t)q * kp];
    part[h * NV + q][i] = s;
    for (int e = threadIdx.x; e < B * B; e += RT) d[e & 31][e >> 5] = L[(j0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp];
    __syncthreads();
    if (h == 0) cs[i][q] = V[(j0 + i) + (int64_t)q * kp] - (part[q][i] + part[NV + q][i]);
    __syncthreads();
    backward_sub_t(d, cs);
    if (h == 0) V[(j0 + i) + (int64_t)q * kp] = cs[i][q];
    __syncthreads();
  }
}

// modified Gram-Schmidt on the NV columns of V (rows < kp)
__device__ void orthonormalize(double *V, int kp, double *red) {
  for (int q = 0; q < NV; ++q) {
    for (int p = 0; p < q; ++p) {
      double s = 0.0;
      for (int i = threadIdx.x; i < kp; i += RT) s += V[i + (int64_t)p * kp] * V[i + (int64_t)q * kp];
      s = wg_sum(s, red);
      for (int i = threadIdx.x; i < kp; i += RT) V[i + (int64_t)q * kp] -= s * V[i + (int64_t)p * kp];
      __syncthreads();
    }
    double s = 0.0;
    for (int i = threadIdx.x; i < kp; i += RT) s += V[i + (int64_t)q * kp] * V[i + (int64_t)q * kp];
    s = wg_sum(s, red);
    const double inv = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
    for (int i = threadIdx.x; i < kp; i += RT) V[i + (int64_t)q * kp] *= inv;
    __syncthreads();
  }
}

// eigenvector of the smallest eigenvalue of the symmetric NV x NV matrix t
// (cyclic Jacobi, one thread); returns the eigenvalue
__device__ double small_eig(double (*t)[NV], double *y) {
  double v[NV][NV];
  for (int i = 0; i < NV; ++i)
    for (int j = 0; j < NV; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < NV; ++i)
      for (int j = 0; j < NV; ++j) {
        tot += t[i][j] * t[i][j];
        if (i != j) off += t[i][j] * t[i][j];
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < NV - 1; ++p)
      for (int q = p + 1; q < NV; ++q) {
        if (t[p][q] == 0.0) continue;
        const double th = (t[q][q] - t[p][p]) / (2.0 * t[p][q]);
        const double tn = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(tn * tn + 1.0), s = tn * c;
        for (int r = 0; r < NV; ++r) {  // columns p, q
          const double a = t[r][p], b = t[r][q];
          t[r][p] = c * a - s * b;
          t[r][q] = s * a + c * b;
        }
        for (int r = 0; r < NV; ++r) {  // rows p, q
          const double a = t[p][r], b = t[q][r];
          t[p][r] = c * a - s * b;
          t[q][r] = s * a + c * b;
        }
        for (int r = 0; r < NV; ++r) {
          const double a = v[r][p], b = v[r][q];
          v[r][p] = c * a - s * b;
          v[r][q] = s * a + c * b;
        }
      }
  }
  int m = 0;
  for (int i = 1; i < NV; ++i)
    if (t[i][i] < t[m][m]) m = i;
  for (int i = 0; i < NV; ++i) y[i] = v[i][m];
  return t[m][m];
}

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
