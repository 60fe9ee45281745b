// k_jacobi.h -- the FP64 Procrustes solve by one-sided (Hestenes) Jacobi SVD,
// shared by the searchlight path's FP64 kernel (k_local.hip) and the template
// chain (k_template.hip), and the FP32 Jacobi's rotations and products.
// Reference: csrc/ha_searchlight.c:578-659 (dgesdd, T = U Vt, the reflection
// check on T), csrc/ha_procrustes.c:17-127.
// Needs NT (threads per workgroup, 256) and wg_sum_f64 from k_gemm.h and
// k_polar.h.

#ifndef K_JACOBI_H
#define K_JACOBI_H

#include <stdint.h>
#include <math.h>

namespace kjac {

// ---------------------------------------------------------------------------
// One-sided Jacobi (Hestenes), FP32 and FP64
// ---------------------------------------------------------------------------
// Pairs follow the round-robin tournament on n = k (+1 if odd) players: in
// round r, pair i is (p(i), p(n-1-i)) with p(0) = 0 and
// p(j) = ((j - 1 + r) mod (n - 1)) + 1. Columns >= k are virtual (skipped).
// 16 lanes cooperate on one pair; 16 pairs run concurrently per workgroup.

__device__ __forceinline__ int rr_player(int j, int r, int n) {
  return j == 0 ? 0 : ((j - 1 + r) % (n - 1)) + 1;
}

template <typename T>
__device__ __forceinline__ T grp16_sum(T v) {
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
  return v;
}

// Rotation threshold |a_p . a_q| > tol * |a_p| |a_q| with tol = 4 eps sqrt(k):
// the rounding level of the computed dot products, so the last sweep finds
// nothing left to rotate instead of chasing noise. A column of rounding noise
// (a zero singular value) never meets the test: its rotations shrink it
// towards underflow, far below the rank cut smax k cut of the completion
// (cut: 2.3e-16 in FP64, 1.2e-7 in FP32), which replaces it; the sweeps then
// end at the limit with the other columns converged.
template <typename T>
struct JacobiTol;
template <>
struct JacobiTol<float> {
  static constexpr double eps = 1.19e-7, cut = 1.2e-7;
  static constexpr int sweeps = 30;
};
template <>
struct JacobiTol<double> {
  static constexpr double eps = 2.22e-16, cut = 2.3e-16;
  static constexpr int sweeps = 30;
};

// Orthogonalises the k columns of Am (col-major, ld kp) while accumulating the
// rotations in Vm (starts as identity). Returns the number of sweeps used
// (negative if not converged).
template <typename T>
__device__ int jacobi_sweeps(T *Am, T *Vm, int k, int kp, int32_t *flag_sh) {
  const int n = k + (k & 1);
  const int grp = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const double tol = 4.0 * JacobiTol<T>::eps * sqrt((double)k);
  int sweep;
  for (sweep = 0; sweep < JacobiTol<T>::sweeps; ++sweep) {
    if (threadIdx.x == 0) *flag_sh = 0;
    __syncthreads();
    for (int r = 0; r < n - 1; ++r) {
      for (int i0 = 0; i0 < n / 2; i0 += NT / 16) {
        int i = i0 + grp;
        if (i < n / 2) {
          int p = rr_player(i, r, n), q = rr_player(n - 1 - i, r, n);
          if (p > q) { int t = p; p = q; q = t; }
          if (q < k) {
            T *ap = Am + (int64_t)p * kp, *aq = Am + (int64_t)q * kp;
            T al = 0, be = 0, ga = 0;
            for (int t = gl; t < k; t += 16) {
              T x = ap[t], y = aq[t];
              al += x * x; be += y * y; ga += x * y;
            }
            al = grp16_sum(al); be = grp16_sum(be); ga = grp16_sum(ga);
            if (fabs((double)ga) > tol * sqrt((double)al * (double)be) && ga != (T)0) {
              T zeta = (be - al) / (2 * ga);
              T tt = (zeta >= 0 ? (T)1 : (T)-1) / (fabs(zeta) + sqrt((T)1 + zeta * zeta));
              T c = (T)1 / sqrt((T)1 + tt * tt), sn = c * tt;
              for (int t = gl; t < k; t += 16) {
                T x = ap[t], y = aq[t];
                ap[t] = c * x - sn * y;
                aq[t] = sn * x + c * y;
              }
              T *vp = Vm + (int64_t)p * kp, *vq = Vm + (int64_t)q * kp;
              for (int t = gl; t < k; t += 16) {
                T x = vp[t], y = vq[t];
                vp[t] = c * x - sn * y;
                vq[t] = sn * x + c * y;
              }
              if (gl == 0) *flag_sh = 1;
            }
          }
        }
      }
      __syncthreads();
    }
    __syncthreads();
    const int rotated = *flag_sh;
    __syncthreads();
    if (!rotated) return sweep + 1;
  }
  return -sweep;
}

// Q(m, n) = sum_{j < k} U(m, j) V(n, j) - [flipj >= 0] 2 U(m, flipj) V(n, flipj)
// for m, n < kp (zero outside k x k). U, V column-major with leading dim kp.
// 32 x 32 output tiles with 32-column chunks of U and V staged in LDS.
template <typename T, typename TO>
__device__ void compose_uvt(const T *U, const T *V, int k, int kp, TO *Q, double *Q64, T (*us)[33], T (*vs)[33],
                            int flipj) {
  const int mi = threadIdx.x & 31, ni = threadIdx.x >> 5;
  for (int tm = 0; tm < kp; tm += 32)
    for (int tn = 0; tn < kp; tn += 32) {
      T acc[4] = {0, 0, 0, 0};
      for (int j0 = 0; j0 < k; j0 += 32) {
        for (int e = threadIdx.x; e < 32 * 32; e += NT) {
          const int c = e & 31, jj = e >> 5;
          const bool okj = j0 + jj < k;
          us[c][jj] = (okj && tm + c < k) ? U[(tm + c) + (int64_t)(j0 + jj) * kp] : (T)0;
          vs[c][jj] = (okj && tn + c < k) ? V[(tn + c) + (int64_t)(j0 + jj) * kp] : (T)0;
        }
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < 32; ++jj) {
          const T u = us[mi][jj];
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] += u * vs[ni + 8 * q][jj];
        }
        __syncthreads();
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = tm + mi, n = tn + ni + 8 * q;
        T v = 0;
        if (m < k && n < k) {
          v = acc[q];
          if (flipj >= 0) v -= (T)2 * U[m + (int64_t)flipj * kp] * V[n + (int64_t)flipj * kp];
        }
        Q[m + (int64_t)n * kp] = (TO)v;
        if (Q64) Q64[m + (int64_t)n * kp] = (double)v;
      }
    }
  __syncthreads();
}

// sign(det M) of the k x k matrix M (column-major, leading dimension kp) by LU
// with partial pivoting of a copy in L; the pivot is the first row of largest
// |value|, as dgetrf takes it. For the orthogonal factors U and V of the FP64
// path |det| = 1, so every pivot is far from zero and the sign is robust.
__device__ inline int lu_det_sign64(const double *M, double *L, int k, int kp, double *pmax_sh, int32_t *pidx_sh,
                             int32_t *piv_sh, int32_t *sign_sh) {
  for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) L[e] = M[e];
  if (threadIdx.x == 0) *sign_sh = 1;
  __syncthreads();
  for (int c = 0; c < k; ++c) {
    // pivot search in column c, rows c..k-1
    double best = -1.0;
    int bi = c;
    for (int r = c + threadIdx.x; r < k; r += NT) {
      double v = fabs(L[r + (int64_t)c * kp]);
      if (v > best) { best = v; bi = r; }
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
      double ob = __shfl_xor(best, o2);
      int oi = __shfl_xor(bi, o2);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { pmax_sh[wave] = best; pidx_sh[wave] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double b0 = pmax_sh[0];
      int i0 = pidx_sh[0];
      for (int q = 1; q < 4; ++q)
        if (pmax_sh[q] > b0 || (pmax_sh[q] == b0 && pidx_sh[q] < i0)) { b0 = pmax_sh[q]; i0 = pidx_sh[q]; }
      *piv_sh = i0;
      if (i0 != c) *sign_sh = -*sign_sh;
    }
    __syncthreads();
    const int pr = *piv_sh;
    if (pr != c)
      for (int j = threadIdx.x; j < k; j += NT) {
        double t = L[c + (int64_t)j * kp];
        L[c + (int64_t)j * kp] = L[pr + (int64_t)j * kp];
        L[pr + (int64_t)j * kp] = t;
      }
    __syncthreads();
    const double d = L[c + (int64_t)c * kp];
    if (threadIdx.x == 0 && d < 0.0) *sign_sh = -*sign_sh;
    if (d != 0.0) {
      for (int64_t e = threadIdx.x; e < (int64_t)(k - c - 1) * (k - c - 1); e += NT) {
        int r = c + 1 + (int)(e % (k - c - 1)), j = c + 1 + (int)(e / (k - c - 1));
        L[r + (int64_t)j * kp] -= L[r + (int64_t)c * kp] / d * L[c + (int64_t)j * kp];
      }
    }
    __syncthreads();
  }
  const int sg = *sign_sh;
  __syncthreads();
  return sg;
}

// LDS of svd_polar64: the 32 x 32 staging tiles of U V^T, the completion's
// dot products (1,024 at a time) and the workgroup's scalars.
struct SvdLds {
  double us[32][33], vs[32][33];
  double wdot[1024];
  double pmax[4];
  int32_t pidx[4], flag, piv, sign;
};

// The FP64 Procrustes solve of the k x k matrix A in Am (column-major, leading
// dimension kp, zero outside the k x k block): one-sided Jacobi A V = U S (Vm
// starts as the identity here), U = the normalised columns, Q = U V^T written
// to Q in FP32 and, when Q64 is given, to Q64 in FP64 (kp x kp, zero outside).
// Columns whose singular value is at most k 2.3e-16 sigma_max (A singular:
// k >= nt samples, or zero columns) are completed to an orthonormal U by
// Gram-Schmidt from fixed pseudo-random starts, so Q is one of the optimal
// transforms; when every singular value is zero U = I, as LAPACK returns for
// the zero matrix (T = I). With det_fix (isReflection = false), det Q < 0
// flips the pair of the smallest singular value (ha_searchlight.c:611-631),
// the sign from LUs of U and V. Lm: kp x kp work (det_fix only); sig: k
// doubles. Returns the sweeps (negative if the sweeps reached their limit,
// which a singular A does: its noise columns never meet the rotation test);
// *sigma_sum = sum(sigma) less twice the flipped one.
__device__ inline int svd_polar64(double *Am, double *Vm, double *Lm, double *sig, int k, int kp, bool det_fix,
                                  float *Q, double *Q64, SvdLds &sv, double *red, double *sigma_sum) {
  for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) {
    const int m = (int)(e % kp), n = (int)(e / kp);
    Vm[e] = (m == n && m < k) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int sweeps = jacobi_sweeps<double>(Am, Vm, k, kp, &sv.flag);
  for (int j = threadIdx.x; j < k; j += NT) {
    double s2 = 0;
    const double *col = Am + (int64_t)j * kp;
    for (int t = 0; t < k; ++t) s2 += col[t] * col[t];
    sig[j] = sqrt(s2);
  }
  __syncthreads();
  double smax = 0;
  for (int j = 0; j < k; ++j) smax = fmax(smax, sig[j]);
  const double cut = smax * k * JacobiTol<double>::cut;
  for (int64_t e = threadIdx.x; e < (int64_t)k * kp; e += NT) {
    const int j = (int)(e / kp), m = (int)(e % kp);
    Am[e] = smax == 0.0 ? (m == j ? 1.0 : 0.0) : (sig[j] > cut) ? Am[e] / sig[j] : 0.0;
  }
  __syncthreads();
  // Complete U for columns with zero singular value: each deficient column
  // starts as a fixed pseudo-random vector, is projected off every finished
  // column twice (classical Gram-Schmidt, all dot products in parallel), and
  // normalised. Finished = full-rank columns plus the deficient columns
  // completed before it.
  for (int j = 0; j < k && smax > 0.0; ++j) {
    if (sig[j] > cut) continue;
    double *uj = Am + (int64_t)j * kp;
    for (int attempt = 0; attempt < 4; ++attempt) {
      for (int t = threadIdx.x; t < k; t += NT) {
        uint64_t z = ((uint64_t)j << 32) ^ ((uint64_t)t * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)attempt << 20);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        uj[t] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
      }
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int i0 = 0; i0 < k; i0 += 1024) {
          // dot products with the finished columns i0 .. i0+1023
          for (int i = i0 + threadIdx.x; i < k && i < i0 + 1024; i += NT) {
            double d = 0.0;
            const bool finished = (i != j) && (sig[i] > cut || i < j);
            if (finished) {
              const double *ui = Am + (int64_t)i * kp;
              for (int t = 0; t < k; ++t) d += ui[t] * uj[t];
            }
            sv.wdot[i - i0] = d;
          }
          __syncthreads();
          for (int t = threadIdx.x; t < k; t += NT) {
            double v = uj[t];
            for (int i = i0; i < k && i < i0 + 1024; ++i) v -= sv.wdot[i - i0] * Am[(int64_t)i * kp + t];
            uj[t] = v;
          }
          __syncthreads();
        }
      }
      double nn = 0.0;
      for (int t = threadIdx.x; t < k; t += NT) nn += uj[t] * uj[t];
      nn = wg_sum_f64(nn, red);
      if (nn > 1e-20) {
        const double is = 1.0 / sqrt(nn);
        for (int t = threadIdx.x; t < k; t += NT) uj[t] *= is;
        __syncthreads();
        break;
      }
    }
    if (threadIdx.x == 0) sig[j] = 0.0;
    __syncthreads();
  }
  if (smax == 0.0) {
    for (int j = threadIdx.x; j < k; j += NT) sig[j] = 0.0;
    __syncthreads();
  }
  // reflection fix: if det(Q) < 0, flip the pair of the smallest singular
  // value. sign(det Q) = sign(det U) sign(det V), from the finished factors
  // as the reference takes it from T = U V^T: for a singular A, det A is
  // rounding noise and the completed columns of U decide the sign.
  int jmin = 0;
  bool flip = false;
  if (det_fix) {
    for (int j = 1; j < k; ++j)
      if (sig[j] <= sig[jmin]) jmin = j;
    const int su = lu_det_sign64(Am, Lm, k, kp, sv.pmax, sv.pidx, &sv.piv, &sv.sign);
    const int sv2 = lu_det_sign64(Vm, Lm, k, kp, sv.pmax, sv.pidx, &sv.piv, &sv.sign);
    flip = su * sv2 < 0;
  }
  compose_uvt<double, float>(Am, Vm, k, kp, Q, Q64, sv.us, sv.vs, flip ? jmin : -1);
  double ss = 0.0;
  for (int j = threadIdx.x; j < k; j += NT) ss += sig[j];
  ss = wg_sum_f64(ss, red);
  *sigma_sum = flip ? ss - 2.0 * sig[jmin] : ss;
  return sweeps;
}

}  // namespace kjac

#endif  // K_JACOBI_H
