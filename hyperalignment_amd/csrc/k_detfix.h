// k_detfix.h -- device building blocks of the reflection fix (isReflection =
// false; ha_searchlight.c:612-632, ha_procrustes.c:56-95), shared by the
// searchlight path's fix (k_detfix.hip) and the template kernel
// (k_template.hip): the sign of det Q from a blocked FP32 LU, and the smallest
// eigenvector of an SPD Gram matrix from its FP64 Cholesky factor by block
// inverse iteration and a Rayleigh-Ritz step (one 256-thread workgroup).
#ifndef K_DETFIX_H
#define K_DETFIX_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "k_chol.h"

namespace kdet {

using namespace kchol;

constexpr int NV = 4;      // inverse iteration vectors
constexpr int NITER = 6;   // inverse iteration steps

__device__ inline double wg_sum(double v, double *red) {
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
__device__ inline int lu_sign(float *M, int k, int kp, float *pmax, int *pidx, int *piv_sh, int *sign_sh,
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

// Wave 0 applies d^-1 (TR: d^-T) to rows 0..31 of the NV vectors at vb
// (leading dim kp), less the eight partial sums in part (part[h NV + q][row]),
// d the 32 x 32 lower triangular diagonal block. Lane l holds row r = l % 32 of
// vectors l / 32 and l / 32 + 2; the rows are solved in order (TR: from the
// last), the solved value of each reaching the other lanes by readlane, so the
// block needs no barrier.
template <bool TR>
__device__ __forceinline__ void block_solve(const double (*d)[B + 1], const double (*part)[B + 1], double *vb,
                                            int kp) {
  const int r = threadIdx.x & 31, qa = (threadIdx.x >> 5) & 1;
  double t0 = vb[r + (int64_t)qa * kp], t1 = vb[r + (int64_t)(qa + 2) * kp];
#pragma unroll
  for (int hh = 0; hh < 8; ++hh) {
    t0 -= part[hh * NV + qa][r];
    t1 -= part[hh * NV + qa + 2][r];
  }
  const double dinv = 1.0 / d[r][r];
#pragma unroll
  for (int s = 0; s < B; ++s) {
    const int ii = TR ? B - 1 - s : s;
    const double y0 = t0 * dinv, y1 = t1 * dinv;  // final on the lanes of row ii
    const double a0 = lane_f64(y0, ii), b0 = lane_f64(y0, B + ii);
    const double a1 = lane_f64(y1, ii), b1 = lane_f64(y1, B + ii);
    const double x0 = qa ? b0 : a0, x1 = qa ? b1 : a1;
    const bool below = TR ? r < ii : r > ii;
    const double l = below ? (TR ? d[ii][r] : d[r][ii]) : 0.0;
    t0 = r == ii ? x0 : t0 - l * x0;
    t1 = r == ii ? x1 : t1 - l * x1;
  }
  vb[r + (int64_t)qa * kp] = t0;
  vb[r + (int64_t)(qa + 2) * kp] = t1;
}

// V <- L^-1 V for the NV columns of V (kp x NV, column-major, leading dim kp;
// in LDS where the caller has room), L lower triangular (kp x kp): 32-row
// blocks. The rows above a block are folded in by all 256 threads: thread
// (row i = t % 32, range h = t / 32) sums L(j0 + i, c) V(c, :) over c = h,
// h + 8, ... < j0 (the 32 rows of one c are one coalesced load, eight of them in
// flight), the eight partial sums meet in LDS, and wave 0 solves the block.
__device__ inline void solve_lower_nv(const double *L, double *V, int kp, double (*d)[B + 1], double (*part)[B + 1]) {
  const int i = threadIdx.x & 31, h = threadIdx.x >> 5;
  for (int j0 = 0; j0 < kp; j0 += B) {
    double sq[NV] = {0.0, 0.0, 0.0, 0.0};
    const double *Lr = L + (j0 + i);
#pragma unroll 8
    for (int c = h; c < j0; c += 8) {
      const double l = Lr[(int64_t)c * kp];
#pragma unroll
      for (int q = 0; q < NV; ++q) sq[q] += l * V[c + (int64_t)q * kp];
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) part[h * NV + q][i] = sq[q];
    for (int e = threadIdx.x; e < B * B; e += RT) d[e & 31][e >> 5] = L[(j0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp];
    __syncthreads();
    if (threadIdx.x < 64) block_solve<false>(d, part, V + j0, kp);
    __syncthreads();
  }
}

// V <- L^-T V: blocks from the bottom. The rows below a block are folded in
// down the block's own columns of L: thread (column i = t / 8, offset h = t % 8)
// sums L(r, j0 + i) V(r, :) over r = j0 + 32 + h, + 8, ... (eight consecutive r
// of one column are one 64-byte load), the partial sums meet in LDS.
__device__ inline void solve_lower_t_nv(const double *L, double *V, int kp, double (*d)[B + 1],
                                        double (*part)[B + 1]) {
  const int h = threadIdx.x & 7, ic = threadIdx.x >> 3;
  for (int j0 = kp - B; j0 >= 0; j0 -= B) {
    double sq[NV] = {0.0, 0.0, 0.0, 0.0};
    const double *Lc = L + (int64_t)(j0 + ic) * kp;
#pragma unroll 8
    for (int r = j0 + B + h; r < kp; r += 8) {
      const double l = Lc[r];
#pragma unroll
      for (int q = 0; q < NV; ++q) sq[q] += l * V[r + (int64_t)q * kp];
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) part[h * NV + q][ic] = sq[q];
    for (int e = threadIdx.x; e < B * B; e += RT) d[e & 31][e >> 5] = L[(j0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp];
    __syncthreads();
    if (threadIdx.x < 64) block_solve<true>(d, part, V + j0, kp);
    __syncthreads();
  }
}

// W = L^T V (kp x NV) for L lower triangular: W(i, :) = sum over r >= i of
// L(r, i) V(r, :), 32 rows of W at a time, summed down the columns of L as in
// solve_lower_t_nv (thread: column i0 + t / 8, rows r = i + t % 8, + 8, ...).
__device__ inline void lower_t_times(const double *L, const double *V, double *W, int kp, double (*part)[B + 1]) {
  const int h = threadIdx.x & 7, ic = threadIdx.x >> 3;
  for (int i0 = 0; i0 < kp; i0 += B) {
    double sq[NV] = {0.0, 0.0, 0.0, 0.0};
    const int ci = i0 + ic;
    const double *Lc = L + (int64_t)ci * kp;
#pragma unroll 4
    for (int r = ci + h; r < kp; r += 8) {
      const double l = Lc[r];
#pragma unroll
      for (int q = 0; q < NV; ++q) sq[q] += l * V[r + (int64_t)q * kp];
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) part[h * NV + q][ic] = sq[q];
    __syncthreads();
    if (threadIdx.x < B * NV) {
      const int ii = threadIdx.x & 31, q = threadIdx.x >> 5;
      double t = 0.0;
      for (int hh = 0; hh < 8; ++hh) t += part[hh * NV + q][ii];
      W[(i0 + ii) + (int64_t)q * kp] = t;
    }
    __syncthreads();
  }
}

// tsh = W^T W for the NV columns of W (kp x NV): its ten distinct entries in
// one pass, summed over each wave by shuffles and over the four waves in LDS
// (red10, 40 doubles) in a fixed order; ends with a barrier.
__device__ inline void gram_nv(const double *W, int kp, double *red10, double (*tsh)[NV]) {
  constexpr int NE = NV * (NV + 1) / 2;
  double sm[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) sm[e] = 0.0;
  for (int i = threadIdx.x; i < kp; i += RT) {
    double w[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) w[q] = W[i + (int64_t)q * kp];
    int e = 0;
#pragma unroll
    for (int p = 0; p < NV; ++p)
#pragma unroll
      for (int q = p; q < NV; ++q) sm[e++] += w[p] * w[q];
  }
#pragma unroll
  for (int e = 0; e < NE; ++e)
    for (int o = 32; o > 0; o >>= 1) sm[e] += __shfl_xor(sm[e], o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int e = 0; e < NE; ++e) red10[wave * NE + e] = sm[e];
  __syncthreads();
  if (threadIdx.x == 0) {
    int e = 0;
    for (int p = 0; p < NV; ++p)
      for (int q = p; q < NV; ++q, ++e) {
        const double v = (red10[e] + red10[NE + e]) + (red10[2 * NE + e] + red10[3 * NE + e]);
        tsh[p][q] = tsh[q][p] = v;
      }
  }
  __syncthreads();
}

// modified Gram-Schmidt on the NV columns of V (rows < kp)
__device__ inline void orthonormalize(double *V, int kp, double *red) {
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
// (cyclic Jacobi, one thread) among its first nvalid rows and columns; the
// others (vectors that k < NV dimensions cannot hold: the block iteration
// leaves them zero or rounding noise in the same span, so they would carry a
// zero eigenvalue) are set aside and get y = 0. Returns the eigenvalue.
__device__ inline double small_eig(double (*t)[NV], double *y, int nvalid) {
  for (int i = nvalid; i < NV; ++i)
    for (int j = 0; j < NV; ++j) {
      t[i][j] = t[j][i] = 0.0;
    }
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
  // the set-aside rows stay decoupled (their off-diagonal entries are zero, so
  // no rotation touches them), and the minimum runs over the valid ones
  int m = 0;
  for (int i = 1; i < nvalid; ++i)
    if (t[i][i] < t[m][m]) m = i;
  for (int i = 0; i < NV; ++i) y[i] = i < nvalid ? v[i][m] : 0.0;
  return t[m][m];
}

}  // namespace kdet

#endif  // K_DETFIX_H
