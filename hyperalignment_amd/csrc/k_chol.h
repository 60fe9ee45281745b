// k_chol.h -- FP64 building blocks on 32 x 32 LDS tiles for one 256-thread
// workgroup: gathered cross-products over the time points, a blocked
// Cholesky factorization, and blocked triangular solves with kp right-hand
// sides (kp a multiple of 32). Used by the searchlight ridge (k_ridge.hip)
// and the reflection fix of the polar path (k_detfix.hip).
#ifndef K_CHOL_H
#define K_CHOL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

namespace kchol {

constexpr int RT = 256;  // threads per workgroup
constexpr int B = 32;    // block size

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// 32 x 32 product on v_mfma_f64_16x16x4f64: wave w takes the 16 x 16 quadrant
// (rows 16 (w >> 1), columns 16 (w & 1)); lane l supplies A(l % 16, l / 16) and
// B(l / 16, l % 16) of each K = 4 step and holds D(l / 16 + 4 r, l % 16), r < 4.
// acc += sum_c X(i, c) Y(c, j) over c < 32 with X(i, c) = xa[i][c] and
// Y(c, j) = yb(c, j) (yb(c, j) = yt[j][c] when ytr, else yt[c][j]).
static __device__ __forceinline__ f64x4_t mfma32(const double (*xa)[B + 1], const double (*yt)[B + 1], bool ytr,
                                          f64x4_t acc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  const int i = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < B; kk += 4) {
    const double av = xa[wm + i][kk + kq];
    const double bv = ytr ? yt[wn + i][kk + kq] : yt[kk + kq][wn + i];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// out[m + n kp] = sum_t P[cols_p[m] nt + t] Q[cols_q[n] nt + t] over t < nt
// (m, n < k; zero elsewhere) + diag on m == n < k and pad_diag on m == n >= k.
// QDENSE: Q's columns in order with leading dimension ldq (cols_q unused).
// lower: only the tiles on and below the diagonal (the Cholesky reads no others).
template <bool QDENSE = false>
static __device__ void crossprod64(const double *__restrict__ P, const int32_t *cols_p, const double *__restrict__ Q,
                            const int32_t *cols_q, int nt, int k, int kp, double diag, double pad_diag, double *out,
                            bool lower, double (*xs)[B + 1], double (*ys)[B + 1], int64_t ldq = 0) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  for (int tm = 0; tm < kp; tm += B)
    for (int tn = 0; tn < (lower ? tm + B : kp); tn += B) {
      f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
      // each thread stages the same four rows of each operand in every chunk:
      // their offsets once per tile, and the next chunk's values loaded into
      // registers while the current chunk's MFMAs run
      constexpr int U = B * B / RT;
      int64_t ox[U], oy[U];
      double px[U], py[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int c = (threadIdx.x + u * RT) >> 5;
        ox[u] = tm + c < k ? (int64_t)cols_p[tm + c] * nt : -1;
        if constexpr (QDENSE)
          oy[u] = tn + c < k ? (int64_t)(tn + c) * ldq : -1;
        else
          oy[u] = tn + c < k ? (int64_t)cols_q[tn + c] * nt : -1;
      }
      auto load = [&](int t0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = t0 + ((threadIdx.x + u * RT) & 31);
          px[u] = (ox[u] >= 0 && t < nt) ? P[ox[u] + t] : 0.0;
          py[u] = (oy[u] >= 0 && t < nt) ? Q[oy[u] + t] : 0.0;
        }
      };
      load(0);
      for (int t0 = 0; t0 < nt; t0 += B) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = threadIdx.x + u * RT;
          xs[e >> 5][e & 31] = px[u];
          ys[e >> 5][e & 31] = py[u];
        }
        __syncthreads();
        if (t0 + B < nt) load(t0 + B);
        acc = mfma32(xs, ys, true, acc);  // Y(t, n) = ys[n][t]
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm + wm + (lane >> 4) + 4 * r, n = tn + wn + (lane & 15);
        double v = acc[r];
        if (m == n) v += (m < k) ? diag : pad_diag;
        out[m + (int64_t)n * kp] = v;
      }
    }
  __syncthreads();
}

// The row Gram matrix of the nt x k block X_sl = P[:, cols] (column j at
// P + cols[j] nt): out[m + n nq] = sum_{j < k} X_sl(m, j) X_sl(n, j) for m, n
// < nt, + diag on m == n < nt and pad_diag on m == n >= nt, over the tiles on
// and below the diagonal of the nq x nq result (nq = roundup(nt, 32)). Each
// chunk stages 32 columns of the block, read down their rows (coalesced).
static __device__ void gram_rows64(const double *__restrict__ P, const int32_t *cols, int nt, int k, int nq,
                                   double diag, double pad_diag, double *out, double (*xs)[B + 1],
                                   double (*ys)[B + 1]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  constexpr int U = B * B / RT;
  for (int tm = 0; tm < nq; tm += B)
    for (int tn = 0; tn <= tm; tn += B) {
      f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
      double px[U], py[U];
      auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = threadIdx.x + u * RT, r = e & 31, j = j0 + (e >> 5);
          const int64_t base = j < k ? (int64_t)cols[j] * nt : -1;
          px[u] = (base >= 0 && tm + r < nt) ? P[base + tm + r] : 0.0;
          py[u] = (base >= 0 && tn + r < nt) ? P[base + tn + r] : 0.0;
        }
      };
      load(0);
      for (int j0 = 0; j0 < k; j0 += B) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = threadIdx.x + u * RT;
          xs[e & 31][e >> 5] = px[u];  // (row, column of the chunk)
          ys[e & 31][e >> 5] = py[u];
        }
        __syncthreads();
        if (j0 + B < k) load(j0 + B);
        acc = mfma32(xs, ys, true, acc);
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm + wm + (lane >> 4) + 4 * r, n = tn + wn + (lane & 15);
        double v = acc[r];
        if (m == n) v += (m < nt) ? diag : pad_diag;
        out[m + (int64_t)n * nq] = v;
      }
    }
  __syncthreads();
}

// v on lane l, read by every lane of the wave (l uniform).
static __device__ __forceinline__ double lane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __hiloint2double(hi, lo);
}

// cs(:, n) <- D^-1 cs(:, n) (TR: D^-T cs(:, n)) for the 32 right-hand sides n,
// D = d lower triangular. Wave w takes the right-hand sides 8w .. 8w + 7, lane l
// row r = l % 32 of four of them (8w + 4 (l / 32) + 0..3). The rows are solved
// in order (TR: from the last), each solved row reaching the other lanes by
// readlane, so no step waits on LDS or a barrier. Ends with a barrier.
template <bool TR>
static __device__ __forceinline__ void subst32(const double (*d)[B + 1], double (*cs)[B + 1]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, hi = lane >> 5, n0 = wave * 8 + hi * 4;
  double t[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) t[j] = cs[r][n0 + j];
  const double dinv = 1.0 / d[r][r];
#pragma unroll
  for (int s = 0; s < B; ++s) {
    const int ii = TR ? B - 1 - s : s;
    const bool below = TR ? r < ii : r > ii;
    const double l = below ? (TR ? d[ii][r] : d[r][ii]) : 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double y = t[j] * dinv;  // final on the lanes of row ii
      const double a = lane_f64(y, ii), b = lane_f64(y, B + ii);
      const double x = hi ? b : a;
      t[j] = r == ii ? x : t[j] - l * x;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) cs[r][n0 + j] = t[j];
  __syncthreads();
}

static_assert(RT == 4 * 64 && B == 32, "subst32: four waves of eight right-hand sides");

static __device__ __forceinline__ void forward_sub(double (*d)[B + 1], double (*cs)[B + 1]) { subst32<false>(d, cs); }
static __device__ __forceinline__ void backward_sub_t(double (*d)[B + 1], double (*cs)[B + 1]) { subst32<true>(d, cs); }

// In-place lower Cholesky of the kp x kp block G (kp a multiple of 32). Returns
// false if a pivot is not positive (G not positive definite). Left-looking by
// 32-column panels: each block (r0, j0) of the panel first takes its update
// -sum_p L(r0, p) L(j0, p)^T over the final panels p before j0 (one pipelined
// MFMA accumulation, the next block's operands loaded into registers while the
// current one's MFMAs run), then the diagonal block is factored in LDS and the
// blocks below it are solved 32 rows at a time with the rows as right-hand
// sides. Only the lower triangle of G is read.
static __device__ bool cholesky64(double *G, int kp, double (*d)[B + 1], double (*cs)[B + 1], double (*as)[B + 1],
                                  double (*bs)[B + 1], int32_t *ok_sh) {
  constexpr int U = B * B / RT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  if (threadIdx.x == 0) *ok_sh = 1;
  for (int j0 = 0; j0 < kp; j0 += B) {
    for (int r0 = j0; r0 < kp; r0 += B) {
      if (j0 > 0) {
        f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
        double pa[U], pb[U];
        auto load = [&](int p0) {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int e = threadIdx.x + u * RT, lo = e & 31, hi = e >> 5;
            pa[u] = G[(r0 + lo) + (int64_t)(p0 + hi) * kp];  // A(i, c) = L(r0 + i, p0 + c)
            pb[u] = G[(j0 + lo) + (int64_t)(p0 + hi) * kp];  // B(c, n) = L(j0 + n, p0 + c)
          }
        };
        load(0);
        for (int p0 = 0; p0 < j0; p0 += B) {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int e = threadIdx.x + u * RT, lo = e & 31, hi = e >> 5;
            as[lo][hi] = pa[u];
            bs[hi][lo] = pb[u];
          }
          __syncthreads();
          if (p0 + B < j0) load(p0 + B);
          acc = mfma32(as, bs, false, acc);
          __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = wm + (lane >> 4) + 4 * r, n = wn + (lane & 15);
          G[(r0 + i) + (int64_t)(j0 + n) * kp] -= acc[r];
        }
        __syncthreads();
      }
      if (r0 == j0) {
        // diagonal block
        for (int e = threadIdx.x; e < B * B; e += RT) d[e & 31][e >> 5] = G[(j0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp];
        __syncthreads();
        for (int c = 0; c < B; ++c) {
          if (threadIdx.x == 0) {
            const double p = d[c][c];
            if (!(p > 0.0)) *ok_sh = 0;
            d[c][c] = sqrt(p > 0.0 ? p : 1.0);
          }
          __syncthreads();
          for (int r = c + 1 + threadIdx.x; r < B; r += RT) d[r][c] /= d[c][c];
          __syncthreads();
          const int n = B - c - 1;
          for (int e = threadIdx.x; e < n * n; e += RT) {
            const int r = c + 1 + e % n, q = c + 1 + e / n;
            if (q <= r) d[r][q] -= d[r][c] * d[q][c];
          }
          __syncthreads();
        }
        for (int e = threadIdx.x; e < B * B; e += RT) {
          const int r = e & 31, q = e >> 5;
          G[(j0 + r) + (int64_t)(j0 + q) * kp] = q <= r ? d[r][q] : 0.0;
        }
        __syncthreads();
      } else {
        // L(r0, j0) = G(r0, j0) L11^-T: L11 L(r0, :)^T = G(r0, :)^T, the 32 rows as right-hand sides
        for (int e = threadIdx.x; e < B * B; e += RT) cs[e >> 5][e & 31] = G[(r0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp];
        __syncthreads();
        forward_sub(d, cs);
        for (int e = threadIdx.x; e < B * B; e += RT) G[(r0 + (e & 31)) + (int64_t)(j0 + (e >> 5)) * kp] = cs[e >> 5][e & 31];
        __syncthreads();
      }
    }
  }
  __syncthreads();
  return *ok_sh != 0;
}

// C <- L^-1 C (trans false) or C <- L^-T C (trans true), L lower triangular
// kp x kp and C kp x nc (both multiples of 32; nc = 0: kp) with leading
// dimension ldc (0: kp), left-looking by 32 x 32 blocks: block
// (r0, n0) of the solution is X = D^-1 (C(r0, n0) - sum_j S(j) X(j, n0)) with
// D the diagonal block of L (D^T when trans), S(j) = L(r0, j) over the blocks j
// before r0 (L(j, r0)^T over the blocks after it when trans). The sum runs on
// MFMA tiles with the next block's operands loaded into registers while the
// current one's MFMAs run; blocks go in the order that leaves every X(j, n0)
// final before it is read, and X overwrites C.
static __device__ void trsm_ll(const double *L, double *C, int kp, bool trans, double (*d)[B + 1], double (*cs)[B + 1],
                               double (*as)[B + 1], double (*bs)[B + 1], int nc = 0, int64_t ldc = 0) {
  constexpr int U = B * B / RT;
  const int ncol = nc ? nc : kp;       // C: kp x ncol (a multiple of 32), leading dimension ldc
  const int64_t lc = ldc ? ldc : kp;
  const int nb = kp / B;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  for (int bi = 0; bi < nb; ++bi) {
    const int r0 = (trans ? nb - 1 - bi : bi) * B;
    for (int e = threadIdx.x; e < B * B; e += RT) d[e & 31][e >> 5] = L[(r0 + (e & 31)) + (int64_t)(r0 + (e >> 5)) * kp];
    const int jbeg = trans ? r0 + B : 0, jend = trans ? kp : r0;
    for (int n0 = 0; n0 < ncol; n0 += B) {
      f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
      double pa[U], pb[U];
      auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = threadIdx.x + u * RT, lo = e & 31, hi = e >> 5;
          // A(i, c): L(r0 + i, j0 + c), or L(j0 + c, r0 + i) when trans; both read down a column
          pa[u] = trans ? L[(j0 + lo) + (int64_t)(r0 + hi) * kp] : L[(r0 + lo) + (int64_t)(j0 + hi) * kp];
          pb[u] = C[(j0 + lo) + (int64_t)(n0 + hi) * lc];  // B(c, n) = X(j0 + c, n0 + n)
        }
      };
      if (jbeg < jend) load(jbeg);
      for (int j0 = jbeg; j0 < jend; j0 += B) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = threadIdx.x + u * RT, lo = e & 31, hi = e >> 5;
          if (trans)
            as[hi][lo] = pa[u];  // (i, c) = (hi, lo)
          else
            as[lo][hi] = pa[u];  // (i, c) = (lo, hi)
          bs[lo][hi] = pb[u];
        }
        __syncthreads();
        if (j0 + B < jend) load(j0 + B);
        acc = mfma32(as, bs, false, acc);
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm + (lane >> 4) + 4 * r, n = wn + (lane & 15);
        cs[i][n] = C[(r0 + i) + (int64_t)(n0 + n) * lc] - acc[r];
      }
      __syncthreads();
      if (trans)
        backward_sub_t(d, cs);
      else
        forward_sub(d, cs);
      for (int e = threadIdx.x; e < B * B; e += RT) C[(r0 + (e & 31)) + (int64_t)(n0 + (e >> 5)) * lc] = cs[e & 31][e >> 5];
      __syncthreads();
    }
  }
}

}  // namespace kchol

#endif  // K_CHOL_H
