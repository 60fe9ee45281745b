// k_polar64.h -- FP64 products and the scheduled polar iteration in FP64 for
// one 256-thread workgroup, on v_mfma_f64_16x16x4f64. For the template chain
// (k_template.hip): its cross-covariances reach sigma_min / sigma_max of 1e-8,
// below what an FP32 iteration resolves (the polar factor's smallest singular
// pair then comes out wrong, and the local template with it).
//
// gemm64: C = A B over M x N (column-major outputs through an epilogue), in
// 128 x 128 output tiles and K chunks of 16 staged in LDS as S[k][m] (row
// stride 136 doubles: the four k rows of one fragment read fall on distinct
// halves of the banks). Wave (wr, wc) owns the 16 x 16 blocks (wr + 2i,
// wc + 2j), i, j < 4, so the blocks of a smaller tile spread evenly; blocks
// outside M x N skip their MFMAs. The next chunk's values are loaded into
// registers while the current chunk's MFMAs run.
#ifndef K_POLAR64_H
#define K_POLAR64_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

namespace kp64 {

constexpr int NTH = 256;      // threads per workgroup
constexpr int OTL = 128;      // output tile edge
constexpr int KCH = 16;       // K chunk
constexpr int LSTR = OTL + 8;   // LDS row stride (doubles)
constexpr int STG = KCH * LSTR; // doubles per staged operand

typedef double f64x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) double lds_f64;

// A value the same in every lane, as a scalar. The arguments of a function that
// is not inlined arrive in vector registers, so branches on them would be
// per-lane masks; these make the sizes and pointers below scalar.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename T>
__device__ __forceinline__ T *uni(T *p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}

// Operand element (m, k) of a chunk: TR = false: src[m + k ld] (m contiguous);
// TR = true: src[k + m ld] (k contiguous). Rows m >= mlim and k >= klim read 0.
template <bool TR, typename S>
struct Chunk {
  double v[OTL * KCH / NTH];
  __device__ __forceinline__ void load(const S *src, int ld, int m0, int k0, int mlim, int klim) {
    const int t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < OTL * KCH / NTH; ++u) {
      int mm, kk;
      if (TR) {
        kk = t & (KCH - 1);
        mm = (t >> 4) + 16 * u;
      } else {
        mm = t & (OTL - 1);
        kk = (t >> 7) + 2 * u;
      }
      const int m = m0 + mm, k = k0 + kk;
      v[u] = (m < mlim && k < klim) ? (double)src[TR ? (k + (int64_t)m * ld) : (m + (int64_t)k * ld)] : 0.0;
    }
  }
  __device__ __forceinline__ void store(double *s) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < OTL * KCH / NTH; ++u) {
      int mm, kk;
      if (TR) {
        kk = t & (KCH - 1);
        mm = (t >> 4) + 16 * u;
      } else {
        mm = t & (OTL - 1);
        kk = (t >> 7) + 2 * u;
      }
      s[kk * LSTR + mm] = v[u];
    }
  }
};

// C(m, n) = sum_k A(m, k) B(k, n) for m < M, n < N, k < K, handed to
// epi(m, n, value) once per element (m + 16 consecutive lanes: consecutive m).
// A(m, k) = a[m + k lda] (TA = false) or a[k + m lda] (TA = true); B(k, n) =
// b[k + n ldb]. sym (M == N, C symmetric): only the tiles and 16 x 16 blocks on
// and above the diagonal are formed, and the others delivered from them. same:
// B(k, n) = A(n, k) (B = A^T: X^T X, or G G with G symmetric), so a diagonal
// tile's B chunk is its A chunk, staged once. as, bs: STG doubles of LDS each.
// Ends with a barrier.
template <bool TA, typename SA, typename SB, class Epi>
__device__ __forceinline__ void gemm64(const SA *a, int lda, const SB *b, int ldb, int M, int N, int K, bool sym,
                                       bool same, double *as, double *bs, Epi epi) {
  a = uni(a);
  b = uni(b);
  lda = uni(lda);
  ldb = uni(ldb);
  M = uni(M);
  N = uni(N);
  K = uni(K);
  // the wave index as a scalar: the block tests below are then uniform branches
  // (not per-lane masks with an LDS wait before every MFMA)
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  for (int tm = 0; tm < M; tm += OTL)
    for (int tn = sym ? tm : 0; tn < N; tn += OTL) {
      bool bm[4], bn[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        bm[i] = tm + (wr + 2 * i) * 16 < M;
        bn[i] = tn + (wc + 2 * i) * 16 < N;
      }
      // blocks to form: inside M x N, and on or above the diagonal when sym
      const bool diag_tile = sym && tm == tn;
      auto live = [&](int i, int j) { return bm[i] && bn[j] && !(diag_tile && wr + 2 * i > wc + 2 * j); };

      f64x4_t acc[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f64x4_t{0.0, 0.0, 0.0, 0.0};
      Chunk<TA, SA> ca;
      Chunk<true, SB> cb;
      const bool one = same && tm == tn;  // the B chunk is the A chunk
      const double *bsrc = one ? as : bs;
      ca.load(a, lda, tm, 0, M, K);
      if (!one) cb.load(b, ldb, tn, 0, N, K);
      for (int k0 = 0; k0 < K; k0 += KCH) {
        ca.store(as);
        if (!one) cb.store(bs);
        __syncthreads();
        if (k0 + KCH < K) {
          ca.load(a, lda, tm, k0 + KCH, M, K);
          if (!one) cb.load(b, ldb, tn, k0 + KCH, N, K);
        }
#pragma unroll
        for (int s = 0; s < KCH / 4; ++s) {
          double fa[4], fb[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            fa[i] = as[(4 * s + lk) * LSTR + (wr + 2 * i) * 16 + li];
            fb[i] = bsrc[(4 * s + lk) * LSTR + (wc + 2 * i) * 16 + li];
          }
          // operands swapped: lane l then holds C(m0 + l % 16, n0 + l / 16 + 4 r)
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
              if (live(i, j)) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (!live(i, j)) continue;
          const bool mirror = sym && !(diag_tile && wr + 2 * i == wc + 2 * j);
          const int m = tm + (wr + 2 * i) * 16 + li;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int n = tn + (wc + 2 * j) * 16 + lk + 4 * r;
            if (m < M && n < N) {
              epi(m, n, acc[i][j][r]);
              if (mirror) epi(n, m, acc[i][j][r]);
            }
          }
        }
    }
  __syncthreads();
}

__device__ __forceinline__ double wsum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  const double r = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return r;
}

// The schedule of the polar iteration (the FP32 path's table, k_polar.h).
struct Sched {
  const float *co;        // [n_buckets][max_sched][3]
  const int32_t *len;     // steps of each bucket's growth phase
  int max_sched, n_buckets, max_steps;
  float bucket_per_decade, ell;
};

// Polar factor of the k x k block of A (kp x kp, FP64, zero outside the block)
// by the scheduled odd-polynomial iteration X <- X (ca I + cb G + cc G^2), G =
// X^T X, from X0 = A / ||A||_F, all in FP64; the restart rule and the schedule
// of the FP32 path, then quintic Newton-Schulz steps until ||G - I||_F < 1e-3
// and one more quintic step (a singular value 1 + d maps to 1 + O(d^3), so Q is
// then orthogonal to about 1e-10, below the FP32 rounding of its later use).
// X, G, P: kp x kp FP64 buffers; Q ends in X.
// Returns the steps taken, or -1 (A zero or not finite, or no convergence).
__device__ __noinline__ int polar64(Sched sc, const double *A, double *X, double *G, double *P, int k, int kp,
                                    lds_f64 *stg, lds_f64 *redp) {
  double *as = (double *)stg, *bs = (double *)stg + STG, *red = (double *)redp;
  A = uni(A);
  X = uni(X);
  G = uni(G);
  P = uni(P);
  k = uni(k);
  kp = uni(kp);
  const int64_t kk2 = (int64_t)kp * kp;
  double f2 = 0.0;
  for (int64_t e = threadIdx.x; e < kk2; e += NTH) f2 += A[e] * A[e];
  const double fro = sqrt(wsum(f2, red));
  if (!(fro > 0.0) || !isfinite(fro)) return -1;
  const double inv = 1.0 / fro;
  for (int64_t e = threadIdx.x; e < kk2; e += NTH) X[e] = A[e] * inv;
  __syncthreads();
  int bucket;
  {
    const int b = (int)ceil(-log10(fmax((double)sc.ell, 1e-30)) * sc.bucket_per_decade - 1e-3);
    bucket = b < 0 ? 0 : (b >= sc.n_buckets ? sc.n_buckets - 1 : b);
  }
  double *Xb = X, *Gb = G;
  int restart = 0, restarts = 0;
  for (int step = 0; step < sc.max_steps; ++step) {
    // G = X^T X and r = ||G - I||_F over the k x k block
    double r2 = 0.0;
    gemm64<true>(Xb, kp, Xb, kp, kp, kp, kp, true, true, as, bs, [&](int m, int n, double v) {
      Gb[m + (int64_t)n * kp] = v;
      if (m < k && n < k) {
        const double d = v - (m == n ? 1.0 : 0.0);
        r2 += d * d;
      }
    });
    const double r = sqrt(wsum(r2, red));
    if (!isfinite(r)) return -1;
    const bool finish = r < 1e-3;
    int ts = step - restart;
    if (!finish && ts >= sc.len[bucket] && r > 0.7 && restarts < 2) {
      // growth phase over with some singular values far behind: sqrt(1 - r)
      // bounds them from below; restart the schedule from that bucket
      const double s_lo = r < 1.0 ? sqrt(1.0 - r) : 1e-3;
      const int b = (int)ceil(-log10(fmax(0.5 * s_lo, 1e-9)) * sc.bucket_per_decade);
      bucket = b < 1 ? 1 : (b >= sc.n_buckets ? sc.n_buckets - 1 : b);
      restart = step;
      ++restarts;
      ts = 0;
    }
    double ca = 1.875, cb = -1.25, cc = 0.375;  // quintic Newton-Schulz
    if (!finish && ts < sc.max_sched) {
      const float *co = sc.co + ((int64_t)bucket * sc.max_sched + ts) * 3;
      ca = co[0];
      cb = co[1];
      cc = co[2];
    }
    // P = ca I + cb G + cc G^2 on the k x k block, zero outside
    const double *Gr = Gb;
    gemm64<false>(Gr, kp, Gr, kp, kp, kp, kp, true, true, as, bs, [&](int m, int n, double v) {
      const int64_t i = m + (int64_t)n * kp;
      P[i] = (m < k && n < k) ? (m == n ? ca : 0.0) + cb * Gr[i] + cc * v : 0.0;
    });
    // X P goes over G (dead after P), and the two trade roles
    double *Xn = Gb;
    gemm64<false>(Xb, kp, P, kp, kp, kp, kp, false, false, as, bs,
                  [&](int m, int n, double v) { Xn[m + (int64_t)n * kp] = v; });
    Gb = Xb;
    Xb = Xn;
    if (finish) {
      if (Xb != X)
        for (int64_t e = threadIdx.x; e < kk2; e += NTH) X[e] = Xb[e];
      __syncthreads();
      return step + 1;
    }
  }
  return -1;
}

}  // namespace kp64

#endif  // K_POLAR64_H
