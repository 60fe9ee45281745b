// k_polar.h -- device code shared by the searchlight kernels: workgroup
// reductions, vectorised elementwise passes and the polar iteration.
//
// Matrices are column-major kp x kp (kp = roundup(k, 32)) FP32 buffers or bf16
// hi/lo split planes of the same bytes; rows/cols k..kp-1 are exactly zero.
#ifndef K_POLAR_H
#define K_POLAR_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"
#include "k_gemm.h"

__device__ __forceinline__ double wg_sum_f64(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return r;
}

// Sum of squares of a kp x kp FP32 block (kp a multiple of 32, 16-byte aligned),
// float4 loads four deep per thread so the loop is not latency-bound.
__device__ __forceinline__ double frob2(const float *M, int kp, double *red) {
  const float4 *m4 = reinterpret_cast<const float4 *>(M);
  const int64_t n4 = (int64_t)kp * kp / 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t e = threadIdx.x;
  for (; e + 3 * NT < n4; e += 4 * NT) {
    float4 x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = m4[e + c * NT];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      acc[c] += (double)x[c].x * x[c].x + (double)x[c].y * x[c].y + (double)x[c].z * x[c].z +
                (double)x[c].w * x[c].w;
  }
  for (; e < n4; e += NT) {
    const float4 x = m4[e];
    acc[0] += (double)x.x * x.x + (double)x.y * x.y + (double)x.z * x.z + (double)x.w * x.w;
  }
  return wg_sum_f64((acc[0] + acc[1]) + (acc[2] + acc[3]), red);
}

// x = split(M * f) for a kp x kp FP32 block, float4 loads four deep per thread
__device__ __forceinline__ void split_scaled(const float *M, const SplitM &x, int kp, float f) {
  const float4 *m4 = reinterpret_cast<const float4 *>(M);
  uint2 *h4 = reinterpret_cast<uint2 *>(x.hi), *l4 = reinterpret_cast<uint2 *>(x.lo);
  const int64_t n4 = (int64_t)kp * kp / 4;
  auto put4 = [&](int64_t e, float4 v) {
    const float z[4] = {v.x * f, v.y * f, v.z * f, v.w * f};
    uint16_t h[4], l[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      h[c] = f2bf(z[c]);
      l[c] = f2bf(z[c] - bf2f(h[c]));
    }
    h4[e] = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
    l4[e] = make_uint2((uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16));
  };
  int64_t e = threadIdx.x;
  for (; e + 3 * NT < n4; e += 4 * NT) {
    float4 v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = m4[e + c * NT];
#pragma unroll
    for (int c = 0; c < 4; ++c) put4(e + c * NT, v[c]);
  }
  for (; e < n4; e += NT) put4(e, m4[e]);
}

// ---------------------------------------------------------------------------
// polar iteration
// ---------------------------------------------------------------------------

struct WorkSlots {
  float *A, *X, *X2, *G, *P;
};

__device__ __forceinline__ WorkSlots slot_buffers(const HakLocalArgs &a, int slot) {
  WorkSlots w;
  float *b = a.work + (int64_t)slot * 5 * a.slot_stride;
  w.A = b;
  w.X = b + a.slot_stride;
  w.X2 = b + 2 * a.slot_stride;
  w.G = b + 3 * a.slot_stride;
  w.P = b + 4 * a.slot_stride;
  return w;
}

__device__ __forceinline__ int next_item(int32_t *queue, int32_t *sh) {
  __syncthreads();
  if (threadIdx.x == 0) *sh = atomicAdd(queue, 1);
  __syncthreads();
  return *sh;
}

// x = M * f for a kp x kp FP32 block (FP32 storage of the iterates)
__device__ __forceinline__ void scale_block(const float *M, float *x, int kp, float f) {
  const float4 *m4 = reinterpret_cast<const float4 *>(M);
  float4 *x4 = reinterpret_cast<float4 *>(x);
  const int64_t n4 = (int64_t)kp * kp / 4;
  for (int64_t e = threadIdx.x; e < n4; e += NT) {
    const float4 v = m4[e];
    x4[e] = make_float4(v.x * f, v.y * f, v.z * f, v.w * f);
  }
}

// Storage of a kp x kp iterate in one precision mode: bf16 hi/lo planes for
// bf16x3 products, plain FP32 for FP32 products. KK / KM: the operand kinds
// reading the matrix with k contiguous (column m of the matrix is row m of
// the operand) or with m contiguous.
template <int PREC>
struct MatS;
template <>
struct MatS<PREC_BF16X3> {
  static constexpr int KK = OP_SPLIT_K, KM = OP_SPLIT_M;
  SplitM s;
  __device__ MatS(float *buf, int kp) : s(as_split(buf, kp)) {}
  __device__ __forceinline__ void put(int64_t i, float v) const { put_split(s, i, v); }
  __device__ __forceinline__ float get(int64_t i) const { return get_split(s, i); }
  __device__ __forceinline__ Opnd opnd(int kind, int ld, int mlimit) const { return opnd_split(s, kind, ld, mlimit); }
  __device__ __forceinline__ void load_scaled(const float *A, int kp, float f) const { split_scaled(A, s, kp, f); }
};
template <>
struct MatS<PREC_F32> {
  static constexpr int KK = OP_F32_GATHER, KM = OP_F32_M;
  float *f;
  __device__ MatS(float *buf, int) : f(buf) {}
  __device__ __forceinline__ void put(int64_t i, float v) const { f[i] = v; }
  __device__ __forceinline__ float get(int64_t i) const { return f[i]; }
  __device__ __forceinline__ Opnd opnd(int kind, int ld, int mlimit) const { return opnd_f32(f, kind, ld, mlimit); }
  __device__ __forceinline__ void load_scaled(const float *A, int kp, float v) const { scale_block(A, f, kp, v); }
};

// Result of one polar factorisation: steps taken (all attempts) and whether it converged.
struct PolarOut {
  int steps;
  bool done;
};

// Polar factor Q of the k x k block of w.A (FP32, kp x kp) by the scheduled
// odd-polynomial iteration X <- X (ca I + cb X^T X + cc (X^T X)^2) on bf16x3
// MFMA products, from X0 = A / scale with sigma_min(X0) >= ell assumed. The
// schedule table picks (ca, cb, cc) per step for the bucket of ell; when the
// growth phase ends with a large residual the schedule restarts from the bound
// the residual gives; once ||X^T X - I||_F < tol a final quintic Newton-Schulz
// step writes Q through fin(m, n, v) for every (m, n) of the kp x kp block.
// learned: (scale, ell) came from a prediction; an attempt that does not
// converge is redone once from (fro, plain_ell), which always holds
// sigma_max(X0) <= 1. w.X, w.X2, w.G, w.P are overwritten.
template <int TILE, int PREC, class Fin>
__device__ __forceinline__ PolarOut polar_solve(GemmSmem<TILE> &sm, const HakLocalArgs &a, const WorkSlots &w,
                                                int k, int kp, float fro, float scale, float ell, bool learned,
                                                double *red, float *trace, Fin fin) {
  using M = MatS<PREC>;
  const M G(w.G, kp), P(w.P, kp);
  int step = 0, total = 0;
  bool done = false;
  // A learned attempt whose scale fell below sigma_max diverges within a few
  // steps (the schedules blow up singular values above about 1.01); it is
  // redone once as the plain iteration from A.
  for (int attempt = 0; attempt < (learned ? 2 : 1) && !done; ++attempt) {
    if (attempt == 1) {
      scale = fro;
      ell = a.plain_ell;
    }
    int bucket;
    {
      // (the 1e-3 keeps an ell of exactly 10^(-b/4) in bucket b despite float rounding)
      const int b = (int)ceilf(-log10f(fmaxf(ell, 1e-30f)) * a.bucket_per_decade - 1e-3f);
      bucket = b < 0 ? 0 : (b >= a.n_buckets ? a.n_buckets - 1 : b);
    }
    // X0 = A / scale, stored as bf16 hi/lo planes
    const float inv = 1.0f / scale;
    M(w.X, kp).load_scaled(w.A, kp, inv);
    __syncthreads();
    float *Xb = w.X, *Xnb = w.X2;
    int restart = 0, restarts = 0;
    for (step = 0; step < a.max_steps; ++step) {
      const M X(Xb, kp), Xn(Xnb, kp);
      // G = X^T X and r = ||G - I||_F over the k x k block
      double r2 = 0.0;
      {
        const Opnd ox = X.opnd(M::KK, kp, kp);
        wg_gemm<TILE, M::KK, M::KK, PREC>(sm, ox, ox, kp, kp, true, [&](int m, int n, float v) {
          G.put(m + n * kp, v);
          if (m < k && n < k) {
            const float d = v - (m == n ? 1.f : 0.f);
            r2 += (double)d * d;
          }
        });
      }
      r2 = wg_sum_f64(r2, red);
      const double r = sqrt(r2);
      if (trace && threadIdx.x == 0) trace[step] = (float)r;
      if (!isfinite(r)) {
        ++step;
        break;
      }
      const bool finish = (r < a.tol);
      int ts = step - restart;
      if (!finish && ts >= a.sched_len[bucket] && r > 0.7 && restarts < 2) {
        // The growth phase of the schedule is over but some singular values lag
        // far behind: with r = ||G - I||_F >= 1 - s^2 for every lagging value s,
        // sqrt(1 - r) bounds them from below. Restart from that bucket. Smaller
        // residuals (every s above 0.55) are left to the Newton-Schulz tail.
        const double s_lo = r < 1.0 ? sqrt(1.0 - r) : 1e-3;
        const int b = (int)ceil(-log10(fmax(0.5 * s_lo, 1e-9)) * a.bucket_per_decade);
        bucket = b < 1 ? 1 : (b >= a.n_buckets ? a.n_buckets - 1 : b);
        restart = step;
        ++restarts;
        ts = 0;
      }
      float ca, cb, cc;
      if (finish) {
        // quintic Newton-Schulz: a singular value 1 + d maps to 1 + O(d^3), so the
        // residual tol leaves no error above the rounding of the products
        ca = 1.875f; cb = -1.25f; cc = 0.375f;
      } else if (ts < a.max_sched) {
        const float *co = a.sched + ((int64_t)bucket * a.max_sched + ts) * 3;
        ca = co[0]; cb = co[1]; cc = co[2];
      } else {
        ca = 1.875f; cb = -1.25f; cc = 0.375f;  // quintic Newton-Schulz
      }
      // P = ca I + cb G + cc G^2 (G^2 only when cc != 0), zero outside the k x k block
      if (cc != 0.f) {
        const Opnd og = G.opnd(M::KK, kp, kp);
        wg_gemm<TILE, M::KK, M::KK, PREC>(sm, og, og, kp, kp, true, [&](int m, int n, float v) {
          const int i = m + n * kp;
          float pv = 0.f;
          if (m < k && n < k) pv = (m == n ? ca : 0.f) + cb * G.get(i) + cc * v;
          P.put(i, pv);
        });
      } else {
        for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) {
          const int m = (int)(e % kp), n = (int)(e / kp);
          float pv = 0.f;
          if (m < k && n < k) pv = (m == n ? ca : 0.f) + cb * G.get(e);
          P.put(e, pv);
        }
      }
      __syncthreads();
      // Xn = X P; the finishing product goes to fin(m, n, v)
      {
        const Opnd ox = X.opnd(M::KM, kp, kp);
        const Opnd op = P.opnd(M::KK, kp, kp);
        if (finish) {
          wg_gemm<TILE, M::KM, M::KK, PREC>(sm, ox, op, kp, kp, false, [&](int m, int n, float v) {
            fin(m, n, v);
          });
        } else {
          wg_gemm<TILE, M::KM, M::KK, PREC>(sm, ox, op, kp, kp, false, [&](int m, int n, float v) {
            Xn.put(m + n * kp, v);
          });
        }
      }
      __syncthreads();
      if (finish) {
        done = true;
        ++step;
        break;
      }
      float *t = Xb;
      Xb = Xnb;
      Xnb = t;
    }
    total += step;
  }
  return PolarOut{total, done};
}

#endif  // K_POLAR_H
