// k_template.hip -- searchlight Procrustes template on gfx950.
//
// One 256-thread workgroup per searchlight (persistent, atomic queue) builds
// the searchlight's local template from the subjects' columns in it, following
// ha_template.c:14-114 (spec local_template.py:165-221):
//
//   common = D_0; aligned_0 = D_0
//   for s = 1 .. ns-1:    T = procrustes(D_s, common); aligned_s = z(D_s T)
//                         common = z((common + aligned_s) / 2)
//   level2 times:         sum = sum_s aligned_s
//                         for s: T = procrustes(D_s, z((sum - aligned_s) / (ns-1)))
//                                aligned_s = D_s T
//   local = z(sum_s aligned_s)          (z = one-pass column z-score,
//                                        ha_linalg.c:263-291; without z-scoring
//                                        the plain averages of the C code)
//
// Every Procrustes forms A = D_s^T R (K = time points) and its polar factor
// in FP64 on v_mfma_f64_16x16x4f64 (k_polar64.h), and the product D_s Q on the
// MFMA engine with FP32 products (v_mfma_f32_32x32x2_f32). The chain's
// cross-covariances reach sigma_min / sigma_max of 1e-8 (20 mm, n = 400), where
// an FP32 iteration loses the smallest singular pair: one local template then
// came out 0.9 from the FP64 reference, and FP32 products of A alone left 1e-4.
// A solve the iteration does not finish (a zero or singular A: k >= nt; or
// sigma_min / sigma_max below its schedules' 1e-9) takes the FP64 one-sided
// Jacobi SVD of k_jacobi.h, which completes U over zero singular values.
// Work slot layout (FP32 words, column-major, leading dim ld = roundup(nt, 32),
// kp = roundup(k, 32), padding rows/columns exactly zero):
//   dsp[ns]   subject columns, FP32 ld x kp each
//   al[ns]    aligned subjects, FP32 ld x kp
//   com, sum  FP32 ld x kp; one more ld x kp block is spare
//   A, X, X2, G, P: FP32 kp x kp buffers (kpmax x kpmax each): Q in X2, the
//             reflection fix's work in X, G, P
//   A64, X64, G64, P64: the FP64 iteration's kp x kp buffers
// The local templates (nt x k per searchlight) are combined per vertex by
// k_template_gather with the searchlight weights (searchlight.py:215-236:
// tpl[:, sl] += local * w), in searchlight order, in FP64.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"
#include "k_gemm.h"
#include "k_polar.h"
#include "k_detfix.h"
#include "k_polar64.h"
#include "k_jacobi.h"

#define LAUNCH_RET() return (int)hipGetLastError()

namespace {

constexpr int TT = 64;  // GEMM tile edge of the template kernel (k is about 100 at 20 mm, order 5)

struct TplSlot {
  float *dsp;      // subject s at dsp + s ld kpmax
  float *al;       // subject s at al + s ld kpmax
  float *com, *sum;
  WorkSlots w;
  double *A64, *X64, *G64, *P64;
};

__device__ __forceinline__ TplSlot tpl_slot(const HakTplArgs &t, int slot) {
  const int64_t lk = (int64_t)t.ld * t.kpmax, kk = (int64_t)t.kpmax * t.kpmax;
  float *b = t.work + (int64_t)slot * t.slot_stride;
  TplSlot s;
  s.dsp = b;
  s.al = b + t.ns * lk;
  s.com = s.al + t.ns * lk;
  s.sum = s.com + lk;
  float *pw = s.sum + 2 * lk;
  s.w.A = pw;
  s.w.X = pw + kk;
  s.w.X2 = pw + 2 * kk;
  s.w.G = pw + 3 * kk;
  s.w.P = pw + 4 * kk;
  double *p64 = reinterpret_cast<double *>(pw + 5 * kk);  // 5 kk is even: 8-byte aligned
  s.A64 = p64;
  s.X64 = p64 + kk;
  s.G64 = p64 + 2 * kk;
  s.P64 = p64 + 3 * kk;
  return s;
}

// One-pass column z-score of the nt x k block (ld x kp buffer) in FP64:
// mean = sum/nt, var = sum2/nt - mean^2, columns with var < 1e-30 become 0
// (ha_linalg.c:263-291). One wave per column.
__device__ void zscore_cols(float *B, int nt, int ld, int k) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = wave; j < k; j += NT / 64) {
    float *c = B + (int64_t)j * ld;
    double s1 = 0.0, s2 = 0.0;
    for (int t = lane; t < nt; t += 64) {
      const double x = c[t];
      s1 += x;
      s2 += x * x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      s1 += __shfl_xor(s1, o);
      s2 += __shfl_xor(s2, o);
    }
    const double inv_n = 1.0 / nt;
    const double mean = s1 * inv_n;
    const double var = s2 * inv_n - mean * mean;
    const double is = var < 1e-30 ? 0.0 : 1.0 / sqrt(var);
    for (int t = lane; t < nt; t += 64) c[t] = (float)(((double)c[t] - mean) * is);
  }
  __syncthreads();
}

// LDS of the reflection fix, kept apart from the GEMM staging only by time: the
// four FP64 32 x 33 tiles alias GemmSmem<TT>, which is idle while the fix runs.
struct FixLds {
  float pmax[4];
  int pidx[4], piv, sign, ok;
  double tsh[kdet::NV][kdet::NV], ysh[kdet::NV], lam;
};
static_assert(sizeof(GemmSmem<TT>) >= 4 * sizeof(double) * kchol::B * (kchol::B + 1), "fix tiles fit the staging");

// Reflection fix of one solve (isReflection = false; ha_procrustes.c:56-95):
// when det Q < 0, Q <- Q - 2 (Q v) v^T with v the smallest right singular vector
// of A. The sign comes from an FP32 LU of a copy of Q (|det Q| = 1, so it is
// robust); v from the FP64 Gram matrix A^T A of the FP64 A (not from Q^T A,
// whose rounding would reach v amplified by sigma_max over the gap below
// sigma_min), its Cholesky factor, block inverse iteration and Rayleigh-Ritz,
// as in k_detfix.hip. Work, all free after the iteration (Q is in X2): X (an FP32
// copy of Q for the LU, then the FP64 vectors), G and P (the FP64 Gram matrix).
// Returns false when the Gram matrix is not positive definite (A singular);
// *sigma_min receives the smallest singular value removed from trace(Q^T A)
// twice (0 when nothing is flipped).
template <typename TA>
__device__ bool fix_reflection(GemmSmem<TT> &sm, FixLds &fx, const TplSlot &sl, float *q, const TA *A, int k,
                               int kp, double *red, double *sigma_min) {
  using namespace kdet;
  double (*t0)[B + 1] = reinterpret_cast<double (*)[B + 1]>(&sm);
  double (*t1)[B + 1] = t0 + B, (*t2)[B + 1] = t0 + 2 * B, (*t3)[B + 1] = t0 + 3 * B;
  *sigma_min = 0.0;
  float *M = sl.w.X;
  for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) M[e] = q[e];
  __syncthreads();
  const int sgn = lu_sign(M, k, kp, fx.pmax, fx.pidx, &fx.piv, &fx.sign, reinterpret_cast<float (*)[B + 1]>(t0),
                          reinterpret_cast<float (*)[B + 1]>(t1));
  if (sgn >= 0) return true;
  // G = A^T A in FP64 (lower tiles; identity on the padding diagonal)
  double *G = reinterpret_cast<double *>(sl.w.G);  // G and P: kp^2 doubles
  {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
    for (int tm = 0; tm < kp; tm += B)
      for (int tn = 0; tn <= tm; tn += B) {
        f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
        for (int r0 = 0; r0 < kp; r0 += B) {
          for (int e = threadIdx.x; e < B * B; e += NT) {
            const int c = e >> 5, r = e & 31, row = r0 + r;
            t0[c][r] = (row < k && tm + c < k) ? (double)A[row + (int64_t)(tm + c) * kp] : 0.0;
            t1[c][r] = (row < k && tn + c < k) ? (double)A[row + (int64_t)(tn + c) * kp] : 0.0;
          }
          __syncthreads();
          acc = mfma32(t0, t1, true, acc);  // G(i, j) += sum_r A(r, tm + i) A(r, tn + j)
          __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = tm + wm + (lane >> 4) + 4 * r, n = tn + wn + (lane & 15);
          double v = acc[r];
          if (m == n && m >= k) v = 1.0;
          G[m + (int64_t)n * kp] = v;
        }
      }
    __syncthreads();
  }
  if (!cholesky64(G, kp, t0, t3, t1, t2, &fx.ok)) return false;
  // block inverse iteration on NV vectors (padding rows stay zero), then Rayleigh-Ritz
  double *V = reinterpret_cast<double *>(sl.w.X), *W = V + (int64_t)NV * kp;
  for (int64_t e = threadIdx.x; e < (int64_t)kp * NV; e += NT) {
    const int i = (int)(e % kp), p = (int)(e / kp);
    V[e] = i < k ? sin(0.7548776662466927 * (i + 1) * (p + 1) + 0.5698402909980532 * p) : 0.0;
  }
  __syncthreads();
  orthonormalize(V, kp, red);
  // the chain's cross-covariances can have clustered smallest singular values
  // (0.95 apart in the tests): 6 steps left 2e-3 in the template, 24 reach the
  // FP32 floor (CPU emulation of this chain against the FP64 template)
  constexpr int NITER_TPL = 24;
  for (int step = 0; step < NITER_TPL; ++step) {
    solve_lower_nv(G, V, kp, t0, t1);
    solve_lower_t_nv(G, V, kp, t0, t1);
    orthonormalize(V, kp, red);
  }
  lower_t_times(G, V, W, kp, t1);
  for (int p = 0; p < NV; ++p)
    for (int r = p; r < NV; ++r) {
      double acc = 0.0;
      for (int i = threadIdx.x; i < kp; i += NT) acc += W[i + (int64_t)p * kp] * W[i + (int64_t)r * kp];
      acc = kdet::wg_sum(acc, red);
      if (threadIdx.x == 0) fx.tsh[p][r] = fx.tsh[r][p] = acc;
    }
  __syncthreads();
  if (threadIdx.x == 0) fx.lam = small_eig(fx.tsh, fx.ysh, k < NV ? k : NV);
  __syncthreads();
  // v = V y (W's first column), u = Q v (W's second), Q -= 2 u v^T
  double *vv = W, *uu = W + kp;
  for (int i = threadIdx.x; i < kp; i += NT) {
    double acc = 0.0;
    for (int p = 0; p < NV; ++p) acc += V[i + (int64_t)p * kp] * fx.ysh[p];
    vv[i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kp; i += NT) {
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += (double)q[i + (int64_t)j * kp] * vv[j];
    uu[i] = acc;
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) {
    const int i = (int)(e % kp), j = (int)(e / kp);
    q[e] = (float)((double)q[e] - 2.0 * uu[i] * vv[j]);
  }
  *sigma_min = sqrt(fmax(fx.lam, 0.0));
  __syncthreads();
  return true;
}

// A64 = D^T R (kp x kp, zero outside the k x k block) from the FP32 columns of
// D and R (ld x kp each, leading dim ld, K = ld with zero padding rows), in
// FP64. Not inlined, like the iteration: their registers stay out of the
// kernel's.
__device__ __noinline__ void cross64(const float *D, const float *R, int ld, int k, int kp, double *A64,
                                     kp64::lds_f64 *stg) {
  double *as = (double *)stg, *bs = (double *)stg + kp64::STG;
  kp64::gemm64<true>(D, ld, R, ld, kp, kp, ld, false, false, as, bs, [&](int m, int n, double v) {
    A64[m + (int64_t)n * kp] = (m < k && n < k) ? v : 0.0;
  });
}

// The FP64 Jacobi solve (k_jacobi.h) of one Procrustes problem of the chain,
// for an A the iteration does not finish: its schedules reach sigma_min /
// sigma_max of about 1e-9, and a singular A (k >= nt z-scored samples leave
// rank <= nt - 1; the FP64 A's rounding-level singular values are about 1e-16
// of ||A||) lies far below that; also for an A whose reflection fix found
// A^T A singular. Q = U V^T with U completed over the zero singular values
// (D_s Q does not depend on the completion when the null space of A is that
// of D_s, as for z-scored data of full rank nt - 1), and Q = I for A = 0, as
// LAPACK's SVD gives. Without reflections det Q = +1. A64 is copied to X64
// (then U), G64 holds V, P64 the LU work and A64 the singular values. Returns
// (the sweeps, or -1 with Q = I for a non-finite A; sum(sigma) less twice a
// flipped one, the trace term of the scaling). Not inlined, like the iteration.
__device__ __noinline__ double2 solve_svd64(double *A64, double *X64, double *G64, double *P64, float *q, int k,
                                            int kp, int fix, kp64::lds_f64 *stg, kp64::lds_f64 *redp) {
  A64 = kp64::uni(A64);
  X64 = kp64::uni(X64);
  G64 = kp64::uni(G64);
  P64 = kp64::uni(P64);
  q = kp64::uni(q);
  k = kp64::uni(k);
  kp = kp64::uni(kp);
  fix = kp64::uni(fix);
  double *red = (double *)redp;
  const int64_t kk2 = (int64_t)kp * kp;
  double f2 = 0.0;
  for (int64_t e = threadIdx.x; e < kk2; e += NT) {
    const double v = A64[e];
    X64[e] = v;
    f2 += v * v;
  }
  f2 = wg_sum_f64(f2, red);
  if (!isfinite(f2)) {
    for (int64_t e = threadIdx.x; e < kk2; e += NT) {
      const int m = (int)(e % kp), n = (int)(e / kp);
      q[e] = (m == n && m < k) ? 1.f : 0.f;
    }
    __syncthreads();
    return make_double2(-1.0, 0.0);
  }
  double ss = 0.0;
  const int sw = kjac::svd_polar64(X64, G64, P64, A64, k, kp, fix != 0, q, nullptr,
                                   *reinterpret_cast<kjac::SvdLds *>((double *)stg), red, &ss);
  // sweeps at their limit (sw < 0) are the singular case: the noise columns
  // were shrunk below the rank cut and completed, as in the searchlight FP64
  // path, which takes the same result
  return make_double2((double)(sw > 0 ? sw : -sw), ss);
}
static_assert(sizeof(kjac::SvdLds) <= 2 * kp64::STG * sizeof(double), "the Jacobi solve's LDS fits the staging");

// out = D_s T with T = procrustes(D_s, R) (reflection allowed unless
// t.reflection is 0, optional
// scaling sum(sigma) / sum_j ||D_s[:, j] - mean||^2). Returns the polar steps,
// or -1 if the iteration did not converge (T = I is used then).
template <bool FIX>
__device__ int solve_one(GemmSmem<TT> &sm, FixLds &fx, const HakTplArgs &t, const HakLocalArgs &pa,
                         const TplSlot &sl, int sidx, const float *R, float *out, int k, int kp, double *red,
                         kp64::lds_f64 *stg) {
  const float *D = sl.dsp + (int64_t)sidx * t.ld * t.kpmax;
  // A = D_s^T R over the time points in FP64 (the data are FP32 values), and
  // its polar factor Q by the FP64 iteration; Q goes to X2 in FP32 for the
  // product D_s Q, and trace(Q^T A) in FP64 for the scaling
  const kp64::Sched sc{pa.sched, pa.sched_len, pa.max_sched, pa.n_buckets, pa.max_steps, pa.bucket_per_decade,
                       pa.plain_ell};
  cross64(D, R, t.ld, k, kp, sl.A64, stg);
  // k > nt: A = D_s^T R has rank <= nt < k, so the iteration cannot finish;
  // such a solve goes to the Jacobi solve directly
  int steps = k > t.nt ? -1 : kp64::polar64(sc, sl.A64, sl.X64, sl.G64, sl.P64, k, kp, stg, (kp64::lds_f64 *)red);
  float *q = sl.w.X2;
  double tr = 0.0;
  if (steps >= 0) {
    for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) {
      const double v = sl.X64[e];
      q[e] = (float)v;
      tr += v * sl.A64[e];
    }
    tr = wg_sum_f64(tr, red);
    if (FIX) {
      double smin = 0.0;
      if (fix_reflection(sm, fx, sl, q, sl.A64, k, kp, red, &smin))
        tr -= 2.0 * smin;  // the flipped pair's singular value changes sign in trace(Q^T A)
      else
        steps = -1;        // A^T A not positive definite: the Jacobi solve below
    }
  }
  if (steps < 0) {
    // zero, singular or too ill-conditioned for the iteration's schedules
    const double2 r = solve_svd64(sl.A64, sl.X64, sl.G64, sl.P64, q, k, kp, FIX ? 1 : 0, stg,
                                  (kp64::lds_f64 *)red);
    steps = (int)r.x;
    tr = r.y;
  }
  float f = 1.f;
  if (t.scaling && steps >= 0) {
    // sum_j ||D_s[:, j] - mean_j||^2 from the FP32 subject data, in FP64
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double vs = 0.0;
    for (int j = wave; j < k; j += NT / 64) {
      const float *c = D + (int64_t)j * t.ld;
      double s1 = 0.0, s2 = 0.0;
      for (int i = lane; i < t.nt; i += 64) {
        s1 += c[i];
        s2 += (double)c[i] * c[i];
      }
      for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
      }
      if (lane == 0) vs += s2 - s1 * s1 / t.nt;
    }
    vs = wg_sum_f64(vs, red);
    f = vs > 0.0 ? (float)(tr / vs) : 1.f;
  }
  __syncthreads();
  // out = D_s Q: out^T(j, t) = sum_i Q(i, j) D_s(t, i)
  {
    const Opnd oq = opnd_f32(q, OP_F32_GATHER, kp, k);
    const Opnd odm = opnd_f32(D, OP_F32_M, t.ld, t.ld);
    const int nt = t.nt, ld = t.ld;
    wg_gemm_mn<TT, OP_F32_GATHER, OP_F32_M, PREC_F32>(sm, oq, odm, kp, ld, kp, false, [&](int j, int tt, float4 v) {
      const float z[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int c = 0; c < 4; ++c) out[tt + (int64_t)(j + c) * ld] = (j + c < k && tt < nt) ? z[c] * f : 0.f;
    });
  }
  __syncthreads();
  return steps;
}

// FIX: the instance with the reflection fix (t.reflection == 0); the other one
// is the kernel without it, whose registers it does not touch.
template <bool FIX>
__global__ __launch_bounds__(NT, 2) void k_template(HakTplArgs t, HakLocalArgs pa) {
  __shared__ GemmSmem<TT> sm;
  __shared__ double red[4];
  __shared__ int32_t item_sh;
  __shared__ FixLds fx;
  __shared__ double stg64[2 * kp64::STG];  // the FP64 products' staging
  kp64::lds_f64 *stg = (kp64::lds_f64 *)stg64;
  const TplSlot sl = tpl_slot(t, blockIdx.x);
  for (;;) {
    const int it = next_item(t.queue, &item_sh);
    if (it >= t.count) break;
    const int s = t.order ? t.order[it] : it;
    const int32_t *idx = t.sl_idx + t.sl_off[s];
    const int k = t.sl_off[s + 1] - t.sl_off[s];
    const int kp = (k + 31) / 32 * 32;
    const int ld = t.ld, nt = t.nt, ns = t.ns;
    const int64_t lk = (int64_t)ld * kp, lkmax = (int64_t)ld * t.kpmax;
    // the subjects' columns of this searchlight (subject 0 also starts the common space)
    for (int sub = 0; sub < ns; ++sub) {
      const float *D = t.data + (int64_t)sub * t.nv * ld;
      float *d = sl.dsp + sub * lkmax;
      for (int64_t e = threadIdx.x; e < lk; e += NT) {
        const int tt = (int)(e % ld), i = (int)(e / ld);
        const float v = (i < k && tt < nt) ? D[(int64_t)idx[i] * ld + tt] : 0.f;
        d[e] = v;
        if (sub == 0) {
          sl.al[e] = v;
          sl.com[e] = v;
        }
      }
    }
    __syncthreads();
    int total = 0;
    bool failed = false;
    auto acc_steps = [&](int st) {
      if (st < 0) failed = true;
      else total += st;
    };
    // sequential alignment to the running common space
    for (int sub = 1; sub < ns; ++sub) {
      float *a = sl.al + sub * lkmax;
      acc_steps(solve_one<FIX>(sm, fx, t, pa, sl, sub, sl.com, a, k, kp, red, stg));
      if (t.zscore) zscore_cols(a, nt, ld, k);
      for (int64_t e = threadIdx.x; e < lk; e += NT) sl.com[e] = (sl.com[e] + a[e]) * 0.5f;
      __syncthreads();
      if (t.zscore) zscore_cols(sl.com, nt, ld, k);
    }
    // leave-one-out refinement
    const float inv_n1 = 1.0f / (float)(ns - 1);
    for (int lv = 0; lv < t.level2; ++lv) {
      for (int64_t e = threadIdx.x; e < lk; e += NT) {
        float acc = 0.f;
        for (int sub = 0; sub < ns; ++sub) acc += sl.al[sub * lkmax + e];
        sl.sum[e] = acc;
      }
      __syncthreads();
      for (int sub = 0; sub < ns; ++sub) {
        float *a = sl.al + sub * lkmax;
        for (int64_t e = threadIdx.x; e < lk; e += NT) sl.com[e] = (sl.sum[e] - a[e]) * inv_n1;
        __syncthreads();
        if (t.zscore) zscore_cols(sl.com, nt, ld, k);
        acc_steps(solve_one<FIX>(sm, fx, t, pa, sl, sub, sl.com, a, k, kp, red, stg));
      }
    }
    // local template
    for (int64_t e = threadIdx.x; e < lk; e += NT) {
      float acc = 0.f;
      for (int sub = 0; sub < ns; ++sub) acc += sl.al[sub * lkmax + e];
      sl.com[e] = t.zscore ? acc : acc * (1.0f / (float)ns);
    }
    __syncthreads();
    if (t.zscore) zscore_cols(sl.com, nt, ld, k);
    float *loc = t.loc + t.loc_off[s];
    for (int64_t e = threadIdx.x; e < (int64_t)nt * k; e += NT) {
      const int tt = (int)(e % nt), i = (int)(e / nt);
      loc[e] = sl.com[tt + (int64_t)i * ld];
    }
    if (threadIdx.x == 0) t.status[s] = failed ? -1 : total;
    __syncthreads();
  }
}

__global__ void k_template_gather(const float *__restrict__ loc, const int64_t *__restrict__ loc_off,
                                  const int32_t *__restrict__ sl_off, const int32_t *__restrict__ sl_of,
                                  const int64_t *__restrict__ occ_ptr, const int32_t *__restrict__ occ,
                                  const double *__restrict__ w, int32_t nt, int32_t nv, double *__restrict__ out) {
  for (int v = blockIdx.x; v < nv; v += gridDim.x) {
    const int64_t o0 = occ_ptr[v], o1 = occ_ptr[v + 1];
    for (int tt = threadIdx.x; tt < nt; tt += blockDim.x) {
      double acc = 0.0;
      for (int64_t o = o0; o < o1; ++o) {
        const int32_t f = occ[o];
        const int32_t s = sl_of[f];
        const int32_t i = f - sl_off[s];
        acc += (double)loc[loc_off[s] + tt + (int64_t)i * nt] * w[f];
      }
      out[tt + (int64_t)v * nt] = acc;
    }
  }
}

}  // namespace

extern "C" int64_t hak_template_slot_floats(int32_t ns, int32_t ld, int32_t kpmax) {
  const int64_t lk = (int64_t)ld * kpmax;
  return 2 * (int64_t)ns * lk + 3 * lk + (5 + 8) * (int64_t)kpmax * kpmax;  // FP32 buffers, then 4 FP64 ones
}

extern "C" int hak_template(const HakTplArgs *t, const HakLocalArgs *pa, void *stream) {
  if (t->count <= 0) return 0;
  if (t->reflection)
    hipLaunchKernelGGL(k_template<false>, dim3(t->n_slots), dim3(NT), 0, (hipStream_t)stream, *t, *pa);
  else
    hipLaunchKernelGGL(k_template<true>, dim3(t->n_slots), dim3(NT), 0, (hipStream_t)stream, *t, *pa);
  LAUNCH_RET();
}

extern "C" int hak_template_gather(const float *loc, const int64_t *loc_off, const int32_t *sl_off,
                                   const int32_t *sl_of, const int64_t *occ_ptr, const int32_t *occ, const double *w,
                                   int32_t nt, int32_t nv, double *out, void *stream) {
  if (nv <= 0) return 0;
  const int grid = nv < 65535 ? nv : 65535;
  hipLaunchKernelGGL(k_template_gather, dim3(grid), dim3(256), 0, (hipStream_t)stream, loc, loc_off, sl_off, sl_of,
                     occ_ptr, occ, w, nt, nv, out);
  LAUNCH_RET();
}
