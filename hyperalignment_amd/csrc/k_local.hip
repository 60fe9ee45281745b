// k_local.hip -- local transforms of searchlight Procrustes on gfx950.
//
// For every searchlight s with vertex lists slx, sly (k entries each):
//   A_s = X[:, slx]^T Y[:, sly]             (k x k, inner dimension nt)
//   Q_s = polar factor of A_s = U V^T       (A = U S V^T)
// Reference: csrc/ha_searchlight.c:578-659 (A = X^T Y, dgesdd, T = U Vt),
// csrc/ha_procrustes.c:17-127.
//
// Three solvers share the persistent one-workgroup-per-searchlight layout:
//  * hak_polar   : bf16x3 MFMA cross-covariance, then the odd-polynomial polar
//                  iteration X <- X (a I + b G + c G^2), G = X^T X, every GEMM in
//                  bf16x3 (hi*hi + lo*hi + hi*lo, FP32 accumulation) on
//                  v_mfma_f32_32x32x8_bf16, finishing with one Newton-Schulz
//                  step X (3I - G)/2 once ||G - I||_F < tol.
//  * hak_jacobi32: the plain FP32 one-sided (Hestenes) Jacobi SVD, Q = U V^T.
//  * hak_jacobi64: FP64 one-sided Jacobi from an FP64 cross-covariance, used as
//                  fallback for rank-deficient / ill-conditioned / unconverged
//                  searchlights, and for the det-sign (reflection) correction.
//
// Matrices live in per-workgroup slot buffers in HBM, column-major with leading
// dimension kp = roundup(k, 32); rows/cols k..kp-1 are kept exactly zero.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"
#include "k_mlp.h"

#define LAUNCH_RET() return (int)hipGetLastError()

#include "k_gemm.h"
#include "k_polar.h"
#include "k_jacobi.h"

using namespace kjac;

// cross-covariance A = X[:, slx]^T Y[:, sly] into the FP32 buffer Ab (kp x kp)
template <int TILE>
__device__ void crosscov(GemmSmem<TILE> &sm, const HakLocalArgs &a, int s, int k, int kp, float *Ab) {
  const int o = a.sl_off[s];
  // the pre-split planes of X and Y: the same products as on the FP32 data,
  // without splitting every chunk while staging it
  Opnd oa = opnd_gather_split(a.Xh, a.Xl, a.slx_idx + o, a.ld, k);
  Opnd ob = opnd_gather_split(a.Yh, a.Yl, a.sly_idx + o, a.ld, k);
  wg_gemm<TILE, OP_SPLIT_GATHER, OP_SPLIT_GATHER>(sm, oa, ob, kp, a.ld, false, [&](int m, int n, float4 v) {
    *reinterpret_cast<float4 *>(Ab + m + (int64_t)n * kp) = v;
  });
}

// The same cross-covariance with FP32 operands and products (v_mfma_f32_32x32x2_f32),
// for the nearly rank-deficient searchlights (HAK_NEAR_DEFICIENT): there A is
// as ill-conditioned as the samples allow, and the bf16x3 products' 2^-16
// relative error in A reaches the polar factor multiplied by that conditioning.
// Out of line, so the registers of the bf16x3 kernel stay as they are.
template <int TILE>
__device__ __noinline__ void crosscov_f32(GemmSmem<TILE> &sm, const HakLocalArgs &a, int s, int k, int kp,
                                          float *Ab) {
  const int o = a.sl_off[s];
  Opnd oa = opnd_gather(a.X, a.slx_idx + o, a.ld, k);
  Opnd ob = opnd_gather(a.Y, a.sly_idx + o, a.ld, k);
  wg_gemm<TILE, OP_F32_GATHER, OP_F32_GATHER, PREC_F32>(sm, oa, ob, kp, a.ld, false, [&](int m, int n, float4 v) {
    *reinterpret_cast<float4 *>(Ab + m + (int64_t)n * kp) = v;
  });
}

// Learned-scaling features of A (12 per searchlight), written when a.features != NULL.
// fro2 = ||A||_F^2 as frob2 gives it.
__device__ __noinline__ void features_of(const HakLocalArgs &a, int s, int k, int kp, const float *Ab, float *Pv,
                                         float *feat_sh, double *red, double fro2) {
  // trace, diagonal statistics
  double tr = 0.0, dabs = 0.0, d2 = 0.0;
  float dmin = 3.0e38f, dmax = 0.f;
  for (int i = threadIdx.x; i < k; i += NT) {
    float d = Ab[i + (int64_t)i * kp];
    tr += d;
    dabs += fabsf(d);
    d2 += (double)d * d;
  }
  tr = wg_sum_f64(tr, red);
  dabs = wg_sum_f64(dabs, red);
  d2 = wg_sum_f64(d2, red);
  // min / max |diag| via the same reduction scratch (sequential over waves)
  for (int i = threadIdx.x; i < k; i += NT) {
    float d = fabsf(Ab[i + (int64_t)i * kp]);
    dmin = fminf(dmin, d);
    dmax = fmaxf(dmax, d);
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) {
    dmin = fminf(dmin, __shfl_xor(dmin, o2));
    dmax = fmaxf(dmax, __shfl_xor(dmax, o2));
  }
  __shared__ float mm_sh[8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { mm_sh[wave] = dmin; mm_sh[4 + wave] = dmax; }
  __syncthreads();
  dmin = fminf(fminf(mm_sh[0], mm_sh[1]), fminf(mm_sh[2], mm_sh[3]));
  dmax = fmaxf(fmaxf(mm_sh[4], mm_sh[5]), fmaxf(mm_sh[6], mm_sh[7]));
  __syncthreads();
  // power iteration on A^T A with v0 = ones: estimates after 2, 4, 8 steps.
  // A (kp x kp, column-major, zero outside k x k, kp a multiple of 32) is read
  // as float4 row quads, four columns per group and FPG groups per batch: a
  // lane keeps up to 8 FPG independent 16-byte loads in flight (the passes are
  // bound by memory latency), and sums in the order of one group at a time.
  // u = A v splits the columns over the four waves (partial sums per wave);
  // v = A^T u gives each wave whole columns.
  float *v = Pv, *u = Pv + kp, *part = Pv + 2 * kp;  // part: 4 x kp partial sums
  for (int i = threadIdx.x; i < kp; i += NT) v[i] = (i < k) ? 1.0f : 0.f;
  __syncthreads();
  double est[3] = {0, 0, 0};
  int ei = 0;
  const int nq = kp / 4;  // row quads per column
  constexpr int FPG = 2;
  for (int it = 1; it <= 8; ++it) {
    // u = A v: lane owns row quads q = lane + 64 c
    for (int q0 = 0; q0 < nq; q0 += 128) {
      float4 acc[2] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
      const bool in0 = q0 + lane < nq, in1 = q0 + 64 + lane < nq;
      for (int jb = wave * 4; jb < k; jb += 16 * FPG) {
        // columns j >= k read column 0 (in bounds) and weigh it by 0
        float4 x0[FPG][4], x1[FPG][4];
        float vj[FPG][4];
#pragma unroll
        for (int g = 0; g < FPG; ++g)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = jb + 16 * g + c;
            vj[g][c] = (j < k) ? v[j] : 0.f;
            const float4 *col = reinterpret_cast<const float4 *>(Ab + (int64_t)(j < k ? j : 0) * kp);
            x0[g][c] = in0 ? col[q0 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            x1[g][c] = in1 ? col[q0 + 64 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
          }
#pragma unroll
        for (int g = 0; g < FPG; ++g)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            acc[0].x += x0[g][c].x * vj[g][c]; acc[0].y += x0[g][c].y * vj[g][c];
            acc[0].z += x0[g][c].z * vj[g][c]; acc[0].w += x0[g][c].w * vj[g][c];
            acc[1].x += x1[g][c].x * vj[g][c]; acc[1].y += x1[g][c].y * vj[g][c];
            acc[1].z += x1[g][c].z * vj[g][c]; acc[1].w += x1[g][c].w * vj[g][c];
          }
      }
      float4 *pw = reinterpret_cast<float4 *>(part + wave * kp);
      if (in0) pw[q0 + lane] = acc[0];
      if (in1) pw[q0 + 64 + lane] = acc[1];
    }
    __syncthreads();
    double nu = 0.0;
    for (int i = threadIdx.x; i < kp; i += NT) {
      const float x = (part[i] + part[kp + i]) + (part[2 * kp + i] + part[3 * kp + i]);
      u[i] = x;
      nu += (double)x * x;
    }
    nu = wg_sum_f64(nu, red);  // its barriers publish u
    // v = A^T u: wave w takes columns w*4 .. w*4+3, then +16
    double nv = 0.0;
    for (int jb = wave * 4; jb < k; jb += 16 * FPG) {
      float dot[FPG][4];
#pragma unroll
      for (int g = 0; g < FPG; ++g)
#pragma unroll
        for (int c = 0; c < 4; ++c) dot[g][c] = 0.f;
      for (int q0 = 0; q0 < nq; q0 += 64) {
        const int q = q0 + lane;
        const float4 uq = (q < nq) ? reinterpret_cast<const float4 *>(u)[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        float4 x[FPG][4];
#pragma unroll
        for (int g = 0; g < FPG; ++g)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = jb + 16 * g + c;
            const float4 *col = reinterpret_cast<const float4 *>(Ab + (int64_t)(j < k ? j : 0) * kp);
            x[g][c] = (q < nq && j < k) ? col[q] : make_float4(0.f, 0.f, 0.f, 0.f);
          }
#pragma unroll
        for (int g = 0; g < FPG; ++g)
#pragma unroll
          for (int c = 0; c < 4; ++c)
            dot[g][c] += x[g][c].x * uq.x + x[g][c].y * uq.y + x[g][c].z * uq.z + x[g][c].w * uq.w;
      }
#pragma unroll
      for (int g = 0; g < FPG; ++g)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          float d = dot[g][c];
          for (int o2 = 32; o2 > 0; o2 >>= 1) d += __shfl_xor(d, o2);
          const int j = jb + 16 * g + c;
          if (lane == 0 && j < k) {
            v[j] = d;
            nv += (double)d * d;
          }
        }
    }
    nv = wg_sum_f64(nv, red);
    // |A^T A v| / |A v| with |v| = 1 tends to sigma_max from below
    double sig = (nu > 0) ? sqrt(nv / nu) : 0.0;
    if (it == 2 || it == 4 || it == 8) est[ei++] = sig;
    double inv = (nv > 0) ? 1.0 / sqrt(nv) : 0.0;
    for (int j = threadIdx.x; j < k; j += NT) v[j] = (float)(v[j] * inv);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double fro = sqrt(fro2) + 1e-30;
    double sk = fro / sqrt((double)k);
    feat_sh[0] = (float)log((double)k);
    feat_sh[1] = (float)log(fro);
    feat_sh[2] = (float)(tr / fro);
    feat_sh[3] = (float)(dabs / k / sk);
    feat_sh[4] = (float)(dmin / sk);
    feat_sh[5] = (float)(dmax / sk);
    double dm = dabs / k;
    double dv = d2 / k - dm * dm;
    feat_sh[6] = (float)(sqrt(dv > 0 ? dv : 0) / sk);
    feat_sh[7] = (float)log(est[0] / fro + 1e-30);
    feat_sh[8] = (float)log(est[1] / fro + 1e-30);
    feat_sh[9] = (float)log(est[2] / fro + 1e-30);
    feat_sh[10] = (float)log((double)a.nt);
    feat_sh[11] = (float)(d2 / (fro2 + 1e-30));
  }
  __syncthreads();
}

// predictions of the MLP for one searchlight (out of line: runs once per searchlight)
__device__ __noinline__ void mlp_predict(const HakLocalArgs &a, const float *feat_sh, float *mlp_sh, float *pred) {
  hak_mlp_forward_wg(a.mlp, a.mlp_norm, feat_sh, mlp_sh, pred);
}

// Stage 1 of the polar path: per searchlight, the gathered cross-covariance
// A = X[:, sl]^T Y[:, sl] (bf16x3 MFMA) written into the searchlight's place in
// the Q store, ||A||_F, and the starting scale and schedule bound: plain
// (scale = ||A||_F, ell = plain_ell) or learned (features + MLP prediction).
// sl_par[3 s] = (scale, ell, ||A||_F); scale = 0 marks a searchlight left to
// the FP64 path (k >= nt, or A zero / non-finite).
template <int TILE>
__global__ __launch_bounds__(NT, 2) void k_crosscov(HakLocalArgs a) {
  __shared__ GemmSmem<TILE> sm;
  __shared__ double red[4];
  __shared__ int32_t item_sh;
  __shared__ float feat_sh[HAK_NFEAT];
  __shared__ float mlp_sh[2 * 64 + 2];
  const int slot = blockIdx.x;
  float *scratch = a.work + (int64_t)slot * a.slot_stride;  // features: 6 kp floats
  for (;;) {
    const int it = next_item(a.queue, &item_sh);
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    const int k = a.sl_off[s + 1] - a.sl_off[s];
    const int kp = a.qld[s];
    float *A = a.q + a.qoff[s];
    // k >= nt: z-scored columns leave A of rank <= nt - 1 < k, so its polar
    // factor is not unique and its smallest singular values are rounding noise
    // (a few 1e-9 of ||A||_F). The plain scale with the schedule of the smallest
    // bucket takes that noise to 1 as well: one orthogonal completion, a valid
    // Procrustes solution. Nearly square searchlights (k / (nt - 1) >= 0.9,
    // HAK_NEAR_DEFICIENT) are full rank but about as ill-conditioned as the
    // data allow, and take the same route, with A itself from FP32 products
    // (bf16x3 products measured 1.4e-4 from the FP64 factor there). One that
    // does not converge goes to the FP64 path.
    const bool deficient = a.rank_check_nt && HAK_NEAR_DEFICIENT(k, a.nt);
    const uint64_t t0 = a.clk ? wall_clock64() : 0;
    if (deficient)
      crosscov_f32<TILE>(sm, a, s, k, kp, A);
    else
      crosscov<TILE>(sm, a, s, k, kp, A);
    __syncthreads();
    const uint64_t t1 = a.clk ? wall_clock64() : 0;
    const double fro2 = frob2(A, kp, red);
    const float fro = (float)sqrt(fro2);
    // plain: X0 = A / ||A||_F (sigma_max(X0) <= 1 always), sigma_min(X0) >= plain_ell
    // assumed (0 for a rank-deficient A: the smallest bucket's schedule)
    float scale = fro, ell = deficient ? 0.f : a.plain_ell;
    if (!(fro > 0.f) || !isfinite(fro)) {
      scale = 0.f;
      if (threadIdx.x == 0) {
        a.flag[s] = 1;
        if (a.steps) a.steps[s] = 0;
      }
    } else if (deficient) {
      if (a.pred_out && threadIdx.x == 0) {  // no prediction is made for it
        a.pred_out[2 * s] = __int_as_float(0x7fc00000);
        a.pred_out[2 * s + 1] = __int_as_float(0x7fc00000);
      }
    } else if (a.features || a.mlp) {
      features_of(a, s, k, kp, A, scratch, feat_sh, red, fro2);
      if (a.features && threadIdx.x < HAK_NFEAT)
        a.features[(int64_t)s * HAK_NFEAT + threadIdx.x] = feat_sh[threadIdx.x];
      if (a.mlp) {
        float pred[2];
        mlp_predict(a, feat_sh, mlp_sh, pred);
        // pred[0] = ln(sigma_max / ||A||_F), pred[1] = log10(sigma_min / sigma_max).
        // scale = predicted sigma_max raised by its margin (never above ||A||_F);
        // ell = predicted sigma_min / scale lowered by its margin. Outside the
        // sizes the model was trained on its outputs can be anything, so the
        // prediction is first held to what A itself proves: sigma_max lies in
        // [max(power estimate, ||A||_F / sqrt k), ||A||_F] (the power-iteration
        // estimate of feature 9 is a lower bound), sigma_min <= sigma_max; a
        // non-finite prediction leaves the plain start.
        if (isfinite(pred[0]) && isfinite(pred[1])) {
          const float lo = fmaxf(feat_sh[9], -0.5f * logf((float)k));
          const float p0 = fminf(fmaxf(pred[0], lo), 0.f);
          const float lf = fminf(p0 + a.mlp_margin_scale, 0.f);
          scale = fro * expf(lf);
          ell = exp10f(fminf(pred[1], 0.f) - a.mlp_margin_ell) * expf(p0 - lf);
        }
        if (a.pred_out && threadIdx.x == 0) {
          a.pred_out[2 * s] = pred[0];
          a.pred_out[2 * s + 1] = pred[1];
        }
      }
    }
    if (threadIdx.x == 0) {
      a.sl_par[3 * s] = scale;
      a.sl_par[3 * s + 1] = ell;
      a.sl_par[3 * s + 2] = fro;
      if (a.clk) {
        const uint64_t t2 = wall_clock64();
        a.clk[4 * slot + 0] += t1 - t0;  // gather + cross-covariance
        a.clk[4 * slot + 1] += t2 - t1;  // scale, features, MLP
        a.clk[4 * slot + 3] += 1;
      }
    }
  }
}

// Stage 2: the polar iteration of every searchlight stage 1 kept. A is read
// from the Q store and Q overwrites it in the finishing product (each element
// is read before the same thread writes it). Without the gather and the
// learned prologue the kernel fits HAK_POLAR_OCC workgroups per CU, so that
// many slots' operand loads are in flight per CU.
template <int TILE, int PREC>
__global__ __launch_bounds__(NT, TILE == 128 ? HAK_POLAR_OCC128 : HAK_POLAR_OCC) void k_polar(HakLocalArgs a) {
  __shared__ GemmSmem<TILE> sm;
  __shared__ double red[4];
  __shared__ int32_t item_sh;
  const int slot = blockIdx.x;
  WorkSlots w;
  {
    float *b = a.work + (int64_t)slot * 4 * a.slot_stride;
    w.X = b;
    w.X2 = b + a.slot_stride;
    w.G = b + 2 * a.slot_stride;
    w.P = b + 3 * a.slot_stride;
  }
  for (;;) {
    const int it = next_item(a.queue, &item_sh);
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    const float scale = a.sl_par[3 * s];
    if (!(scale > 0.f)) continue;  // left to the FP64 path by stage 1
    if (a.skip_done && a.steps[s] != 0) continue;  // finished (or given up) by the team kernel
    if (a.skip_done && threadIdx.x == 0 && a.team_count) atomicAdd(&a.team_count[1], 1);  // taken over
    const float ell = a.sl_par[3 * s + 1], fro = a.sl_par[3 * s + 2];
    const int k = a.sl_off[s + 1] - a.sl_off[s];
    const int kp = a.qld[s];
    float *Q = a.q + a.qoff[s];
    w.A = Q;
    const uint64_t t0 = a.clk ? wall_clock64() : 0;
    double tr = 0.0;
    const bool want_tr = a.sigma_sum != nullptr;
    SoloTeam team;
    const PolarOut po = polar_solve<TILE, PREC>(sm, a, w, k, kp, fro, scale, ell, a.learned_scale != 0,
                                                       red, a.trace ? a.trace + (int64_t)s * a.max_steps : nullptr,
                                                       team, [&](int m, int n, float4 v) {
                                                         float4 *q4 = reinterpret_cast<float4 *>(Q + m + (int64_t)n * kp);
                                                         if (want_tr) {
                                                           const float4 o = *q4;
                                                           tr += (double)v.x * o.x + (double)v.y * o.y +
                                                                 (double)v.z * o.z + (double)v.w * o.w;
                                                         }
                                                         *q4 = v;
                                                       });
    if (want_tr) tr = wg_sum_f64(tr, red);
    if (threadIdx.x == 0) {
      if (!po.done) a.flag[s] = 1;  // budget exhausted or non-finite: the FP64 pass writes Q
      else if (want_tr) a.sigma_sum[s] = tr;
      if (a.steps) a.steps[s] = po.done ? po.steps : -po.steps;
      if (a.clk) a.clk[4 * slot + 2] += wall_clock64() - t0;  // iteration
    }
    __syncthreads();
  }
}

// Stage 2 with teams: HAK_GRP workgroups of one XCD (workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup b runs on XCD b % 8) iterate one
// searchlight together, sharing panel reads through that XCD's L2. The
// finishing product writes Q over A in the Q store. Before it starts, the
// searchlight is marked for the FP64 pass (flag 1, steps -1, published by the
// barrier ahead of the product), and the mark is cleared once every member has
// written its tiles: a team that gives up part way leaves a marked searchlight,
// and one that gives up earlier leaves A intact and steps 0 for the
// one-workgroup kernel that follows. Two instances: 128 x 128 bf16x3 tiles for
// the 128-tile searchlights, 64 x 64 FP32 tiles for the rank-deficient ones.
template <int TILE, int PREC>
__global__ __launch_bounds__(NT, TILE == 128 ? HAK_POLAR_OCC128 : HAK_POLAR_OCC) void k_polar_grp(HakLocalArgs a) {
  __shared__ GemmSmem<TILE> sm;
  __shared__ double red[4];
  __shared__ int ok_sh;
  const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int team_id = (j / HAK_GRP) * 8 + xcd, member = j % HAK_GRP;
  GrpState *st = reinterpret_cast<GrpState *>(a.grp) + team_id;
  // wait budget: grp_spin microseconds (100 ticks of the 100 MHz clock each)
  GroupTeam team{st, member, 0u, 100ull * (a.grp_spin ? a.grp_spin : HAK_GRP_WAIT_US), &ok_sh};
  // counts this team once in team_count[0] if its barrier wait runs out
  auto gave_up = [&]() {
    if (member == 0 && threadIdx.x == 0 && a.team_count) atomicAdd(&a.team_count[0], 1);
  };
  WorkSlots w;
  {
    float *b = a.work + (int64_t)team_id * 4 * a.slot_stride;
    w.X = b;
    w.X2 = b + a.slot_stride;
    w.G = b + 2 * a.slot_stride;
    w.P = b + 3 * a.slot_stride;
  }
  const bool lead = member == 0 && threadIdx.x == 0;
  // the barrier relies on the members sharing one L2: each publishes its XCD
  if (threadIdx.x == 0) {
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    __hip_atomic_fetch_or(&st->xcc_mask, 1u << (xcc & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!team.sync()) {
    gave_up();
    return;
  }
  if (__popc(__hip_atomic_load(&st->xcc_mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 1)
    return;  // members on different XCDs: every searchlight is left to the one-workgroup kernel
  // Turn t takes its searchlight from item[t & 1]: a member that skips a
  // searchlight without a barrier (scale <= 0) cannot overwrite the slot another
  // member has yet to read, since that slot is written again only after the
  // next turn's barrier, which every member reaches after its read.
  for (int turn = 0;; ++turn) {
    if (lead) st->item[turn & 1] = atomicAdd(a.queue, 1);
    if (!team.sync()) {
      gave_up();
      break;
    }
    const int it = __hip_atomic_load(&st->item[turn & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    const float scale = a.sl_par[3 * s];
    if (!(scale > 0.f)) continue;  // left to the FP64 path by stage 1
    const float ell = a.sl_par[3 * s + 1], fro = a.sl_par[3 * s + 2];
    const int k = a.sl_off[s + 1] - a.sl_off[s];
    const int kp = a.qld[s];
    float *Q = a.q + a.qoff[s];
    w.A = Q;
    const bool want_tr = a.sigma_sum != nullptr;
    double tr = 0.0;
    // Q over A: each element of A is read (for tr(Q^T A)) and then replaced by
    // the same thread, and A is not read again for this searchlight
    const PolarOut po = polar_solve<TILE, PREC>(
        sm, a, w, k, kp, fro, scale, ell, a.learned_scale != 0, red,
        a.trace ? a.trace + (int64_t)s * a.max_steps : nullptr, team,
        [&](int m, int n, float4 v) {
          float4 *q4 = reinterpret_cast<float4 *>(Q + m + (int64_t)n * kp);
          if (want_tr) {
            const float4 o = *q4;
            tr += (double)v.x * o.x + (double)v.y * o.y + (double)v.z * o.z + (double)v.w * o.w;
          }
          *q4 = v;
        },
        [&]() {
          if (lead) {
            a.flag[s] = 1;
            a.steps[s] = -1;
          }
        });
    if (po.aborted) {
      gave_up();
      break;
    }
    if (!po.done) {
      if (lead) {
        a.flag[s] = 1;  // budget exhausted or non-finite: the FP64 pass writes Q
        a.steps[s] = -po.steps;
      }
      continue;
    }
    // done: polar_solve's last barrier followed every member's Q tiles
    if (want_tr) {
      bool ok;
      tr = team.sum(tr, red, &ok);
      if (!ok) {
        gave_up();
        break;
      }
    }
    if (lead) {
      a.flag[s] = 0;
      if (want_tr) a.sigma_sum[s] = tr;
      a.steps[s] = po.steps;
    }
  }
}

extern "C" int hak_polar_grp(const HakLocalArgs *a, int n_teams, void *stream) {
  if (a->n_order <= 0) return 0;
  if (n_teams <= 0 || n_teams % 8 != 0 || !a->grp || !a->steps || !a->flag) return (int)hipErrorInvalidValue;
  if (a->prec_f32)
    hipLaunchKernelGGL((k_polar_grp<64, PREC_F32>), dim3(n_teams * HAK_GRP), dim3(NT), 0, (hipStream_t)stream, *a);
  else
    hipLaunchKernelGGL((k_polar_grp<128, PREC_BF16X3>), dim3(n_teams * HAK_GRP), dim3(NT), 0, (hipStream_t)stream,
                       *a);
  LAUNCH_RET();
}

extern "C" int hak_crosscov(const HakLocalArgs *a, int n_slots, void *stream) {
  if (a->n_order <= 0) return 0;
  if (a->tile == 128)
    hipLaunchKernelGGL(k_crosscov<128>, dim3(n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  else
    hipLaunchKernelGGL(k_crosscov<64>, dim3(n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  LAUNCH_RET();
}

// prec_f32: FP32 products on 64 x 64 tiles, for (nearly) rank-deficient
// searchlights (HAK_NEAR_DEFICIENT): their rounding-noise singular values are raised from about 1e-9
// to 1 over a long schedule, which carries the bf16x3 products' 2^-16 error
// into the range part of Q (about 5e-5 against 1e-6 with FP32 products).
extern "C" int hak_polar(const HakLocalArgs *a, void *stream) {
  if (a->n_order <= 0) return 0;
  if (a->prec_f32)
    hipLaunchKernelGGL((k_polar<64, PREC_F32>), dim3(a->n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  else if (a->tile == 64)
    hipLaunchKernelGGL((k_polar<64, PREC_BF16X3>), dim3(a->n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  else
    hipLaunchKernelGGL((k_polar<128, PREC_BF16X3>), dim3(a->n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  LAUNCH_RET();
}

// Q = sum_j (a_j / |a_j|) v_j^T over columns with |a_j| > cut; returns sum |a_j|.
// Columns with |a_j| <= cut are reported through *rank_def.
template <typename T, typename TO>
__device__ double jacobi_compose(T *Am, const T *Vm, int k, int kp, TO *Q, int qld, T *sig, int32_t *rank_sh,
                                 double *red, T (*us)[33], T (*vs)[33]) {
  // singular values and normalised U columns
  if (threadIdx.x == 0) *rank_sh = 0;
  __syncthreads();
  T smax = 0;
  for (int j = threadIdx.x; j < k; j += NT) {
    T s2 = 0;
    const T *col = Am + (int64_t)j * kp;
    for (int t = 0; t < k; ++t) s2 += col[t] * col[t];
    sig[j] = sqrt(s2);
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) smax = sig[j] > smax ? sig[j] : smax;
  const T cut = smax * (T)k * (T)JacobiTol<T>::cut;
  for (int j = threadIdx.x; j < k; j += NT) {
    if (sig[j] <= cut) atomicExch(rank_sh, 1);
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)k * kp; e += NT) {
    int j = (int)(e / kp);
    T sj = sig[j];
    Am[e] = (sj > cut) ? Am[e] / sj : (T)0;
  }
  __syncthreads();
  compose_uvt<T, TO>(Am, Vm, k, kp, Q, nullptr, us, vs, -1);
  double ssum = 0.0;
  for (int j = threadIdx.x; j < k; j += NT) ssum += (double)sig[j];
  return wg_sum_f64(ssum, red);
}

template <int TILE>
__global__ __launch_bounds__(NT) void k_jacobi32(HakLocalArgs a) {
  __shared__ GemmSmem<TILE> sm;
  __shared__ double red[4];
  __shared__ int32_t item_sh, flag_sh, rank_sh;
  __shared__ float feat_sh[HAK_NFEAT];
  __shared__ float us32[32][33], vs32[32][33];
  const int slot = blockIdx.x;
  WorkSlots w = slot_buffers(a, slot);
  for (;;) {
    const int it = next_item(a.queue, &item_sh);
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    const int k = a.sl_off[s + 1] - a.sl_off[s];
    const int kp = a.qld[s];
    if (a.rank_check_nt && k >= a.nt) {
      if (threadIdx.x == 0) { a.flag[s] = 1; if (a.steps) a.steps[s] = 0; }
      continue;
    }
    // cross-covariance in bf16x3 MFMA (same kernel as the polar path), then Jacobi in FP32
    crosscov<TILE>(sm, a, s, k, kp, w.A);
    __syncthreads();
    if (a.features) {
      features_of(a, s, k, kp, w.A, w.P, feat_sh, red, frob2(w.A, kp, red));
      if (threadIdx.x < HAK_NFEAT) a.features[(int64_t)s * HAK_NFEAT + threadIdx.x] = feat_sh[threadIdx.x];
      __syncthreads();
    }
    if (a.sigma_sum)  // keep A for trace(Q^T A)
      for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) w.X2[e] = w.A[e];
    for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) {
      int m = (int)(e % kp), n = (int)(e / kp);
      w.G[e] = (m == n && m < k) ? 1.f : 0.f;
    }
    __syncthreads();
    int sweeps = jacobi_sweeps<float>(w.A, w.G, k, kp, &flag_sh);
    float *Q = a.q + a.qoff[s];
    jacobi_compose<float, float>(w.A, w.G, k, kp, Q, kp, w.P, &rank_sh, red, us32, vs32);
    if (a.sig_out && threadIdx.x == 0) {
      float smax = 0.f, smin = 3.0e38f;
      for (int j = 0; j < k; ++j) {
        smax = fmaxf(smax, w.P[j]);
        smin = fminf(smin, w.P[j]);
      }
      a.sig_out[2 * s] = smax;
      a.sig_out[2 * s + 1] = smin;
    }
    if (a.sigma_sum) {
      double tr = 0.0;
      for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += NT) tr += (double)Q[e] * w.X2[e];
      tr = wg_sum_f64(tr, red);
      if (threadIdx.x == 0) a.sigma_sum[s] = tr;
    }
    if (threadIdx.x == 0) {
      if (a.steps) a.steps[s] = sweeps;
      if ((rank_sh || sweeps < 0) && a.fallback_enabled) a.flag[s] = 1;
    }
    __syncthreads();
  }
}

extern "C" int hak_jacobi32(const HakLocalArgs *a, void *stream) {
  if (a->n_order <= 0) return 0;
  if (a->tile == 64)
    hipLaunchKernelGGL(k_jacobi32<64>, dim3(a->n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  else
    hipLaunchKernelGGL(k_jacobi32<128>, dim3(a->n_slots), dim3(NT), 0, (hipStream_t)stream, *a);
  LAUNCH_RET();
}

// FP64 path: one workgroup per flagged searchlight (all of them when
// a.flag == NULL). Work per slot (4 * stride64 doubles, stride64 = kpmax^2):
// A (then U), V, the LU copy of the sign test, and sig (kp doubles).
__global__ __launch_bounds__(NT) void k_jacobi64(HakLocalArgs a, int want_det_fix, double *work64,
                                                 int64_t stride64, int32_t count) {
  __shared__ double red[4];
  __shared__ int32_t item_sh;
  __shared__ SvdLds sv;
  double *Am = work64 + (int64_t)blockIdx.x * 4 * stride64;
  double *Vm = Am + stride64;
  double *Lm = Am + 2 * stride64;
  double *sig = Am + 3 * stride64;
  for (;;) {
    const int s = next_item(a.queue, &item_sh);
    if (s >= count) break;
    if (a.flag && a.flag[s] == 0) continue;
    const int o = a.sl_off[s];
    const int k = a.sl_off[s + 1] - o;
    const int kp = a.qld[s];
    // A = X64[:, slx]^T Y64[:, sly] in FP64: 32 x 32 output tiles, 32-sample
    // chunks of the gathered columns staged in LDS (coalesced column reads)
    for (int tm = 0; tm < kp; tm += 32)
      for (int tn = 0; tn < kp; tn += 32) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const int mi = threadIdx.x & 31, ni = threadIdx.x >> 5;  // ni in 0..7
        for (int t0 = 0; t0 < a.nt; t0 += 32) {
          for (int e = threadIdx.x; e < 32 * 32; e += NT) {
            const int c = e >> 5, t = t0 + (e & 31);
            const int gm = tm + c, gn = tn + c;
            sv.us[c][e & 31] = (gm < k && t < a.nt) ? a.X64[(int64_t)a.slx_idx[o + gm] * a.nt + t] : 0.0;
            sv.vs[c][e & 31] = (gn < k && t < a.nt) ? a.Y64[(int64_t)a.sly_idx[o + gn] * a.nt + t] : 0.0;
          }
          __syncthreads();
#pragma unroll 4
          for (int t = 0; t < 32; ++t) {
            const double xv = sv.us[mi][t];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += xv * sv.vs[ni + 8 * q][t];
          }
          __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) Am[(tm + mi) + (int64_t)(tn + ni + 8 * q) * kp] = acc[q];
      }
    __syncthreads();
    double ss = 0.0;
    const int sweeps = svd_polar64(Am, Vm, Lm, sig, k, kp, want_det_fix != 0, a.q + a.qoff[s],
                                   a.q64 ? a.q64 + a.qoff[s] : nullptr, sv, red, &ss);
    if (threadIdx.x == 0) {
      if (a.sigma_sum) a.sigma_sum[s] = ss;
      if (a.steps) a.steps[s] = 1000 + (sweeps > 0 ? sweeps : -sweeps);
    }
    __syncthreads();
  }
}

extern "C" int hak_jacobi64(const HakLocalArgs *a, int want_det_fix, double *work64, int64_t stride64,
                            int32_t n_slots, int32_t count, void *stream) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_jacobi64, dim3(n_slots), dim3(NT), 0, (hipStream_t)stream, *a, want_det_fix, work64,
                     stride64, count);
  LAUNCH_RET();
}

// ---------------------------------------------------------------------------
// data layout and scaling helpers
// ---------------------------------------------------------------------------

// FP64 input -> FP32 (leading dim ld), optional FP64 copy, and optional bf16
// hi/lo planes of the FP32 values split as the bf16x3 staging splits FP32
// operands (hi = bf16(x), lo = bf16(x - hi)), so a product on the planes equals
// one on the FP32 data bit for bit
__global__ void k_convert(const double *__restrict__ src, int32_t nt, int32_t nv, int col_major,
                          float *__restrict__ d32, int32_t ld, double *__restrict__ d64, uint16_t *__restrict__ hi,
                          uint16_t *__restrict__ lo) {
  const int64_t n = (int64_t)ld * nv;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t v = e / ld;
    int32_t t = (int32_t)(e % ld);
    double x = 0.0;
    if (t < nt) {
      x = col_major ? src[t + v * nt] : src[(int64_t)t * nv + v];
      if (d64) d64[t + v * nt] = x;
    }
    const float f = (float)x;
    d32[e] = f;
    if (hi) {
      const uint16_t h = f2bf(f);
      hi[e] = h;
      lo[e] = f2bf(f - bf2f(h));
    }
  }
}
extern "C" int hak_convert_data(const double *src, int32_t nt, int32_t nv, int col_major, float *dst32,
                                int32_t ld, double *dst64, uint16_t *hi, uint16_t *lo, void *stream) {
  if ((hi == nullptr) != (lo == nullptr)) return (int)hipErrorInvalidValue;
  int64_t n = (int64_t)ld * nv;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_convert, dim3((int)g), dim3(256), 0, (hipStream_t)stream, src, nt, nv, col_major, dst32,
                     ld, dst64, hi, lo);
  LAUNCH_RET();
}

// sum over the searchlight's X columns of the population variance, in column
// order (ha_searchlight.c:641-653)
__global__ __launch_bounds__(NT) void k_var_sum(const double *__restrict__ X64, int32_t nt,
                                                const int32_t *__restrict__ idx, const int32_t *__restrict__ off,
                                                int32_t count, double *__restrict__ var_sum) {
  __shared__ double vs[1024];
  for (int s = blockIdx.x; s < count; s += gridDim.x) {
    const int o = off[s], k = off[s + 1] - o;
    double total = 0.0;
    for (int j0 = 0; j0 < k; j0 += 1024) {
      for (int j = j0 + threadIdx.x; j < k && j < j0 + 1024; j += NT) {
        const double *col = X64 + (int64_t)idx[o + j] * nt;
        double mean = 0.0;
        for (int t = 0; t < nt; ++t) mean += col[t];
        mean /= nt;
        double var = 0.0;
        for (int t = 0; t < nt; ++t) {
          double d = col[t] - mean;
          var += d * d;
        }
        vs[j - j0] = var / nt;
      }
      __syncthreads();
      if (threadIdx.x == 0)
        for (int j = j0; j < k && j < j0 + 1024; ++j) total += vs[j - j0];
      __syncthreads();
    }
    if (threadIdx.x == 0) var_sum[s] = total;
  }
}
extern "C" int hak_var_sum(const double *X64, int32_t nt, const int32_t *slx_idx, const int32_t *sl_off,
                           int32_t count, double *var_sum, void *stream) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_var_sum, dim3(count < 65536 ? count : 65536), dim3(NT), 0, (hipStream_t)stream, X64, nt,
                     slx_idx, sl_off, count, var_sum);
  LAUNCH_RET();
}

__global__ void k_scale(const double *__restrict__ ss, const double *__restrict__ vs, int32_t nt, int32_t count,
                        double *__restrict__ scale) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < count; s += gridDim.x * blockDim.x)
    scale[s] = ss[s] / (vs[s] * nt);
}
extern "C" int hak_scale(const double *sigma_sum, const double *var_sum, int32_t nt, int32_t count,
                         double *scale, void *stream) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_scale, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, sigma_sum, var_sum,
                     nt, count, scale);
  LAUNCH_RET();
}
