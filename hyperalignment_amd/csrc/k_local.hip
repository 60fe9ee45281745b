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
// nothing left to rotate instead of chasing noise.
template <typename T>
struct JacobiTol;
template <>
struct JacobiTol<float> { static constexpr double eps = 1.19e-7; static constexpr int sweeps = 30; };
template <>
struct JacobiTol<double> { static constexpr double eps = 2.22e-16; static constexpr int sweeps = 30; };

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
  const T cut = smax * (T)k * (sizeof(T) == 8 ? (T)2.3e-16 : (T)1.2e-7);
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

// sign(det M) of the k x k matrix M (column-major, leading dimension kp) by LU
// with partial pivoting of a copy in L; the pivot is the first row of largest
// |value|, as dgetrf takes it. For the orthogonal factors U and V of the FP64
// path |det| = 1, so every pivot is far from zero and the sign is robust.
__device__ int lu_det_sign64(const double *M, double *L, int k, int kp, double *pmax_sh, int32_t *pidx_sh,
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

// FP64 path: one workgroup per flagged searchlight (all of them when
// a.flag == NULL). Work per slot (4 * stride64 doubles, stride64 = kpmax^2):
// A (then U), V, the LU copy of the sign test, and sig (kp doubles).
__global__ __launch_bounds__(NT) void k_jacobi64(HakLocalArgs a, int want_det_fix, double *work64,
                                                 int64_t stride64, int32_t count) {
  __shared__ double red[4];
  __shared__ int32_t item_sh, flag_sh, rank_sh, sign_sh, piv_sh;
  __shared__ double pmax_sh[4];
  __shared__ int32_t pidx_sh[4];
  __shared__ double xs64[32][33], ys64[32][33];
  __shared__ double wdot[1024];
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
            xs64[c][e & 31] = (gm < k && t < a.nt) ? a.X64[(int64_t)a.slx_idx[o + gm] * a.nt + t] : 0.0;
            ys64[c][e & 31] = (gn < k && t < a.nt) ? a.Y64[(int64_t)a.sly_idx[o + gn] * a.nt + t] : 0.0;
          }
          __syncthreads();
#pragma unroll 4
          for (int t = 0; t < 32; ++t) {
            const double xv = xs64[mi][t];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += xv * ys64[ni + 8 * q][t];
          }
          __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int m = tm + mi, n = tn + ni + 8 * q;
          Am[m + (int64_t)n * kp] = acc[q];
          Vm[m + (int64_t)n * kp] = (m == n && m < k) ? 1.0 : 0.0;
        }
      }
    __syncthreads();
    int sweeps = jacobi_sweeps<double>(Am, Vm, k, kp, &flag_sh);
    // U columns, rank completion for zero singular values
    if (threadIdx.x == 0) rank_sh = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += NT) {
      double s2 = 0;
      const double *col = Am + (int64_t)j * kp;
      for (int t = 0; t < k; ++t) s2 += col[t] * col[t];
      sig[j] = sqrt(s2);
    }
    __syncthreads();
    double smax = 0;
    for (int j = 0; j < k; ++j) smax = fmax(smax, sig[j]);
    const double cut = smax * k * 2.3e-16;
    for (int64_t e = threadIdx.x; e < (int64_t)k * kp; e += NT) {
      int j = (int)(e / kp);
      Am[e] = (sig[j] > cut) ? Am[e] / sig[j] : 0.0;
    }
    __syncthreads();
    // Complete U for columns with zero singular value (rank-deficient A): each
    // deficient column starts as a fixed pseudo-random vector, is projected off
    // every finished column twice (classical Gram-Schmidt, all dot products in
    // parallel), and normalised. Finished = full-rank columns plus the deficient
    // columns completed before it.
    for (int j = 0; j < k; ++j) {
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
              wdot[i - i0] = d;
            }
            __syncthreads();
            for (int t = threadIdx.x; t < k; t += NT) {
              double v = uj[t];
              for (int i = i0; i < k && i < i0 + 1024; ++i) v -= wdot[i - i0] * Am[(int64_t)i * kp + t];
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
    // reflection fix (ha_searchlight.c:611-631): if det(Q) < 0, flip the pair
    // of the smallest singular value. sign(det Q) = sign(det U) sign(det V),
    // from the finished factors as the reference takes it from Q = U V^T: for
    // a rank-deficient A (k >= nt) det A is rounding noise and the completed
    // columns of U decide the sign.
    int jmin = 0;
    bool flip = false;
    if (want_det_fix) {
      for (int j = 1; j < k; ++j)
        if (sig[j] <= sig[jmin]) jmin = j;
      const int su = lu_det_sign64(Am, Lm, k, kp, pmax_sh, pidx_sh, &piv_sh, &sign_sh);
      const int sv = lu_det_sign64(Vm, Lm, k, kp, pmax_sh, pidx_sh, &piv_sh, &sign_sh);
      flip = su * sv < 0;
    }
    float *Q = a.q + a.qoff[s];
    double *Q64 = a.q64 ? a.q64 + a.qoff[s] : nullptr;
    compose_uvt<double, float>(Am, Vm, k, kp, Q, Q64, xs64, ys64, flip ? jmin : -1);
    if (a.sigma_sum) {
      double ss = 0.0;
      for (int j = threadIdx.x; j < k; j += NT) ss += sig[j];
      ss = wg_sum_f64(ss, red);
      if (threadIdx.x == 0) a.sigma_sum[s] = flip ? ss - 2.0 * sig[jmin] : ss;
    }
    if (threadIdx.x == 0) {
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
