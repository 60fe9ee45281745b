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

typedef __attribute__((address_space(3))) double lds_f64;

// The vector phase of the fix for one searchlight whose Cholesky factor L of
// G = A64^T A64 is in G: block inverse iteration on NV vectors V, Rayleigh-Ritz,
// then Q <- Q - 2 (Q v) v^T. V lives at Vp (LDS, or global for large kp), W =
// L^T V at Wg (global), v and u at vs and us (kp doubles each; beside V, or
// over W, which is dead by then). Returns the smallest eigenvalue of G
// (sigma_min^2). Not inlined, so that its registers do not crowd the
// products' loops in the kernel; the LDS arguments keep their address space.
template <typename VP>
__device__ __noinline__ double fix_vector(const double *G, float *Q, int k, int kp, VP *Vp, double *Wg, VP *vsp,
                                          VP *usp, lds_f64 *dp, lds_f64 *partp, lds_f64 *redp, lds_f64 *red10p,
                                          lds_f64 *tshp, lds_f64 *yshp, lds_f64 *lamp) {
  double *V = (double *)Vp, *vs = (double *)vsp, *us = (double *)usp;
  double (*d)[B + 1] = (double (*)[B + 1])(double *)dp, (*part)[B + 1] = (double (*)[B + 1])(double *)partp;
  double *red = (double *)redp, *red10 = (double *)red10p, *ysh = (double *)yshp, *lam_sh = (double *)lamp;
  double (*tsh)[NV] = (double (*)[NV])(double *)tshp;
  for (int64_t e = threadIdx.x; e < (int64_t)kp * NV; e += RT) {
    const int i = (int)(e % kp), q = (int)(e / kp);
    V[e] = i < k ? sin(0.7548776662466927 * (i + 1) * (q + 1) + 0.5698402909980532 * q) : 0.0;
  }
  __syncthreads();
  orthonormalize(V, kp, red);
  for (int step = 0; step < NITER; ++step) {
    solve_lower_nv(G, V, kp, d, part);
    solve_lower_t_nv(G, V, kp, d, part);
    orthonormalize(V, kp, red);
  }
  // Rayleigh-Ritz: T = (L^T V)^T (L^T V), its ten distinct entries summed in one pass
  lower_t_times(G, V, Wg, kp, part);
  gram_nv(Wg, kp, red10, tsh);
  if (threadIdx.x == 0) *lam_sh = small_eig(tsh, ysh, k < NV ? k : NV);
  __syncthreads();
  // v = V y, u = Q v, Q -= 2 u v^T
  for (int i = threadIdx.x; i < kp; i += RT) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < NV; ++q) acc += V[i + (int64_t)q * kp] * ysh[q];
    vs[i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kp; i += RT) {
    double acc = 0.0;
#pragma unroll 8
    for (int j = 0; j < k; ++j) acc += (double)Q[i + (int64_t)j * kp] * vs[j];
    us[i] = acc;
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)kp * kp; e += RT) {
    const int i = (int)(e % kp), j = (int)(e / kp);
    Q[e] = (float)((double)Q[e] - 2.0 * us[i] * vs[j]);
  }
  const double lam = *lam_sh;
  __syncthreads();
  return lam;
}

__global__ __launch_bounds__(RT, 4) void k_detfix(HakLocalArgs a, double *work64, int64_t stride64,
                                               const int32_t *iota) {
  // t[0..3]: the FP64 tiles of the products and the Cholesky factorization;
  // in the vector phase t[0] and t[1] are the solves' diagonal block and
  // partial sums, and t[2..3] hold V (kp <= 528; then v and u go to t[0..1])
  __shared__ double tl[4][B][B + 1];
  __shared__ double red[4], red10[4 * 10];
  __shared__ float pmax[4];
  __shared__ int pidx[4], piv_sh, sign_sh, ok_sh, item_sh;
  __shared__ double tsh[NV][NV], ysh[NV], lam_sh;
  double (*t0)[B + 1] = tl[0], (*t1)[B + 1] = tl[1], (*t2)[B + 1] = tl[2], (*t3)[B + 1] = tl[3];
  constexpr int VCAP = 2 * B * (B + 1);  // doubles of t[2..3]
  double *A64 = work64 + (int64_t)blockIdx.x * 4 * stride64;
  double *G = A64 + stride64;
  float *M = reinterpret_cast<float *>(A64 + 2 * stride64);
  double *Vg = A64 + 3 * stride64;  // V when it does not fit in LDS, then W = L^T V (kp x NV each)
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
    // the vector phase; the padding rows of V stay zero (G is block diagonal)
    double *W = Vg + (int64_t)NV * kp;
    lds_f64 *l0 = (lds_f64 *)&t0[0][0], *l1 = (lds_f64 *)&t1[0][0], *lr = (lds_f64 *)red,
            *lr10 = (lds_f64 *)red10, *lt = (lds_f64 *)&tsh[0][0], *ly = (lds_f64 *)ysh, *ll = (lds_f64 *)&lam_sh;
    const double lam = kp * NV <= VCAP
                           ? fix_vector<lds_f64>(G, Q, k, kp, (lds_f64 *)&t2[0][0], W, l0, l1, l0, l1, lr, lr10, lt,
                                                 ly, ll)
                           : fix_vector<double>(G, Q, k, kp, Vg, W, W, W + kp, l0, l1, lr, lr10, lt, ly, ll);
    if (threadIdx.x == 0 && a.sigma_sum) a.sigma_sum[s] -= 2.0 * sqrt(fmax(lam, 0.0));
  }
}

typedef float f32x16_lu __attribute__((ext_vector_type(16)));
constexpr int PW = 32;  // LU panel width

// LDS of k_lu_sign beyond its static part: the panel (kpmax rows of PW + 1
// floats), the row permutation and one 32 x 33 staging tile per wave.
__host__ __device__ constexpr size_t lu_lds_bytes(int kpmax) {
  return (size_t)kpmax * (PW + 1) * sizeof(float) + (size_t)kpmax * sizeof(int) + 4 * 32 * 33 * sizeof(float);
}

// One column step of the register-panel LU of k_lu_sign (column C of the
// panel, a compile-time constant so that the panel rows stay in registers):
// pivot search, interchange through LDS, multipliers and rank-1 update of this
// thread's rows. Two barriers; a singular pivot sets `singular` (uniformly).
template <int XS>
struct LuCol {
  float (&r)[XS][PW];
  int tid, lane, wave, nb, m, j0;
  bool singular;
  float *pmax;
  int *pidx;
  float *prow, *crow;
  int *perm;
  int *sign_sh;
};

template <int C, int XS>
__device__ __forceinline__ void lu_column(LuCol<XS> &L) {
  if (C >= L.nb || L.singular) return;
  const int tid = L.tid;
  float best = -1.f;
  int bi = C;
#pragma unroll
  for (int x = 0; x < XS; ++x) {
    const int i = tid + RT * x;
    const float v = fabsf(L.r[x][C]);
    if (i >= C && i < L.m && v > best) {
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
  if (L.lane == 0) {
    L.pmax[L.wave] = best;
    L.pidx[L.wave] = bi;
  }
  __syncthreads();
  // every thread picks the same winner
  float b0 = L.pmax[0];
  int pr = L.pidx[0];
#pragma unroll
  for (int q = 1; q < 4; ++q)
    if (L.pmax[q] > b0 || (L.pmax[q] == b0 && L.pidx[q] < pr)) {
      b0 = L.pmax[q];
      pr = L.pidx[q];
    }
  // the pivot row to prow (from its owner), row C to crow (thread C, slot 0)
  const int tp = pr % RT, xp = pr / RT;
  if (tid == tp) {
#pragma unroll
    for (int x = 0; x < XS; ++x)
      if (x == xp) {
#pragma unroll
        for (int j = 0; j < PW; ++j) L.prow[j] = L.r[x][j];
      }
  }
  if (tid == C && pr != C) {
#pragma unroll
    for (int j = 0; j < PW; ++j) L.crow[j] = L.r[0][j];
  }
  if (tid == 0 && pr != C) {
    *L.sign_sh = -*L.sign_sh;
    const int t = L.perm[L.j0 + C];
    L.perm[L.j0 + C] = L.perm[L.j0 + pr];
    L.perm[L.j0 + pr] = t;
  }
  __syncthreads();
  if (pr != C) {
    if (tid == C) {
#pragma unroll
      for (int j = 0; j < PW; ++j) L.r[0][j] = L.prow[j];
    }
    if (tid == tp) {
#pragma unroll
      for (int x = 0; x < XS; ++x)
        if (x == xp) {
#pragma unroll
          for (int j = 0; j < PW; ++j) L.r[x][j] = L.crow[j];
        }
    }
  }
  const float d = L.prow[C];
  if (d == 0.f) {
    L.singular = true;
    return;
  }
  if (tid == 0 && d < 0.f) *L.sign_sh = -*L.sign_sh;
  // the rows below: thread t keeps its rows t, t + RT, ... for every column
#pragma unroll
  for (int x = 0; x < XS; ++x) {
    const int i = tid + RT * x;
    if (i > C && i < L.m) {
      const float l = L.r[x][C] / d;
      L.r[x][C] = l;
#pragma unroll
      for (int j = C + 1; j < PW; ++j) L.r[x][j] -= l * L.prow[j];
    }
  }
}

template <int XS, int C = 0>
__device__ __forceinline__ void lu_columns(LuCol<XS> &L) {
  if constexpr (C < PW) {
    lu_column<C, XS>(L);
    lu_columns<XS, C + 1>(L);
  }
}

// sign(det Q_s) of every searchlight of the order with flag 0, one workgroup
// each from the queue. LU with partial pivoting (largest magnitude, the lowest
// row on ties) of Q in FP32 (Q is orthogonal, so |det Q| = 1 and the sign is
// robust), Q left unchanged: the first panel's updates go to M in this
// workgroup's work slot, and the later panels work on M:
//  * rows are never moved in M: logical row i lives in physical row perm[i],
//    and every interchange swaps two perm entries and flips the sign;
//  * each 32-column panel (logical rows j0 .. k-1) is factored in registers:
//    thread t holds the panel rows t, t + 256, ... (XS of them). Per column,
//    the waves' pivot candidates meet in LDS and every thread picks the same
//    winner; the owners of the pivot row and of row c swap them through LDS,
//    where the pivot row stays for the rank-1 update of every thread's rows:
//    two barriers per column. The factored panel then goes to LDS;
//  * U12 = L11^-1 A12 is solved one column per thread, into U12T (32 values
//    per column, contiguous);
//  * the trailing matrix is updated in place in 32 x 32 tiles, each wave on
//    its own tiles (column by column) without block barriers: A22 -= L21 U12
//    as 16 FP32 MFMAs (v_mfma_f32_32x32x2f32), L21 from the LDS panel, the U12
//    block staged by the wave in its own LDS tile once per tile column, and the
//    next tile's A22 loaded while the current one's MFMAs run.
// det = sign(perm) * prod(pivots). An exactly zero pivot (singular block)
// records 0 (no fix), like lu_sign. XS * 256 >= kpmax.
template <int XS>
__global__ __launch_bounds__(RT, 2) void k_lu_sign(HakLocalArgs a, float *work, int64_t slot_floats, int32_t kpmax,
                                                int32_t *neg) {
  extern __shared__ float lds_dyn[];
  float (*pan)[PW + 1] = reinterpret_cast<float (*)[PW + 1]>(lds_dyn);
  int *perm = reinterpret_cast<int *>(lds_dyn + (size_t)kpmax * (PW + 1));
  float (*ubs)[32][33] = reinterpret_cast<float (*)[32][33]>(perm + kpmax);
  __shared__ float pmax[4];
  __shared__ int pidx[4], sign_sh, item_sh;
  __shared__ float prow[PW], crow[PW];
  float *M = work + (int64_t)blockIdx.x * slot_floats;
  float *U12T = M + (int64_t)kpmax * kpmax;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (;;) {
    __syncthreads();
    if (tid == 0) item_sh = atomicAdd(a.queue, 1);
    __syncthreads();
    const int it = item_sh;
    if (it >= a.n_order) break;
    const int s = a.order ? a.order[it] : it;
    if (a.flag[s] != 0) {
      if (tid == 0) neg[s] = 0;
      continue;
    }
    const int k = a.sl_off[s + 1] - a.sl_off[s], kp = a.qld[s];
    // Q itself is only read: the first panel, its U12 and its trailing update
    // read Q, and the trailing update writes M, which every later phase reads
    const float *Qs = a.q + a.qoff[s];
    for (int i = tid; i < k; i += RT) perm[i] = i;
    if (tid == 0) sign_sh = 1;
    __syncthreads();
    bool singular = false;
    for (int j0 = 0; j0 < k && !singular; j0 += PW) {
      const int nb = k - j0 < PW ? k - j0 : PW, m = k - j0, je = j0 + nb;
      const float *Ms = j0 == 0 ? Qs : M;
      // the panel rows of this thread: relative rows tid + RT x
      float r[XS][PW];
#pragma unroll
      for (int x = 0; x < XS; ++x) {
        const int i = tid + RT * x;
        const int64_t prw = i < m ? perm[j0 + i] : 0;
#pragma unroll
        for (int c = 0; c < PW; ++c) r[x][c] = (i < m && c < nb) ? Ms[prw + (int64_t)(j0 + c) * kp] : 0.f;
      }
      {
        LuCol<XS> col{r, tid, lane, wave, nb, m, j0, false, pmax, pidx, prow, crow, perm, &sign_sh};
        lu_columns<XS>(col);
        singular = col.singular;
      }
      if (singular) break;
      // the factored panel (L11 \ U11 over L21) to LDS
#pragma unroll
      for (int x = 0; x < XS; ++x) {
        const int i = tid + RT * x;
        if (i < m) {
#pragma unroll
          for (int c = 0; c < PW; ++c) pan[i][c] = r[x][c];
        }
      }
      __syncthreads();
      if (je >= k) break;
      const int n2 = k - je;
      // U12 = L11^-1 A12 (unit lower L11 in the panel's first nb rows)
      for (int jc = tid; jc < n2; jc += RT) {
        const int64_t col = (int64_t)(je + jc) * kp;
        float u[PW];
#pragma unroll
        for (int i = 0; i < PW; ++i) u[i] = i < nb ? Ms[perm[j0 + i] + col] : 0.f;
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
      // A22 -= L21 U12 in 32 x 32 tiles: wave w takes tiles w, w + 4, ... in
      // column-major order (tile t: row block t % ntr, column block t / ntr)
      const int ntr = (n2 + 31) / 32, ntile = ntr * ntr;
      const int mrow = lane & 31, h = lane >> 5;
      float (*ub)[33] = ubs[wave];
      auto tile_load = [&](int t, f32x16_lu &acc) {
        const int R = je + 32 * (t % ntr) + mrow, c0 = je + 32 * (t / ntr);
        const bool rv = R < k;
        const int64_t prw = rv ? perm[R] : 0;
#pragma unroll
        for (int r2 = 0; r2 < 16; ++r2) {
          const int C = c0 + (r2 >> 2) * 8 + h * 4 + (r2 & 3);
          acc[r2] = (rv && C < k) ? Ms[prw + (int64_t)C * kp] : 0.f;
        }
      };
      f32x16_lu accA, accB;
      int cur_tj = -1;
      if (wave < ntile) tile_load(wave, accA);
      for (int t = wave; t < ntile; t += 4) {
        const int tj = t / ntr, c0 = je + 32 * tj;
        const int R = je + 32 * (t % ntr) + mrow;
        const bool rv = R < k;
        if (tj != cur_tj) {
          // the U12 block of this tile column into the wave's own LDS tile; only
          // this wave writes and reads it, and a wave's LDS accesses complete
          // in order
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int x = 0; x < 16; ++x) {
            const int e = lane + 64 * x, n = e >> 5, kk = e & 31, col = c0 + n;
            ub[kk][n] = (col < k && kk < nb) ? U12T[(int64_t)(col - je) * PW + kk] : 0.f;
          }
          __builtin_amdgcn_wave_barrier();
          cur_tj = tj;
        }
        if (t + 4 < ntile) tile_load(t + 4, accB);
#pragma unroll
        for (int st = 0; st < PW / 2; ++st) {
          const int kk = 2 * st + h;
          const float av = (rv && kk < nb) ? -pan[R - j0][kk] : 0.f;
          accA = __builtin_amdgcn_mfma_f32_32x32x2f32(ub[kk][mrow], av, accA, 0, 0, 0);
        }
        if (rv) {
          const int64_t prw = perm[R];
#pragma unroll
          for (int r2 = 0; r2 < 16; ++r2) {
            const int C = c0 + (r2 >> 2) * 8 + h * 4 + (r2 & 3);
            if (C < k) M[prw + (int64_t)C * kp] = accA[r2];
          }
        }
        accA = accB;
      }
      __syncthreads();
    }
    __syncthreads();
    if (tid == 0) neg[s] = (singular || sign_sh >= 0) ? 0 : 1;
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
    hipError_t e = hipFuncSetAttribute((const void *)k_lu_sign<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lu_lds_bytes(2 * RT));
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void *)k_lu_sign<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lu_lds_bytes(HAK_LU_KPMAX));
    if (e != hipSuccess) return (int)e;
    lds_raised = true;
  }
  const size_t lds = lu_lds_bytes(kpmax);
  if (kpmax <= 2 * RT)
    hipLaunchKernelGGL(k_lu_sign<2>, dim3(n_wg), dim3(RT), lds, (hipStream_t)stream, *a, work, slot_floats, kpmax,
                       neg);
  else
    hipLaunchKernelGGL(k_lu_sign<4>, dim3(n_wg), dim3(RT), lds, (hipStream_t)stream, *a, work, slot_floats, kpmax,
                       neg);
  LAUNCH_RET();
}

extern "C" int hak_detfix(const HakLocalArgs *a, double *work64, int64_t stride64, const int32_t *iota, int n_wg,
                          void *stream) {
  if (a->n_order <= 0) return 0;
  if (n_wg <= 0 || (!a->flag && !a->det_neg) || !iota) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_detfix, dim3(n_wg), dim3(RT), 0, (hipStream_t)stream, *a, work64, stride64, iota);
  LAUNCH_RET();
}
