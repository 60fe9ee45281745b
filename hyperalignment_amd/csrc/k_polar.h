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
  // sixteen loads in flight, summed four at a time in the same order
  for (; e + 15 * NT < n4; e += 16 * NT) {
    float4 x[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) x[c] = m4[e + c * NT];
#pragma unroll
    for (int c = 0; c < 16; ++c)
      acc[c & 3] += (double)x[c].x * x[c].x + (double)x[c].y * x[c].y + (double)x[c].z * x[c].z +
                    (double)x[c].w * x[c].w;
  }
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

// x = split(M * f) for a kp x kp FP32 block, float4 loads four deep per thread;
// a workgroup of a team handles float4 e0, e0 + es, ... (es a multiple of NT)
__device__ __forceinline__ void split_scaled(const float *M, const SplitM &x, int kp, float f,
                                             int64_t e0 = -1, int64_t es = NT) {
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
  int64_t e = e0 < 0 ? threadIdx.x : e0;
  if (es == NT) {
    for (; e + 3 * NT < n4; e += 4 * NT) {
      float4 v[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] = m4[e + c * NT];
#pragma unroll
      for (int c = 0; c < 4; ++c) put4(e + c * NT, v[c]);
    }
  }
  for (; e < n4; e += es) put4(e, m4[e]);
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
__device__ __forceinline__ void scale_block(const float *M, float *x, int kp, float f, int64_t e0 = -1,
                                            int64_t es = NT) {
  const float4 *m4 = reinterpret_cast<const float4 *>(M);
  float4 *x4 = reinterpret_cast<float4 *>(x);
  const int64_t n4 = (int64_t)kp * kp / 4;
  for (int64_t e = e0 < 0 ? threadIdx.x : e0; e < n4; e += es) {
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
  template <bool COH = false>
  __device__ __forceinline__ float get(int64_t i) const { return get_split<COH>(s, i); }
  __device__ __forceinline__ void put4(int64_t i, float4 v) const { put_split4(s, i, v); }
  template <bool COH = false>
  __device__ __forceinline__ float4 get4(int64_t i) const { return get_split4<COH>(s, i); }
  __device__ __forceinline__ Opnd opnd(int kind, int ld, int mlimit) const { return opnd_split(s, kind, ld, mlimit); }
  __device__ __forceinline__ void load_scaled(const float *A, int kp, float f, int64_t e0 = -1,
                                              int64_t es = NT) const {
    split_scaled(A, s, kp, f, e0, es);
  }
};
template <>
struct MatS<PREC_F32> {
  static constexpr int KK = OP_F32_GATHER, KM = OP_F32_M;
  float *f;
  __device__ MatS(float *buf, int) : f(buf) {}
  __device__ __forceinline__ void put(int64_t i, float v) const { f[i] = v; }
  template <bool COH = false>
  __device__ __forceinline__ float get(int64_t i) const {
    return __uint_as_float(ld_b32<COH>(f, 4 * i));
  }
  __device__ __forceinline__ void put4(int64_t i, float4 v) const { *reinterpret_cast<float4 *>(f + i) = v; }
  template <bool COH = false>
  __device__ __forceinline__ float4 get4(int64_t i) const {
    const uint4 u = ld_b128<COH>(f, 4 * i);
    return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
  }
  __device__ __forceinline__ Opnd opnd(int kind, int ld, int mlimit) const { return opnd_f32(f, kind, ld, mlimit); }
  __device__ __forceinline__ void load_scaled(const float *A, int kp, float v, int64_t e0 = -1,
                                              int64_t es = NT) const {
    scale_block(A, f, kp, v, e0, es);
  }
};

// Result of one polar factorisation: steps taken (all attempts), whether it
// converged, and whether a team barrier gave up (the team's state is then lost).
struct PolarOut {
  int steps;
  bool done;
  bool aborted;
};

// The workgroups that run one factorisation. SoloTeam: one workgroup does it all.
// GroupTeam: 4 workgroups of one XCD share it. Tile (I, J) of every product goes
// to member (I + J) mod 4 (the members work on tiles of one row at a time, so
// its row panel is read from HBM once and from L2 by the others), and each does
// a quarter of the elementwise passes; they meet at a barrier on a counter in
// global memory after every product. Reductions go through per-member partial sums added in member
// order, so the result does not depend on timing.
struct SoloTeam {
  static constexpr int nw = 1;
  static constexpr bool coh = false;  // plain loads: nothing is shared with other workgroups
  __device__ int member_id() const { return 0; }
  __device__ bool leader() const { return true; }
  __device__ int64_t e0() const { return threadIdx.x; }
  __device__ int64_t es() const { return NT; }
  __device__ bool sync() {
    __syncthreads();
    return true;
  }
  __device__ double sum(double v, double *red, bool *ok) {
    *ok = true;
    return wg_sum_f64(v, red);
  }
};

#define HAK_GRP_WAIT_US 100000u  // microseconds a team barrier waits before it gives up

// Per-team state in global memory, zeroed before the launch.
struct GrpState {
  uint32_t count;     // barrier arrivals (monotonic)
  int32_t item[2];    // the team's searchlight positions, by loop turn parity
  uint32_t xcc_mask;  // bit x set by a member running on XCD x
  double part[HAK_GRP];
  double pad2[16 - 2 - HAK_GRP];
};
static_assert(sizeof(GrpState) == HAK_GRP_STATE_BYTES, "team state size");

struct GroupTeam {
  static constexpr int nw = HAK_GRP;
  // every load of what another member wrote (operands of the products, G in
  // the elementwise passes) is a device-scope load: see sync()
  static constexpr bool coh = true;
  GrpState *st;
  int member;
  uint32_t gen;     // barriers passed by this workgroup
  uint64_t limit;   // wait budget of one barrier in ticks of the 100 MHz wall clock
  int *ok_sh;       // LDS flag broadcast by thread 0
  __device__ int member_id() const { return member; }
  __device__ bool leader() const { return member == 0; }
  __device__ int64_t e0() const { return (int64_t)member * NT + threadIdx.x; }
  __device__ int64_t es() const { return (int64_t)HAK_GRP * NT; }
  // Barrier of the team, valid because all members run on one XCD and so share
  // its L2 (checked by the kernel through xcc_mask). Release: each thread waits
  // until its global stores have reached L2 (the vector L1 is write-through);
  // the L2 write-back of an agent-scope release fence is left out on purpose,
  // since the members read through the same L2. The arrival counter is an L2
  // atomic polled past L1. Acquire: none in the cache. The members read each
  // other's data only with device-scope loads (coh, sc1), which skip the L1 and
  // so cannot see a stale line; an L1 invalidate here (an agent-scope acquire
  // fence) costs about 5% of the iteration. The workgroup-scope invalidate
  // (buffer_inv sc0) is not enough: it left stale L1 lines behind. Any new load
  // of memory another member writes must be a device-scope load too. The asm
  // statements keep the compiler from moving memory accesses across the wait.
  // Returns false (in every thread) when the wait gives up.
  __device__ bool sync() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    ++gen;
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(&st->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t target = gen * HAK_GRP;
      int ok = 1;
      if (__hip_atomic_load(&st->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        // bounded by time, not by polls: the wait ends after limit ticks whatever the poll rate
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(&st->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
          if (wall_clock64() - t0 > limit) {
            ok = 0;
            break;
          }
        }
      }
      *ok_sh = ok;
    }
    __syncthreads();
    asm volatile("" ::: "memory");
    const bool ok = *ok_sh != 0;
    __syncthreads();
    return ok;
  }
  __device__ double sum(double v, double *red, bool *ok) {
    const double p = wg_sum_f64(v, red);
    if (threadIdx.x == 0) st->part[member] = p;
    *ok = sync();
    double t = 0.0;
#pragma unroll
    for (int m = 0; m < HAK_GRP; ++m) t += __hip_atomic_load(&st->part[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return t;
  }
};

// Polar factor Q of the k x k block of w.A (FP32, kp x kp) by the scheduled
// odd-polynomial iteration X <- X (ca I + cb X^T X + cc (X^T X)^2) on bf16x3
// MFMA products, from X0 = A / scale with sigma_min(X0) >= ell assumed. The
// schedule table picks (ca, cb, cc) per step for the bucket of ell; when the
// growth phase ends with a large residual the schedule restarts from the bound
// the residual gives; once ||X^T X - I||_F < tol a final quintic Newton-Schulz
// step writes Q through fin(m, n, v), v = Q(m..m+3, n), for every (m, n) of
// the kp x kp block (m a multiple of 4).
// learned: (scale, ell) came from a prediction; an attempt that does not
// converge is redone once from (fro, plain_ell), which always holds
// sigma_max(X0) <= 1. w.X, w.G, w.P are overwritten (w.X2 is not used: each
// step's X P goes over G, dead by then, and the two buffers trade roles).
// on_fin() runs in every thread once the finishing step is decided, before the
// team's barrier that precedes the finishing product (so before any fin call
// of any member).
struct NoFinMark {
  __device__ void operator()() const {}
};
template <int TILE, int PREC, class Team, class Fin, class OnFin = NoFinMark>
__device__ __forceinline__ PolarOut polar_solve(GemmSmem<TILE> &sm, const HakLocalArgs &a, const WorkSlots &w,
                                                int k, int kp, float fro, float scale, float ell, bool learned,
                                                double *red, float *trace, Team &team, Fin fin,
                                                OnFin on_fin = OnFin()) {
  using M = MatS<PREC>;
  const M P(w.P, kp);
  int step = 0, total = 0;
  bool done = false;
  constexpr int nw = Team::nw;
  const int tw = team.member_id();
  if (!team.leader()) trace = nullptr;
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
    M(w.X, kp).load_scaled(w.A, kp, inv, team.e0(), team.es());
    if (!team.sync()) return PolarOut{total + step, false, true};
    float *Xb = w.X, *Gb = w.G;
    int restart = 0, restarts = 0;
    for (step = 0; step < a.max_steps; ++step) {
      // Xn goes over G: after the barrier that follows P, nothing reads G
      const M X(Xb, kp), G(Gb, kp), Xn(Gb, kp);
      // G = X^T X and r = ||G - I||_F over the k x k block
      double r2 = 0.0;
      {
        const Opnd ox = X.opnd(M::KK, kp, kp);
        wg_gemm<TILE, M::KK, M::KK, PREC, Team::coh>(
            sm, ox, ox, kp, kp, true,
            [&](int m, int n, float4 v) {
              G.put4(m + n * kp, v);
              if (n < k) {
                const float z[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int c = 0; c < 4; ++c)
                  if (m + c < k) {
                    const float d = z[c] - (m + c == n ? 1.f : 0.f);
                    r2 += (double)d * d;
                  }
              }
            },
            nw, tw);
      }
      {
        bool ok;
        r2 = team.sum(r2, red, &ok);
        if (!ok) return PolarOut{total + step, false, true};
      }
      const double r = sqrt(r2);
      if (trace && threadIdx.x == 0) trace[step] = (float)r;
      if (!isfinite(r)) {
        ++step;
        break;
      }
      const bool finish = (r < a.tol);
      if (finish) on_fin();
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
      // P = ca I + cb G + cc G^2, zero outside the k x k block. In a team, with
      // cc != 0, the product is G (G + (cb / cc) I), whose G term comes from the
      // operand rows already staged in LDS, so the epilogue reads nothing (a cc
      // so small that cb / cc overflows leaves G^2 below the rounding of cb G);
      // one workgroup alone reads G back in the epilogue, which measured faster
      // there.
      if (Team::nw > 1 && cc != 0.f && isfinite(cb / cc)) {
        const Opnd og = G.opnd(M::KK, kp, kp);
        wg_gemm<TILE, M::KK, M::KK, PREC, Team::coh>(
            sm, og, og, kp, kp, true,
            [&](int m, int n, float4 v) {
              float4 pv = make_float4(0.f, 0.f, 0.f, 0.f);
              if (n < k && m < k) {
                pv.x = (m == n ? ca : 0.f) + cc * v.x;
                pv.y = (m + 1 < k) ? (m + 1 == n ? ca : 0.f) + cc * v.y : 0.f;
                pv.z = (m + 2 < k) ? (m + 2 == n ? ca : 0.f) + cc * v.z : 0.f;
                pv.w = (m + 3 < k) ? (m + 3 == n ? ca : 0.f) + cc * v.w : 0.f;
              }
              P.put4(m + n * kp, pv);
            },
            nw, tw, cb / cc);
      } else if (cc != 0.f) {
        const Opnd og = G.opnd(M::KK, kp, kp);
        wg_gemm<TILE, M::KK, M::KK, PREC, Team::coh>(
            sm, og, og, kp, kp, true,
            [&](int m, int n, float4 v) {
              const int i = m + n * kp;
              float4 pv = make_float4(0.f, 0.f, 0.f, 0.f);
              if (n < k && m < k) {
                const float4 g = G.template get4<Team::coh>(i);
                pv.x = (m == n ? ca : 0.f) + cb * g.x + cc * v.x;
                pv.y = (m + 1 < k) ? (m + 1 == n ? ca : 0.f) + cb * g.y + cc * v.y : 0.f;
                pv.z = (m + 2 < k) ? (m + 2 == n ? ca : 0.f) + cb * g.z + cc * v.z : 0.f;
                pv.w = (m + 3 < k) ? (m + 3 == n ? ca : 0.f) + cb * g.w + cc * v.w : 0.f;
              }
              P.put4(i, pv);
            },
            nw, tw);
      } else {
        for (int64_t e = team.e0(); e < (int64_t)kp * kp; e += team.es()) {
          const int m = (int)(e % kp), n = (int)(e / kp);
          float pv = 0.f;
          if (m < k && n < k) pv = (m == n ? ca : 0.f) + cb * G.template get<Team::coh>(e);
          P.put(e, pv);
        }
      }
      if (!team.sync()) return PolarOut{total + step, false, true};
      // Xn = X P; the finishing product goes to fin(m, n, v)
      {
        const Opnd ox = X.opnd(M::KM, kp, kp);
        const Opnd op = P.opnd(M::KK, kp, kp);
        if (finish) {
          wg_gemm<TILE, M::KM, M::KK, PREC, Team::coh>(
              sm, ox, op, kp, kp, false, [&](int m, int n, float4 v) { fin(m, n, v); }, nw, tw);
        } else {
          wg_gemm<TILE, M::KM, M::KK, PREC, Team::coh>(
              sm, ox, op, kp, kp, false, [&](int m, int n, float4 v) { Xn.put4(m + n * kp, v); }, nw, tw);
        }
      }
      if (!team.sync()) return PolarOut{total + step, false, true};
      if (finish) {
        done = true;
        ++step;
        break;
      }
      float *t = Xb;
      Xb = Gb;
      Gb = t;
    }
    total += step;
  }
  return PolarOut{total, done, false};
}

#endif  // K_POLAR_H
