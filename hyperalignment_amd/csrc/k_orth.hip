// k_orth.hip -- FP64 orthogonality check of the local transforms.
//
// For a local transform Q_s (k x k, FP32, column-major at q + qoff[s] with
// leading dimension qld[s]) the check measures e_s = ||Q_s^T Q_s - I||_F in
// FP64 arithmetic (products of FP32 values are exact in FP64):
//   * exact: every entry of the Gram matrix G = Q^T Q on v_mfma_f64_16x16x4f64
//     32 x 32 tiles (only tiles I <= J; an off-diagonal tile counts twice);
//   * probe: the estimate e^2 ~ (1/4) sum_p ||Q^T (Q z_p) - z_p||^2 over four
//     Rademacher vectors z_p, whose expectation is ||Q^T Q - I||_F^2 for any
//     Q. Two streaming reads of Q; cheap enough to run in every pass.
// A searchlight whose e_s exceeds tol (or is not finite) is flagged for the
// FP64 Jacobi path when flag is given. Searchlights already flagged get
// e_s = -1 (not checked: their Q store does not hold a result yet).
//
// The reference computes Q = U V^T in FP64 LAPACK and has no such check
// (csrc/ha_searchlight.c:593-609); the polar iteration here produces Q in
// bf16x3 arithmetic, and this check is what sends a transform that is not
// orthogonal to working precision to the FP64 path.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "hak.h"
#include "k_chol.h"

#define LAUNCH_RET() return (int)hipGetLastError()

using kchol::B;
using kchol::RT;
using kchol::f64x4_t;

namespace {

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  const double r = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return r;
}

// Rademacher signs of the four probes at row/column j (bit p: probe p is +1),
// from a splitmix64-style mix of j
__device__ __forceinline__ uint8_t probe_bits(int j) {
  uint64_t z = (uint64_t)(uint32_t)j + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint8_t)(z & 15u);
}

// exact ||Q^T Q - I||_F^2 of one k x k block (ld = kp, rows/cols >= k ignored)
__device__ double orth_exact(const float *__restrict__ Q, int k, int kp, double (*xs)[B + 1], double (*ys)[B + 1],
                             double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  double part = 0.0;
  for (int tm = 0; tm < kp; tm += B)
    for (int tn = tm; tn < kp; tn += B) {
      f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
      for (int r0 = 0; r0 < kp; r0 += B) {
        // xs[c][r] = Q(r0 + r, tm + c), ys[c][r] = Q(r0 + r, tn + c): 32 contiguous rows per column
        for (int e = threadIdx.x; e < B * B; e += RT) {
          const int c = e >> 5, r = e & 31;
          const int row = r0 + r;
          xs[c][r] = (row < k && tm + c < k) ? (double)Q[row + (int64_t)(tm + c) * kp] : 0.0;
          ys[c][r] = (row < k && tn + c < k) ? (double)Q[row + (int64_t)(tn + c) * kp] : 0.0;
        }
        __syncthreads();
        // G(i, j) += sum_r Q(r, tm + i) Q(r, tn + j): X(i, r) = xs[i][r], Y(r, j) = ys[j][r]
        acc = kchol::mfma32(xs, ys, true, acc);
        __syncthreads();
      }
      const double wgt = tm == tn ? 1.0 : 2.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm + wm + (lane >> 4) + 4 * r, n = tn + wn + (lane & 15);
        if (m < k && n < k) {
          const double d = acc[r] - (m == n ? 1.0 : 0.0);
          part += wgt * d * d;
        }
      }
    }
  return block_sum(part, red);
}

__device__ __forceinline__ double sgn(uint32_t bits, int p) { return ((bits >> p) & 1u) ? 1.0 : -1.0; }

// probe estimate of ||Q^T Q - I||_F^2; y: LDS for k x 4 doubles, zb: the four
// probe signs of every row as bits 0..3
__device__ double orth_probe(const float *__restrict__ Q, int k, int kp, double *y, const uint8_t *zb,
                             double *red) {
  // y_p = Q z_p: one row per thread, columns streamed (coalesced across the workgroup)
  for (int i0 = 0; i0 < k; i0 += RT) {
    const int i = i0 + threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (i < k) {
      for (int j = 0; j < k; ++j) {
        const double v = (double)Q[i + (int64_t)j * kp];
        const uint32_t b = zb[j];
        a0 += v * sgn(b, 0);
        a1 += v * sgn(b, 1);
        a2 += v * sgn(b, 2);
        a3 += v * sgn(b, 3);
      }
      y[4 * i + 0] = a0;
      y[4 * i + 1] = a1;
      y[4 * i + 2] = a2;
      y[4 * i + 3] = a3;
    }
  }
  __syncthreads();
  // w_p(j) = sum_i Q(i, j) y_p(i): one column per wave at a time, rows over the lanes
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double part = 0.0;
  for (int j = wave; j < k; j += RT / 64) {
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
    for (int i = lane; i < k; i += 64) {
      const double v = (double)Q[i + (int64_t)j * kp];
      w0 += v * y[4 * i + 0];
      w1 += v * y[4 * i + 1];
      w2 += v * y[4 * i + 2];
      w3 += v * y[4 * i + 3];
    }
    for (int o = 32; o > 0; o >>= 1) {
      w0 += __shfl_xor(w0, o);
      w1 += __shfl_xor(w1, o);
      w2 += __shfl_xor(w2, o);
      w3 += __shfl_xor(w3, o);
    }
    if (lane == 0) {
      const uint32_t b = zb[j];
      const double d0 = w0 - sgn(b, 0), d1 = w1 - sgn(b, 1);
      const double d2 = w2 - sgn(b, 2), d3 = w3 - sgn(b, 3);
      part += d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3;
    }
  }
  return 0.25 * block_sum(part, red);
}

}  // namespace

template <bool EXACT>
__global__ __launch_bounds__(RT) void k_orth(const float *__restrict__ q, const int64_t *__restrict__ qoff,
                                             const int32_t *__restrict__ qld, const int32_t *__restrict__ sl_off,
                                             const int32_t *__restrict__ order, int32_t n, int32_t kpmax,
                                             double *__restrict__ out, int32_t *flag, double tol) {
  extern __shared__ double ysh[];  // probe: kpmax x 4 doubles, then kpmax sign bytes
  __shared__ double red[4];
  uint8_t *zb = reinterpret_cast<uint8_t *>(ysh + (int64_t)kpmax * 4);
  if constexpr (!EXACT) {
    for (int j = threadIdx.x; j < kpmax; j += RT) zb[j] = probe_bits(j);
    __syncthreads();
  }
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int s = order ? order[it] : it;
    if (flag && flag[s] != 0) {
      if (threadIdx.x == 0) out[s] = -1.0;
      continue;
    }
    const int k = sl_off[s + 1] - sl_off[s];
    const int kp = qld[s];
    const float *Q = q + qoff[s];
    double e2;
    if constexpr (EXACT) {
      __shared__ double xs[B][B + 1], ys[B][B + 1];
      e2 = orth_exact(Q, k, kp, xs, ys, red);
    } else {
      e2 = orth_probe(Q, k, kp, ysh, zb, red);
    }
    if (threadIdx.x == 0) {
      const double e = sqrt(fmax(e2, 0.0));
      out[s] = e;
      if (flag && !(e <= tol)) flag[s] = 1;
    }
    __syncthreads();
  }
}

extern "C" int hak_orth(const float *q, const int64_t *qoff, const int32_t *qld, const int32_t *sl_off,
                        const int32_t *order, int32_t n, int32_t kpmax, int exact, double *out, int32_t *flag,
                        double tol, int32_t n_wg, void *stream) {
  if (n <= 0) return 0;
  if (!exact && (kpmax <= 0 || kpmax > HAK_ORTH_PROBE_MAX_KP)) return (int)hipErrorInvalidValue;
  if (n_wg <= 0 || n_wg > n) n_wg = n;
  if (exact)
    hipLaunchKernelGGL(k_orth<true>, dim3(n_wg), dim3(RT), 0, (hipStream_t)stream, q, qoff, qld, sl_off, order, n,
                       kpmax, out, flag, tol);
  else
    hipLaunchKernelGGL(k_orth<false>, dim3(n_wg), dim3(RT), (size_t)kpmax * (4 * sizeof(double) + 1),
                       (hipStream_t)stream, q, qoff, qld, sl_off, order, n, kpmax, out, flag, tol);
  LAUNCH_RET();
}
