// k_noalloc.hip -- ha_procrustes on the caller's device buffers, without any
// allocation: for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The work of one call (the column-major FP64 copies of X and Y, the FP64 and
// FP32 local transform, the FP64 Jacobi workspace and the index arrays of one
// searchlight) lives in the file-scope device array `scratch` below, of
// HA_NOALLOC_SCRATCH_BYTES. Each GPU holds its own copy of it until the
// process ends. Its device address comes from hipGetSymbolAddress, once per
// device with that device current (in ha_hip_init for device 0, else at the
// first call on a device), and only that address is used on the host.
//
// Every call on one GPU shares that GPU's array: two calls that overlap on one
// GPU (on different streams, or from different host threads) race. Calls in
// order on one stream are safe: each call rewrites what it reads first.
//
// Same arithmetic as ha_procrustes (the FP64 one-sided Jacobi kernel of the
// searchlight path on one N-vertex searchlight), so its results equal that
// entry point's bit for bit.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <stddef.h>
#include "hak.h"
#include "ha_common.h"
#include "ha_procrustes.h"

#define LAUNCH_RET() return (int)hipGetLastError()

namespace {

constexpr int KPMAX = HA_NOALLOC_MAX_N;  // a multiple of 32
// byte offsets into scratch (all multiples of 256)
constexpr int64_t OFF_X = 0;                                   // FP64 X, column-major M x N
constexpr int64_t OFF_Y = OFF_X + (int64_t)HA_NOALLOC_MAX_MN * 8;  // FP64 Y, column-major
constexpr int64_t OFF_Q64 = OFF_Y + (int64_t)HA_NOALLOC_MAX_MN * 8;  // kp x kp FP64 transform
constexpr int64_t OFF_Q = OFF_Q64 + (int64_t)KPMAX * KPMAX * 8;     // kp x kp FP32 transform
constexpr int64_t OFF_WORK = OFF_Q + (int64_t)KPMAX * KPMAX * 4;    // 4 kp x kp FP64 Jacobi buffers
constexpr int64_t OFF_SMALL = OFF_WORK + 4 * (int64_t)KPMAX * KPMAX * 8;
constexpr int64_t SMALL_BYTES = 64 * 1024;
static_assert(OFF_SMALL + SMALL_BYTES == HA_NOALLOC_SCRATCH_BYTES, "scratch layout");
static_assert(OFF_Q % 256 == 0 && OFF_WORK % 256 == 0 && OFF_SMALL % 256 == 0, "scratch alignment");

// index arrays and per-call scalars inside the small area
struct Small {
  int32_t iota[KPMAX];  // vertex list of the one searchlight: 0 .. N-1
  int32_t off[2];       // searchlight offsets {0, N}
  int32_t qld[1];       // kp
  int32_t steps[1];
  int32_t queue[4];     // work counter of the Jacobi launch
  int64_t qoff[1];      // 0
  double sigma_sum[1], var_sum[1], scale[1];
};
static_assert(sizeof(Small) <= SMALL_BYTES, "small area");

}  // namespace

static __device__ __align__(256) unsigned char scratch[HA_NOALLOC_SCRATCH_BYTES];

// Device address of scratch per device (host side): set by hak_noalloc_attach.
static void *g_scratch[HA_NOALLOC_MAX_DEVICES];

extern "C" int hak_noalloc_attach(void) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= HA_NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
  if (g_scratch[dev]) return 0;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return (int)e;
  g_scratch[dev] = p;
  return 0;
}

// Row-major M x N inputs -> column-major FP64 copies; the index arrays and the
// scalars that must start at zero for this call.
__global__ void k_noalloc_prepare(const double *__restrict__ X, const double *__restrict__ Y, int32_t M, int32_t N,
                                  int32_t kp, double *__restrict__ Xc, double *__restrict__ Yc, Small *sm) {
  const int64_t n = (int64_t)M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / N, j = e % N;  // element (t, j) of the row-major input
    Xc[t + j * M] = X[e];
    Yc[t + j * M] = Y[e];
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) sm->iota[i] = i;
    if (threadIdx.x == 0) {
      sm->off[0] = 0;
      sm->off[1] = N;
      sm->qld[0] = kp;
      sm->qoff[0] = 0;
      sm->steps[0] = 0;
      for (int q = 0; q < 4; ++q) sm->queue[q] = 0;
      sm->sigma_sum[0] = 0.0;
      sm->var_sum[0] = 0.0;
      sm->scale[0] = 1.0;
    }
  }
}

// T (row-major N x N) = Q64 (column-major, leading dim kp), times the scale
// when scaling (the product the host applies in ha_procrustes)
__global__ void k_noalloc_out(const double *__restrict__ q64, int32_t kp, int32_t N, const Small *sm, int scaling,
                              double *__restrict__ T) {
  const int64_t n = (int64_t)N * N;
  const double sc = sm->scale[0];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / N, c = e % N;
    const double v = q64[r + c * kp];
    T[e] = scaling ? v * sc : v;
  }
}

extern "C" int ha_procrustes_noalloc(const double *X, const double *Y, int32_t M, int32_t N, double *T,
                                     bool isReflection, bool isScaling, void *stream) {
  if (!X || !Y || !T || M <= 0 || N <= 0) return kHaErrorArg;
  if (N > HA_NOALLOC_MAX_N || (int64_t)M * N > HA_NOALLOC_MAX_MN) {
    fprintf(stderr, "ha_hip: ha_procrustes_noalloc holds N <= %d vertices and M * N <= %d values "
            "(got M = %d, N = %d)\n", HA_NOALLOC_MAX_N, HA_NOALLOC_MAX_MN, M, N);
    return kHaErrorArg;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= HA_NOALLOC_MAX_DEVICES) {
    (void)hipGetLastError();
    fprintf(stderr, "ha_hip: ha_procrustes_noalloc: no usable current device\n");
    return kHaErrorInternal;
  }
  if (!g_scratch[dev] && hak_noalloc_attach() != 0) {
    (void)hipGetLastError();
    fprintf(stderr, "ha_hip: ha_procrustes_noalloc: hipGetSymbolAddress of the scratch array failed\n");
    return kHaErrorInternal;
  }
  unsigned char *base = static_cast<unsigned char *>(g_scratch[dev]);
  double *Xc = reinterpret_cast<double *>(base + OFF_X);
  double *Yc = reinterpret_cast<double *>(base + OFF_Y);
  double *q64 = reinterpret_cast<double *>(base + OFF_Q64);
  float *q = reinterpret_cast<float *>(base + OFF_Q);
  double *work = reinterpret_cast<double *>(base + OFF_WORK);
  // device addresses inside the small area, from offsets only (nothing here
  // is dereferenced on the host)
  unsigned char *small = base + OFF_SMALL;
  Small *sm = reinterpret_cast<Small *>(small);
  int32_t *iota = reinterpret_cast<int32_t *>(small + offsetof(Small, iota));
  int32_t *off = reinterpret_cast<int32_t *>(small + offsetof(Small, off));
  int32_t *qld = reinterpret_cast<int32_t *>(small + offsetof(Small, qld));
  int32_t *steps = reinterpret_cast<int32_t *>(small + offsetof(Small, steps));
  int32_t *queue = reinterpret_cast<int32_t *>(small + offsetof(Small, queue));
  int64_t *qoff = reinterpret_cast<int64_t *>(small + offsetof(Small, qoff));
  double *sigma_sum = reinterpret_cast<double *>(small + offsetof(Small, sigma_sum));
  double *var_sum = reinterpret_cast<double *>(small + offsetof(Small, var_sum));
  double *scale = reinterpret_cast<double *>(small + offsetof(Small, scale));
  const int32_t kp = (N + 31) / 32 * 32;
  hipStream_t st = (hipStream_t)stream;

  const int64_t n = (int64_t)M * N;
  int blocks = (int)((n + 255) / 256);
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_noalloc_prepare, dim3(blocks), dim3(256), 0, st, X, Y, M, N, kp, Xc, Yc, sm);
  if (hipGetLastError() != hipSuccess) return kHaErrorInternal;

  HakLocalArgs a;
  memset(&a, 0, sizeof(a));
  a.X64 = Xc;
  a.Y64 = Yc;
  a.nt = M;
  a.ld = M;
  a.slx_idx = iota;
  a.sly_idx = iota;
  a.sl_off = off;
  a.n_order = 1;
  a.q = q;
  a.qoff = qoff;
  a.qld = qld;
  a.q64 = q64;
  a.sigma_sum = isScaling ? sigma_sum : nullptr;
  a.flag = nullptr;
  a.steps = steps;
  a.queue = queue;
  if (hak_jacobi64(&a, isReflection ? 0 : 1, work, (int64_t)kp * kp, 1, 1, st) != 0) return kHaErrorInternal;
  if (isScaling) {
    if (hak_var_sum(Xc, M, iota, off, 1, var_sum, st) != 0) return kHaErrorInternal;
    if (hak_scale(sigma_sum, var_sum, M, 1, scale, st) != 0) return kHaErrorInternal;
  }
  const int64_t nn = (int64_t)N * N;
  blocks = (int)((nn + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_noalloc_out, dim3(blocks), dim3(256), 0, st, q64, kp, N, sm, isScaling ? 1 : 0, T);
  if (hipGetLastError() != hipSuccess) return kHaErrorInternal;
  return kHaSuccess;
}
