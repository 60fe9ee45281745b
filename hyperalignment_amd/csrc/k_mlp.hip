// k_mlp.hip -- training and batch inference of the spectral-scaling MLP
// (12 -> 64 -> 64 -> 2, exact GELU) on gfx950.
//
// Loss per sample: sum_j wgt_j * (out_j - y_j)^2, averaged over the batch.
// Gradients are reduced deterministically: each workgroup writes the partial
// gradient of its chunk of samples, and one pass sums the partials in a fixed
// order before the Adam update.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hak.h"
#include "k_mlp.h"

#define LAUNCH_RET() return (int)hipGetLastError()
#define MT 64    // samples per workgroup chunk (= threads per workgroup)
#define IN HAK_MLP_IN
#define HID HAK_MLP_HID
#define OUTN HAK_MLP_OUT

// Per-chunk activations, one sample per thread (row t).
struct MlpSmem {
  float x[MT][IN];        // normalised inputs
  float h1[MT][HID + 1];  // gelu(z1)
  float d1[MT][HID + 1];  // dL/dz1
  float h2[MT][HID + 1];  // gelu(z2)
  float d2[MT][HID + 1];  // dL/dz2
  float dout[MT][OUTN];   // dL/dout
};

// Forward of one row; returns z1, z2 in registers through h1/h2 slots (pre-activations).
__device__ __forceinline__ void row_forward(const float *__restrict__ P, int t, MlpSmem &sm, float *out) {
  for (int j = 0; j < HID; ++j) {
    float acc = P[HAK_MLP_OFF_B1 + j];
    for (int i = 0; i < IN; ++i) acc += sm.x[t][i] * P[HAK_MLP_OFF_W1 + i * HID + j];
    sm.d1[t][j] = acc;                  // z1 (kept for gelu')
    sm.h1[t][j] = hak_gelu(acc);
  }
  for (int j = 0; j < HID; ++j) {
    float acc = P[HAK_MLP_OFF_B2 + j];
    for (int i = 0; i < HID; ++i) acc += sm.h1[t][i] * P[HAK_MLP_OFF_W2 + i * HID + j];
    sm.d2[t][j] = acc;                  // z2 (kept for gelu')
    sm.h2[t][j] = hak_gelu(acc);
  }
  for (int j = 0; j < OUTN; ++j) {
    float acc = P[HAK_MLP_OFF_B3 + j];
    for (int i = 0; i < HID; ++i) acc += sm.h2[t][i] * P[HAK_MLP_OFF_W3 + i * OUTN + j];
    out[j] = acc;
  }
}

__device__ __forceinline__ void load_row(const float *__restrict__ X, const float *__restrict__ norm, int64_t row,
                                         int t, MlpSmem &sm) {
  for (int i = 0; i < IN; ++i) sm.x[t][i] = (X[row * IN + i] - norm[i]) / norm[IN + i];
}

__global__ __launch_bounds__(MT) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ norm,
                                                    const float *__restrict__ X, int64_t N, float *__restrict__ Y) {
  __shared__ MlpSmem sm;
  const int t = threadIdx.x;
  for (int64_t base = (int64_t)blockIdx.x * MT; base < N; base += (int64_t)gridDim.x * MT) {
    if (base + t < N) {
      float out[OUTN];
      load_row(X, norm, base + t, t, sm);
      row_forward(P, t, sm, out);
      for (int j = 0; j < OUTN; ++j) Y[(base + t) * OUTN + j] = out[j];
    }
  }
}

extern "C" int hak_mlp_forward(const float *params, const float *norm, const float *X, int64_t N, float *Y,
                               void *stream) {
  if (N <= 0) return 0;
  int64_t g = (N + MT - 1) / MT;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_mlp_forward, dim3((int)g), dim3(MT), 0, (hipStream_t)stream, params, norm, X, N, Y);
  LAUNCH_RET();
}

// Partial gradient of chunk c: rows idx[c*MT + t] (identity order when idx is
// NULL) of the nb-row batch. partial[c][NPARAM] and loss_part[c] (loss of the
// chunk, already divided by nb).
__global__ __launch_bounds__(MT) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ norm,
                                                 const float *__restrict__ X, const float *__restrict__ Yt,
                                                 const int32_t *__restrict__ idx, int64_t nb, float w0, float w1,
                                                 float inv_batch, float *__restrict__ partial,
                                                 float *__restrict__ loss_part) {
  __shared__ MlpSmem sm;
  __shared__ float lred[MT];
  const int c = blockIdx.x, t = threadIdx.x;
  const int64_t base = (int64_t)c * MT;
  const int n = (int)min((int64_t)MT, nb - base);
  if (t < n) {
    const int64_t row = idx ? (int64_t)idx[base + t] : base + t;
    load_row(X, norm, row, t, sm);
    float out[OUTN];
    row_forward(P, t, sm, out);
    const float wj[OUTN] = {w0, w1};
    float l = 0.f;
    for (int j = 0; j < OUTN; ++j) {
      const float e = out[j] - Yt[row * OUTN + j];
      l += wj[j] * e * e;
      sm.dout[t][j] = 2.f * wj[j] * e * inv_batch;
    }
    lred[t] = l * inv_batch;
    // d2 = (dout W3^T) * gelu'(z2); z2 is in d2's slot
    for (int i = 0; i < HID; ++i) {
      float acc = 0.f;
      for (int j = 0; j < OUTN; ++j) acc += sm.dout[t][j] * P[HAK_MLP_OFF_W3 + i * OUTN + j];
      sm.d2[t][i] = acc * hak_gelu_grad(sm.d2[t][i]);
    }
    // d1 = (d2 W2^T) * gelu'(z1); z1 is in d1's slot
    for (int i = 0; i < HID; ++i) {
      float acc = 0.f;
      for (int j = 0; j < HID; ++j) acc += sm.d2[t][j] * P[HAK_MLP_OFF_W2 + i * HID + j];
      sm.d1[t][i] = acc * hak_gelu_grad(sm.d1[t][i]);
    }
  } else {
    lred[t] = 0.f;
  }
  __syncthreads();
  float *out = partial + (int64_t)c * HAK_MLP_NPARAM;
  for (int p = t; p < HID * OUTN; p += MT) {  // dW3 = h2^T dout
    const int i = p / OUTN, j = p % OUTN;
    float acc = 0.f;
    for (int s2 = 0; s2 < n; ++s2) acc += sm.h2[s2][i] * sm.dout[s2][j];
    out[HAK_MLP_OFF_W3 + p] = acc;
  }
  for (int j = t; j < OUTN; j += MT) {  // db3
    float acc = 0.f;
    for (int s2 = 0; s2 < n; ++s2) acc += sm.dout[s2][j];
    out[HAK_MLP_OFF_B3 + j] = acc;
  }
  for (int p = t; p < HID * HID; p += MT) {  // dW2 = h1^T d2
    const int i = p / HID, j = p % HID;
    float acc = 0.f;
    for (int s2 = 0; s2 < n; ++s2) acc += sm.h1[s2][i] * sm.d2[s2][j];
    out[HAK_MLP_OFF_W2 + p] = acc;
  }
  for (int j = t; j < HID; j += MT) {  // db2
    float acc = 0.f;
    for (int s2 = 0; s2 < n; ++s2) acc += sm.d2[s2][j];
    out[HAK_MLP_OFF_B2 + j] = acc;
  }
  for (int p = t; p < IN * HID; p += MT) {  // dW1 = x^T d1
    const int q = p / HID, i = p % HID;
    float acc = 0.f;
    for (int s2 = 0; s2 < n; ++s2) acc += sm.x[s2][q] * sm.d1[s2][i];
    out[HAK_MLP_OFF_W1 + p] = acc;
  }
  for (int i = t; i < HID; i += MT) {  // db1
    float acc = 0.f;
    for (int s2 = 0; s2 < n; ++s2) acc += sm.d1[s2][i];
    out[HAK_MLP_OFF_B1 + i] = acc;
  }
  if (t == 0) {
    float l = 0.f;
    for (int s2 = 0; s2 < MT; ++s2) l += lred[s2];
    loss_part[c] = l;
  }
}

// Sums the partial gradients of nc chunks in chunk order, then applies Adam
// (beta1, beta2, eps; bias-corrected with step count `step`). grad_out (may be
// NULL) receives the summed gradient; lr == 0 skips the update.
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ partial, int nc, float lr, float b1, float b2, float eps,
                           int step, float *__restrict__ grad_out) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < HAK_MLP_NPARAM; p += gridDim.x * blockDim.x) {
    float g = 0.f;
    for (int c = 0; c < nc; ++c) g += partial[(int64_t)c * HAK_MLP_NPARAM + p];
    if (grad_out) grad_out[p] = g;
    if (lr != 0.f) {
      const float mm = b1 * m[p] + (1.f - b1) * g;
      const float vv = b2 * v[p] + (1.f - b2) * g * g;
      m[p] = mm;
      v[p] = vv;
      const float mh = mm / (1.f - powf(b1, (float)step));
      const float vh = vv / (1.f - powf(b2, (float)step));
      P[p] -= lr * mh / (sqrtf(vh) + eps);
    }
  }
}

__global__ void k_sum_f32(const float *__restrict__ x, int n, float *out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += x[i];
    *out = s;
  }
}

extern "C" int hak_mlp_grad_step(float *params, float *adam_m, float *adam_v, const float *norm, const float *X,
                                 const float *Y, const int32_t *idx, int64_t nb, float w0, float w1, float lr,
                                 int step, float *partial, float *loss_part, float *loss_out, float *grad_out,
                                 void *stream) {
  if (nb <= 0) return 0;
  const int nc = (int)((nb + MT - 1) / MT);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nc), dim3(MT), 0, (hipStream_t)stream, params, norm, X, Y, idx, nb, w0, w1,
                     1.0f / (float)nb, partial, loss_part);
  hipLaunchKernelGGL(k_mlp_adam, dim3((HAK_MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, params,
                     adam_m, adam_v, partial, nc, lr, 0.9f, 0.999f, 1e-8f, step, grad_out);
  if (loss_out) hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(64), 0, (hipStream_t)stream, loss_part, nc, loss_out);
  LAUNCH_RET();
}
