// k_mlp.h -- device code of the spectral-scaling MLP (12 -> 64 -> 64 -> 2, GELU).
//
// Parameter layout (FP32, contiguous): W1[IN][HID], b1[HID], W2[HID][HID],
// b2[HID], W3[HID][OUT], b3[OUT]; W[i][j] connects input unit i to output j.
// Inputs are normalised as (f - mean) / std with mean[IN], std[IN].
#ifndef K_MLP_H
#define K_MLP_H

#include <hip/hip_runtime.h>

#define HAK_MLP_IN 12
#define HAK_MLP_HID 64
#define HAK_MLP_OUT 2
#define HAK_MLP_OFF_W1 0
#define HAK_MLP_OFF_B1 (HAK_MLP_OFF_W1 + HAK_MLP_IN * HAK_MLP_HID)
#define HAK_MLP_OFF_W2 (HAK_MLP_OFF_B1 + HAK_MLP_HID)
#define HAK_MLP_OFF_B2 (HAK_MLP_OFF_W2 + HAK_MLP_HID * HAK_MLP_HID)
#define HAK_MLP_OFF_W3 (HAK_MLP_OFF_B2 + HAK_MLP_HID)
#define HAK_MLP_OFF_B3 (HAK_MLP_OFF_W3 + HAK_MLP_HID * HAK_MLP_OUT)
#define HAK_MLP_NPARAM (HAK_MLP_OFF_B3 + HAK_MLP_OUT)

// exact GELU: x * Phi(x)
__device__ __forceinline__ float hak_gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
// d/dx GELU = Phi(x) + x * phi(x)
__device__ __forceinline__ float hak_gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Forward pass for one feature row by a whole workgroup (blockDim >= 64).
// feat: HAK_MLP_IN floats in LDS; scratch: 2*HID+2 floats in LDS; out: 2 floats
// (identical in every thread on return).
__device__ __forceinline__ void hak_mlp_forward_wg(const float *__restrict__ P, const float *__restrict__ norm,
                                                   const float *feat, float *scratch, float *out) {
  float *h1 = scratch, *h2 = scratch + HAK_MLP_HID, *o = scratch + 2 * HAK_MLP_HID;
  const int t = threadIdx.x;
  if (t < HAK_MLP_HID) {
    float acc = P[HAK_MLP_OFF_B1 + t];
    for (int i = 0; i < HAK_MLP_IN; ++i) {
      float x = (feat[i] - norm[i]) / norm[HAK_MLP_IN + i];
      acc += x * P[HAK_MLP_OFF_W1 + i * HAK_MLP_HID + t];
    }
    h1[t] = hak_gelu(acc);
  }
  __syncthreads();
  if (t < HAK_MLP_HID) {
    float acc = P[HAK_MLP_OFF_B2 + t];
    for (int i = 0; i < HAK_MLP_HID; ++i) acc += h1[i] * P[HAK_MLP_OFF_W2 + i * HAK_MLP_HID + t];
    h2[t] = hak_gelu(acc);
  }
  __syncthreads();
  if (t < HAK_MLP_OUT) {
    float acc = P[HAK_MLP_OFF_B3 + t];
    for (int i = 0; i < HAK_MLP_HID; ++i) acc += h2[i] * P[HAK_MLP_OFF_W3 + i * HAK_MLP_OUT + t];
    o[t] = acc;
  }
  __syncthreads();
  out[0] = o[0];
  out[1] = o[1];
  __syncthreads();
}

#endif  // K_MLP_H
