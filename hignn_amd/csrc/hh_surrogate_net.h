// The far-field surrogate net shared by its training kernels
// (hh_surrogate.hip) and the ACA entry and judge kernels (hh_kernels.hip).
//
//   s(r) ~ |r| S(M(r)),  z = r / r_scale,
//   3 -> 32 -> 32 -> 6, SiLU hidden layers (silu x = x / (1 + e^-x)),
//
// outputs (xx, yy, zz, xy, yz, zx) of the symmetrised block times |r|.
// Parameter vector (1,382 floats): W1 [32][3], b1 [32], W2 [32][32], b2 [32],
// W3 [6][32], b3 [6].  One sample per thread, FP32 on the VALU; the
// parameters are read from LDS (the same address in every lane: broadcast).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hh {
namespace sur {

constexpr int kIn = 3, kH = 32, kOut = 6;
constexpr int kOW1 = 0, kOB1 = kOW1 + kH * kIn, kOW2 = kOB1 + kH, kOB2 = kOW2 + kH * kH,
              kOW3 = kOB2 + kH, kOB3 = kOW3 + kOut * kH, kNP = kOB3 + kOut;
static_assert(kNP == 1382, "surrogate parameter count");

// (row, col) of output k; weight of its squared error in the block's
// Frobenius norm (off-diagonal outputs stand for two entries)
__device__ __forceinline__ int entry_row(int k) { return k < 3 ? k : (k == 3 ? 0 : (k == 4 ? 1 : 2)); }
__device__ __forceinline__ int entry_col(int k) { return k < 3 ? k : (k == 3 ? 1 : (k == 4 ? 2 : 0)); }
__device__ __forceinline__ float entry_weight(int k) { return k < 3 ? 1.0f : 2.0f; }

// sigma(x) = 1 / (1 + 2^(-x log2 e)); 2^t overflows to inf for very negative
// x and gives sigma = 0, as the limit
__device__ __forceinline__ float sigmoid(float x) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * x));
}
__device__ __forceinline__ float silu(float x) { return x * sigmoid(x); }
__device__ __forceinline__ float dsilu(float x) {
  const float s = sigmoid(x);
  return s * __builtin_fmaf(x, 1.0f - s, 1.0f);
}

// Forward pass of one sample; keeps the pre-activations for the backward pass.
__device__ __forceinline__ void forward(const float *P, const float (&z)[kIn], float (&x1)[kH],
                                        float (&x2)[kH], float (&y)[kOut]) {
  float a1[kH];
#pragma unroll
  for (int j = 0; j < kH; j++) {
    float v = P[kOB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; k++) v = __builtin_fmaf(P[kOW1 + j * kIn + k], z[k], v);
    x1[j] = v;
    a1[j] = silu(v);
  }
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float v = P[kOB2 + i];
#pragma unroll
    for (int j = 0; j < kH; j++) v = __builtin_fmaf(P[kOW2 + i * kH + j], a1[j], v);
    x2[i] = v;
  }
#pragma unroll
  for (int o = 0; o < kOut; o++) {
    float v = P[kOB3 + o];
#pragma unroll
    for (int i = 0; i < kH; i++) v = __builtin_fmaf(P[kOW3 + o * kH + i], silu(x2[i]), v);
    y[o] = v;
  }
}

// Outputs only (the ACA entry and judge kernels).
__device__ __forceinline__ void eval(const float *P, float rx, float ry, float rz, float inv_scale,
                                     float (&y)[kOut]) {
  const float z[kIn] = {rx * inv_scale, ry * inv_scale, rz * inv_scale};
  float x1[kH], x2[kH];
  forward(P, z, x1, x2, y);
}

// Copies the parameters into LDS (all threads of the block take part).
__device__ __forceinline__ void stage(const float *params, float *P) {
  for (int i = threadIdx.x; i < kNP; i += blockDim.x) P[i] = params[i];
  __syncthreads();
}

}  // namespace sur
}  // namespace hh
