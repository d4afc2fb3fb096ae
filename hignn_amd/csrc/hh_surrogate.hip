// Distilled far-field surrogate of the two-body pair function: a small net
// s(r) ~ |r| M(r) with inputs (r/|r|, ln(|r| / r0)), trained on the GPU by
// the kernels below (forward, backward, Adam) on pairs labelled by the full
// MLP.  4 -> 32 -> 32 -> 9, tanh hidden layers, linear output; mean squared
// error over the 9 outputs.
//
// Parameter vector (1,513 floats): W1 [32][4], b1 [32], W2 [32][32], b2 [32],
// W3 [9][32], b3 [9].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hh.h"
#include "../../include/hh_surrogate.h"
#include "hh_host.h"

namespace hh {
namespace {

constexpr int kIn = 4, kH = 32, kOut = 9;
constexpr int kOW1 = 0, kOB1 = kOW1 + kH * kIn, kOW2 = kOB1 + kH, kOB2 = kOW2 + kH * kH,
              kOW3 = kOB2 + kH, kOB3 = kOW3 + kOut * kH, kNP = kOB3 + kOut;
static_assert(kNP == HH_SURROGATE_PARAMS, "parameter count");
constexpr int kTB = 128;  // samples (threads) per training block

__device__ __forceinline__ void features(const float *r, float r0, float (&x)[kIn]) {
  const float d = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  const float inv = d > 0.0f ? 1.0f / d : 0.0f;
  x[0] = r[0] * inv;
  x[1] = r[1] * inv;
  x[2] = r[2] * inv;
  x[3] = logf(fmaxf(d, 1e-30f) / r0);
}

// Forward pass of one sample; a1, a2: hidden activations.
__device__ __forceinline__ void forward(const float *P, const float (&x)[kIn], float (&a1)[kH],
                                        float (&a2)[kH], float (&y)[kOut]) {
#pragma unroll
  for (int j = 0; j < kH; j++) {
    float z = P[kOB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; k++) z = fmaf(P[kOW1 + j * kIn + k], x[k], z);
    a1[j] = tanhf(z);
  }
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float z = P[kOB2 + i];
#pragma unroll
    for (int j = 0; j < kH; j++) z = fmaf(P[kOW2 + i * kH + j], a1[j], z);
    a2[i] = tanhf(z);
  }
#pragma unroll
  for (int o = 0; o < kOut; o++) {
    float z = P[kOB3 + o];
#pragma unroll
    for (int i = 0; i < kH; i++) z = fmaf(P[kOW3 + o * kH + i], a2[i], z);
    y[o] = z;
  }
}

__global__ __launch_bounds__(256) void surr_eval_kernel(const float *params, const float *r,
                                                        int64_t n, float r0, float *out) {
  __shared__ float P[kNP];
  for (int i = threadIdx.x; i < kNP; i += blockDim.x) P[i] = params[i];
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn], a1[kH], a2[kH], y[kOut];
  features(r + 3 * s, r0, x);
  forward(P, x, a1, a2, y);
#pragma unroll
  for (int o = 0; o < kOut; o++) out[kOut * s + o] = y[o];
}

// Gradient of the loss L = (1 / (9 n)) sum_s |y_s - t_s|^2 over the n samples,
// accumulated into grad (zeroed by the caller), and L into loss[0].  One
// sample per thread; the block's activations and output errors go through
// LDS, each thread sums a share of the weight gradients over the block's
// samples and adds them to grad once.
__global__ __launch_bounds__(kTB) void surr_grad_kernel(const float *params, const float *r,
                                                        const float *t, int64_t n, float r0,
                                                        float *grad, float *loss) {
  __shared__ float P[kNP];
  __shared__ float sx[kTB][kIn], sa1[kTB][kH + 1], sd1[kTB][kH + 1], sa2[kTB][kH + 1],
      sd2[kTB][kH + 1], sdy[kTB][kOut];
  __shared__ float sl[kTB];
  for (int i = threadIdx.x; i < kNP; i += blockDim.x) P[i] = params[i];
  __syncthreads();
  const int tid = threadIdx.x;
  const int64_t s = (int64_t)blockIdx.x * kTB + tid;
  const float scale = 2.0f / (9.0f * (float)n);
  float x[kIn] = {0, 0, 0, 0}, a1[kH], a2[kH], y[kOut], dy[kOut];
  float l = 0.0f;
  if (s < n) {
    features(r + 3 * s, r0, x);
    forward(P, x, a1, a2, y);
#pragma unroll
    for (int o = 0; o < kOut; o++) {
      const float e = y[o] - t[kOut * s + o];
      l = fmaf(e, e, l);
      dy[o] = scale * e;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kH; j++) a1[j] = a2[j] = 0.0f;
#pragma unroll
    for (int o = 0; o < kOut; o++) dy[o] = 0.0f;
  }
  // back through layer 3 and 2
  float d2[kH], d1[kH];
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float g = 0.0f;
#pragma unroll
    for (int o = 0; o < kOut; o++) g = fmaf(P[kOW3 + o * kH + i], dy[o], g);
    d2[i] = g * (1.0f - a2[i] * a2[i]);
  }
#pragma unroll
  for (int j = 0; j < kH; j++) {
    float g = 0.0f;
#pragma unroll
    for (int i = 0; i < kH; i++) g = fmaf(P[kOW2 + i * kH + j], d2[i], g);
    d1[j] = g * (1.0f - a1[j] * a1[j]);
  }
#pragma unroll
  for (int k = 0; k < kIn; k++) sx[tid][k] = x[k];
#pragma unroll
  for (int j = 0; j < kH; j++) sa1[tid][j] = a1[j], sd1[tid][j] = d1[j];
#pragma unroll
  for (int i = 0; i < kH; i++) sa2[tid][i] = a2[i], sd2[tid][i] = d2[i];
#pragma unroll
  for (int o = 0; o < kOut; o++) sdy[tid][o] = dy[o];
  sl[tid] = l;
  __syncthreads();
  // weight gradients: parameter p summed over the block's samples
  for (int p = tid; p < kNP; p += kTB) {
    float g = 0.0f;
    if (p < kOB1) {  // W1[j][k]
      const int j = p / kIn, k = p % kIn;
      for (int q = 0; q < kTB; q++) g = fmaf(sd1[q][j], sx[q][k], g);
    } else if (p < kOW2) {  // b1[j]
      const int j = p - kOB1;
      for (int q = 0; q < kTB; q++) g += sd1[q][j];
    } else if (p < kOB2) {  // W2[i][j]
      const int i = (p - kOW2) / kH, j = (p - kOW2) % kH;
      for (int q = 0; q < kTB; q++) g = fmaf(sd2[q][i], sa1[q][j], g);
    } else if (p < kOW3) {  // b2[i]
      const int i = p - kOB2;
      for (int q = 0; q < kTB; q++) g += sd2[q][i];
    } else if (p < kOB3) {  // W3[o][i]
      const int o = (p - kOW3) / kH, i = (p - kOW3) % kH;
      for (int q = 0; q < kTB; q++) g = fmaf(sdy[q][o], sa2[q][i], g);
    } else {  // b3[o]
      const int o = p - kOB3;
      for (int q = 0; q < kTB; q++) g += sdy[q][o];
    }
    atomicAdd(&grad[p], g);
  }
  if (tid == 0) {
    float b = 0.0f;
    for (int q = 0; q < kTB; q++) b += sl[q];
    atomicAdd(loss, b / (9.0f * (float)n));
  }
}

__global__ void surr_zero_kernel(float *g, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = 0.0f;
}

// Adam (Kingma & Ba 2015) with bias correction; step is 1-based.
__global__ void surr_adam_kernel(float *params, const float *grad, float *m, float *v, int n,
                                 float lr, float b1, float b2, float eps, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)step));
  const float vh = vi / (1.0f - powf(b2, (float)step));
  params[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace
}  // namespace hh

using namespace hh;

struct hh_surrogate {
  float r0 = 1.0f;
  DBuf params, grad, m, v, loss;
  int step = 0;
  hipStream_t st = nullptr;
  ~hh_surrogate() {
    for (DBuf *b : {&params, &grad, &m, &v, &loss}) b->release();
    if (st) (void)hipStreamDestroy(st);
  }
  void run_grad(const float *r, const float *t, int64_t n) {
    hipLaunchKernelGGL(surr_zero_kernel, dim3((kNP + 255) / 256), dim3(256), 0, st,
                       grad.as<float>(), kNP);
    hipLaunchKernelGGL(surr_zero_kernel, dim3(1), dim3(64), 0, st, loss.as<float>(), 1);
    hipLaunchKernelGGL(surr_grad_kernel, dim3((unsigned)((n + kTB - 1) / kTB)), dim3(kTB), 0, st,
                       params.as<float>(), r, t, n, r0, grad.as<float>(), loss.as<float>());
    HIPCHK(hipGetLastError());
  }
};

namespace {
hh_surrogate *need(hh_surrogate *s) {
  if (!s) throw HHError(HH_ERR_ARG, "NULL surrogate");
  return s;
}
void check_n(int64_t n) {
  if (n <= 0 || n > ((int64_t)1 << 31) - 1) throw HHError(HH_ERR_ARG, "sample count out of range");
}
}  // namespace

extern "C" {

int hh_surrogate_create(float r0, hh_surrogate **out) {
  return guarded([&]() {
    if (!out) throw HHError(HH_ERR_ARG, "NULL argument");
    if (!(r0 > 0)) throw HHError(HH_ERR_ARG, "r0 must be positive");
    ensure_device();
    hh_surrogate *s = new hh_surrogate();
    s->r0 = r0;
    try {
      HIPCHK(hipStreamCreate(&s->st));
      for (DBuf *b : {&s->params, &s->grad, &s->m, &s->v}) {
        b->reserve(4 * kNP);
        HIPCHK(hipMemset(b->p, 0, 4 * kNP));  // all-zero net and Adam state
      }
      s->loss.reserve(4);
      HIPCHK(hipMemset(s->loss.p, 0, 4));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

void hh_surrogate_destroy(hh_surrogate *s) { delete s; }

int hh_surrogate_set_params(hh_surrogate *s, const float *params) {
  return guarded([&]() {
    if (!params) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    HIPCHK(hipMemcpy(s->params.p, params, 4 * kNP, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->m.p, 0, 4 * kNP));
    HIPCHK(hipMemset(s->v.p, 0, 4 * kNP));
    s->step = 0;
  });
}

int hh_surrogate_get_params(hh_surrogate *s, float *params) {
  return guarded([&]() {
    if (!params) throw HHError(HH_ERR_ARG, "NULL argument");
    HIPCHK(hipStreamSynchronize(need(s)->st));
    HIPCHK(hipMemcpy(params, s->params.p, 4 * kNP, hipMemcpyDeviceToHost));
  });
}

int hh_surrogate_eval_device(hh_surrogate *s, const float *r_dev, int64_t n, float *out_dev) {
  return guarded([&]() {
    if (!r_dev || !out_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    check_n(n);
    hipLaunchKernelGGL(surr_eval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->st,
                       s->params.as<float>(), r_dev, n, s->r0, out_dev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
  });
}

int hh_surrogate_gradient_device(hh_surrogate *s, const float *r_dev, const float *t_dev,
                                 int64_t n, float *grad, float *loss) {
  return guarded([&]() {
    if (!r_dev || !t_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    check_n(n);
    s->run_grad(r_dev, t_dev, n);
    HIPCHK(hipStreamSynchronize(s->st));
    if (grad) HIPCHK(hipMemcpy(grad, s->grad.p, 4 * kNP, hipMemcpyDeviceToHost));
    if (loss) HIPCHK(hipMemcpy(loss, s->loss.p, 4, hipMemcpyDeviceToHost));
  });
}

int hh_surrogate_train_device(hh_surrogate *s, const float *r_dev, const float *t_dev, int64_t n,
                              int steps, float lr, float *final_loss) {
  return guarded([&]() {
    if (!r_dev || !t_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    check_n(n);
    if (steps < 0 || !(lr > 0)) throw HHError(HH_ERR_ARG, "steps >= 0 and lr > 0 required");
    if (steps == 0) s->run_grad(r_dev, t_dev, n);  // the loss at the current parameters
    for (int k = 0; k < steps; k++) {
      s->run_grad(r_dev, t_dev, n);
      s->step++;
      hipLaunchKernelGGL(surr_adam_kernel, dim3((kNP + 255) / 256), dim3(256), 0, s->st,
                         s->params.as<float>(), s->grad.as<float>(), s->m.as<float>(),
                         s->v.as<float>(), kNP, lr, 0.9f, 0.999f, 1e-8f, s->step);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s->st));
    if (final_loss) HIPCHK(hipMemcpy(final_loss, s->loss.p, 4, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
