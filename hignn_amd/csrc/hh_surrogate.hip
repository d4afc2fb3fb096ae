// Distilled far-field surrogate of the two-body pair function (the net is
// in hh_surrogate_net.h): training kernels and the C ABI of
// include/hh_surrogate.h.
//
// Training is minibatch Adam on samples labelled by the full MLP: each step
// runs one gradient pass over its batch (forward, backward and per-block
// weight-gradient sums, no atomics) and one reduction that sums the block
// partials in block order, so the gradient does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hh.h"
#include "../../include/hh_surrogate.h"
#include "hh_host.h"
#include "hh_surrogate_internal.h"
#include "hh_surrogate_net.h"

namespace hh {
namespace {

using namespace sur;

static_assert(kNP == HH_SURROGATE_PARAMS, "parameter count");
constexpr int kTB = 128;         // samples per chunk of a gradient block
constexpr int kMaxGradBlocks = 512;
constexpr int kPartStride = kNP + 1;  // per block: gradient sums, then the loss sum

// Workspace of hh_surrogate_gradient_noalloc: the gradient pass's per-block
// partial sums (512 x 1,383 floats = 2.8 MB), held by every GPU the library
// runs on until the process ends.  Host code never passes this name to a
// kernel (on the host it is only a stand-in for the symbol): it uses the
// device address from hipGetSymbolAddress, one per device (scratch_ptr).
static __device__ __align__(256) float scratch[kMaxGradBlocks * kPartStride];

__global__ __launch_bounds__(256) void surr_eval_kernel(const float *params, const float *r,
                                                        int64_t n, float inv_scale, float *out) {
  __shared__ float P[kNP];
  stage(params, P);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float y[kOut];
  eval(P, r[3 * s], r[3 * s + 1], r[3 * s + 2], inv_scale, y);
#pragma unroll
  for (int o = 0; o < kOut; o++) out[kOut * s + o] = y[o];
}

// Gradient partials of L = (1/n) sum_s sum_k w_k (y_sk - t_sk)^2 over samples
// [s0, s0 + count): block b takes chunks b, b + gridDim.x, ... of kTB
// samples (one per thread), keeps each chunk's activations and output
// errors in LDS, and each thread adds its share of the parameters' sums over
// the chunk to registers.  Block b writes part[b * kPartStride + p] (p <
// kNP) and its loss sum at kNP.  scale = 1/n of the whole loss.
__global__ __launch_bounds__(kTB) void surr_grad_kernel(const float *params, const float *r,
                                                        const float *t, int64_t s0, int64_t count,
                                                        float inv_scale, float scale, float *part) {
  __shared__ float P[kNP];
  __shared__ float sz[kTB][kIn], sa1[kTB][kH + 1], sd1[kTB][kH + 1], sa2[kTB][kH + 1],
      sd2[kTB][kH + 1], sdy[kTB][kOut];
  __shared__ float sl[kTB];
  stage(params, P);
  const int tid = threadIdx.x;
  constexpr int kPer = (kNP + kTB - 1) / kTB;  // parameters per thread
  float acc[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) acc[q] = 0.0f;
  float lsum = 0.0f;
  const int64_t nchunks = (count + kTB - 1) / kTB;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t s = c * kTB + tid;
    float z[kIn] = {0, 0, 0}, x1[kH], x2[kH], y[kOut], dy[kOut];
    float l = 0.0f;
    const bool valid = s < count;
    if (valid) {
      const int64_t g = s0 + s;
#pragma unroll
      for (int k = 0; k < kIn; k++) z[k] = r[3 * g + k] * inv_scale;
      forward(P, z, x1, x2, y);
#pragma unroll
      for (int o = 0; o < kOut; o++) {
        const float e = y[o] - t[kOut * g + o];
        const float we = entry_weight(o) * e;
        l = __builtin_fmaf(we, e, l);
        dy[o] = 2.0f * scale * we;
      }
    } else {
#pragma unroll
      for (int j = 0; j < kH; j++) x1[j] = x2[j] = 0.0f;
#pragma unroll
      for (int o = 0; o < kOut; o++) dy[o] = 0.0f;
    }
    // back through layers 3 and 2 (zero errors for padded samples)
    float d2[kH], d1[kH];
#pragma unroll
    for (int i = 0; i < kH; i++) {
      float gsum = 0.0f;
#pragma unroll
      for (int o = 0; o < kOut; o++) gsum = __builtin_fmaf(P[kOW3 + o * kH + i], dy[o], gsum);
      d2[i] = gsum * dsilu(x2[i]);
    }
#pragma unroll
    for (int j = 0; j < kH; j++) {
      float gsum = 0.0f;
#pragma unroll
      for (int i = 0; i < kH; i++) gsum = __builtin_fmaf(P[kOW2 + i * kH + j], d2[i], gsum);
      d1[j] = gsum * dsilu(x1[j]);
    }
    __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
    for (int k = 0; k < kIn; k++) sz[tid][k] = z[k];
#pragma unroll
    for (int j = 0; j < kH; j++) sa1[tid][j] = valid ? silu(x1[j]) : 0.0f, sd1[tid][j] = d1[j];
#pragma unroll
    for (int i = 0; i < kH; i++) sa2[tid][i] = valid ? silu(x2[i]) : 0.0f, sd2[tid][i] = d2[i];
#pragma unroll
    for (int o = 0; o < kOut; o++) sdy[tid][o] = dy[o];
    sl[tid] = l;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int p = tid + q * kTB;
      if (p >= kNP) break;
      float g = 0.0f;
      if (p < kOB1) {  // W1[j][k]
        const int j = p / kIn, k = p % kIn;
        for (int u = 0; u < kTB; u++) g = __builtin_fmaf(sd1[u][j], sz[u][k], g);
      } else if (p < kOW2) {  // b1[j]
        const int j = p - kOB1;
        for (int u = 0; u < kTB; u++) g += sd1[u][j];
      } else if (p < kOB2) {  // W2[i][j]
        const int i = (p - kOW2) / kH, j = (p - kOW2) % kH;
        for (int u = 0; u < kTB; u++) g = __builtin_fmaf(sd2[u][i], sa1[u][j], g);
      } else if (p < kOW3) {  // b2[i]
        const int i = p - kOB2;
        for (int u = 0; u < kTB; u++) g += sd2[u][i];
      } else if (p < kOB3) {  // W3[o][i]
        const int o = (p - kOW3) / kH, i = (p - kOW3) % kH;
        for (int u = 0; u < kTB; u++) g = __builtin_fmaf(sdy[u][o], sa2[u][i], g);
      } else {  // b3[o]
        const int o = p - kOB3;
        for (int u = 0; u < kTB; u++) g += sdy[u][o];
      }
      acc[q] += g;
    }
    if (tid == 0) {
      float b = 0.0f;
      for (int u = 0; u < kTB; u++) b += sl[u];
      lsum += b;
    }
  }
  float *dst = part + (size_t)blockIdx.x * kPartStride;
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int p = tid + q * kTB;
    if (p < kNP) dst[p] = acc[q];
  }
  if (tid == 0) dst[kNP] = lsum * scale;
}

// grad[p] = sum over blocks of part[b][p], in block order; loss likewise.
__global__ void surr_reduce_kernel(const float *part, int nblocks, float *grad, float *loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kNP) return;
  float s = 0.0f;
  for (int b = 0; b < nblocks; b++) s += part[(size_t)b * kPartStride + p];
  if (p < kNP)
    grad[p] = s;
  else
    loss[0] = s;
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

int surrogate_grad_blocks(int64_t n) {
  return (int)std::min<int64_t>((n + kTB - 1) / kTB, kMaxGradBlocks);
}

// Gradient of the loss on samples [s0, s0 + count) (scale = 1/n of the
// loss being differentiated) into grad_dev / loss_dev, with part holding
// surrogate_grad_blocks(count) * (HH_SURROGATE_PARAMS + 1) floats.
void surrogate_gradient(const float *params, const float *r, const float *t, int64_t s0,
                        int64_t count, float inv_scale, float scale, float *part, float *grad_dev,
                        float *loss_dev, hipStream_t st) {
  const int nb = surrogate_grad_blocks(count);
  hipLaunchKernelGGL(surr_grad_kernel, dim3(nb), dim3(kTB), 0, st, params, r, t, s0, count,
                     inv_scale, scale, part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(surr_reduce_kernel, dim3((kNP + 1 + 255) / 256), dim3(256), 0, st,
                     (const float *)part, nb, grad_dev, loss_dev);
  HIPCHK(hipGetLastError());
}

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
constexpr int kMaxDevices = 64;
static float *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

float *scratch_ptr() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw HHError(HH_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HHError(HH_ERR_DEVICE, "no device address for the gradient workspace");
    g_scratch[dev] = (float *)p;
  }
  return g_scratch[dev];
}

}  // namespace hh

using namespace hh;

hh_surrogate::~hh_surrogate() {
  for (DBuf *b : {&params, &grad, &m, &v, &loss, &part}) b->release();
}

void hh_surrogate::eval(const float *r, int64_t n, float *out, hipStream_t st) {
  hipLaunchKernelGGL(surr_eval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     params.as<float>(), r, n, 1.0f / r_hi, out);
  HIPCHK(hipGetLastError());
}

namespace {
hh_surrogate *need(hh_surrogate *s) {
  if (!s) throw HHError(HH_ERR_ARG, "NULL surrogate");
  return s;
}
void check_n(int64_t n) {
  if (n <= 0 || n > ((int64_t)1 << 31) - 1) throw HHError(HH_ERR_ARG, "sample count out of range");
}
hipStream_t stream_of(void *stream) { return (hipStream_t)stream; }
}  // namespace

extern "C" {

int hh_surrogate_create(float r_lo, float r_hi, hh_surrogate **out) {
  return guarded([&]() {
    if (!out) throw HHError(HH_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (!(r_lo > 0) || !(r_hi > r_lo)) throw HHError(HH_ERR_ARG, "need 0 < r_lo < r_hi");
    ensure_device();
    hh_surrogate *s = new hh_surrogate();
    s->r_lo = r_lo;
    s->r_hi = r_hi;
    try {
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

int hh_surrogate_range(hh_surrogate *s, float *r_lo, float *r_hi) {
  return guarded([&]() {
    need(s);
    if (r_lo) *r_lo = s->r_lo;
    if (r_hi) *r_hi = s->r_hi;
  });
}

int hh_surrogate_set_params(hh_surrogate *s, const float *params) {
  return guarded([&]() {
    if (!params) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(s->params.p, params, 4 * kNP, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->m.p, 0, 4 * kNP));
    HIPCHK(hipMemset(s->v.p, 0, 4 * kNP));
    s->step = 0;
  });
}

int hh_surrogate_get_params(hh_surrogate *s, float *params) {
  return guarded([&]() {
    if (!params) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(params, s->params.p, 4 * kNP, hipMemcpyDeviceToHost));
  });
}

int hh_surrogate_eval_device(hh_surrogate *s, const float *r_dev, int64_t n, float *out_dev,
                             void *stream) {
  return guarded([&]() {
    if (!r_dev || !out_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    check_n(n);
    s->eval(r_dev, n, out_dev, stream_of(stream));
    HIPCHK(hipStreamSynchronize(stream_of(stream)));
  });
}

int hh_surrogate_gradient_device(hh_surrogate *s, const float *r_dev, const float *t_dev,
                                 int64_t n, float *grad, float *loss, void *stream) {
  return guarded([&]() {
    if (!r_dev || !t_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    check_n(n);
    hipStream_t st = stream_of(stream);
    s->part.reserve((size_t)surrogate_grad_blocks(n) * kPartStride * 4);
    surrogate_gradient(s->params.as<float>(), r_dev, t_dev, 0, n, 1.0f / s->r_hi,
                       1.0f / (float)n, s->part.as<float>(), s->grad.as<float>(),
                       s->loss.as<float>(), st);
    HIPCHK(hipStreamSynchronize(st));
    if (grad) HIPCHK(hipMemcpy(grad, s->grad.p, 4 * kNP, hipMemcpyDeviceToHost));
    if (loss) HIPCHK(hipMemcpy(loss, s->loss.p, 4, hipMemcpyDeviceToHost));
  });
}

int hh_surrogate_gradient_noalloc(hh_surrogate *s, const float *r_dev, const float *t_dev,
                                  int64_t n, float *grad_dev, float *loss_dev, void *stream) {
  return guarded([&]() {
    if (!r_dev || !t_dev || !grad_dev || !loss_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    if (n <= 0 || n > HH_SURROGATE_NOALLOC_MAX_SAMPLES)
      throw HHError(HH_ERR_ARG, "sample count must be in 1.." +
                                    std::to_string((long long)HH_SURROGATE_NOALLOC_MAX_SAMPLES));
    // the partials of at most kMaxGradBlocks blocks fit the workspace
    static_assert(sizeof(scratch) == (size_t)kMaxGradBlocks * kPartStride * sizeof(float),
                  "workspace size");
    if (surrogate_grad_blocks(n) > kMaxGradBlocks) throw HHError(HH_ERR_ARG, "workspace too small");
    surrogate_gradient(s->params.as<float>(), r_dev, t_dev, 0, n, 1.0f / s->r_hi, 1.0f / (float)n,
                       scratch_ptr(), grad_dev, loss_dev, stream_of(stream));
  });
}

int hh_surrogate_train_device(hh_surrogate *s, const float *r_dev, const float *t_dev, int64_t n,
                              int steps, int64_t batch, float lr0, float lr1, float *final_loss,
                              void *stream) {
  return guarded([&]() {
    if (!r_dev || !t_dev) throw HHError(HH_ERR_ARG, "NULL argument");
    need(s);
    check_n(n);
    if (steps < 0 || !(lr0 > 0) || !(lr1 > 0)) throw HHError(HH_ERR_ARG, "steps >= 0 and lr > 0 required");
    if (batch <= 0 || batch > n) batch = n;
    hipStream_t st = stream_of(stream);
    s->part.reserve((size_t)surrogate_grad_blocks(batch) * kPartStride * 4);
    const float inv_scale = 1.0f / s->r_hi;
    if (steps == 0)  // the loss at the current parameters over all samples
      surrogate_gradient(s->params.as<float>(), r_dev, t_dev, 0, n, inv_scale, 1.0f / (float)n,
                         s->part.as<float>(), s->grad.as<float>(), s->loss.as<float>(), st);
    int64_t off = 0;
    for (int k = 0; k < steps; k++) {
      // consecutive batches, wrapping round the samples
      if (off + batch > n) off = 0;
      surrogate_gradient(s->params.as<float>(), r_dev, t_dev, off, batch, inv_scale,
                         1.0f / (float)batch, s->part.as<float>(), s->grad.as<float>(),
                         s->loss.as<float>(), st);
      off += batch;
      s->step++;
      // cosine schedule from lr0 to lr1 over the call's steps
      const double c = steps > 1 ? 0.5 * (1.0 + std::cos(M_PI * (double)k / (double)(steps - 1))) : 1.0;
      const float lr = (float)(lr1 + (lr0 - lr1) * c);
      hipLaunchKernelGGL(surr_adam_kernel, dim3((kNP + 255) / 256), dim3(256), 0, st,
                         s->params.as<float>(), (const float *)s->grad.as<float>(),
                         s->m.as<float>(), s->v.as<float>(), kNP, lr, 0.9f, 0.999f, 1e-8f,
                         s->step);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    if (final_loss) HIPCHK(hipMemcpy(final_loss, s->loss.p, 4, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
