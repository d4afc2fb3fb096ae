// The surrogate handle, shared by its C ABI (hh_surrogate.hip) and the
// model's far field (hh_model.cpp), which reads the parameters and the
// training range of an attached surrogate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hh_host.h"

struct hh_surrogate {
  float r_lo = 1.0f, r_hi = 2.0f;  // training range of |r|; inputs are r / r_hi
  hh::DBuf params, grad, m, v, loss, part;
  int step = 0;
  ~hh_surrogate();
  // out (n,6) = s(r) for n device displacements, on stream st
  void eval(const float *r, int64_t n, float *out, hipStream_t st);
};

namespace hh {
// Device address of the no-allocation gradient's workspace on the current
// device (looked up once per device).
float *scratch_ptr();
int surrogate_grad_blocks(int64_t n);
void surrogate_gradient(const float *params, const float *r, const float *t, int64_t s0,
                        int64_t count, float inv_scale, float scale, float *part, float *grad_dev,
                        float *loss_dev, hipStream_t st);
}  // namespace hh
