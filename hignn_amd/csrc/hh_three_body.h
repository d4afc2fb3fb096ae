// Launch descriptors of the three-body / self MLP kernels (hh_three_body.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hh {

// One batch of edges [first, first + count) of an edge list (3 int32 per
// three-body edge, 2 per self edge), with per-edge records written by layer 1
// and read by layer 4: rec = (r1, r2, -, -), frc = (f, target index bits).
struct EdgeBatch {
  const float *x, *f;     // (n,3) positions and forces
  const int32_t *edges;
  int64_t first, count;
  float4 *rec, *frc;
};

// Device weights of one net (Linear(d,256) Tanh Linear(256,1024) Tanh
// Linear(1024,512) Tanh Linear(512,9)); layers 1-3 pre-scaled by 2 log2(e).
struct NetDev {
  const float *w1s, *b1s;          // 256 x d, 256
  const float *w2s, *b2s;          // 1024 x 256 (FP32 path), 1024
  const float *w3s, *b3s;          // 512 x 1024 (FP32 path), 512
  const void *w2h, *w3h;           // FP16 copies of W2, W3
  const float *w4, *b4;            // 9 x 512, 9 (not scaled)
};

hipError_t launch_triplet_count(const uint32_t *off, const int32_t *idx, int64_t n, uint32_t *cnt,
                                hipStream_t st);
hipError_t launch_triplet_fill(const uint32_t *off, const int32_t *idx, int64_t n,
                               const uint32_t *toff, int32_t *trip, hipStream_t st);
// u (FP64, n x 3) += the batch's velocity contributions; h1/h2/h3 hold
// ceil(count / 256) * 256 rows of 256 / 1024 / 512 activations (FP16 if
// half, else FP32).
hipError_t launch_edge_mlp(const EdgeBatch &B, const NetDev &net, int din, int half, void *h1,
                           void *h2, void *h3, double *u, hipStream_t st);
hipError_t launch_to_float(const double *u, float *out, int64_t n3, int accumulate,
                           hipStream_t st);

}  // namespace hh
