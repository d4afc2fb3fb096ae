// HIP kernels of libhh for gfx950 (CDNA4, 64-wide waves, MFMA).
//
// The pair function (python/convert.py:11-20)
//     y = W3 tanh(W2 tanh(W1 r + b1) + b2) + b3,  M = (y + I)/|r|,
// is evaluated for 16 displacements per MFMA tile with the displacement index
// on the MFMA column axis and the hidden features on the row axis, so that
// each layer's output tile is directly the next layer's B operand (register
// chaining, no round trip through memory):
//   * FP32 / FP64 path: v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64.  Lane l
//     (g = l/16, p = l%16) holds A[p][g], B[g][p], and D[4g+r][p] (r = 0..3).
//     Layer 1 (K = 3) runs on the VALU in FP32/FP64 and produces feature
//     4s+g of K-step s; layer 2's D register r of row tile mt is feature
//     16mt+4g+r, which is the K-slot g of a layer-3 step whose A operand holds
//     W3 columns 16mt+4g+r (host-side fragment layout).
//   * 16-bit path: v_mfma_f32_16x16x32_{bf16,f16}.  Lane holds
//     A[p][8g+j], B[8g+j][p] (j = 0..7), D[4g+r][p].  Layer 1 in FP32 on the
//     VALU, rounded to 16 bit; two layer-2 row tiles (32 features) form one
//     layer-3 K-step (K-slot 8g+j <-> feature 32q+4g+j for j<4 and
//     32q+16+4g+j-4 for j>=4).  W2/W3 fragments are staged in LDS once per
//     persistent workgroup.
// The epilogue (+I, /|r|, |r| < 1e-5 -> I, M f) is lane-local: the lane that
// holds output e = 3 row + col of a pair adds its share of M f directly, and
// the per-pair sums are completed by cross-lane reductions, so the 3x3 block
// never has to be gathered into one lane.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include <cstdint>

#include "hh_internal.h"
#include "hh_kernels.h"
#include "hh_surrogate_net.h"

namespace hh {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mma(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mma(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma(f16x8 a, f16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Hidden activation.  The FP32 and 16-bit paths receive W1, b1, W2, b2
// scaled by 2 log2 e on the host, so their pre-activation is x' = 2x log2 e
// and tanh x = 1 - 2/(2^x' + 1): one v_exp_f32 and one v_rcp_f32, no scaling
// multiply.  Saturates correctly at +-1 (2^x' = inf -> 1, 0 -> -1); absolute
// error a few ulp of 1.  The FP64 path keeps the weights unscaled.
__device__ __forceinline__ float act(float xs) {
  return __builtin_fmaf(-2.0f, __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(xs) + 1.0f), 1.0f);
}
// FP64 (the check path): tanh x = 1 - 2 / (e^{2x} + 1), one exponential and
// one division instead of the library tanh (about a quarter of its
// instructions).  Both are correct to a few ulp, so the result is within a
// few 1e-16 of tanh x in absolute terms (relative error grows where
// tanh x ~ x is small, but the next layer sums absolute contributions);
// e^{2x} = inf gives 1 and e^{2x} = 0 gives -1.
__device__ __forceinline__ double act(double x) { return 1.0 - 2.0 / (exp(2.0 * x) + 1.0); }

__device__ __forceinline__ float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_r(double x) { return 1.0 / sqrt(x); }

template <class R>
struct Vec4;
template <>
struct Vec4<float> {
  typedef f32x4 T;
};
template <>
struct Vec4<double> {
  typedef f64x4 T;
};

// Row of the 16x16 MFMA output held by lane group g in register r:
// 4g + r for the f32 (and 16-bit) instructions, 4r + g for
// v_mfma_f64_16x16x4f64 (its lane 16(i%4)+j, register i/4 holds D[i][j]).
// The host orders the rows of W3 (and b3) so that the layer-3 output row
// out_row(g, r) is block entry M[g][r] for g, r < 3 and zero otherwise: lane
// group g < 3 holds row g of the 3x3 block of its pair in registers 0..2.
template <class R>
__device__ __forceinline__ int out_row(int g, int r) {
  return 4 * g + r;
}
template <>
__device__ __forceinline__ int out_row<double>(int g, int r) {
  return 4 * r + g;
}

// ---------------------------------------------------------------------------
// FP32 / FP64 MLP on NT tiles of 16 displacements.
// ---------------------------------------------------------------------------
template <class R>
struct WF {
  const R *w1b;    // [128][4] = W1 row k (3) + b1[k]
  const R *w2f;    // [32 mt][32 s][64 lanes]: W2[16mt + p][4s + g]
  const R *b2;     // [512]
  const R *w3f;    // [32 mt][4 r][64 lanes]: W3[p][16mt + 4g + r] (0 for p >= 9)
  const R *b3;     // [16] (0 beyond 9)
};

template <class R, int NT>
__device__ __forceinline__ void mlp_tiles_f(const WF<R> &W, const R (&rx)[NT], const R (&ry)[NT],
                                            const R (&rz)[NT],
                                            typename Vec4<R>::T (&y)[NT]) {
  typedef typename Vec4<R>::T V;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  R h1[NT][32];
#pragma unroll
  for (int s = 0; s < 32; s++) {
    const R *w = W.w1b + 4 * (4 * s + g);
    const R w0 = w[0], w1 = w[1], w2 = w[2], b = w[3];
#pragma unroll
    for (int t = 0; t < NT; t++) h1[t][s] = act(fma(w2, rz[t], fma(w1, ry[t], fma(w0, rx[t], b))));
  }
#pragma unroll
  for (int t = 0; t < NT; t++) y[t] = V{0, 0, 0, 0};
  for (int mt = 0; mt < 32; mt++) {
    const R *bb = W.b2 + 16 * mt;
    const V b2v = V{bb[out_row<R>(g, 0)], bb[out_row<R>(g, 1)], bb[out_row<R>(g, 2)],
                    bb[out_row<R>(g, 3)]};
    V acc[NT];
    const R *a2 = W.w2f + (size_t)mt * 32 * 64 + lane;
    {
      const R a = a2[0];
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t] = mma(a, h1[t][0], b2v);  // bias enters as C
    }
#pragma unroll
    for (int s = 1; s < 32; s++) {
      const R a = a2[s * 64];
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t] = mma(a, h1[t][s], acc[t]);
    }
    const R *a3 = W.w3f + (size_t)mt * 4 * 64 + lane;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const R a = a3[r * 64];
#pragma unroll
      for (int t = 0; t < NT; t++) y[t] = mma(a, act(acc[t][r]), y[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) y[t][r] += W.b3[out_row<R>(g, r)];
}

// ---------------------------------------------------------------------------
// 16-bit MLP (BF16 or FP16 operands, FP32 accumulation).
// ---------------------------------------------------------------------------
template <class H>
struct H16;
template <>
struct H16<bf16x8> {
  typedef __bf16 E;
};
template <>
struct H16<f16x8> {
  typedef _Float16 E;
};

struct WH {
  const float *w1b;  // [4 s][2 half][64 lanes]: layer-1 A fragments (below)
  const float *b2;   // [512]
  const float *b3;   // [16]
  const void *w2f;   // global: [32 mt][4 s][64 lanes] x 8 halves
  const void *w3f;   // global: [16 q][64 lanes] x 8 halves
};

constexpr int kW2FragBytes = 32 * 4 * 64 * 16;  // 128 KiB
constexpr int kW3FragBytes = 16 * 64 * 16;      // 16 KiB

template <class H, int NT>
__device__ __forceinline__ void mlp_tiles_h(const WH &W, const H *lw2, const H *lw3,
                                            const float (&rx)[NT], const float (&ry)[NT],
                                            const float (&rz)[NT], f32x4 (&y)[NT]) {
  typedef typename H16<H>::E E;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  // Layer 1 on MFMA (v_mfma_f32_16x16x4f32, the VALU being the busy pipe):
  // A = 16 rows of [W1 | b1] per (s, half), B = (rx, ry, rz, 1) of the
  // tile's 16 pairs.  The host orders the rows so that output row 4g + r of
  // (s, half) is hidden unit 32s + 8g + 4 half + r: lane group g then holds
  // units 32s + 8g + j (j = 0..7) of its pair, layer 2's B operand.
  float wa[4][2];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int half = 0; half < 2; half++) wa[s][half] = W.w1b[(2 * s + half) * 64 + lane];
  H h1[NT][4];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const float b = g == 0 ? rx[t] : g == 1 ? ry[t] : g == 2 ? rz[t] : 1.0f;
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const f32x4 d = mma(wa[s][half], b, f32x4{0, 0, 0, 0});
#pragma unroll
        for (int r = 0; r < 4; r++) h1[t][s][4 * half + r] = (E)act(d[r]);
      }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) y[t] = f32x4{0, 0, 0, 0};
  // Layer 2 on MFMA (bias as C), tanh on the VALU and layer 3 on MFMA, per
  // block q of 32 hidden units.  Software pipelined: the layer-2 MFMAs of
  // tile t+1 (after the last tile: of tile 0 of block q+1) are issued beside
  // the activation of tile t, so that the matrix cores and the VALU work at
  // the same time.  A tile's accumulators are refilled only after they have
  // been consumed, so the pipeline costs no registers; every output keeps
  // the same operation order as an unpipelined loop.
  H fa0[4], fa1[4];
  f32x4 bq0, bq1;
  f32x4 acc0[NT], acc1[NT];
  auto load_block = [&](int q) {
    const H *a0p = lw2 + (size_t)(2 * q) * 4 * 64 + lane;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      fa0[s] = a0p[s * 64];
      fa1[s] = a0p[(4 + s) * 64];
    }
    const float *bb0 = W.b2 + 32 * q + 4 * g;
    bq0 = f32x4{bb0[0], bb0[1], bb0[2], bb0[3]};
    bq1 = f32x4{bb0[16], bb0[17], bb0[18], bb0[19]};
  };
  auto layer2 = [&](int t) {
    f32x4 c0 = mma(fa0[0], h1[t][0], bq0);
    f32x4 c1 = mma(fa1[0], h1[t][0], bq1);
#pragma unroll
    for (int s = 1; s < 4; s++) {
      c0 = mma(fa0[s], h1[t][s], c0);
      c1 = mma(fa1[s], h1[t][s], c1);
    }
    acc0[t] = c0;
    acc1[t] = c1;
  };
  auto layer3 = [&](int t, const H &a3) {
    H b;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      b[r] = (E)act(acc0[t][r]);
      b[4 + r] = (E)act(acc1[t][r]);
    }
    y[t] = mma(a3, b, y[t]);
  };
  load_block(0);
  layer2(0);
  for (int q = 0; q < 15; q++) {
    const H a3 = lw3[q * 64 + lane];
#pragma unroll
    for (int t = 0; t + 1 < NT; t++) {
      layer2(t + 1);
      layer3(t, a3);
    }
    load_block(q + 1);
    layer2(0);
    layer3(NT - 1, a3);
    // interleave: one MFMA, two transcendental, two other VALU instructions
#pragma unroll
    for (int i = 0; i < 9 * NT; i++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x400, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
    }
  }
  {
    const H a3 = lw3[15 * 64 + lane];
#pragma unroll
    for (int t = 0; t + 1 < NT; t++) {
      layer2(t + 1);
      layer3(t, a3);
    }
    layer3(NT - 1, a3);
#pragma unroll
    for (int i = 0; i < 9 * NT; i++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x400, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
    }
  }
  const float *b3 = W.b3 + 4 * g;
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) y[t][r] += b3[r];
}

// ---------------------------------------------------------------------------
// Shared epilogue helpers.
// ---------------------------------------------------------------------------
// Row g of the 3x3 block of a pair with displacement r from the raw network
// output y (registers 0..2 of lane group g): (y + I)/|r|, or I when
// |r| < 1e-5 (python/convert.py:14-19).  One reciprocal square root per pair
// instead of a division per entry; lanes of group 3 get zeros.
template <class R>
__device__ __forceinline__ void block_row(const typename Vec4<R>::T &y, R rx, R ry, R rz, int g,
                                          R (&m)[3]) {
  const R nr2 = rx * rx + ry * ry + rz * rz;
  const bool small = nr2 < (R)1e-10;
  const R inv = small ? (R)0 : rsqrt_r(nr2);
  const R d = small ? (R)1 : inv;
#pragma unroll
  for (int c = 0; c < 3; c++) m[c] = fma((R)y[c], inv, c == g ? d : (R)0);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Near field / dense row blocks: u_i += sum_j M(x_j - x_i) f_j  (CloseDot,
// src/CloseDot.cpp:245-299; DenseDot, src/DenseDot.cpp:230-246).
// One wave handles one task (4 consecutive targets of a row block) and walks
// the block's sources 16 at a time; tile t of the MLP is target t.
// ---------------------------------------------------------------------------
template <class R>
struct NearSrc {
  const float4 *xp;    // tree-ordered positions (w unused)
  const float4 *fp;    // tree-ordered forces
  const int32_t *tlist;  // explicit targets (BF_EXPLICIT_I)
  const RowBlock *blocks;
  const NearTask *tasks;
  int64_t ntasks;
  double *u;           // (.,3) accumulators
  int32_t *next;       // work queue of the persistent kernel (zero before the
                       // launch), or null: tasks dealt round-robin
};

// Epilogue of one MLP call: lane (g < 3, p) holds row g of the block of pair
// (target t, source p) and accumulates
//   v[t]     += sum_c M[g][c] f_p[c]           (target t, component g)
//   v2[t][c] += M[g][c] f_p[g]                 (target t, component c; with
//               BF_SYMM both halves are weighted 1/2: u_i += (M + M^T)/2 f_p)
//   w        += sum_t sum_c M[g][c] f_t[c]     (BF_REUSE: source p,
//               component g, the same block un-transposed)
// fj is zero for an invalid source and fi[t] for an invalid target, so padded
// pairs add nothing.  Products and three-term sums in R, accumulation in FP64.
template <class R, int NT>
__device__ __forceinline__ void near_epilogue(const typename Vec4<R>::T (&y)[NT], const R (&rx)[NT],
                                              const R (&ry)[NT], const R (&rz)[NT], int g,
                                              const float4 &fj, float fjg,
                                              const float4 (&fi)[NT], int flags, double (&v)[NT],
                                              double (&v2)[NT][3], double &w) {
#pragma unroll
  for (int t = 0; t < NT; t++) {
    R m[3];
    block_row<R>(y[t], rx[t], ry[t], rz[t], g, m);
    const R s = m[0] * (R)fj.x + m[1] * (R)fj.y + m[2] * (R)fj.z;
    if (flags & BF_SYMM) {
      v[t] += 0.5 * (double)s;
#pragma unroll
      for (int c = 0; c < 3; c++) v2[t][c] += 0.5 * (double)(m[c] * (R)fjg);
    } else {
      v[t] += (double)s;
    }
    if (flags & BF_REUSE) w += (double)(m[0] * (R)fi[t].x + m[1] * (R)fi[t].y + m[2] * (R)fi[t].z);
  }
}

// One wave handles one task (4 consecutive targets of a row block) and walks
// the block's sources 16 at a time; tile t of an MLP call is target tg + t.
// NT tiles per MLP call, 4/NT passes over the sources.
template <class R, int NT, class Core>
__device__ __forceinline__ void near_body(const NearSrc<R> &S, Core core) {
  static_assert(4 % NT == 0, "NT must divide 4");
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, p = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  // First task: the wave's index; then, with a work queue, the next task not
  // yet taken (tasks differ in cost with their blocks' source counts), else
  // round-robin.
  auto next_task = [&](int64_t task) -> int64_t {
    if (!S.next) return task + nwaves;
    int32_t t = 0;
    if (lane == 0) t = atomicAdd(S.next, 1);
    return nwaves + (int64_t)__builtin_amdgcn_readfirstlane(t);
  };
  for (int64_t task = wave; task < S.ntasks; task = next_task(task)) {
    const NearTask tk = S.tasks[task];
    const RowBlock B = S.blocks[tk.block];
    for (int tg = 0; tg < 4; tg += NT) {
      if (tk.ioff + tg >= B.m) break;
      int ti[NT];
      bool iv[NT];
      float4 xi[NT], fi[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) {
        iv[t] = tk.ioff + tg + t < B.m;
        const int li = B.i0 + tk.ioff + (iv[t] ? tg + t : 0);
        ti[t] = li;
        const int xi_idx = (B.flags & BF_EXPLICIT_I) ? S.tlist[li] : li;
        xi[t] = S.xp[xi_idx];
        fi[t] = iv[t] ? S.fp[xi_idx] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      double v[NT], v2[NT][3];
#pragma unroll
      for (int t = 0; t < NT; t++) v[t] = v2[t][0] = v2[t][1] = v2[t][2] = 0.0;
      for (int j0 = 0; j0 < B.n; j0 += 16) {
        const bool jv = j0 + p < B.n;
        const int j = B.j0 + j0 + (jv ? p : 0);
        const float4 xj = S.xp[j];
        const float4 fj = jv ? S.fp[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float fjg = g == 0 ? fj.x : (g == 1 ? fj.y : (g == 2 ? fj.z : 0.f));
        R rx[NT], ry[NT], rz[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
          // (R)x - (R)y: the float difference for R = float (as the
          // reference, src/CloseDot.cpp:206), the exact one for R = double
          if (B.flags & BF_NEG) {
            rx[t] = (R)xi[t].x - (R)xj.x, ry[t] = (R)xi[t].y - (R)xj.y, rz[t] = (R)xi[t].z - (R)xj.z;
          } else {
            rx[t] = (R)xj.x - (R)xi[t].x, ry[t] = (R)xj.y - (R)xi[t].y, rz[t] = (R)xj.z - (R)xi[t].z;
          }
        }
        typename Vec4<R>::T y[NT];
        core(rx, ry, rz, y);
        double w = 0.0;
        near_epilogue<R, NT>(y, rx, ry, rz, g, fj, fjg, fi, B.flags, v, v2, w);
        if ((B.flags & BF_REUSE) && g < 3 && jv) unsafeAtomicAdd(&S.u[3 * (size_t)j + g], w);
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double s = wave_sum((g == c ? v[t] : 0.0) + v2[t][c]);
          if (lane == 0 && iv[t]) unsafeAtomicAdd(&S.u[3 * (size_t)ti[t] + c], s);
        }
    }
  }
}

__global__ __launch_bounds__(256) void near_fp32_kernel(NearSrc<float> S, WF<float> W) {
  near_body<float, 4>(S, [&](const float(&rx)[4], const float(&ry)[4], const float(&rz)[4],
                             f32x4(&y)[4]) { mlp_tiles_f<float, 4>(W, rx, ry, rz, y); });
}

// FP64 variant: two tiles per MLP call keep the double registers in budget.
__global__ __launch_bounds__(256) void near_fp64_kernel(NearSrc<double> S, WF<double> W) {
  near_body<double, 2>(S, [&](const double(&rx)[2], const double(&ry)[2], const double(&rz)[2],
                              f64x4(&y)[2]) { mlp_tiles_f<double, 2>(W, rx, ry, rz, y); });
}

// Stage the 16-bit W2/W3 fragments of a persistent workgroup in LDS.
template <class H>
__device__ __forceinline__ void stage_fragments(const WH &W, H *lw2, H *lw3) {
  const uint4 *src2 = (const uint4 *)W.w2f;
  const uint4 *src3 = (const uint4 *)W.w3f;
  uint4 *d2 = (uint4 *)lw2, *d3 = (uint4 *)lw3;
  for (int i = threadIdx.x; i < kW2FragBytes / 16; i += blockDim.x) d2[i] = src2[i];
  for (int i = threadIdx.x; i < kW3FragBytes / 16; i += blockDim.x) d3[i] = src3[i];
  __syncthreads();
}

// 16-bit near and ACA MLP kernels: persistent workgroups of kNearHThreads
// threads (two waves per SIMD) share the LDS-staged W2/W3 fragments.
constexpr int kNearHThreads = 512;
constexpr int kNearHTiles = 4;

// Tile shape of the near kernel: NT 16-pair tiles per MLP call (NT
// targets of a task at a time), THREADS per persistent workgroup (one per
// compute unit: THREADS/256 waves per SIMD).  The launch picks one of the
// shapes below (NearLaunch::hcfg), tuned by hh_model_tune.
template <class H, int NT, int THREADS>
__global__ __launch_bounds__(THREADS) void near_h_kernel(NearSrc<float> S, WH W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  H *lw2 = (H *)lds;
  H *lw3 = (H *)(lds + kW2FragBytes);
  stage_fragments<H>(W, lw2, lw3);
  near_body<float, NT>(S, [&](const float(&rx)[NT], const float(&ry)[NT], const float(&rz)[NT],
                              f32x4(&y)[NT]) { mlp_tiles_h<H, NT>(W, lw2, lw3, rx, ry, rz, y); });
}

// ---------------------------------------------------------------------------
// ACA: batched block cross approximation (src/FarDot.cpp:186-1388).
// ---------------------------------------------------------------------------
struct AcaTile {
  int32_t slot;  // index into the active list
  int32_t off;   // first item (row or column index inside the block)
};

// Tile or chunk q of a list in which slot s owns the units first[s] ..
// first[s + 1] - 1 (first ascending, every slot at least one unit) and unit
// u of a slot starts at item u * items.
__device__ __forceinline__ AcaTile locate(const int32_t *first, int nslots, int32_t q, int items) {
  int lo = 0, hi = nslots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= q)
      lo = mid;
    else
      hi = mid - 1;
  }
  return AcaTile{lo, (q - first[lo]) * items};
}

// One 16-item tile of an ACA column or row pass, resolved on the device
// before the pass: the pivot particle, the first item particle, the number
// of valid items and the first row of the entry buffer Y.
struct AcaPassTile {
  int32_t piv;    // tree-order index of the pivot particle
  int32_t item0;  // tree-order index of the tile's first item
  int32_t nvalid; // items in the tile (1..16)
  int32_t pad;
  int64_t yrow0;  // row of Y of the tile's first item
};
static_assert(sizeof(AcaPassTile) == 24, "AcaPassTile layout (24 B per tile on the host)");

struct AcaMlpArgs {
  const float4 *xp;
  const AcaPassTile *tiles;
  int64_t ntiles;
  int row_pass;           // 0: column pass (items i in I), 1: row pass (j in J)
  float *Y;               // (items, 9) block entries M (not symmetrised)
};

// Tile lists of an ACA pass (src/FarDot.cpp:346-458, 667-782): active slot s
// writes its ceil(items / 16) tiles at toff[s] of the list of its kind; tile
// u of a slot starts at item 16u.  Column pass: pivot x[js + jpiv[k]], items
// x[is + i]; row pass: pivot x[is + ipiv[k]], items x[js + j].
// One wave per active slot: the lanes write the slot's consecutive tiles
// (coalesced 24-byte records), grid-stride over the slots.
__global__ __launch_bounds__(256) void expand_tiles_kernel(AcaExpandLaunch A) {
  const int na = A.cnt->na;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < na; s += nwaves) {
    const int b = A.active[s];
    const FarBlock fb = A.blk[b];
    const int k = A.it[b];
    const int piv = A.row_pass ? fb.is + A.ipiv[(size_t)b * A.max_iter + k]
                               : fb.js + A.jpiv[(size_t)b * (A.max_iter + 1) + k];
    const int base = A.row_pass ? fb.js : fb.is;
    const int lim = A.row_pass ? fb.n : fb.m;
    const int64_t y0 = A.ybase[s];
    AcaPassTile *tl = (AcaPassTile *)((A.kind && A.kind[b]) ? A.tiles_sur : A.tiles_full);
    const int t0 = A.toff[s], nt = (lim + 15) / 16;
    for (int t = lane; t < nt; t += 64) {
      const int off = 16 * t;
      tl[t0 + t] = AcaPassTile{piv, base + off, min(16, lim - off), 0, y0 + off};
    }
  }
}

// Displacements of the four tiles q0..q0+3 of a wave (lane p: item p of each
// tile), loaded one iteration ahead of their use.
struct AcaPassIn {
  float4 xi[4], xv[4];
  int64_t yrow[4];
  bool valid[4];
};

__device__ __forceinline__ void aca_pass_load(const AcaMlpArgs &A, int64_t q0, int p,
                                              AcaPassIn &in) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int64_t q = q0 + t < A.ntiles ? q0 + t : A.ntiles - 1;
    const AcaPassTile T = A.tiles[q];
    in.valid[t] = (q0 + t < A.ntiles) && p < T.nvalid;
    const int pc = in.valid[t] ? p : 0;
    in.xi[t] = A.xp[T.item0 + pc];
    in.xv[t] = A.xp[T.piv];
    in.yrow[t] = T.yrow0 + pc;
  }
}

template <class Core>
__device__ __forceinline__ void aca_mlp_body(const AcaMlpArgs &A, Core core) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, p = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  int64_t q0 = 4 * wave;
  if (q0 >= A.ntiles) return;
  AcaPassIn cur;
  aca_pass_load(A, q0, p, cur);
  for (; q0 < A.ntiles; q0 += 4 * nwaves) {
    float rx[4], ry[4], rz[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (!A.row_pass) {  // pivot - item
        rx[t] = cur.xv[t].x - cur.xi[t].x, ry[t] = cur.xv[t].y - cur.xi[t].y,
        rz[t] = cur.xv[t].z - cur.xi[t].z;
      } else {  // item - pivot
        rx[t] = cur.xi[t].x - cur.xv[t].x, ry[t] = cur.xi[t].y - cur.xv[t].y,
        rz[t] = cur.xi[t].z - cur.xv[t].z;
      }
    }
    bool valid[4];
    int64_t yrow[4];
#pragma unroll
    for (int t = 0; t < 4; t++) valid[t] = cur.valid[t], yrow[t] = cur.yrow[t];
    // the next tiles' loads are in flight while this MLP runs
    if (q0 + 4 * nwaves < A.ntiles) aca_pass_load(A, q0 + 4 * nwaves, p, cur);
    f32x4 y[4];
    core(rx, ry, rz, y);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (!valid[t] || g >= 3) continue;
      float m[3];
      block_row<float>(y[t], rx[t], ry[t], rz[t], g, m);
#pragma unroll
      for (int c = 0; c < 3; c++) A.Y[9 * yrow[t] + 3 * g + c] = m[c];
    }
  }
}

__global__ __launch_bounds__(256, 2) void aca_mlp_fp32_kernel(AcaMlpArgs A, const int32_t *dntiles,
                                                             WF<float> W) {
  A.ntiles = *dntiles;
  aca_mlp_body(A, [&](const float(&rx)[4], const float(&ry)[4], const float(&rz)[4],
                      f32x4(&y)[4]) { mlp_tiles_f<float, 4>(W, rx, ry, rz, y); });
}

template <class H>
__global__ __launch_bounds__(kNearHThreads) void aca_mlp_h_kernel(AcaMlpArgs A, const int32_t *dntiles,
                                                                 WH W) {
  A.ntiles = *dntiles;
  // a workgroup whose first tile is past the count has no work: it leaves
  // before staging the fragments (uniform)
  if ((int64_t)blockIdx.x * (blockDim.x >> 6) * 4 >= A.ntiles) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  H *lw2 = (H *)lds;
  H *lw3 = (H *)(lds + kW2FragBytes);
  stage_fragments<H>(W, lw2, lw3);
  aca_mlp_body(A, [&](const float(&rx)[4], const float(&ry)[4], const float(&rz)[4],
                      f32x4(&y)[4]) { mlp_tiles_h<H, 4>(W, lw2, lw3, rx, ry, rz, y); });
}

// ---- surrogate-generated ACA entries -------------------------------------
// The tiles of blocks that passed the surrogate judge.  One wave per 16-item
// tile, with the tile's items on the column axis of the MFMAs and the hidden
// features on the row axis, so each layer's output registers are the next
// layer's B operand:
//   layer 1 (K = 3 inputs + bias) on v_mfma_f32_16x16x4f32: A = [W1 | b1]
//            rows 16t..16t+15 (lane g, p: W1[16t+p][g], b1 for g = 3), B =
//            (z, 1) of item p; D lane (g, p) register r = feature 16t + 4g +
//            r, t = 0, 1;
//   layers 2 and 3 (K = 32) on v_mfma_f32_16x16x32_f16: lane (g, p) holds
//            K-slots 8g + j, j = 0..7, which are the features f(g, j) = 4g + j
//            (j < 4) and 16 + 4g + j - 4 (j >= 4) -- exactly the eight
//            activations the lane holds after the layer before -- so A =
//            W[16mt + p][f(g, j)] (W3 rows 6..15 zero) and the bias enters as C.
// The ACA needs entries without 16-bit noise, so the 16-bit operands are
// split: w = w_hi + w_lo and a = a_hi + a_lo with w_hi = fp16(w), w_lo =
// fp16(w - w_hi), and each product is w_hi a_hi + w_hi a_lo + w_lo a_hi in
// FP32 accumulation (three MFMAs): ~2^-22 of each term (FP16 subnormals are
// kept), about a quarter of the MFMA cycles of the FP32 instruction.  Each
// lane gathers and splits its weight fragments and biases once.  Y receives
// the symmetric block s(r)/|r| in the full MLP's row-major layout (the step
// kernels' symmetrisation leaves it unchanged).  Same displacement
// convention as aca_mlp_body: x_j - x_i.
__device__ __forceinline__ int sur_feature(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + j - 4; }

__device__ __forceinline__ void split16(float v, _Float16 &hi, _Float16 &lo) {
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}

// The hidden layers carry scaled pre-activations z' = -z log2 e (W1, b1 and
// b2 scaled on loading; W2 unchanged, as log2 e ln 2 = 1) and the scaled
// activation s' = z' sigma(z) = z' / (1 + 2^z'), so that silu z = -ln 2 s'
// costs one exponential, one add and one reciprocal for sigma; the product
// z' sigma is formed inside the two mixed-precision FMAs that split it into
// 16-bit operands.  The factor -ln 2 is folded into W3.
constexpr float kSurScale = -1.4426950408889634f;  // -log2 e
constexpr float kSurUnscale = -0.6931471805599453f;  // -ln 2
// sigma(z) of a scaled pre-activation z': s' = z' sig_scaled(z')
__device__ __forceinline__ float sig_scaled(float zs) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(zs));
}

struct SurMfma {
  float w1[2];
  f16x8 w2h[2], w2l[2], w3h, w3l;
  f32x4 b2v[2], b3v;

  // The lane's weight fragments and biases, gathered from the parameters.
  __device__ __forceinline__ void init(const float *params, int g, int p) {
    using namespace sur;
#pragma unroll
    for (int t = 0; t < 2; t++)
      w1[t] = kSurScale * (g < 3 ? params[kOW1 + (16 * t + p) * kIn + g] : params[kOB1 + 16 * t + p]);
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        _Float16 hi, lo;
        split16(params[kOW2 + (16 * mt + p) * kH + sur_feature(g, j)], hi, lo);
        w2h[mt][j] = hi;
        w2l[mt][j] = lo;
      }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      _Float16 hi, lo;
      split16(p < kOut ? kSurUnscale * params[kOW3 + p * kH + sur_feature(g, j)] : 0.0f, hi, lo);
      w3h[j] = hi;
      w3l[j] = lo;
    }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int r = 0; r < 4; r++) b2v[mt][r] = kSurScale * params[kOB2 + 16 * mt + 4 * g + r];
#pragma unroll
    for (int r = 0; r < 4; r++) b3v[r] = 4 * g + r < kOut ? params[kOB3 + 4 * g + r] : 0.0f;
  }

  // The lane's eight activations v = z sigma (two output registers of four
  // of the layer before, given as z and sigma) as split 16-bit B operands:
  // K-slot 8g + j <- v[j / 4][j % 4].  hi = fp16(z sigma) and lo = fp16(z
  // sigma - hi), each one mixed-precision FMA (z sigma - hi is exact in the
  // FMA, hi read as a 16-bit operand), so hi + lo = z sigma to ~2^-22.
  __device__ __forceinline__ void operands(const f32x4 (&z)[2], const f32x4 (&sg)[2], f16x8 &hi,
                                           f16x8 &lo) const {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float zz = z[j / 4][j % 4], ss = sg[j / 4][j % 4];
      const _Float16 h = (_Float16)__builtin_fmaf(zz, ss, 0.0f);
      hi[j] = h;
      lo[j] = (_Float16)__builtin_fmaf(zz, ss, -(float)h);
    }
  }

  // Outputs 4g + r (r = 0..3) of the tile's item p; bz is z_g of item p
  // (g < 3) or 1 (g = 3).  Every lane of the wave takes part.
  __device__ __forceinline__ f32x4 eval(float bz) const {
    f32x4 a1[2], s1[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      a1[t] = mma(w1[t], bz, f32x4{0, 0, 0, 0});
#pragma unroll
      for (int r = 0; r < 4; r++) s1[t][r] = sig_scaled(a1[t][r]);
    }
    f16x8 h1, l1;
    operands(a1, s1, h1, l1);
    f32x4 a2[2], s2[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
      f32x4 c = mma(w2h[mt], h1, b2v[mt]);
      c = mma(w2h[mt], l1, c);
      c = mma(w2l[mt], h1, c);
#pragma unroll
      for (int r = 0; r < 4; r++) s2[mt][r] = sig_scaled(c[r]);
      a2[mt] = c;
    }
    f16x8 h2, l2;
    operands(a2, s2, h2, l2);
    f32x4 y = mma(w3h, h2, b3v);
    y = mma(w3h, l2, y);
    y = mma(w3l, h2, y);
    return y;
  }
};

__global__ __launch_bounds__(256) void aca_sur_kernel(const float4 *xp, const AcaPassTile *tiles,
                                                      const int32_t *dntiles, int row_pass, float *Y,
                                                      const float *params, float inv_scale) {
  const int64_t ntiles = *dntiles;
  if (ntiles <= 0) return;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, p = lane & 15;
  SurMfma S;
  S.init(params, g, p);
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < ntiles; q += nwaves) {
    const AcaPassTile T = tiles[q];
    const bool valid = p < T.nvalid;
    const float4 xi = xp[T.item0 + (valid ? p : 0)], xv = xp[T.piv];
    float rx, ry, rz;
    if (!row_pass) {
      rx = xv.x - xi.x, ry = xv.y - xi.y, rz = xv.z - xi.z;
    } else {
      rx = xi.x - xv.x, ry = xi.y - xv.y, rz = xi.z - xv.z;
    }
    const float bz = g == 0 ? rx * inv_scale : g == 1 ? ry * inv_scale : g == 2 ? rz * inv_scale : 1.0f;
    const f32x4 y = S.eval(bz);
    if (!valid || g > 1) continue;
    const float inv = rsqrt_r(rx * rx + ry * ry + rz * rz);
    float *o = Y + 9 * (T.yrow0 + p);
    if (g == 0) {
      o[0] = y[0] * inv, o[4] = y[1] * inv, o[8] = y[2] * inv;
      o[1] = o[3] = y[3] * inv;
    } else {
      o[5] = o[7] = y[0] * inv;
      o[2] = o[6] = y[1] * inv;
    }
  }
}

// Surrogate outputs (n,6) for n displacements, 16 per wave on the MFMA path
// of aca_sur_kernel (the judge's samples).
__global__ __launch_bounds__(256) void sur_pairs_kernel(const float *r, int64_t n, float *out,
                                                        const float *params, float inv_scale) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, p = lane & 15;
  SurMfma S;
  S.init(params, g, p);
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; 16 * q < n; q += nwaves) {
    const int64_t k = 16 * q + p;
    const bool valid = k < n;
    const int64_t kc = valid ? k : n - 1;
    const float bz = g < 3 ? r[3 * kc + g] * inv_scale : 1.0f;
    const f32x4 y = S.eval(bz);
    if (!valid || g > 1) continue;
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
      if (4 * g + rr < sur::kOut) out[sur::kOut * k + 4 * g + rr] = y[rr];
  }
}

// ---- surrogate judge --------------------------------------------------------
// Sample t of far block b: pair (i, j) = (h(bS+t) mod m, h'(bS+t) mod n), a
// fixed integer hash spreading the samples over the block; r = x_j - x_i.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__global__ void judge_pairs_kernel(const FarBlock *blk, int64_t nblk, const float4 *xp,
                                   int samples, float *r) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk * samples) return;
  const FarBlock B = blk[k / samples];
  const uint32_t h = (uint32_t)k;
  const int i = (int)(mix32(2u * h) % (uint32_t)B.m);
  const int j = (int)(mix32(2u * h + 1u) % (uint32_t)B.n);
  const float4 a = xp[B.is + i], c = xp[B.js + j];
  r[3 * k] = c.x - a.x;
  r[3 * k + 1] = c.y - a.y;
  r[3 * k + 2] = c.z - a.z;
}

// One wave per block: lane t < samples compares the surrogate with the full
// MLP's block m9 at its sample (both as 3x3 blocks, the MLP's symmetrised
// as the ACA uses it); the block passes when the sampled Frobenius error is
// at most tol times the sampled Frobenius norm and every sampled |r| lies in
// [r_lo, r_hi].  err[b] receives the relative error (inf when out of range).
__global__ __launch_bounds__(256) void judge_kernel(const float *r, const float *m9, const float *s6,
                                                    int64_t nblk, int samples, float r_lo,
                                                    float r_hi, double tol, int32_t *flag,
                                                    float *err) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; b < nblk; b += nwaves) {
  double e2 = 0.0, n2 = 0.0;
  int out = 0;
  for (int t = lane; t < samples; t += 64) {
    const int64_t k = b * samples + t;
    const float rx = r[3 * k], ry = r[3 * k + 1], rz = r[3 * k + 2];
    const float d = sqrtf(rx * rx + ry * ry + rz * rz);
    if (!(d >= r_lo && d <= r_hi)) out = 1;
    const float *y = s6 + sur::kOut * k;
    const float *M = m9 + 9 * k;
#pragma unroll
    for (int o = 0; o < sur::kOut; o++) {
      const int a = sur::entry_row(o), c = sur::entry_col(o);
      const double full = 0.5 * (double)(M[3 * a + c] + M[3 * c + a]);
      const double s = (double)y[o] / (double)d;
      const double w = sur::entry_weight(o);
      e2 += w * (s - full) * (s - full);
      n2 += w * full * full;
    }
  }
  e2 = wave_sum(e2);
  n2 = wave_sum(n2);
  for (int m = 32; m >= 1; m >>= 1) out |= __shfl_xor(out, m, 64);
  if (lane == 0) {
    const double rel = n2 > 0.0 ? sqrt(e2 / n2) : (e2 > 0.0 ? INFINITY : 0.0);
    const bool ok = !out && rel <= tol;
    flag[b] = ok ? 1 : 0;
    err[b] = out ? INFINITY : (float)rel;
  }
  }
}

// ---- surrogate training samples -------------------------------------------
// Sample k: |r| = r_lo (r_hi / r_lo)^u0 and a uniform direction (cos theta =
// 2 u1 - 1, phi = 2 pi u2), the u from a splitmix64 counter hash of
// (seed, k).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double unit(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

__global__ void sur_samples_kernel(int64_t n, float r_lo, float r_hi, uint64_t seed, float *r) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t base = splitmix64(seed ^ 0x5851f42d4c957f2dull) + 3ull * (uint64_t)k;
  const double u0 = unit(splitmix64(base)), u1 = unit(splitmix64(base + 1)),
               u2 = unit(splitmix64(base + 2));
  const double d = (double)r_lo * exp(u0 * log((double)r_hi / (double)r_lo));
  const double ct = 2.0 * u1 - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct));
  const double ph = 6.283185307179586 * u2;
  r[3 * k] = (float)(d * st * cos(ph));
  r[3 * k + 1] = (float)(d * st * sin(ph));
  r[3 * k + 2] = (float)(d * ct);
}

// t = |r| S(M) as the six surrogate outputs (xx, yy, zz, xy, yz, zx).
__global__ void sur_targets_kernel(const float *r, const float *m9, int64_t n, float *t) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float rx = r[3 * k], ry = r[3 * k + 1], rz = r[3 * k + 2];
  const float d = sqrtf(rx * rx + ry * ry + rz * rz);
  const float *M = m9 + 9 * k;
#pragma unroll
  for (int o = 0; o < sur::kOut; o++) {
    const int a = sur::entry_row(o), c = sur::entry_col(o);
    t[sur::kOut * k + o] = d * (0.5f * (M[3 * a + c] + M[3 * c + a]));
  }
}

// ---- chunked per-block steps ------------------------------------------------
// The items of the active blocks (rows i of I for the column side, columns j
// of J for the row side) are cut into chunks of at most kAcaChunk items and one
// wave handles one chunk, so a block of 10^5 items spreads over the whole chip
// instead of one CU.  Per-chunk partial results (pivot arg-max, Gram sums,
// cross-term sums) are combined per block, in chunk order, by one-workgroup
// kernels.
constexpr int kAcaIpt = kAcaChunk / 64;  // items per lane
constexpr double kNegHuge = -1.7976931348623157e308;

struct AcaPart {  // per chunk
  double val;     // arg-max of det over the chunk's eligible items
  int32_t idx;    // -1: none
  int32_t pad;
  double g[6];    // Gram sums: GC (column side) or GQ (row side)
};
static_assert(sizeof(AcaPart) == kAcaPartBytes, "AcaPart layout");

struct AcaStepArgs {
  const FarBlock *blk;
  const int32_t *active;    // slot -> block
  const int64_t *ybase_c;   // slot -> first row of Y in the column pass
  const int64_t *ybase_r;   // slot -> first row of Y in the row pass
  const float *Y;
  int32_t *it;
  int32_t *ipiv;            // per block [max_iter]
  int32_t *jpiv;            // per block [max_iter + 1]
  double *mu2;
  double *nu2;
  int32_t *stop;            // per slot
  int max_iter;
  double eps2;
  int nactive;
  const int32_t *cfirst_c;  // slot -> first column-side chunk (nactive + 1 entries)
  const int32_t *cfirst_r;
  AcaPart *part_c;          // per column-side chunk
  AcaPart *part_r;
  double *x_c;              // per chunk: max_iter x 9 column cross-term sums
  double *x_r;              // per chunk: max_iter x 9 row cross-term sums
  double *inv;              // per slot: inverse of the pivot block (9)
  int chunk;                // items per chunk, 1..kAcaChunk
  const AcaCounts *cnt;     // na, ncc, ncr of the iteration
};

template <class S, class P>
__device__ __forceinline__ void load9(const S *src, P (&d)[9]) {
#pragma unroll
  for (int e = 0; e < 9; e++) d[e] = (P)src[e];
}

// Batched ACA factor pools, addressed by item (one 3x3 entry block each).
// FP32 / FP64 pools interleave: entry e of item i at 9 i + e (a grouped FP32
// layout like the BF16 one below measured 1.4% slower per matvec).  BF16
// pools hold items in groups of 64: the group's entries 0-7 as 16-byte
// vectors (64 x 16 B), then its entries 8 (64 x 2 B), 1,152 bytes per group,
// so a wave's loads of consecutive items are aligned 16-byte loads plus one
// coalesced 2-byte load (interleaved 18-byte items load at ~3/4 the rate).
template <class S>
struct PoolAt {
  __device__ static S get(const S *p, int64_t i, int e) { return p[9 * i + e]; }
  template <class P>
  __device__ static void load(const S *p, int64_t i, P (&d)[9]) {
    load9(p + 9 * i, d);
  }
  __device__ static void store(S *p, int64_t i, const S (&v)[9]) {
#pragma unroll
    for (int e = 0; e < 9; e++) p[9 * i + e] = v[e];
  }
};
template <>
struct PoolAt<__bf16> {
  __device__ static int64_t at8(int64_t i) { return (i >> 6) * 576 + (i & 63) * 8; }
  __device__ static int64_t at1(int64_t i) { return (i >> 6) * 576 + 512 + (i & 63); }
  __device__ static __bf16 get(const __bf16 *p, int64_t i, int e) {
    return e < 8 ? p[at8(i) + e] : p[at1(i)];
  }
  template <class P>
  __device__ static void load(const __bf16 *p, int64_t i, P (&d)[9]) {
    const uint4 w = *(const uint4 *)(p + at8(i));
    const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {  // bf16 -> FP32 is exact: the bits move up
      d[2 * q] = (P)__uint_as_float(ww[q] << 16);
      d[2 * q + 1] = (P)__uint_as_float(ww[q] & 0xffff0000u);
    }
    d[8] = (P)p[at1(i)];
  }
  __device__ static void store(__bf16 *p, int64_t i, const __bf16 (&v)[9]) {
    uint32_t ww[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
      ww[q] = (uint32_t)__builtin_bit_cast(uint16_t, v[2 * q]) |
              ((uint32_t)__builtin_bit_cast(uint16_t, v[2 * q + 1]) << 16);
    *(uint4 *)(p + at8(i)) = make_uint4(ww[0], ww[1], ww[2], ww[3]);
    p[at1(i)] = v[8];
  }
};

template <class P>
__device__ __forceinline__ P det3(const P (&m)[9]) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// symmetrised entries of the cMat/qMat copy (src/FarDot.cpp:441-456): the
// off-diagonal pair is added in float, then halved.
template <class P>
__device__ __forceinline__ void sym_load(const float *y, P (&d)[9]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      d[3 * r + c] = (r == c) ? (P)y[3 * r + c] : (P)(0.5 * (double)(y[3 * r + c] + y[3 * c + r]));
}

struct ArgMax {
  double val;
  int idx;
};

__device__ __forceinline__ ArgMax better(ArgMax a, ArgMax b) {
  // larger value wins; equal values: smaller index (first index wins)
  if (b.idx < 0) return a;
  if (a.idx < 0) return b;
  if (b.val > a.val || (b.val == a.val && b.idx < a.idx)) return b;
  return a;
}

__device__ __forceinline__ ArgMax wave_argmax(ArgMax v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    ArgMax o;
    o.val = __shfl_xor(v.val, m, 64);
    o.idx = __shfl_xor(v.idx, m, 64);
    v = better(v, o);
  }
  return v;
}

__device__ ArgMax block_argmax(ArgMax v, ArgMax *sh) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_argmax(v);
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  ArgMax r = sh[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); k++) r = better(r, sh[k]);
  __syncthreads();
  return r;
}

__device__ double block_sum(double v, double *sh) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
  __syncthreads();
  return r;
}

// Wave sums (in FP64) of N <= 32 per-lane values of a one-wave workgroup,
// the first nvalid stored to dst.  The values are transposed through LDS (sh:
// N rows of 65 values of type T, the odd stride spreads the rows over the
// banks); lanes e and e + 32 each add half of row e and one exchange
// completes the sum.
// wave_sum_lds: the same once the lanes have written their values to
// sh[e * 65 + lane] themselves.
template <class T, int N>
__device__ __forceinline__ void wave_sum_lds(T *sh, double *dst, int nvalid) {
  static_assert(N <= 32, "at most 32 values");
  const int lane = threadIdx.x & 63;
  __syncthreads();
  const int e = lane & 31, h = lane >> 5;
  double t = 0.0;
  if (e < N) {
#pragma unroll
    for (int q = 0; q < 32; q++) t += (double)sh[e * 65 + 32 * h + q];
  }
  t += __shfl_xor(t, 32, 64);
  if (lane < nvalid) dst[lane] = t;
  __syncthreads();
}

// Sums each of the N values x over the 64 lanes (FP64 across lanes) and
// stores the sums to dst[0..nvalid).  For workgroups of one wave only: the
// wave's LDS instructions execute in order, so the exchange through sh needs
// wave-scope ordering, not a workgroup barrier (whose release fence would
// also wait for the wave's outstanding global loads).
template <class T, int N>
__device__ __forceinline__ void wave_sum_store(const T (&x)[N], T *sh, double *dst, int nvalid) {
  static_assert(N <= 32, "at most 32 values");
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int e = 0; e < N; e++) sh[e * 65 + lane] = x[e];
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int e = lane & 31, h = lane >> 5;
  double t = 0.0;
  if (e < N) {
#pragma unroll
    for (int q = 0; q < 32; q++) t += (double)sh[e * 65 + 32 * h + q];
  }
  t += __shfl_xor(t, 32, 64);
  if (lane < nvalid) dst[lane] = t;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// 3x3 Gram sums of one item's block a (row-major): column side GC = a^T a,
// row side GQ = a a^T; entries (00, 11, 22, 01, 12, 20).
__device__ __forceinline__ void gram_col(const double (&c)[9], double (&g)[6]) {
#pragma unroll
  for (int kk = 0; kk < 3; kk++) g[kk] += c[kk] * c[kk] + c[kk + 3] * c[kk + 3] + c[kk + 6] * c[kk + 6];
#pragma unroll
  for (int kk = 3; kk < 6; kk++) {
    const int a = kk % 3, bq = (kk + 1) % 3;
    g[kk] += c[a] * c[bq] + c[a + 3] * c[bq + 3] + c[a + 6] * c[bq + 6];
  }
}
__device__ __forceinline__ void gram_row(const double (&q)[9], double (&g)[6]) {
  g[0] += q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  g[1] += q[3] * q[3] + q[4] * q[4] + q[5] * q[5];
  g[2] += q[6] * q[6] + q[7] * q[7] + q[8] * q[8];
  g[3] += q[0] * q[3] + q[1] * q[4] + q[2] * q[5];
  g[4] += q[3] * q[6] + q[4] * q[7] + q[5] * q[8];
  g[5] += q[6] * q[0] + q[7] * q[1] + q[8] * q[2];
}

// ---- residual passes with the cross terms of the previous iteration --------
// The stop test of iteration k needs the cross terms C_k^T C_l and Q_k Q_l^T,
// l < k-1 (src/FarDot.cpp:1041-1078), i.e. one more pass over all previous
// factors.  The residual passes of iteration k+1 read exactly those factors,
// so they take the cross-term sums of iteration k as well, and the soft stop
// test of iteration k (nu2 < mu2 eps^2) is decided one iteration late.  A
// block that stops there has computed one column and row pass too many; its
// rank and every decision are the reference's.  The hard stops
// (iteration count, block size) are taken at once, and so is the soft stop
// whenever nu2 < (mu2_{k-1} + nu2) eps^2 already holds without the cross
// terms (they only add to mu2).

constexpr int kCrossN = 9;  // cross sums per wave reduction (one factor)

// Residual sums of a lane's kAcaIpt items, sum = sum_{l<k} L_l R_l with the
// 3x3 products of the reference's dot loops (src/FarDot.cpp:487-511, 812-836;
// the caller subtracts them from the new column or row),
// summed in the pool precision, plus the cross-term sums of iteration k-1
// for l < k-2 against a = factor k-1 of the same item (per lane in the pool
// precision over at most kAcaIpt items, across the wave in FP64):
//   column side: L_l = C_l(i) from the pool, R_l = Q_l(jk) from LDS,
//                x[9l + 3d1 + d2] = sum_i sum_row C_{k-1}(i)[row][d1] C_l(i)[row][d2]
//   row side:    L_l = C_l(ik) from LDS, R_l = Q_l(j) from the pool,
//                x[9l + 3d1 + d2] = sum_j sum_col Q_{k-1}(j)[d1][col] Q_l(j)[d2][col]
// One wave per chunk (the kernels launch workgroups of one wave).  item0 is
// the lane's first item of C_0 (Q_0) in the pool, factor l starts stride
// items later; the lane's items are 64 apart.
template <class P, class S, int SIDE>
__device__ __forceinline__ void residual_cross(P (&sum)[kAcaIpt][9], const bool (&valid)[kAcaIpt],
                                               const S *pool, int64_t item0, int64_t stride,
                                               int k, const P (*piv)[9], P *sh, double *X) {
  const int ncross = k - 2;  // cross terms of iteration k-1: l < k-2
  // Loads are unconditional, so that all of a lane's items are in flight at
  // once: an item past the chunk's end reads the chunk's first item (always
  // valid) and is zeroed after the load.
  int64_t off[kAcaIpt];
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) off[s] = item0 + (valid[s] ? 64 * s : -(int64_t)threadIdx.x);
  P a[kAcaIpt][9];
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    if (ncross > 0)  // uniform
      PoolAt<S>::load(pool, (int64_t)(k - 1) * stride + off[s], a[s]);
#pragma unroll
    for (int e = 0; e < 9; e++) {
      sum[s][e] = (P)0;
      a[s][e] = (valid[s] && ncross > 0) ? a[s][e] : (P)0;
    }
  }
  // sum += L_l R_l for the lane's items from factor w = factor l; with the
  // cross terms of factor l against a when cross
  auto accum = [&](const P (&w)[kAcaIpt][9], int l, bool cross, P (&x)[9]) {
#pragma unroll
    for (int s = 0; s < kAcaIpt; s++) {
      if (!valid[s]) continue;
      const P *cl = SIDE == 0 ? w[s] : piv[l];
      const P *ql = SIDE == 0 ? piv[l] : w[s];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
          P &t = sum[s][3 * r + cc];
          if (sizeof(P) == 8) {  // FP64 pools: the oracle's rounding, no contraction
            t += cl[3 * r] * ql[cc];
            t += cl[3 * r + 1] * ql[3 + cc];
            t += cl[3 * r + 2] * ql[6 + cc];
          } else {
            t = fma(cl[3 * r], ql[cc], t);
            t = fma(cl[3 * r + 1], ql[3 + cc], t);
            t = fma(cl[3 * r + 2], ql[6 + cc], t);
          }
        }
      if (!cross) continue;
#pragma unroll
      for (int d1 = 0; d1 < 3; d1++)
#pragma unroll
        for (int d2 = 0; d2 < 3; d2++) {
          P &t = x[3 * d1 + d2];
          if (SIDE == 0) {
            t = fma(a[s][d1], w[s][d2], t);
            t = fma(a[s][d1 + 3], w[s][d2 + 3], t);
            t = fma(a[s][d1 + 6], w[s][d2 + 6], t);
          } else {
            t = fma(a[s][3 * d1], w[s][3 * d2], t);
            t = fma(a[s][3 * d1 + 1], w[s][3 * d2 + 1], t);
            t = fma(a[s][3 * d1 + 2], w[s][3 * d2 + 2], t);
          }
        }
    }
  };
  // Factors 0 .. kp-1 come from the pool; with cross terms factor k-1 is
  // already in a (and has no cross terms: k-1 >= ncross).  The loads of
  // factor l+1 go out as soon as factor l has been used, into the same
  // registers, so that they are in flight during the wave reduction of
  // factor l's cross terms.
  const int kp = ncross > 0 ? k - 1 : k;
  P v[kAcaIpt][9];
  if (kp > 0) {
#pragma unroll
    for (int s = 0; s < kAcaIpt; s++) PoolAt<S>::load(pool, off[s], v[s]);
  }
  for (int l = 0; l < kp; l++) {
    const bool cross = l < ncross;  // uniform
    P x[9];
#pragma unroll
    for (int e = 0; e < 9; e++) x[e] = (P)0;
    accum(v, l, cross, x);
    if (l + 1 < kp) {
#pragma unroll
      for (int s = 0; s < kAcaIpt; s++) PoolAt<S>::load(pool, (int64_t)(l + 1) * stride + off[s], v[s]);
    }
    if (cross) wave_sum_store<P, 9>(x, sh, X + 9 * l, 9);
  }
  if (ncross > 0) {
    P x[9];
    accum(a, k - 1, false, x);
  }
}

// Column pass residual R_k, the column cross-term sums of iteration k-1 and
// the row pivot candidates of one chunk (src/FarDot.cpp:474-580).  One wave
// per chunk.
template <class P, class SC, class SQ>
__device__ __forceinline__ void col_resid_chunk(const AcaStepArgs &A, SC *Cp, const SQ *Qp, int q,
                                                int na, P (*piv)[9], int32_t *used, P *sh) {
  const AcaTile ch = locate(A.cfirst_c, na, q, A.chunk);
  const int b = A.active[ch.slot];
  const FarBlock fb = A.blk[b];
  const int k = A.it[b];
  const int jk = A.jpiv[(size_t)b * (A.max_iter + 1) + k];
  for (int t = threadIdx.x; t < 9 * k; t += 64)
    piv[t / 9][t % 9] = (P)PoolAt<SQ>::get(Qp, fb.qoff + (int64_t)(t / 9) * fb.n + jk, t % 9);
  for (int l = threadIdx.x; l < k; l += 64) used[l] = A.ipiv[(size_t)b * A.max_iter + l];
  __syncthreads();
  const int iend = min(ch.off + A.chunk, fb.m);
  bool valid[kAcaIpt];
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) valid[s] = ch.off + (int)threadIdx.x + 64 * s < iend;
  P c[kAcaIpt][9];
  residual_cross<P, SC, 0>(c, valid, Cp, fb.coff + ch.off + threadIdx.x, fb.m, k, piv, sh,
                           A.x_c + (size_t)q * A.max_iter * 9);
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    if (!valid[s]) continue;
    P raw[9];
    sym_load<P>(A.Y + 9 * (A.ybase_c[ch.slot] + ch.off + threadIdx.x + 64 * s), raw);
#pragma unroll
    for (int e = 0; e < 9; e++) c[s][e] = raw[e] - c[s][e];
  }
  const int64_t ck0 = fb.coff + (int64_t)k * fb.m;  // first item of C_k
  // 16-bit C pools: the un-normalised residual column stays in FP32, in place
  // of the pass's raw entries, so the pivot inverse and the normalisation see
  // it unrounded; only the normalised factor is stored in 16 bit
  float *Rk = const_cast<float *>(A.Y) + 9 * A.ybase_c[ch.slot];
  ArgMax best{kNegHuge, -1};
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    const int i = ch.off + threadIdx.x + 64 * s;
    if (!valid[s]) continue;
    if constexpr (sizeof(SC) == 2) {
#pragma unroll
      for (int e = 0; e < 9; e++) Rk[9 * (int64_t)i + e] = (float)c[s][e];
    } else {
      SC v[9];
#pragma unroll
      for (int e = 0; e < 9; e++) v[e] = (SC)c[s][e];
      PoolAt<SC>::store(Cp, ck0 + i, v);
    }
    bool u = false;
    for (int l = 0; l < k; l++) u |= (used[l] == i);
    const double dt = (double)det3(c[s]);
    if (!u && dt > best.val) best = ArgMax{dt, i};
  }
  const ArgMax am = wave_argmax(best);
  if (threadIdx.x == 0) {
    A.part_c[q].val = am.val;
    A.part_c[q].idx = am.idx;
  }
}

// One wave per chunk, grid-stride over the iteration's chunks.
// LDS of a residual-pass workgroup, sized by max_iter at launch (so that
// more of these one-wave workgroups fit beside the near kernel's LDS): the
// pivot factors piv[max_iter][9], the used pivots used[max_iter + 1] and the
// cross-term exchange sh[kCrossN * 65].
template <class P>
__host__ __device__ constexpr size_t resid_lds_used_off(int max_iter) {
  return sizeof(P) * 9 * (size_t)max_iter;
}
template <class P>
__host__ __device__ constexpr size_t resid_lds_sh_off(int max_iter) {
  return (resid_lds_used_off<P>(max_iter) + 4 * ((size_t)max_iter + 1) + 15) / 16 * 16;
}
template <class P>
__host__ __device__ constexpr size_t resid_lds_bytes(int max_iter) {
  return resid_lds_sh_off<P>(max_iter) + sizeof(P) * kCrossN * 65;
}

template <class P, class SC, class SQ>
__global__ __launch_bounds__(64) void aca_col_resid_kernel(AcaStepArgs A, SC *Cp, const SQ *Qp) {
  extern __shared__ __attribute__((aligned(16))) unsigned char resid_lds[];
  P(*piv)[9] = (P(*)[9])resid_lds;
  int32_t *used = (int32_t *)(resid_lds + resid_lds_used_off<P>(A.max_iter));
  P *sh = (P *)(resid_lds + resid_lds_sh_off<P>(A.max_iter));
  const int na = A.cnt->na, nc = A.cnt->ncc;
  for (int q = blockIdx.x; q < nc; q += gridDim.x) {
    col_resid_chunk<P, SC, SQ>(A, Cp, Qp, q, na, piv, used, sh);
    __syncthreads();  // piv / used of this chunk are read until here
  }
}

// Row pivot of each active block and the inverse of its pivot block
// (src/FarDot.cpp:537-656).  One wave per block.
template <class SC>
__device__ __forceinline__ void col_pivot_slot(const AcaStepArgs &A, const SC *Cp, int slot) {
  const int b = A.active[slot];
  const FarBlock fb = A.blk[b];
  const int k = A.it[b];
  ArgMax best{kNegHuge, -1};
  for (int c = A.cfirst_c[slot] + (int)threadIdx.x; c < A.cfirst_c[slot + 1]; c += 64)
    best = better(best, ArgMax{A.part_c[c].val, A.part_c[c].idx});
  best = wave_argmax(best);
  if (threadIdx.x == 0) {
    const int ik = best.idx < 0 ? 0 : best.idx;
    A.ipiv[(size_t)b * A.max_iter + k] = ik;
    double m[9];
    if constexpr (sizeof(SC) == 2) {  // the FP32 residual column (see the column pass)
      const float *Rk = A.Y + 9 * (A.ybase_c[slot] + (int64_t)ik);
#pragma unroll
      for (int e = 0; e < 9; e++) m[e] = (double)Rk[e];
    } else {
      PoolAt<SC>::load(Cp, fb.coff + (int64_t)k * fb.m + ik, m);
    }
    const double a = m[0], bb = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6],
                 h = m[7], ii = m[8];
    const double det = a * (e * ii - f * h) - bb * (d * ii - f * g) + c * (d * h - e * g);
    double iv[9];
    iv[0] = (e * ii - f * h) / det;
    iv[1] = -(bb * ii - c * h) / det;
    iv[2] = (bb * f - c * e) / det;
    iv[3] = -(d * ii - f * g) / det;
    iv[4] = (a * ii - c * g) / det;
    iv[5] = -(a * f - c * d) / det;
    iv[6] = (d * h - e * g) / det;
    iv[7] = -(a * h - bb * g) / det;
    iv[8] = (a * e - bb * d) / det;
    // An exhausted residual can leave an exactly singular pivot block (more
    // likely with 16-bit factor storage); its term is then set to zero, the
    // norm test stops the block, and no inf/NaN enters the pools.
    bool finite = det != 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) finite = finite && isfinite(iv[q]);
    double *inv = A.inv + 9 * (size_t)slot;
#pragma unroll
    for (int q = 0; q < 9; q++) inv[q] = finite ? iv[q] : 0.0;
  }
}

template <class SC>
__global__ __launch_bounds__(64) void aca_col_pivot_kernel(AcaStepArgs A, const SC *Cp) {
  const int na = A.cnt->na;
  for (int slot = blockIdx.x; slot < na; slot += gridDim.x) col_pivot_slot<SC>(A, Cp, slot);
}

// Column normalisation C_k = R_k inv (src/FarDot.cpp:630-656) and the Gram
// sums of C_k for one chunk.  One wave per chunk.
template <class P, class SC>
__device__ __forceinline__ void col_norm_chunk(const AcaStepArgs &A, SC *Cp, int q, int na) {
  const AcaTile ch = locate(A.cfirst_c, na, q, A.chunk);
  const int b = A.active[ch.slot];
  const FarBlock fb = A.blk[b];
  const int k = A.it[b];
  P inv[9];
#pragma unroll
  for (int e = 0; e < 9; e++) inv[e] = (P)A.inv[9 * (size_t)ch.slot + e];
  const int64_t ck0 = fb.coff + (int64_t)k * fb.m;  // first item of C_k
  const int iend = min(ch.off + A.chunk, fb.m);
  double g[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    const int i = ch.off + threadIdx.x + 64 * s;
    if (i >= iend) continue;
    P c[9], o[9];
    if constexpr (sizeof(SC) == 2)
      load9(A.Y + 9 * (A.ybase_c[ch.slot] + (int64_t)i), c);
    else
      PoolAt<SC>::load(Cp, ck0 + i, c);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++)
        o[3 * r + cc] = c[3 * r] * inv[cc] + c[3 * r + 1] * inv[3 + cc] + c[3 * r + 2] * inv[6 + cc];
    double a[9];
    SC st[9];  // the norms see the stored factor
#pragma unroll
    for (int e = 0; e < 9; e++) {
      st[e] = (SC)o[e];
      a[e] = (double)st[e];
    }
    PoolAt<SC>::store(Cp, ck0 + i, st);
    gram_col(a, g);
  }
#pragma unroll
  for (int e = 0; e < 6; e++) {
    const double s = wave_sum(g[e]);
    if (threadIdx.x == 0) A.part_c[q].g[e] = s;
  }
}

template <class P, class SC>
__global__ __launch_bounds__(64) void aca_col_norm_kernel(AcaStepArgs A, SC *Cp) {
  const int na = A.cnt->na, nc = A.cnt->ncc;
  for (int q = blockIdx.x; q < nc; q += gridDim.x) col_norm_chunk<P, SC>(A, Cp, q, na);
}

// Row pass residual Q_k, the row cross-term sums of iteration k-1, column
// pivot candidates and the Gram sums of Q_k for one chunk
// (src/FarDot.cpp:797-1078).  One wave per chunk.
template <class P, class SC, class SQ>
__device__ __forceinline__ void row_resid_chunk(const AcaStepArgs &A, const SC *Cp, SQ *Qp, int q,
                                                int na, P (*piv)[9], int32_t *used, P *sh) {
  const AcaTile ch = locate(A.cfirst_r, na, q, A.chunk);
  const int b = A.active[ch.slot];
  const FarBlock fb = A.blk[b];
  const int k = A.it[b];
  const int ik = A.ipiv[(size_t)b * A.max_iter + k];
  for (int t = threadIdx.x; t < 9 * k; t += 64)
    piv[t / 9][t % 9] = (P)PoolAt<SC>::get(Cp, fb.coff + (int64_t)(t / 9) * fb.m + ik, t % 9);
  for (int l = threadIdx.x; l <= k; l += 64) used[l] = A.jpiv[(size_t)b * (A.max_iter + 1) + l];
  __syncthreads();
  const int jend = min(ch.off + A.chunk, fb.n);
  bool valid[kAcaIpt];
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) valid[s] = ch.off + (int)threadIdx.x + 64 * s < jend;
  P qv[kAcaIpt][9];
  residual_cross<P, SQ, 1>(qv, valid, Qp, fb.qoff + ch.off + threadIdx.x, fb.n, k, piv, sh,
                           A.x_r + (size_t)q * A.max_iter * 9);
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    if (!valid[s]) continue;
    P raw[9];
    sym_load<P>(A.Y + 9 * (A.ybase_r[ch.slot] + ch.off + threadIdx.x + 64 * s), raw);
#pragma unroll
    for (int e = 0; e < 9; e++) qv[s][e] = raw[e] - qv[s][e];
  }
  const int64_t qk0 = fb.qoff + (int64_t)k * fb.n;  // first item of Q_k
  ArgMax best{kNegHuge, -1};
  double g[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    const int j = ch.off + threadIdx.x + 64 * s;
    if (!valid[s]) continue;
    double a[9];
    SQ st[9];  // pivot and norms see the stored factor
#pragma unroll
    for (int e = 0; e < 9; e++) {
      st[e] = (SQ)qv[s][e];
      qv[s][e] = (P)st[e];
      a[e] = (double)st[e];
    }
    PoolAt<SQ>::store(Qp, qk0 + j, st);
    bool u = false;
    for (int l = 0; l <= k; l++) u |= (used[l] == j);
    const double dt = (double)det3(qv[s]);
    if (!u && dt > best.val) best = ArgMax{dt, j};
    gram_row(a, g);
  }
  const ArgMax am = wave_argmax(best);
  AcaPart &pt = A.part_r[q];
  if (threadIdx.x == 0) pt.val = am.val, pt.idx = am.idx;
#pragma unroll
  for (int e = 0; e < 6; e++) {
    const double s = wave_sum(g[e]);
    if (threadIdx.x == 0) pt.g[e] = s;
  }
}

template <class P, class SC, class SQ>
__global__ __launch_bounds__(64) void aca_row_resid_kernel(AcaStepArgs A, const SC *Cp, SQ *Qp) {
  extern __shared__ __attribute__((aligned(16))) unsigned char resid_lds[];
  P(*piv)[9] = (P(*)[9])resid_lds;
  int32_t *used = (int32_t *)(resid_lds + resid_lds_used_off<P>(A.max_iter));
  P *sh = (P *)(resid_lds + resid_lds_sh_off<P>(A.max_iter));
  const int na = A.cnt->na, nc = A.cnt->ncr;
  for (int q = blockIdx.x; q < nc; q += gridDim.x) {
    row_resid_chunk<P, SC, SQ>(A, Cp, Qp, q, na, piv, used, sh);
    __syncthreads();  // piv / used of this chunk are read until here
  }
}

// Per block, after the passes of iteration k: the deferred soft stop test of
// iteration k-1 with its cross terms added to mu2 in the reference's order
// (l, d1, d2 ascending); then the next column pivot, nu2 = |C_k Q_k|_F^2 and
// the hard stop tests of iteration k (src/FarDot.cpp:1002-1108).  mu2[b]
// holds mu2 of iteration k-1 without its cross terms, nu2[b] its nu2; it[b]
// ends as the block's rank when it stops.
__device__ __forceinline__ void finalize_slot(const AcaStepArgs &A, int slot, double *pr,
                                              ArgMax *sh_am, double *sh_g) {
  // Every global load below depends only on the slot's block, so the loads of
  // the pivot arg-max (the threads of the first row chunks), the Gram sums
  // (lanes 32-43, one entry each) and the cross terms (threads 64-255) are in
  // flight together before the first barrier.  Each sum keeps its order.
  const int tid = threadIdx.x;
  const int b = A.active[slot];
  const int c0 = A.cfirst_c[slot], c1 = A.cfirst_c[slot + 1];
  const int r0 = A.cfirst_r[slot], r1 = A.cfirst_r[slot + 1];
  const int k = A.it[b];
  double mu2 = 0.0, nu2_prev = 0.0;
  int mn = 0;
  if (tid == 0) {
    mu2 = A.mu2[b];
    nu2_prev = A.nu2[b];
    const FarBlock fb = A.blk[b];
    mn = fb.m < fb.n ? fb.m : fb.n;
  }
  ArgMax best{kNegHuge, -1};
  for (int c = r0 + tid; c < r1; c += blockDim.x)
    best = better(best, ArgMax{A.part_r[c].val, A.part_r[c].idx});
  if (tid >= 32 && tid < 44) {  // column-pass Gram entry tid-32, row-pass entry tid-38
    const int e = tid < 38 ? tid - 32 : tid - 38;
    const AcaPart *P = tid < 38 ? A.part_c : A.part_r;
    const int e0 = tid < 38 ? c0 : r0, e1 = tid < 38 ? c1 : r1;
    double g = 0.0;
    for (int c = e0; c < e1; c++) g += P[c].g[e];
    sh_g[tid - 32] = g;
  }
  const int ncross = 9 * (k > 2 ? k - 2 : 0);
  if (tid >= 64)
    for (int t = tid - 64; t < ncross; t += (int)blockDim.x - 64) {
      double sc = 0.0, sq = 0.0;
      for (int c = c0; c < c1; c++) sc += A.x_c[(size_t)c * A.max_iter * 9 + t];
      for (int c = r0; c < r1; c++) sq += A.x_r[(size_t)c * A.max_iter * 9 + t];
      pr[t] = 2.0 * fabs(sc) * fabs(sq);
    }
  const ArgMax am = block_argmax(best, sh_am);  // its barriers publish pr and sh_g
  if (tid == 0) {
    bool soft = false;
    if (k >= 1) {
      for (int t = 0; t < ncross; t++) mu2 += pr[t];  // the reference's order (l, d1, d2)
      soft = nu2_prev < mu2 * A.eps2;
    }
    if (soft) {  // iteration k-1 met the tolerance: rank k
      A.mu2[b] = mu2;
      A.stop[slot] = 1;
    } else {
      const double *GC = sh_g, *GQ = sh_g + 6;
      const double nu2 = GC[0] * GQ[0] + GC[1] * GQ[1] + GC[2] * GQ[2] + 2.0 * GQ[3] * GC[3] +
                         2.0 * GQ[4] * GC[4] + 2.0 * GQ[5] * GC[5];
      A.jpiv[(size_t)b * (A.max_iter + 1) + k + 1] = am.idx < 0 ? 0 : am.idx;
      const double m2 = mu2 + nu2;
      A.mu2[b] = m2;
      A.nu2[b] = nu2;
      const int itn = k + 1;
      A.it[b] = itn;
      // The tolerance test of this iteration waits for its cross terms (the
      // next residual passes), which add non-negative terms to m2: when
      // nu2 < m2 eps^2 already holds, that test holds too (rounding is
      // monotone), so the block stops now with the same rank and skips the
      // next column and row pass.
      const bool met = nu2 < m2 * A.eps2;
      A.stop[slot] = (met || itn >= A.max_iter || itn >= mn) ? 1 : 0;
    }
  }
  __syncthreads();  // pr / sh_g / sh_am are reused by the next slot
}

// pr: 9 (max_iter - 2) doubles, sized at launch (fewer bytes of LDS than
// max_iter = 128 needs, so that more workgroups fit beside the near kernel)
__global__ __launch_bounds__(256) void aca_finalize_kernel(AcaStepArgs A) {
  extern __shared__ double pr[];
  __shared__ ArgMax sh_am[4];
  __shared__ double sh_g[12];
  const int na = A.cnt->na;
  for (int slot = blockIdx.x; slot < na; slot += gridDim.x) finalize_slot(A, slot, pr, sh_am, sh_g);
}

// ---- chunked apply of stopped blocks (src/FarDot.cpp:1200-1388) -----------
//   u_I += sum_l C_l (sum_j Q_l[j] f_j);  sym: u_J += sum_l Q_l^T (sum_i C_l[i]^T f_i).
// Phase 1 (per source-side chunk): partial w_l over the chunk's items.
// Phase 2 (per target-side chunk): w_l summed over the block's source chunks,
// then u += sum_l (C_l or Q_l^T) w_l.  Side 0: sources J, targets I; side 1
// (symmetry): sources I, targets J.
struct AcaApplyArgs {
  const FarBlock *blk;
  const int32_t *list;       // apply slot -> block
  const int32_t *it;
  const float4 *fp;
  double *u;
  int nlist;
  const int32_t *first_i;    // apply slot -> first I chunk (nlist + 1)
  const int32_t *first_j;
  double *w_j;               // per J chunk: max_iter x 3 (side 0 partials)
  double *w_i;               // per I chunk: max_iter x 3 (side 1 partials)
  int max_iter;
  int chunk;                 // items per chunk, 1..kAcaChunk
  const AcaCounts *cnt;      // nl, nci, ncj
};

template <class SC, class SQ, int SIDE>
__device__ __forceinline__ void apply_reduce_chunk(const AcaApplyArgs &A, const SC *Cp, const SQ *Qp,
                                                   int q, int nl, double *sh) {
  // SIDE 0 reduces over the Q pool, SIDE 1 over the C pool
  typedef typename std::conditional<SIDE == 0, SQ, SC>::type S;
  const AcaTile ch = SIDE == 0 ? locate(A.first_j, nl, q, A.chunk) : locate(A.first_i, nl, q, A.chunk);
  const int b = A.list[ch.slot];
  const FarBlock fb = A.blk[b];
  const int K = min(A.it[b], 128);
  const int ns = SIDE == 0 ? fb.n : fb.m;
  const int s0 = SIDE == 0 ? fb.js : fb.is;
  const S *Src;
  int64_t item0;  // first item of the side's factor 0
  if constexpr (SIDE == 0)
    Src = Qp, item0 = fb.qoff;
  else
    Src = Cp, item0 = fb.coff;
  const int send = min(ch.off + A.chunk, ns);
  double *W = (SIDE == 0 ? A.w_j : A.w_i) + (size_t)q * A.max_iter * 3;
  float4 f[kAcaIpt];
#pragma unroll
  for (int s = 0; s < kAcaIpt; s++) {
    const int it = ch.off + threadIdx.x + 64 * s;
    f[s] = it < send ? A.fp[s0 + it] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  // eight rank indices (24 sums) per wave reduction; each rank index's three
  // partial sums go to LDS as they are formed
  const int lane = threadIdx.x & 63;
  typedef typename std::conditional<sizeof(S) == 8, double, float>::type F;  // exact
  for (int l0 = 0; l0 < K; l0 += 8) {
    const int nl = min(8, K - l0);
    // two rank indices at a time: both factors' loads are in flight before
    // either is used (an item past the chunk, or a rank past nl, reads the
    // chunk's first item of factor 0 and adds nothing)
#pragma unroll 1
    for (int c0 = 0; c0 < 8; c0 += 2) {
      F qv[2][kAcaIpt][9];
#pragma unroll
      for (int cc = 0; cc < 2; cc++)
#pragma unroll
        for (int s = 0; s < kAcaIpt; s++) {
          const int it = ch.off + threadIdx.x + 64 * s;
          const bool ok = it < send && c0 + cc < nl;
          PoolAt<S>::load(Src, item0 + (ok ? (int64_t)(l0 + c0 + cc) * ns + it : (int64_t)ch.off), qv[cc][s]);
        }
#pragma unroll
      for (int cc = 0; cc < 2; cc++) {
        const int c = c0 + cc;
        double x0 = 0.0, x1 = 0.0, x2 = 0.0;
        if (c < nl) {  // uniform
#pragma unroll
          for (int s = 0; s < kAcaIpt; s++) {
            const int it = ch.off + threadIdx.x + 64 * s;
            if (it >= send) continue;
            const F(&q)[9] = qv[cc][s];
            if (SIDE == 0) {
              x0 += (double)q[0] * f[s].x + (double)q[1] * f[s].y + (double)q[2] * f[s].z;
              x1 += (double)q[3] * f[s].x + (double)q[4] * f[s].y + (double)q[5] * f[s].z;
              x2 += (double)q[6] * f[s].x + (double)q[7] * f[s].y + (double)q[8] * f[s].z;
            } else {
              x0 += (double)q[0] * f[s].x + (double)q[3] * f[s].y + (double)q[6] * f[s].z;
              x1 += (double)q[1] * f[s].x + (double)q[4] * f[s].y + (double)q[7] * f[s].z;
              x2 += (double)q[2] * f[s].x + (double)q[5] * f[s].y + (double)q[8] * f[s].z;
            }
          }
        }
        sh[(3 * c) * 65 + lane] = x0;
        sh[(3 * c + 1) * 65 + lane] = x1;
        sh[(3 * c + 2) * 65 + lane] = x2;
      }
    }
    wave_sum_lds<double, 24>(sh, W + 3 * l0, 3 * nl);
  }
}

template <class SC, class SQ, int SIDE>
__global__ __launch_bounds__(64) void aca_apply_reduce_kernel(AcaApplyArgs A, const SC *Cp,
                                                              const SQ *Qp) {
  __shared__ double sh[24 * 65];
  const int nl = A.cnt->nl, nc = SIDE == 0 ? A.cnt->ncj : A.cnt->nci;
  for (int q = blockIdx.x; q < nc; q += gridDim.x) apply_reduce_chunk<SC, SQ, SIDE>(A, Cp, Qp, q, nl, sh);
}

template <class SC, class SQ, int SIDE>
__device__ __forceinline__ void apply_update_chunk(const AcaApplyArgs &A, const SC *Cp, const SQ *Qp,
                                                   int q, int nl, double *wsh) {
  // SIDE 0 updates from the C pool, SIDE 1 from the Q pool
  typedef typename std::conditional<SIDE == 0, SC, SQ>::type S;
  const AcaTile ch = SIDE == 0 ? locate(A.first_i, nl, q, A.chunk) : locate(A.first_j, nl, q, A.chunk);
  const int b = A.list[ch.slot];
  const FarBlock fb = A.blk[b];
  const int K = min(A.it[b], 128);
  const int w0 = SIDE == 0 ? A.first_j[ch.slot] : A.first_i[ch.slot];
  const int w1 = SIDE == 0 ? A.first_j[ch.slot + 1] : A.first_i[ch.slot + 1];
  const double *W = SIDE == 0 ? A.w_j : A.w_i;
  for (int t = threadIdx.x; t < 3 * K; t += blockDim.x) {
    double s = 0.0;
    for (int c = w0; c < w1; c++) s += W[(size_t)c * A.max_iter * 3 + t];
    wsh[t] = s;
  }
  __syncthreads();
  const int nd = SIDE == 0 ? fb.m : fb.n;
  const int d0 = SIDE == 0 ? fb.is : fb.js;
  const S *Dst;
  int64_t item0;  // first item of the side's factor 0
  if constexpr (SIDE == 0)
    Dst = Cp, item0 = fb.coff;
  else
    Dst = Qp, item0 = fb.qoff;
  const int dend = min(ch.off + A.chunk, nd);
  typedef typename std::conditional<sizeof(S) == 8, double, float>::type F;  // exact
  auto add = [&](const F (&c)[9], int l, double &o0, double &o1, double &o2) {
    const double x0 = wsh[3 * l], x1 = wsh[3 * l + 1], x2 = wsh[3 * l + 2];
    if (SIDE == 0) {
      o0 += (double)c[0] * x0 + (double)c[1] * x1 + (double)c[2] * x2;
      o1 += (double)c[3] * x0 + (double)c[4] * x1 + (double)c[5] * x2;
      o2 += (double)c[6] * x0 + (double)c[7] * x1 + (double)c[8] * x2;
    } else {
      o0 += (double)c[0] * x0 + (double)c[3] * x1 + (double)c[6] * x2;
      o1 += (double)c[1] * x0 + (double)c[4] * x1 + (double)c[7] * x2;
      o2 += (double)c[2] * x0 + (double)c[5] * x1 + (double)c[8] * x2;
    }
  };
  for (int d = ch.off + threadIdx.x; d < dend; d += blockDim.x) {
    double o0 = 0, o1 = 0, o2 = 0;
    int l = 0;
    // four factors' loads in flight before the first is used; the sums keep
    // the order l = 0, 1, 2, ...
    for (; l + 4 <= K; l += 4) {
      F c[4][9];
#pragma unroll
      for (int q = 0; q < 4; q++) PoolAt<S>::load(Dst, item0 + (int64_t)(l + q) * nd + d, c[q]);
#pragma unroll
      for (int q = 0; q < 4; q++) add(c[q], l + q, o0, o1, o2);
    }
    for (; l < K; l++) {
      F c[9];
      PoolAt<S>::load(Dst, item0 + (int64_t)l * nd + d, c);
      add(c, l, o0, o1, o2);
    }
    double *u = A.u + 3 * (size_t)(d0 + d);
    unsafeAtomicAdd(u + 0, o0);
    unsafeAtomicAdd(u + 1, o1);
    unsafeAtomicAdd(u + 2, o2);
  }
}

template <class SC, class SQ, int SIDE>
__global__ __launch_bounds__(64) void aca_apply_update_kernel(AcaApplyArgs A, const SC *Cp,
                                                               const SQ *Qp) {
  __shared__ double wsh[128 * 3];
  const int nl = A.cnt->nl, nc = SIDE == 0 ? A.cnt->nci : A.cnt->ncj;
  for (int q = blockIdx.x; q < nc; q += gridDim.x) {
    apply_update_chunk<SC, SQ, SIDE>(A, Cp, Qp, q, nl, wsh);
    __syncthreads();  // wsh of this chunk is read until here
  }
}

// ---- device plan of a batched ACA iteration --------------------------------
// Per input element (a block of the previous active list): the values whose
// exclusive prefix sums place it in the new active list (kept) or in the
// apply list (stopped), and two running totals (entries per kind).
__device__ __forceinline__ void plan_values(const AcaPlanLaunch &L, int e, int64_t (&v)[kPlanFields],
                                            int &b, bool &keep, bool &full) {
  b = L.act_in ? L.act_in[e] : e;
  keep = !(L.stop_in && L.stop_in[e]);
  full = !(L.kind && L.kind[b]);
  // first plan of a loop over one kind: the other blocks are neither kept nor
  // applied
  const bool skip = !L.act_in && L.select_kind >= 0 && (full ? 0 : 1) != L.select_kind;
  const FarBlock fb = L.blk[b];
  const int64_t m = fb.m, n = fb.n;
  const int64_t tm = (m + 15) / 16, tn = (n + 15) / 16;
  const int64_t cm = (m + L.chunk - 1) / L.chunk, cn = (n + L.chunk - 1) / L.chunk;
  const int64_t k = (keep && !skip) ? 1 : 0, d = (keep || skip) ? 0 : 1, f = full ? 1 : 0;
  v[0] = k;
  v[1] = k * m;
  v[2] = k * n;
  v[3] = k * f * tm;
  v[4] = k * (1 - f) * tm;
  v[5] = k * f * tn;
  v[6] = k * (1 - f) * tn;
  v[7] = k * cm;
  v[8] = k * cn;
  v[9] = d;
  v[10] = d * cm;
  v[11] = d * cn;
  v[12] = k * f * (m + n);
  v[13] = k * (1 - f) * (m + n);
}

__device__ __forceinline__ void plan_range(const AcaPlanLaunch &L, int &e0, int &e1) {
  const int nin = L.cnt_in ? L.cnt_in->na : L.nb;
  const int per = (nin + kPlanGroups - 1) / kPlanGroups;
  e0 = min(nin, (int)blockIdx.x * per);
  e1 = min(nin, e0 + per);
}

__global__ __launch_bounds__(256) void aca_plan_reduce_kernel(AcaPlanLaunch L) {
  __shared__ int64_t sh[4][kPlanFields];
  int e0, e1;
  plan_range(L, e0, e1);
  int64_t acc[kPlanFields];
#pragma unroll
  for (int f = 0; f < kPlanFields; f++) acc[f] = 0;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
    int64_t v[kPlanFields];
    int b;
    bool keep, full;
    plan_values(L, e, v, b, keep, full);
#pragma unroll
    for (int f = 0; f < kPlanFields; f++) acc[f] += v[f];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int f = 0; f < kPlanFields; f++) {
    int64_t t = acc[f];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) sh[wv][f] = t;
  }
  __syncthreads();
  if (threadIdx.x < kPlanFields) {
    const int f = threadIdx.x;
    L.part[(size_t)blockIdx.x * kPlanFields + f] = sh[0][f] + sh[1][f] + sh[2][f] + sh[3][f];
  }
}

// One workgroup of kPlanGroups threads: exclusive scan of the group sums,
// totals into cnt_out, sentinels of the prefix arrays.
__global__ __launch_bounds__(kPlanGroups) void aca_plan_scan_kernel(AcaPlanLaunch L) {
  __shared__ int64_t sh[kPlanGroups];
  __shared__ int64_t tot[kPlanFields];
  const int g = threadIdx.x;
  for (int f = 0; f < kPlanFields; f++) {
    const int64_t x = L.part[(size_t)g * kPlanFields + f];
    sh[g] = x;
    __syncthreads();
    for (int o = 1; o < kPlanGroups; o <<= 1) {  // inclusive Hillis-Steele scan
      const int64_t y = g >= o ? sh[g - o] : 0;
      __syncthreads();
      sh[g] += y;
      __syncthreads();
    }
    L.part[(size_t)g * kPlanFields + f] = sh[g] - x;  // exclusive
    if (g == kPlanGroups - 1) tot[f] = sh[g];
    __syncthreads();
  }
  if (g == 0) {
    AcaCounts c;
    c.na = (int32_t)tot[0];
    c.ysum_c = tot[1];
    c.ysum_r = tot[2];
    c.ntc[0] = (int32_t)tot[3];
    c.ntc[1] = (int32_t)tot[4];
    c.ntr[0] = (int32_t)tot[5];
    c.ntr[1] = (int32_t)tot[6];
    c.ncc = (int32_t)tot[7];
    c.ncr = (int32_t)tot[8];
    c.nl = (int32_t)tot[9];
    c.nci = (int32_t)tot[10];
    c.ncj = (int32_t)tot[11];
    c.entries[0] = (L.cnt_in ? L.cnt_in->entries[0] : 0) + tot[12];
    c.entries[1] = (L.cnt_in ? L.cnt_in->entries[1] : 0) + tot[13];
    *L.cnt_out = c;
    L.cfirst_c[c.na] = c.ncc;
    L.cfirst_r[c.na] = c.ncr;
    L.first_i[c.nl] = c.nci;
    L.first_j[c.nl] = c.ncj;
  }
}

// The group's range in order, 256 elements per round: block-wide exclusive
// scan of the placing values, plus the group's offset and the rounds before.
// The two Y-row sums are 64-bit; the other placing fields (slots, tiles,
// chunks; all below 2^31, checked when the batches are laid out) 32-bit.
__device__ __forceinline__ int64_t wave_incl_scan(int64_t t, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(t, o, 64);
    if (lane >= o) t += y;
  }
  return t;
}
__device__ __forceinline__ int32_t wave_incl_scan(int32_t t, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(t, o, 64);
    if (lane >= o) t += y;
  }
  return t;
}

__global__ __launch_bounds__(256) void aca_plan_write_kernel(AcaPlanLaunch L) {
  constexpr int F = 12;  // fields that place an element: 1 and 2 64-bit, the rest 32-bit
  __shared__ int64_t wsum64[4][2];
  __shared__ int32_t wsum32[4][F];
  int e0, e1;
  plan_range(L, e0, e1);
  int64_t carry64[2];
  int32_t carry32[F];
#pragma unroll
  for (int f = 0; f < F; f++) {
    const int64_t c = L.part[(size_t)blockIdx.x * kPlanFields + f];
    if (f == 1 || f == 2) carry64[f - 1] = c;
    else carry32[f] = (int32_t)c;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int r0 = e0; r0 < e1; r0 += blockDim.x) {  // uniform
    const int e = r0 + (int)threadIdx.x;
    int64_t v[kPlanFields];
    int b = 0;
    bool keep = false, full = true;
    if (e < e1) {
      plan_values(L, e, v, b, keep, full);
    } else {
#pragma unroll
      for (int f = 0; f < kPlanFields; f++) v[f] = 0;
    }
    int64_t pre64[2];
    int32_t pre32[F];
#pragma unroll
    for (int f = 0; f < F; f++) {
      if (f == 1 || f == 2) {
        const int64_t t = wave_incl_scan(v[f], lane);
        pre64[f - 1] = t - v[f];
        if (lane == 63) wsum64[wv][f - 1] = t;
      } else {
        const int32_t x = (int32_t)v[f];
        const int32_t t = wave_incl_scan(x, lane);
        pre32[f] = t - x;
        if (lane == 63) wsum32[wv][f] = t;
      }
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < F; f++) {
      if (f == 1 || f == 2) {
        int64_t w = 0;
        for (int q = 0; q < wv; q++) w += wsum64[q][f - 1];
        pre64[f - 1] += carry64[f - 1] + w;
        carry64[f - 1] += wsum64[0][f - 1] + wsum64[1][f - 1] + wsum64[2][f - 1] + wsum64[3][f - 1];
      } else {
        int32_t w = 0;
        for (int q = 0; q < wv; q++) w += wsum32[q][f];
        pre32[f] += carry32[f] + w;
        carry32[f] += wsum32[0][f] + wsum32[1][f] + wsum32[2][f] + wsum32[3][f];
      }
    }
    __syncthreads();  // wsum is rewritten by the next round
    if (e < e1 && (v[0] | v[9])) {  // kept or stopped (not skipped)
      if (keep) {
        const int s = pre32[0];
        L.act_out[s] = b;
        L.ybase_c[s] = pre64[0];
        L.ybase_r[s] = pre64[1];
        L.toff_c[s] = full ? pre32[3] : pre32[4];
        L.toff_r[s] = full ? pre32[5] : pre32[6];
        L.cfirst_c[s] = pre32[7];
        L.cfirst_r[s] = pre32[8];
      } else {
        const int a = pre32[9];
        L.alist[a] = b;
        L.first_i[a] = pre32[10];
        L.first_j[a] = pre32[11];
      }
    }
  }
}

__global__ void aca_init_kernel(const FarBlock *blk, int nb, int max_iter, int32_t *it, int32_t *jpiv,
                                double *mu2, double *nu2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  it[b] = 0;
  jpiv[(size_t)b * (max_iter + 1)] = blk[b].n / 2;
  mu2[b] = 0.0;
  nu2[b] = 0.0;
}

__global__ void aca_entries_kernel(const AcaCounts *cnt, int64_t *acc) {
  if (threadIdx.x < 2) acc[threadIdx.x] += cnt->entries[threadIdx.x];
}

__global__ void copy_ranks_kernel(const int32_t *it, int nb, int32_t *rank_out, const int32_t *kind,
                                  int select_kind) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb && (!kind || (kind[b] ? 1 : 0) == select_kind)) rank_out[b] = it[b];
}

// ---------------------------------------------------------------------------
// Persistent far field: one workgroup (4 waves) takes far blocks from the
// cost-sorted list and runs the whole block ACA of src/FarDot.cpp:186-1388 for
// each of them -- column pass, residual, row pivot and normalisation, row
// pass, residual, column pivot, norms and stop test, then the low-rank apply
// -- with the block's factors in a per-workgroup slot of the C/Q pools (so
// they stay cache-resident between iterations) and no host round trip per
// iteration.  The per-block arithmetic is that of the batched reference.
// ---------------------------------------------------------------------------
struct PAcaArgs {
  const float4 *xp;
  const float4 *fp;
  double *u;
  const FarBlock *blk;   // blocks in processing order
  const int32_t *order;  // position -> far-list index (for the rank output)
  int32_t *rank_out;     // per far-list index
  int64_t nblocks;
  int max_iter;
  double eps2;
  int use_symmetry;
  void *poolC;           // nslots x slotC entries (3x3)
  void *poolQ;
  int64_t slotC, slotQ;  // entries per slot
  float *Y;              // nslots x slotY x 9
  int64_t slotY;
};

// MLP over nitems displacements of one block: column pass (row_pass = 0):
// x[js + piv] - x[is + i]; row pass: x[js + j] - x[is + piv].  Writes the
// 3x3 entries (after +I, /|r|) to Y[item][9].
template <class Core>
__device__ __forceinline__ void pass_entries(const float4 *xp, int is, int js, int piv, int nitems,
                                             int row_pass, float *Y, Core core) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int g = lane >> 4, p = lane & 15;
  const float4 xpiv = row_pass ? xp[is + piv] : xp[js + piv];
  for (int base = 64 * wv; base < nitems; base += 64 * nw) {
    float rx[4], ry[4], rz[4];
    int item[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      item[t] = base + 16 * t + p;
      const int it = item[t] < nitems ? item[t] : 0;
      if (row_pass) {
        const float4 xb = xp[js + it];
        rx[t] = xb.x - xpiv.x, ry[t] = xb.y - xpiv.y, rz[t] = xb.z - xpiv.z;
      } else {
        const float4 xa = xp[is + it];
        rx[t] = xpiv.x - xa.x, ry[t] = xpiv.y - xa.y, rz[t] = xpiv.z - xa.z;
      }
    }
    f32x4 y[4];
    core(rx, ry, rz, y);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (item[t] >= nitems || g >= 3) continue;
      float m[3];
      block_row<float>(y[t], rx[t], ry[t], rz[t], g, m);
#pragma unroll
      for (int c = 0; c < 3; c++) Y[9 * (int64_t)item[t] + 3 * g + c] = m[c];
    }
  }
}

template <class P, class Core>
__device__ void persistent_aca(const PAcaArgs &A, Core core) {
  __shared__ int sh_stop;
  __shared__ int ipiv[128];
  __shared__ int jpiv[129];
  __shared__ double sh_inv[9];
  __shared__ double sh_mu2;
  __shared__ ArgMax sh_am[8];
  __shared__ double sh_s[8];
  __shared__ double sh_x[4][72];
  __shared__ double wsh[128][3];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  P *Cp = (P *)A.poolC + (int64_t)blockIdx.x * A.slotC * 9;
  P *Qp = (P *)A.poolQ + (int64_t)blockIdx.x * A.slotQ * 9;
  float *Y = A.Y + (int64_t)blockIdx.x * A.slotY * 9;
  // Blocks are dealt out statically (b = blockIdx.x + i * gridDim.x over the
  // cost-sorted list) so that the loop bounds are wave-uniform scalars.
  // Every lane-0 region below is followed by a barrier: two such regions on
  // either side of a loop's back edge could otherwise be merged into one
  // divergent detour, letting the other lanes pass a barrier ahead of lane 0
  // (seen in the generated code of an atomic-queue version of this loop).
  for (int64_t bb = blockIdx.x; bb < A.nblocks; bb += gridDim.x) {
    const int b = (int)bb;
    const FarBlock fb = A.blk[b];
    const int m = fb.m, n = fb.n;
    if (tid == 0) {
      jpiv[0] = n / 2;
      sh_mu2 = 0.0;
    }
    __syncthreads();
    int k = 0;
    for (;;) {
      // ---- column pass and residual, row pivot (src/FarDot.cpp:346-611)
      const int jk = jpiv[k];
      pass_entries(A.xp, fb.is, fb.js, jk, m, 0, Y, core);
      __syncthreads();
      P *Ck = Cp + 9 * (int64_t)k * m;
      ArgMax best{-1.7976931348623157e308, -1};
      for (int i = tid; i < m; i += blockDim.x) {
        P c[9];
        sym_load<P>(Y + 9 * (int64_t)i, c);
        P sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int l = 0; l < k; l++) {
          P cl[9], ql[9];
          load9(Cp + 9 * ((int64_t)l * m + i), cl);
          load9(Qp + 9 * ((int64_t)l * n + jk), ql);
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
              P &s = sum[3 * r + cc];
              s += cl[3 * r] * ql[cc];
              s += cl[3 * r + 1] * ql[3 + cc];
              s += cl[3 * r + 2] * ql[6 + cc];
            }
        }
#pragma unroll
        for (int e = 0; e < 9; e++) c[e] -= sum[e];
#pragma unroll
        for (int e = 0; e < 9; e++) Ck[9 * (int64_t)i + e] = c[e];
        bool used = false;
        for (int l = 0; l < k; l++) used |= (ipiv[l] == i);
        const double d = (double)det3(c);
        if (!used && d > best.val) best.val = d, best.idx = i;
      }
      const ArgMax am = block_argmax(best, sh_am);
      const int ik = am.idx < 0 ? 0 : am.idx;
      if (tid == 0) {
        ipiv[k] = ik;
        double mm[9];
#pragma unroll
        for (int e = 0; e < 9; e++) mm[e] = (double)Ck[9 * (int64_t)ik + e];
        const double a = mm[0], bb = mm[1], c = mm[2], d = mm[3], e = mm[4], f = mm[5], g = mm[6],
                     h = mm[7], ii = mm[8];
        const double det = a * (e * ii - f * h) - bb * (d * ii - f * g) + c * (d * h - e * g);
        double iv[9];
        iv[0] = (e * ii - f * h) / det;
        iv[1] = -(bb * ii - c * h) / det;
        iv[2] = (bb * f - c * e) / det;
        iv[3] = -(d * ii - f * g) / det;
        iv[4] = (a * ii - c * g) / det;
        iv[5] = -(a * f - c * d) / det;
        iv[6] = (d * h - e * g) / det;
        iv[7] = -(a * h - bb * g) / det;
        iv[8] = (a * e - bb * d) / det;
        // as the batched path: an exactly singular or non-finite pivot
        // inverse becomes zero, so the new term vanishes, nu2 = 0 ends the
        // block at the norm test, and no inf/NaN reaches the pools or u
        bool finite = det != 0.0;
#pragma unroll
        for (int q = 0; q < 9; q++) finite = finite && isfinite(iv[q]);
#pragma unroll
        for (int q = 0; q < 9; q++) sh_inv[q] = finite ? iv[q] : 0.0;
      }
      __syncthreads();
      {
        P inv[9];
#pragma unroll
        for (int e = 0; e < 9; e++) inv[e] = (P)sh_inv[e];
        for (int i = tid; i < m; i += blockDim.x) {
          P c[9], o[9];
          load9(Ck + 9 * (int64_t)i, c);
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++)
              o[3 * r + cc] = c[3 * r] * inv[cc] + c[3 * r + 1] * inv[3 + cc] + c[3 * r + 2] * inv[6 + cc];
#pragma unroll
          for (int e = 0; e < 9; e++) Ck[9 * (int64_t)i + e] = o[e];
        }
      }
      // ---- row pass and residual, column pivot, norms (src/FarDot.cpp:667-1108)
      pass_entries(A.xp, fb.is, fb.js, ik, n, 1, Y, core);
      __syncthreads();
      P *Qk = Qp + 9 * (int64_t)k * n;
      ArgMax bq{-1.7976931348623157e308, -1};
      double gq[6] = {0, 0, 0, 0, 0, 0};
      for (int j = tid; j < n; j += blockDim.x) {
        P q[9];
        sym_load<P>(Y + 9 * (int64_t)j, q);
        P sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int l = 0; l < k; l++) {
          P cl[9], ql[9];
          load9(Cp + 9 * ((int64_t)l * m + ik), cl);
          load9(Qp + 9 * ((int64_t)l * n + j), ql);
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
              P &s = sum[3 * r + cc];
              s += cl[3 * r] * ql[cc];
              s += cl[3 * r + 1] * ql[3 + cc];
              s += cl[3 * r + 2] * ql[6 + cc];
            }
        }
#pragma unroll
        for (int e = 0; e < 9; e++) q[e] -= sum[e];
#pragma unroll
        for (int e = 0; e < 9; e++) Qk[9 * (int64_t)j + e] = q[e];
        bool used = false;
        for (int l = 0; l <= k; l++) used |= (jpiv[l] == j);
        const double d = (double)det3(q);
        if (!used && d > bq.val) bq.val = d, bq.idx = j;
        double qd[9];
#pragma unroll
        for (int e = 0; e < 9; e++) qd[e] = (double)q[e];
        gq[0] += qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2];
        gq[1] += qd[3] * qd[3] + qd[4] * qd[4] + qd[5] * qd[5];
        gq[2] += qd[6] * qd[6] + qd[7] * qd[7] + qd[8] * qd[8];
        gq[3] += qd[0] * qd[3] + qd[1] * qd[4] + qd[2] * qd[5];
        gq[4] += qd[3] * qd[6] + qd[4] * qd[7] + qd[5] * qd[8];
        gq[5] += qd[6] * qd[0] + qd[7] * qd[1] + qd[8] * qd[2];
      }
      double gc[6] = {0, 0, 0, 0, 0, 0};
      for (int i = tid; i < m; i += blockDim.x) {
        double c[9];
#pragma unroll
        for (int e = 0; e < 9; e++) c[e] = (double)Ck[9 * (int64_t)i + e];
#pragma unroll
        for (int kk = 0; kk < 3; kk++) gc[kk] += c[kk] * c[kk] + c[kk + 3] * c[kk + 3] + c[kk + 6] * c[kk + 6];
#pragma unroll
        for (int kk = 3; kk < 6; kk++) {
          const int a = kk % 3, bb = (kk + 1) % 3;
          gc[kk] += c[a] * c[bb] + c[a + 3] * c[bb + 3] + c[a + 6] * c[bb + 6];
        }
      }
      const ArgMax aq = block_argmax(bq, sh_am);
      double GC[6], GQ[6];
#pragma unroll
      for (int e = 0; e < 6; e++) {
        GC[e] = block_sum(gc[e], sh_s);
        GQ[e] = block_sum(gq[e], sh_s);
      }
      double nu2 = 0.0;
      if (tid == 0) {
        nu2 = GC[0] * GQ[0] + GC[1] * GQ[1] + GC[2] * GQ[2] + 2.0 * GQ[3] * GC[3] +
              2.0 * GQ[4] * GC[4] + 2.0 * GQ[5] * GC[5];
        sh_mu2 += nu2;
      }
      constexpr int LC = 4;
      for (int l0 = 0; l0 + 1 < k; l0 += LC) {
        const int nl = (k - 1 - l0) < LC ? (k - 1 - l0) : LC;
        double cd[LC][9], qd[LC][9];
#pragma unroll
        for (int c = 0; c < LC; c++)
#pragma unroll
          for (int e = 0; e < 9; e++) cd[c][e] = qd[c][e] = 0.0;
        for (int i = tid; i < m; i += blockDim.x) {
          double a[9];
#pragma unroll
          for (int e = 0; e < 9; e++) a[e] = (double)Ck[9 * (int64_t)i + e];
#pragma unroll
          for (int c = 0; c < LC; c++) {
            if (c >= nl) break;
            const P *Cl = Cp + 9 * ((int64_t)(l0 + c) * m + i);
            double v[9];
#pragma unroll
            for (int e = 0; e < 9; e++) v[e] = (double)Cl[e];
#pragma unroll
            for (int d1 = 0; d1 < 3; d1++)
#pragma unroll
              for (int d2 = 0; d2 < 3; d2++)
                cd[c][3 * d1 + d2] += a[d1] * v[d2] + a[d1 + 3] * v[d2 + 3] + a[d1 + 6] * v[d2 + 6];
          }
        }
        for (int j = tid; j < n; j += blockDim.x) {
          double a[9];
#pragma unroll
          for (int e = 0; e < 9; e++) a[e] = (double)Qk[9 * (int64_t)j + e];
#pragma unroll
          for (int c = 0; c < LC; c++) {
            if (c >= nl) break;
            const P *Ql = Qp + 9 * ((int64_t)(l0 + c) * n + j);
            double v[9];
#pragma unroll
            for (int e = 0; e < 9; e++) v[e] = (double)Ql[e];
#pragma unroll
            for (int d1 = 0; d1 < 3; d1++)
#pragma unroll
              for (int d2 = 0; d2 < 3; d2++)
                qd[c][3 * d1 + d2] += a[3 * d1] * v[3 * d2] + a[3 * d1 + 1] * v[3 * d2 + 1] +
                                      a[3 * d1 + 2] * v[3 * d2 + 2];
          }
        }
#pragma unroll
        for (int c = 0; c < LC; c++)
#pragma unroll
          for (int e = 0; e < 9; e++) {
            const double s1 = wave_sum(cd[c][e]);
            const double s2 = wave_sum(qd[c][e]);
            if (lane == 0) {
              sh_x[wv][c * 9 + e] = s1;
              sh_x[wv][LC * 9 + c * 9 + e] = s2;
            }
          }
        __syncthreads();
        if (tid == 0) {
          for (int c = 0; c < nl; c++)
            for (int e = 0; e < 9; e++) {
              double C_ = 0.0, Q_ = 0.0;
              for (int q = 0; q < nw; q++) C_ += sh_x[q][c * 9 + e], Q_ += sh_x[q][LC * 9 + c * 9 + e];
              sh_mu2 += 2.0 * fabs(C_) * fabs(Q_);
            }
        }
        __syncthreads();
      }
      if (tid == 0) {
        jpiv[k + 1] = aq.idx < 0 ? 0 : aq.idx;
        const int itn = k + 1;
        const int mn = m < n ? m : n;
        sh_stop = (nu2 < sh_mu2 * A.eps2 || itn >= A.max_iter || itn >= mn) ? 1 : 0;
      }
      __syncthreads();
      k++;
      if (__builtin_amdgcn_readfirstlane(sh_stop)) break;  // uniform exit, see above
    }
    // ---- apply (src/FarDot.cpp:1200-1388)
    const int K = k;
    for (int pass = 0; pass < (A.use_symmetry ? 2 : 1); pass++) {
      const P *Src = pass == 0 ? Qp : Cp;
      const P *Dst = pass == 0 ? Cp : Qp;
      const int ns = pass == 0 ? n : m, nd = pass == 0 ? m : n;
      const int s0 = pass == 0 ? fb.js : fb.is, d0 = pass == 0 ? fb.is : fb.js;
      for (int l = wv; l < K; l += nw) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int s = lane; s < ns; s += 64) {
          const P *q = Src + 9 * ((int64_t)l * ns + s);
          const float4 f = A.fp[s0 + s];
          if (pass == 0) {
            a0 += (double)q[0] * f.x + (double)q[1] * f.y + (double)q[2] * f.z;
            a1 += (double)q[3] * f.x + (double)q[4] * f.y + (double)q[5] * f.z;
            a2 += (double)q[6] * f.x + (double)q[7] * f.y + (double)q[8] * f.z;
          } else {
            a0 += (double)q[0] * f.x + (double)q[3] * f.y + (double)q[6] * f.z;
            a1 += (double)q[1] * f.x + (double)q[4] * f.y + (double)q[7] * f.z;
            a2 += (double)q[2] * f.x + (double)q[5] * f.y + (double)q[8] * f.z;
          }
        }
        a0 = wave_sum(a0), a1 = wave_sum(a1), a2 = wave_sum(a2);
        if (lane == 0) wsh[l][0] = a0, wsh[l][1] = a1, wsh[l][2] = a2;
      }
      __syncthreads();
      for (int d = tid; d < nd; d += blockDim.x) {
        double o0 = 0, o1 = 0, o2 = 0;
        for (int l = 0; l < K; l++) {
          const P *c = Dst + 9 * ((int64_t)l * nd + d);
          const double w0 = wsh[l][0], w1 = wsh[l][1], w2 = wsh[l][2];
          if (pass == 0) {
            o0 += (double)c[0] * w0 + (double)c[1] * w1 + (double)c[2] * w2;
            o1 += (double)c[3] * w0 + (double)c[4] * w1 + (double)c[5] * w2;
            o2 += (double)c[6] * w0 + (double)c[7] * w1 + (double)c[8] * w2;
          } else {
            o0 += (double)c[0] * w0 + (double)c[3] * w1 + (double)c[6] * w2;
            o1 += (double)c[1] * w0 + (double)c[4] * w1 + (double)c[7] * w2;
            o2 += (double)c[2] * w0 + (double)c[5] * w1 + (double)c[8] * w2;
          }
        }
        double *uu = A.u + 3 * (size_t)(d0 + d);
        unsafeAtomicAdd(uu + 0, o0);
        unsafeAtomicAdd(uu + 1, o1);
        unsafeAtomicAdd(uu + 2, o2);
      }
      __syncthreads();
    }
    if (tid == 0) A.rank_out[A.order[b]] = K;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void persistent_aca_fp32_kernel(PAcaArgs A, WF<float> W) {
  persistent_aca<double>(A, [&](const float(&rx)[4], const float(&ry)[4], const float(&rz)[4],
                                f32x4(&y)[4]) { mlp_tiles_f<float, 4>(W, rx, ry, rz, y); });
}

template <class H>
__global__ __launch_bounds__(256) void persistent_aca_h_kernel(PAcaArgs A, WH W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  H *lw2 = (H *)lds;
  H *lw3 = (H *)(lds + kW2FragBytes);
  stage_fragments<H>(W, lw2, lw3);
  persistent_aca<float>(A, [&](const float(&rx)[4], const float(&ry)[4], const float(&rz)[4],
                               f32x4(&y)[4]) { mlp_tiles_h<H, 4>(W, lw2, lw3, rx, ry, rz, y); });
}

// ---------------------------------------------------------------------------
// Pair function on a list of displacements: out9[p] = M(r[p]).
// ---------------------------------------------------------------------------
template <class R, class Core>
__device__ __forceinline__ void pair_body(const float *r, int64_t P, float *out, Core core) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, p = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t base = 64 * wave; base < P; base += 64 * nwaves) {
    R rx[4], ry[4], rz[4];
    int64_t idx[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      idx[t] = base + 16 * t + p;
      const int64_t c = idx[t] < P ? idx[t] : P - 1;
      rx[t] = (R)r[3 * c], ry[t] = (R)r[3 * c + 1], rz[t] = (R)r[3 * c + 2];
    }
    typename Vec4<R>::T y[4];
    core(rx, ry, rz, y);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (idx[t] >= P || g >= 3) continue;
      R m[3];
      block_row<R>(y[t], rx[t], ry[t], rz[t], g, m);
#pragma unroll
      for (int c = 0; c < 3; c++) out[9 * idx[t] + 3 * g + c] = (float)m[c];
    }
  }
}

__global__ __launch_bounds__(256) void pair_f32_kernel(const float *r, int64_t P, float *out,
                                                       WF<float> W) {
  pair_body<float>(r, P, out, [&](const float(&x)[4], const float(&y)[4], const float(&z)[4],
                                  f32x4(&o)[4]) { mlp_tiles_f<float, 4>(W, x, y, z, o); });
}

__global__ __launch_bounds__(256) void pair_f64_kernel(const float *r, int64_t P, float *out,
                                                       WF<double> W) {
  pair_body<double>(r, P, out, [&](const double(&x)[4], const double(&y)[4], const double(&z)[4],
                                   f64x4(&o)[4]) { mlp_tiles_f<double, 4>(W, x, y, z, o); });
}

template <class H>
__global__ __launch_bounds__(256) void pair_h_kernel(const float *r, int64_t P, float *out, WH W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  H *lw2 = (H *)lds;
  H *lw3 = (H *)(lds + kW2FragBytes);
  stage_fragments<H>(W, lw2, lw3);
  pair_body<float>(r, P, out, [&](const float(&x)[4], const float(&y)[4], const float(&z)[4],
                                  f32x4(&o)[4]) { mlp_tiles_h<H, 4>(W, lw2, lw3, x, y, z, o); });
}

// ---------------------------------------------------------------------------
// Permutations (src/HignnModel.cpp:252-320).
// ---------------------------------------------------------------------------
__global__ void gather_forces_kernel(const float *f, int64_t n, const int32_t *perm, float4 *fp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = perm[i];
  fp[i] = make_float4(f[3 * s], f[3 * s + 1], f[3 * s + 2], 0.0f);
}

__global__ void scatter_velocity_kernel(const double *u, int64_t n, const int32_t *perm, float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t d = perm[i];
  out[3 * d] = (float)u[3 * i];
  out[3 * d + 1] = (float)u[3 * i + 1];
  out[3 * d + 2] = (float)u[3 * i + 2];
}

// ---------------------------------------------------------------------------
// Launchers (called from hh_model.cpp).
// ---------------------------------------------------------------------------
static void allow_big_lds() {
  static bool done = false;
  if (done) return;
  const int lds = kW2FragBytes + kW3FragBytes;
  for (const void *k : {(const void *)near_h_kernel<bf16x8, 4, 512>, (const void *)near_h_kernel<f16x8, 4, 512>,
                        (const void *)near_h_kernel<bf16x8, 2, 512>, (const void *)near_h_kernel<f16x8, 2, 512>,
                        (const void *)near_h_kernel<bf16x8, 2, 1024>, (const void *)near_h_kernel<f16x8, 2, 1024>,
                        (const void *)near_h_kernel<bf16x8, 4, 256>, (const void *)near_h_kernel<f16x8, 4, 256>})
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void *)aca_mlp_h_kernel<bf16x8>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void *)aca_mlp_h_kernel<f16x8>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void *)pair_h_kernel<bf16x8>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void *)pair_h_kernel<f16x8>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  done = true;
}

static int grid_for(int64_t waves_needed, int waves_per_block, int max_blocks) {
  int64_t blocks = (waves_needed + waves_per_block - 1) / waves_per_block;
  if (blocks < 1) blocks = 1;
  if (blocks > max_blocks) blocks = max_blocks;
  return (int)blocks;
}

__global__ void zero_i32_kernel(int32_t *p, int64_t n);

hipError_t launch_near(const NearLaunch &L, hipStream_t st) {
  if (L.ntasks <= 0) return hipSuccess;
  if (L.prec == PREC_FP32) {
    NearSrc<float> S{L.xp, L.fp, L.tlist, L.blocks, L.tasks, L.ntasks, L.u, nullptr};
    WF<float> W{(const float *)L.wf.w1b, (const float *)L.wf.w2f, (const float *)L.wf.b2,
                (const float *)L.wf.w3f, (const float *)L.wf.b3};
    const int grid = grid_for(L.ntasks, 4, L.max_blocks);
    hipLaunchKernelGGL(near_fp32_kernel, dim3(grid), dim3(256), 0, st, S, W);
  } else if (L.prec == PREC_FP64) {
    NearSrc<double> S{L.xp, L.fp, L.tlist, L.blocks, L.tasks, L.ntasks, L.u, nullptr};
    WF<double> W{(const double *)L.wf.w1b, (const double *)L.wf.w2f, (const double *)L.wf.b2,
                 (const double *)L.wf.w3f, (const double *)L.wf.b3};
    const int grid = grid_for(L.ntasks, 4, L.max_blocks);
    hipLaunchKernelGGL(near_fp64_kernel, dim3(grid), dim3(256), 0, st, S, W);
  } else {
    allow_big_lds();
    if (L.next) {
      hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, st, L.next, 1);
    }
    NearSrc<float> S{L.xp, L.fp, L.tlist, L.blocks, L.tasks, L.ntasks, L.u, L.next};
    WH W{L.wh.w1b, L.wh.b2, L.wh.b3, L.wh.w2f, L.wh.w3f};
    const size_t lds = kW2FragBytes + kW3FragBytes;
    const bool b16 = L.prec == PREC_BF16;
    const int threads = kNearShapes[L.hcfg].threads;
    const int grid = grid_for(L.ntasks, threads / 64, L.cus);
#define HH_NEAR_H(NT, TH)                                                                         \
  do {                                                                                            \
    if (b16)                                                                                      \
      hipLaunchKernelGGL((near_h_kernel<bf16x8, NT, TH>), dim3(grid), dim3(TH), lds, st, S, W);  \
    else                                                                                          \
      hipLaunchKernelGGL((near_h_kernel<f16x8, NT, TH>), dim3(grid), dim3(TH), lds, st, S, W);   \
  } while (0)
    switch (L.hcfg) {
      case 1: HH_NEAR_H(2, 512); break;
      case 2: HH_NEAR_H(2, 1024); break;
      case 3: HH_NEAR_H(4, 256); break;
      default: HH_NEAR_H(4, 512);
    }
#undef HH_NEAR_H
  }
  return hipGetLastError();
}

hipError_t launch_aca_plan(const AcaPlanLaunch &L, hipStream_t st) {
  hipLaunchKernelGGL(aca_plan_reduce_kernel, dim3(kPlanGroups), dim3(256), 0, st, L);
  hipLaunchKernelGGL(aca_plan_scan_kernel, dim3(1), dim3(kPlanGroups), 0, st, L);
  hipLaunchKernelGGL(aca_plan_write_kernel, dim3(kPlanGroups), dim3(256), 0, st, L);
  return hipGetLastError();
}

hipError_t launch_aca_init(const FarBlock *blk, int nb, int max_iter, int32_t *it, int32_t *jpiv,
                           double *mu2, double *nu2, hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(aca_init_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, blk, nb,
                     max_iter, it, jpiv, mu2, nu2);
  return hipGetLastError();
}

hipError_t launch_aca_entries(const AcaCounts *cnt, int64_t *acc, hipStream_t st) {
  hipLaunchKernelGGL(aca_entries_kernel, dim3(1), dim3(64), 0, st, cnt, acc);
  return hipGetLastError();
}

hipError_t launch_copy_ranks(const int32_t *it, int nb, int32_t *rank_out, const int32_t *kind,
                             int select_kind, hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy_ranks_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, it, nb,
                     rank_out, kind, select_kind);
  return hipGetLastError();
}

hipError_t launch_aca_expand(const AcaExpandLaunch &L, hipStream_t st) {
  if (L.grid <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((L.grid + 3) / 4, 8192);  // a wave per slot
  hipLaunchKernelGGL(expand_tiles_kernel, dim3((unsigned)blocks), dim3(256), 0, st, L);
  return hipGetLastError();
}

hipError_t launch_aca_mlp(const AcaMlpLaunch &L, hipStream_t st) {
  if (L.ntiles <= 0) return hipSuccess;
  AcaMlpArgs A{L.xp, (const AcaPassTile *)L.tiles, L.ntiles, L.row_pass, L.Y};
  const int64_t waves = (L.ntiles + 3) / 4;
  if (L.prec == PREC_FP32) {
    WF<float> W{(const float *)L.wf.w1b, (const float *)L.wf.w2f, (const float *)L.wf.b2,
                (const float *)L.wf.w3f, (const float *)L.wf.b3};
    hipLaunchKernelGGL(aca_mlp_fp32_kernel, dim3(grid_for(waves, 4, L.max_blocks)), dim3(256), 0,
                       st, A, L.dntiles, W);
  } else {
    allow_big_lds();
    WH W{L.wh.w1b, L.wh.b2, L.wh.b3, L.wh.w2f, L.wh.w3f};
    const size_t lds = kW2FragBytes + kW3FragBytes;
    const int grid = grid_for(waves, kNearHThreads / 64, L.cus);
    if (L.prec == PREC_BF16)
      hipLaunchKernelGGL(aca_mlp_h_kernel<bf16x8>, dim3(grid), dim3(kNearHThreads), lds, st, A,
                         L.dntiles, W);
    else
      hipLaunchKernelGGL(aca_mlp_h_kernel<f16x8>, dim3(grid), dim3(kNearHThreads), lds, st, A,
                         L.dntiles, W);
  }
  return hipGetLastError();
}

hipError_t launch_aca_sur(const AcaSurLaunch &L, hipStream_t st) {
  if (L.ntiles <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((L.ntiles + 3) / 4, 16 * (int64_t)L.cus);
  hipLaunchKernelGGL(aca_sur_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                     L.xp, (const AcaPassTile *)L.tiles, L.dntiles, L.row_pass, L.Y,
                     L.params, L.inv_scale);
  return hipGetLastError();
}

hipError_t launch_judge(const JudgeLaunch &L, hipStream_t st) {
  if (L.nblk <= 0) return hipSuccess;
  const int64_t k = L.nblk * L.samples;
  hipLaunchKernelGGL(judge_pairs_kernel, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, st,
                     L.blk, L.nblk, L.xp, L.samples, L.r);
  hipError_t e = launch_pair(PREC_FP32, L.r, k, L.m9, L.wf, L.wh, L.cus, st);
  if (e != hipSuccess) return e;
  const int64_t sblocks = std::min<int64_t>((k + 63) / 64, 16 * (int64_t)L.cus);
  hipLaunchKernelGGL(sur_pairs_kernel, dim3((unsigned)sblocks), dim3(256), 0, st, (const float *)L.r,
                     k, L.s6, L.params, 1.0f / L.r_hi);
  hipLaunchKernelGGL(judge_kernel, dim3((unsigned)((L.nblk + 3) / 4)), dim3(256), 0, st, L.r,
                     (const float *)L.m9, (const float *)L.s6, L.nblk, L.samples, L.r_lo, L.r_hi,
                     L.tol, L.flag, L.err);
  return hipGetLastError();
}

hipError_t launch_sur_samples(int64_t n, float r_lo, float r_hi, uint64_t seed, float *r,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(sur_samples_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     r_lo, r_hi, seed, r);
  return hipGetLastError();
}

hipError_t launch_sur_targets(const float *r, const float *m9, int64_t n, float *t,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(sur_targets_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, r,
                     m9, n, t);
  return hipGetLastError();
}

// Pool storage: SC for the C factors, SQ for the Q factors; P is the
// arithmetic type of the steps.
template <class P, class SC, class SQ>
static void step_launch(const AcaStepArgs &A, const AcaStepLaunch &L, int which, dim3 gc, dim3 gr,
                        dim3 ga, hipStream_t st) {
  if (which == 0) {
    hipLaunchKernelGGL((aca_col_resid_kernel<P, SC, SQ>), gc, dim3(64), resid_lds_bytes<P>(L.max_iter), st,
                       A, (SC *)L.C, (const SQ *)L.Q);
    hipLaunchKernelGGL(aca_col_pivot_kernel<SC>, ga, dim3(64), 0, st, A, (const SC *)L.C);
    hipLaunchKernelGGL((aca_col_norm_kernel<P, SC>), gc, dim3(64), 0, st, A, (SC *)L.C);
  } else {
    hipLaunchKernelGGL((aca_row_resid_kernel<P, SC, SQ>), gr, dim3(64), resid_lds_bytes<P>(L.max_iter), st,
                       A, (const SC *)L.C, (SQ *)L.Q);
  }
}

static unsigned capped(int64_t n, int64_t cap) { return (unsigned)std::max<int64_t>(1, std::min(n, cap)); }

hipError_t launch_aca_steps(const AcaStepLaunch &L, int which, hipStream_t st) {
  if (L.nactive <= 0) return hipSuccess;
  AcaStepArgs A{L.blk, L.active, L.ybase_c, L.ybase_r, L.Y, L.it, L.ipiv, L.jpiv, L.mu2, L.nu2,
                L.stop, L.max_iter, L.eps2, (int)L.nactive, L.cfirst_c, L.cfirst_r, (AcaPart *)L.part_c,
                (AcaPart *)L.part_r, L.x_c, L.x_r, L.inv, L.chunk, L.cnt};
  const dim3 gc(capped(L.nchunks_c, L.grid_cap)), gr(capped(L.nchunks_r, L.grid_cap)),
      ga(capped(L.nactive, L.grid_cap));
  typedef __bf16 B;
  if (L.pool_f64)
    step_launch<double, double, double>(A, L, which, gc, gr, ga, st);
  else if (L.pool_c16 && L.pool_q16)
    step_launch<float, B, B>(A, L, which, gc, gr, ga, st);
  else if (L.pool_c16)
    step_launch<float, B, float>(A, L, which, gc, gr, ga, st);
  else if (L.pool_q16)
    step_launch<float, float, B>(A, L, which, gc, gr, ga, st);
  else
    step_launch<float, float, float>(A, L, which, gc, gr, ga, st);
  if (which == 1) {
    const size_t pr_bytes = sizeof(double) * 9 * (size_t)std::max(1, L.max_iter - 2);
    hipLaunchKernelGGL(aca_finalize_kernel, ga, dim3(256), pr_bytes, st, A);
  }
  return hipGetLastError();
}

template <class SC, class SQ>
static void apply_launch(const AcaApplyArgs &A, const AcaApplyLaunch &L, hipStream_t st) {
  const SC *C = (const SC *)L.C;
  const SQ *Q = (const SQ *)L.Q;
  const dim3 gi(capped(L.nchunks_i, L.grid_cap)), gj(capped(L.nchunks_j, L.grid_cap));
  hipLaunchKernelGGL((aca_apply_reduce_kernel<SC, SQ, 0>), gj, dim3(64), 0, st, A, C, Q);
  if (L.use_symmetry)
    hipLaunchKernelGGL((aca_apply_reduce_kernel<SC, SQ, 1>), gi, dim3(64), 0, st, A, C, Q);
  hipLaunchKernelGGL((aca_apply_update_kernel<SC, SQ, 0>), gi, dim3(64), 0, st, A, C, Q);
  if (L.use_symmetry)
    hipLaunchKernelGGL((aca_apply_update_kernel<SC, SQ, 1>), gj, dim3(64), 0, st, A, C, Q);
}

hipError_t launch_aca_apply(const AcaApplyLaunch &L, hipStream_t st) {
  if (L.nlist <= 0) return hipSuccess;
  AcaApplyArgs A{L.blk, L.list, L.it, L.fp, L.u, (int)L.nlist, L.first_i, L.first_j, L.w_j, L.w_i,
                 L.max_iter, L.chunk, L.cnt};
  typedef __bf16 B;
  if (L.pool_f64)
    apply_launch<double, double>(A, L, st);
  else if (L.pool_c16 && L.pool_q16)
    apply_launch<B, B>(A, L, st);
  else if (L.pool_c16)
    apply_launch<B, float>(A, L, st);
  else if (L.pool_q16)
    apply_launch<float, B>(A, L, st);
  else
    apply_launch<float, float>(A, L, st);
  return hipGetLastError();
}

hipError_t launch_pair(int prec, const float *r, int64_t P, float *out, DevWeightsF wf,
                       DevWeightsH wh, int cus, hipStream_t st) {
  if (P <= 0) return hipSuccess;
  const int64_t waves = (P + 63) / 64;
  if (prec == PREC_FP32) {
    WF<float> W{(const float *)wf.w1b, (const float *)wf.w2f, (const float *)wf.b2,
                (const float *)wf.w3f, (const float *)wf.b3};
    hipLaunchKernelGGL(pair_f32_kernel, dim3(grid_for(waves, 4, 8 * cus)), dim3(256), 0, st, r, P,
                       out, W);
  } else if (prec == PREC_FP64) {
    WF<double> W{(const double *)wf.w1b, (const double *)wf.w2f, (const double *)wf.b2,
                 (const double *)wf.w3f, (const double *)wf.b3};
    hipLaunchKernelGGL(pair_f64_kernel, dim3(grid_for(waves, 4, 8 * cus)), dim3(256), 0, st, r, P,
                       out, W);
  } else {
    allow_big_lds();
    WH W{wh.w1b, wh.b2, wh.b3, wh.w2f, wh.w3f};
    const size_t lds = kW2FragBytes + kW3FragBytes;
    if (prec == PREC_BF16)
      hipLaunchKernelGGL(pair_h_kernel<bf16x8>, dim3(grid_for(waves, 4, cus)), dim3(256), lds, st,
                         r, P, out, W);
    else
      hipLaunchKernelGGL(pair_h_kernel<f16x8>, dim3(grid_for(waves, 4, cus)), dim3(256), lds, st,
                         r, P, out, W);
  }
  return hipGetLastError();
}

hipError_t launch_persistent_aca(const PAcaLaunch &L, hipStream_t st) {
  if (L.nblocks <= 0) return hipSuccess;
  PAcaArgs A{L.xp, L.fp, L.u, L.blk, L.order, L.rank_out, L.nblocks, L.max_iter, L.eps2,
             L.use_symmetry, L.poolC, L.poolQ, L.slotC, L.slotQ, L.Y, L.slotY};
  if (L.prec == PREC_FP32) {
    WF<float> W{(const float *)L.wf.w1b, (const float *)L.wf.w2f, (const float *)L.wf.b2,
                (const float *)L.wf.w3f, (const float *)L.wf.b3};
    hipLaunchKernelGGL(persistent_aca_fp32_kernel, dim3(L.nslots), dim3(256), 0, st, A, W);
  } else {
    static bool attr = false;
    const int lds = kW2FragBytes + kW3FragBytes;
    if (!attr) {
      (void)hipFuncSetAttribute((const void *)persistent_aca_h_kernel<bf16x8>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      (void)hipFuncSetAttribute((const void *)persistent_aca_h_kernel<f16x8>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      attr = true;
    }
    WH W{L.wh.w1b, L.wh.b2, L.wh.b3, L.wh.w2f, L.wh.w3f};
    if (L.prec == PREC_BF16)
      hipLaunchKernelGGL(persistent_aca_h_kernel<bf16x8>, dim3(L.nslots), dim3(256), lds, st, A, W);
    else
      hipLaunchKernelGGL(persistent_aca_h_kernel<f16x8>, dim3(L.nslots), dim3(256), lds, st, A, W);
  }
  return hipGetLastError();
}

__global__ void copy16_kernel(uint4 *dst, const uint4 *src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy16(void *dst, const void *src, int64_t n16, hipStream_t st) {
  if (n16 <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, 65536);
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint4 *)dst,
                     (const uint4 *)src, n16);
  return hipGetLastError();
}

__global__ void zero_i32_kernel(int32_t *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void zero_f64_kernel(double *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_gather_forces(const float *f, int64_t n, const int32_t *perm, float4 *fp,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_forces_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f,
                     n, perm, fp);
  return hipGetLastError();
}

hipError_t launch_scatter_velocity(const double *u, int64_t n, const int32_t *perm, float *out,
                                   hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_velocity_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     u, n, perm, out);
  return hipGetLastError();
}

}  // namespace hh
