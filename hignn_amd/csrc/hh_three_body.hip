// Three-body and self corrections of the H-HIGNN velocity on the device.
//
// The reference evaluates them with torch over edge lists
// (python/HIGNN/model_structure.py:74-140, called from
// python/gravity_field.py:34-59 with the lists of
// include/NeighborLists.hpp:512-588):
//   three-body edge (i, n, p), n a neighbour of i, p a neighbour of n, p != i:
//     v_p += net3(x_n - x_i, x_p - x_n) / |x_n - x_i| / |x_p - x_n| . f_i
//   self edge (i, n), n a neighbour of i:
//     v_n += net_self(x_i - x_n) / |x_i - x_n|^2 . f_n
// with net = Linear(d,256) Tanh Linear(256,1024) Tanh Linear(1024,512) Tanh
// Linear(512,9), the 9 outputs read as a row-major 3x3 block.
//
// Edges are processed in batches: layer 1 on the VALU (d = 6 or 3 inputs),
// layers 2 and 3 as tiled MFMA GEMMs with bias and tanh fused into the
// epilogue (activations stay in HBM between layers: 1.8 KB per edge in 16 bit
// against 1.6 MFLOP), layer 4 and the 3x3 product on the VALU with FP64
// atomic accumulation of the velocities.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "hh_three_body.h"

namespace hh {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

// tanh from weights pre-scaled by 2 log2(e): tanh(x) = 1 - 2 / (2^(2 x log2 e) + 1)
__device__ __forceinline__ float act(float xs) {
  return __builtin_fmaf(-2.0f, __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(xs) + 1.0f), 1.0f);
}

template <class T>
__device__ __forceinline__ T to_t(float v);
template <>
__device__ __forceinline__ float to_t<float>(float v) {
  return v;
}
template <>
__device__ __forceinline__ _Float16 to_t<_Float16>(float v) {
  return (_Float16)v;
}

// ---------------------------------------------------------------------------
// Triplet lists (NeighborLists::BuildThreeBodyInfo, NeighborLists.hpp:512-588):
// per target i, for each neighbour n (ascending), each neighbour p of n
// (ascending) other than i.
// ---------------------------------------------------------------------------
__global__ void triplet_count_kernel(const uint32_t *off, const int32_t *idx, int64_t n,
                                     uint32_t *cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = 0;
  for (uint32_t a = off[i]; a < off[i + 1]; a++) {
    const int32_t nb = idx[a];
    for (uint32_t b = off[nb]; b < off[nb + 1]; b++) c += idx[b] != (int32_t)i;
  }
  cnt[i] = c;
}

__global__ void triplet_fill_kernel(const uint32_t *off, const int32_t *idx, int64_t n,
                                    const uint32_t *toff, int32_t *trip) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t o = toff[i];
  for (uint32_t a = off[i]; a < off[i + 1]; a++) {
    const int32_t nb = idx[a];
    for (uint32_t b = off[nb]; b < off[nb + 1]; b++) {
      const int32_t p = idx[b];
      if (p == (int32_t)i) continue;
      trip[3 * o] = (int32_t)i;
      trip[3 * o + 1] = nb;
      trip[3 * o + 2] = p;
      o++;
    }
  }
}

// ---------------------------------------------------------------------------
// Layer 1 (VALU) and the per-edge record of the epilogue.
// Block: 256 threads = the 256 outputs; 32 edges per block.
// ---------------------------------------------------------------------------
constexpr int kL1Edges = 32;

template <int DIN, class T>
__global__ __launch_bounds__(kThreads) void layer1_kernel(EdgeBatch B, const float *w1s,
                                                          const float *b1s, T *h1) {
  __shared__ float xin[kL1Edges][DIN];
  const int t = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * kL1Edges;
  if (t < kL1Edges) {
    const int64_t e = e0 + t;
    float r1 = 1.0f, r2 = 1.0f, f[3] = {0, 0, 0};
    int32_t tgt = -1;
    float d[6] = {0, 0, 0, 0, 0, 0};
    if (e < B.count) {
      const int64_t g = B.first + e;
      if (DIN == 6) {
        const int32_t i = B.edges[3 * g], nb = B.edges[3 * g + 1], p = B.edges[3 * g + 2];
        for (int j = 0; j < 3; j++) {
          d[j] = B.x[3 * nb + j] - B.x[3 * i + j];
          d[3 + j] = B.x[3 * p + j] - B.x[3 * nb + j];
          f[j] = B.f[3 * i + j];
        }
        r1 = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        r2 = sqrtf(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
        tgt = p;
      } else {
        const int32_t i = B.edges[2 * g], nb = B.edges[2 * g + 1];
        for (int j = 0; j < 3; j++) {
          d[j] = B.x[3 * i + j] - B.x[3 * nb + j];
          f[j] = B.f[3 * nb + j];
        }
        r1 = r2 = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        tgt = nb;
      }
      B.rec[e] = make_float4(r1, r2, 0.0f, 0.0f);
      B.frc[e] = make_float4(f[0], f[1], f[2], __int_as_float(tgt));
    }
    for (int j = 0; j < DIN; j++) xin[t][j] = d[j];
  }
  __syncthreads();
  float w[DIN];
  for (int j = 0; j < DIN; j++) w[j] = w1s[t * DIN + j];
  const float b = b1s[t];
  for (int k = 0; k < kL1Edges; k++) {
    float a = b;
    for (int j = 0; j < DIN; j++) a = __builtin_fmaf(w[j], xin[k][j], a);
    h1[(e0 + k) * 256 + t] = to_t<T>(act(a));
  }
}

// ---------------------------------------------------------------------------
// C[e][o] = tanh(sum_k A[e][k] W[o][k] + b[o]) for a batch of edges:
// large tiles (TM x TN, Tiles below) in waves of WM x WN.  K is staged
// through two LDS buffers: while the waves multiply one K step, the next is
// in flight from global memory into registers and then stored to the other
// buffer, one barrier per step.  MFMA 16x16x32 f16 (FP16 path, K steps of 64)
// or 16x16x4 f32 (FP32 path, K steps of 32); lane l holds
// D[4 (l/16) + r][l % 16] of a 16 x 16 tile.
// Requires E a multiple of TM (padded batch), O a multiple of TN and K a
// multiple of the K step.
// ---------------------------------------------------------------------------
template <class T>
struct Gemm;

template <>
struct Gemm<_Float16> {
  static constexpr int kKS = 64;   // K per stage
  static constexpr int kPad = 8;   // row stride 72 halves = 144 B
  static constexpr int kSub = 2;   // MFMAs (K = 32) per stage
  typedef f16x8 Frag;
  __device__ static Frag frag(const _Float16 *row, int sub, int lane) {
    return *(const f16x8 *)(row + 32 * sub + 8 * (lane >> 4));
  }
  __device__ static f32x4 mfma(Frag a, Frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Gemm<float> {
  static constexpr int kKS = 32;
  static constexpr int kPad = 4;   // row stride 36 floats
  static constexpr int kSub = 8;   // MFMAs (K = 4) per stage
  typedef float Frag;
  __device__ static Frag frag(const float *row, int sub, int lane) { return row[4 * sub + (lane >> 4)]; }
  __device__ static f32x4 mfma(Frag a, Frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Workgroup tile TM x TN (edges x outputs) in waves of WM x WN.
template <class T, int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(64 * (TM / WM) * (TN / WN), 256 / ((TM / WM) * (TN / WN) * 64) + 1)
void gemm_act_kernel(
    const T *A, const T *W, const float *bias, T *C, int K, int O, int ntiles) {
  typedef Gemm<T> G;
  constexpr int NTH = 64 * (TM / WM) * (TN / WN);   // threads
  constexpr int MT = WM / 16, NT = WN / 16;         // 16 x 16 tiles per wave
  constexpr int KS = G::kKS;
  constexpr int RS = KS + G::kPad;                  // LDS row stride (elements)
  constexpr int CPR = KS * (int)sizeof(T) / 16;     // 16-byte chunks per row and stage
  constexpr int PA = TM * CPR / NTH, PW = TN * CPR / NTH;  // chunks per thread
  constexpr int EPC = 16 / (int)sizeof(T);          // elements per chunk
  constexpr int SA = TM * RS, SW = TN * RS;         // elements of one stage
  static_assert(PA * NTH == TM * CPR && PW * NTH == TN * CPR, "tile / thread mismatch");
  __shared__ __attribute__((aligned(16))) T lds[2 * (SA + SW)];  // [stage][A | W]

  // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin, so
  // launch index L runs on XCD L % 8; XCD k takes the k-th contiguous run of
  // row-major tiles, and the column tiles of a row block share its L2.
  const int per_xcd = (ntiles + 7) / 8;
  const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const int nx = O / TN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t e0 = (int64_t)(tile / nx) * TM;
  const int o0 = (tile % nx) * TN;
  const int wr = (wave / (TN / WN)) * WM, wc = (wave % (TN / WN)) * WN;
  const T *ga = A + e0 * K, *gw = W + (int64_t)o0 * K;

  i32x4 ra[PA], rw[PW];
  int gao[PA], sao[PA], gwo[PW], swo[PW];
#pragma unroll
  for (int j = 0; j < PA; j++) {
    const int c = t + NTH * j, row = c / CPR, q = c % CPR;
    gao[j] = row * K + q * EPC;
    sao[j] = row * RS + q * EPC;
  }
#pragma unroll
  for (int j = 0; j < PW; j++) {
    const int c = t + NTH * j, row = c / CPR, q = c % CPR;
    gwo[j] = row * K + q * EPC;
    swo[j] = row * RS + q * EPC;
  }

  f32x4 acc[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < NT; b++) acc[a][b] = f32x4{0, 0, 0, 0};

#pragma unroll
  for (int j = 0; j < PA; j++) ra[j] = *(const i32x4 *)(ga + gao[j]);
#pragma unroll
  for (int j = 0; j < PW; j++) rw[j] = *(const i32x4 *)(gw + gwo[j]);
#pragma unroll
  for (int j = 0; j < PA; j++) *(i32x4 *)(lds + sao[j]) = ra[j];
#pragma unroll
  for (int j = 0; j < PW; j++) *(i32x4 *)(lds + SA + swo[j]) = rw[j];
  __syncthreads();

  const int nsteps = K / KS;
  for (int step = 0; step < nsteps; step++) {
    const int cur = step & 1;
    // next stage (the last step reloads itself into the idle buffer)
    const int k1 = min(step + 1, nsteps - 1) * KS;
#pragma unroll
    for (int j = 0; j < PA; j++) ra[j] = *(const i32x4 *)(ga + gao[j] + k1);
#pragma unroll
    for (int j = 0; j < PW; j++) rw[j] = *(const i32x4 *)(gw + gwo[j] + k1);
    const T *sa = lds + cur * (SA + SW), *sw = sa + SA;
#pragma unroll
    for (int s = 0; s < G::kSub; s++) {
      typename G::Frag fa[MT], fb[NT];
#pragma unroll
      for (int m = 0; m < MT; m++) fa[m] = G::frag(sa + (wr + 16 * m + (lane & 15)) * RS, s, lane);
#pragma unroll
      for (int m = 0; m < NT; m++) fb[m] = G::frag(sw + (wc + 16 * m + (lane & 15)) * RS, s, lane);
#pragma unroll
      for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) acc[a][b] = G::mfma(fa[a], fb[b], acc[a][b]);
    }
    T *da = lds + (cur ^ 1) * (SA + SW);
#pragma unroll
    for (int j = 0; j < PA; j++) *(i32x4 *)(da + sao[j]) = ra[j];
#pragma unroll
    for (int j = 0; j < PW; j++) *(i32x4 *)(da + SA + swo[j]) = rw[j];
    __syncthreads();
  }

  // epilogue: bias + tanh into the (now idle) LDS as a row-major tile, then
  // 16-byte coalesced stores of whole rows
  constexpr int TS = TN + EPC;  // tile row stride (elements)
  static_assert(TM * TS <= 2 * (SA + SW), "output tile does not fit the stage buffers");
#pragma unroll
  for (int b = 0; b < NT; b++) {
    const int col = wc + 16 * b + (lane & 15);
    const float bo = bias[o0 + col];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        lds[(wr + 16 * a + 4 * (lane >> 4) + r) * TS + col] = to_t<T>(act(acc[a][b][r] + bo));
  }
  __syncthreads();
  constexpr int OPR = TN / EPC;  // 16-byte chunks per output row
#pragma unroll 4
  for (int c = t; c < TM * OPR; c += NTH) {
    const int row = c / OPR, q = c % OPR;
    *(i32x4 *)(C + (e0 + row) * O + o0 + q * EPC) = *(const i32x4 *)(lds + row * TS + q * EPC);
  }
}

// Tiles of the two layer GEMMs.  A wave tile of WM x WN does WM WN / (WM + WN)
// FLOP per LDS byte; 128 x 64 (43) keeps the MFMA units fed within the CU's
// LDS bandwidth, where 64 x 64 (32) sits right at it.  The batch rows are
// padded to a multiple of 256.
template <class T>
struct Tiles {  // FP16: 256 x 256 in 8 waves of 128 x 64, one workgroup per CU
  static constexpr int TM = 256, TN = 256, WM = 128, WN = 64;
  static constexpr int threads = 64 * (TM / WM) * (TN / WN);
};
template <>
struct Tiles<float> {  // FP32: 128 x 128 in 4 waves of 64 x 64
  static constexpr int TM = 128, TN = 128, WM = 64, WN = 64;
  static constexpr int threads = 64 * (TM / WM) * (TN / WN);
};

// ---------------------------------------------------------------------------
// Layer 4 (512 -> 9) on MFMA, then the 3x3 block over the radii times the
// force: each wave takes 16 edges, y (16 x 16, outputs 9..15 zero) =
// h3 (16 x 512) W4^T with W4^T in LDS; y goes through LDS so that 48 lanes
// form (edge, row a) of M f and add it to the target's FP64 velocity.
// Blocks stride over the batch so W4 is staged once per block.  Rows are read
// up to ceil(count / 64) * 64 <= the padded batch.
// ---------------------------------------------------------------------------
constexpr int kL4Blocks = 2048;

template <class T>
__global__ __launch_bounds__(kThreads) void layer4_kernel(EdgeBatch B, const T *h3,
                                                          const float *w4, const float *b4,
                                                          double *u) {
  typedef Gemm<T> G;
  constexpr int RS = 512 + G::kPad;
  constexpr int KM = G::kKS / G::kSub;  // K per MFMA
  __shared__ __attribute__((aligned(16))) T sw[16 * RS];
  __shared__ float ys[4][16][17];
  for (int k = threadIdx.x; k < 16 * 512; k += kThreads) {
    const int o = k / 512, c = k % 512;
    sw[o * RS + c] = to_t<T>(o < 9 ? w4[o * 512 + c] : 0.0f);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const T *wrow = sw + (lane & 15) * RS;
  // block-uniform trip count (the loop holds barriers); a wave whose 16 edges
  // start past the batch multiplies padded rows (ceil(count/256)*256 exist)
  // and skips the epilogue
  for (int64_t b0 = (int64_t)blockIdx.x * 64; b0 < B.count; b0 += (int64_t)gridDim.x * 64) {
    const int64_t e0 = b0 + 16 * wave;
    const T *arow = h3 + (e0 + (lane & 15)) * 512;
    f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll 4
    for (int s = 0; s < 512 / KM; s++) acc = G::mfma(G::frag(arow, s, lane), G::frag(wrow, s, lane), acc);
    // lane (p, g) holds y[edge 4g + r][output p]
    for (int r = 0; r < 4; r++) ys[wave][4 * (lane >> 4) + r][lane & 15] = acc[r];
    __syncthreads();  // uniform: every wave runs the same number of iterations per block
    if (lane < 48) {
      const int q = lane / 3, a = lane % 3;
      const int64_t e = e0 + q;
      if (e < B.count) {
        const float4 rr = B.rec[e];
        const float4 fr = B.frc[e];
        const float f[3] = {fr.x, fr.y, fr.z};
        float v = 0.0f;
        for (int c = 0; c < 3; c++) {
          const float m = ((ys[wave][q][3 * a + c] + b4[3 * a + c]) / rr.x) / rr.y;
          v = __builtin_fmaf(m, f[c], v);
        }
        atomicAdd(&u[3 * (int64_t)__float_as_int(fr.w) + a], (double)v);
      }
    }
    __syncthreads();
  }
}

__global__ void to_float_kernel(const double *u, float *out, int64_t n3, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  out[i] = accumulate ? (float)((double)out[i] + u[i]) : (float)u[i];
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

}  // namespace

hipError_t launch_triplet_count(const uint32_t *off, const int32_t *idx, int64_t n, uint32_t *cnt,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(triplet_count_kernel, grid_for(n), dim3(kThreads), 0, st, off, idx, n, cnt);
  return hipGetLastError();
}

hipError_t launch_triplet_fill(const uint32_t *off, const int32_t *idx, int64_t n,
                               const uint32_t *toff, int32_t *trip, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(triplet_fill_kernel, grid_for(n), dim3(kThreads), 0, st, off, idx, n, toff,
                     trip);
  return hipGetLastError();
}

hipError_t launch_edge_mlp(const EdgeBatch &B, const NetDev &net, int din, int half, void *h1,
                           void *h2, void *h3, double *u, hipStream_t st) {
  if (B.count <= 0) return hipSuccess;
  const dim3 g4((unsigned)std::min<int64_t>((B.count + 63) / 64, kL4Blocks));
  auto grids = [&](auto tiles, int &t2, int &t3, dim3 &g2, dim3 &g3) {
    typedef decltype(tiles) TL;
    const int64_t epad = (B.count + TL::TM - 1) / TL::TM * TL::TM;
    t2 = (int)(1024 / TL::TN * (epad / TL::TM));
    t3 = (int)(512 / TL::TN * (epad / TL::TM));
    g2 = dim3((unsigned)((t2 + 7) / 8 * 8));
    g3 = dim3((unsigned)((t3 + 7) / 8 * 8));
  };
  const dim3 g1((unsigned)((B.count + kL1Edges - 1) / kL1Edges));
  int t2, t3;
  dim3 g2, g3;
  if (half) {
    typedef _Float16 H;
    typedef Tiles<H> TL;
    grids(TL{}, t2, t3, g2, g3);
    if (din == 6)
      hipLaunchKernelGGL((layer1_kernel<6, H>), g1, dim3(kThreads), 0, st, B, net.w1s, net.b1s,
                         (H *)h1);
    else
      hipLaunchKernelGGL((layer1_kernel<3, H>), g1, dim3(kThreads), 0, st, B, net.w1s, net.b1s,
                         (H *)h1);
    hipLaunchKernelGGL((gemm_act_kernel<H, TL::TM, TL::TN, TL::WM, TL::WN>), g2, dim3(TL::threads), 0, st,
                       (const H *)h1, (const H *)net.w2h, net.b2s, (H *)h2, 256, 1024, t2);
    hipLaunchKernelGGL((gemm_act_kernel<H, TL::TM, TL::TN, TL::WM, TL::WN>), g3, dim3(TL::threads), 0, st,
                       (const H *)h2, (const H *)net.w3h, net.b3s, (H *)h3, 1024, 512, t3);
    hipLaunchKernelGGL(layer4_kernel<H>, g4, dim3(kThreads), 0, st, B, (const H *)h3, net.w4,
                       net.b4, u);
  } else {
    typedef Tiles<float> TL;
    grids(TL{}, t2, t3, g2, g3);
    if (din == 6)
      hipLaunchKernelGGL((layer1_kernel<6, float>), g1, dim3(kThreads), 0, st, B, net.w1s,
                         net.b1s, (float *)h1);
    else
      hipLaunchKernelGGL((layer1_kernel<3, float>), g1, dim3(kThreads), 0, st, B, net.w1s,
                         net.b1s, (float *)h1);
    hipLaunchKernelGGL((gemm_act_kernel<float, TL::TM, TL::TN, TL::WM, TL::WN>), g2, dim3(TL::threads), 0,
                       st, (const float *)h1, net.w2s, net.b2s, (float *)h2, 256, 1024, t2);
    hipLaunchKernelGGL((gemm_act_kernel<float, TL::TM, TL::TN, TL::WM, TL::WN>), g3, dim3(TL::threads), 0,
                       st, (const float *)h2, net.w3s, net.b3s, (float *)h3, 1024, 512, t3);
    hipLaunchKernelGGL(layer4_kernel<float>, g4, dim3(kThreads), 0, st, B, (const float *)h3,
                       net.w4, net.b4, u);
  }
  return hipGetLastError();
}

hipError_t launch_to_float(const double *u, float *out, int64_t n3, int accumulate,
                           hipStream_t st) {
  if (n3 <= 0) return hipSuccess;
  hipLaunchKernelGGL(to_float_kernel, grid_for(n3), dim3(kThreads), 0, st, u, out, n3, accumulate);
  return hipGetLastError();
}

}  // namespace hh
