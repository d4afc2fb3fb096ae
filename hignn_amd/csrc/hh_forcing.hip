// Device forcing and time-integration kernels around the matvec.
//
// Forcing (per particle, gather form: each particle sums its own terms, so no
// atomics touch the result and the sum order is fixed):
//   f = b                                     body force (python/simulate.py:131-132)
//     + k_b * bending(chain)                  python/simulate.py:14-53
//     + tension(chain)                        python/simulate.py:55-88
//     + sum over Morse neighbours             include/PotentialForce.hpp:115-140, 468-580
// Bending and tension terms are evaluated in float in the numpy order and
// summed in double, as numpy adds float32 terms into a float64 array; each
// group is rounded to float when it is added to f (force += ..., float32).
//
// Morse neighbours are the sources within distance epsilon of the target
// (radius search on the kd-tree, include/PointCloudSearch.hpp:386-400:
// squared distance in double < (float)(eps*eps)), found here through a hashed
// uniform cell list of cell size slightly above epsilon.  With a periodic
// domain the sources are the reference's duplicated images
// (PotentialForce.hpp:247-392) and the displacement of a neighbour is taken
// against the source array entry at the neighbour's *target* index, exactly
// as GetPotentialForce does after mapping neighbours to targets
// (PotentialForce.hpp:424-446, 548).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hh_forcing.h"

namespace hh {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t mask) {
  return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) &
         mask;
}

// Cell coordinate, clamped so the conversion to int stays defined for any
// input (coordinates beyond 2^30 cells share the outermost cell).
__device__ __forceinline__ int cell_of(float x, double inv_h) {
  const double c = floor((double)x * inv_h);
  return (int)fmin(fmax(c, -1073741824.0), 1073741824.0);
}

// ---------------------------------------------------------------------------
// Periodic source images (PotentialForce.hpp:247-392).  Axis j of target i is
// duplicated when x_j < lo_j + margin (image shifted by +L_j) or
// x_j > hi_j - margin (shifted by -L_j); the margin is the two-body radius.  Image n of a target with `num` images carries the shift
// of its k-th duplicated axis when bit (num >> (k+1)) of n is set: the
// reference fills halves of blocks of size num, num/2, ... in axis order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int image_axes(const float *x, const ForcingLaunch &L, int dup[3]) {
  int num = 1;
  for (int j = 0; j < 3; j++) {
    const float clo = L.lo[j] + L.margin, chi = L.hi[j] - L.margin;
    dup[j] = x[j] < clo ? -2 : (x[j] > chi ? 2 : 1);
    if (dup[j] != 1) num *= 2;
  }
  return num;
}

__global__ void image_count_kernel(ForcingLaunch L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  int dup[3];
  L.cnt[i] = (uint32_t)image_axes(L.x + 3 * i, L, dup);
}

__global__ void image_fill_kernel(ForcingLaunch L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  const float *x = L.x + 3 * i;
  int dup[3];
  const int num = image_axes(x, L, dup);
  const uint32_t base = L.off[i];
  for (int n = 0; n < num; n++) {
    float p[3];
    int k = 0;
    for (int j = 0; j < 3; j++) {
      float o = 0.0f;
      if (dup[j] != 1) {
        if (n & (num >> (k + 1))) o = dup[j] == 2 ? -(L.hi[j] - L.lo[j]) : (L.hi[j] - L.lo[j]);
        k++;
      }
      p[j] = x[j] + o;
    }
    L.src[base + n] = make_float4(p[0], p[1], p[2], 0.0f);
    L.src_tgt[base + n] = (int32_t)i;
  }
}

__global__ void plain_sources_kernel(ForcingLaunch L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *L.nsrc = (uint32_t)L.n;
  if (i >= L.n) return;
  L.src[i] = make_float4(L.x[3 * i], L.x[3 * i + 1], L.x[3 * i + 2], 0.0f);
  L.src_tgt[i] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// Hashed cell list: count per bucket, exclusive scan, scatter, then a sort of
// each bucket by source index so the visiting order does not depend on the
// order of the scatter's atomics.
// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void cell_count_kernel(ForcingLaunch L) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (int64_t)*L.nsrc) return;
  const float4 p = L.src[m];
  const int cx = cell_of(p.x, L.inv_h), cy = cell_of(p.y, L.inv_h), cz = cell_of(p.z, L.inv_h);
  L.src_cell[m] = make_int4(cx, cy, cz, L.src_tgt[m]);
  atomicAdd(&L.bcnt[cell_hash(cx, cy, cz, L.tmask)], 1u);
}

__global__ void cell_scatter_kernel(ForcingLaunch L) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (int64_t)*L.nsrc) return;
  const int4 c = L.src_cell[m];
  const uint32_t b = cell_hash(c.x, c.y, c.z, L.tmask);
  const uint32_t slot = atomicAdd(&L.bcnt[b], 1u);  // bcnt was cleared again before this pass
  L.sidx[L.bstart[b] + slot] = (int32_t)m;
}

__global__ void bucket_sort_kernel(ForcingLaunch L) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)L.tsize) return;
  const uint32_t s = L.bstart[b], e = L.bstart[b + 1];
  for (uint32_t k = s + 1; k < e; k++) {  // insertion sort by source index
    const int32_t v = L.sidx[k];
    uint32_t q = k;
    while (q > s && L.sidx[q - 1] > v) {
      L.sidx[q] = L.sidx[q - 1];
      q--;
    }
    L.sidx[q] = v;
  }
  for (uint32_t k = s; k < e; k++) {
    const int32_t m = L.sidx[k];
    L.spos[k] = L.src[m];
    L.scell[k] = L.src_cell[m];
  }
}

// Exclusive scan of n counts into out[0..n] (out[n] = total): block-local
// scans of 2048 items, one block scanning the block totals, then the offsets.
constexpr int kScanItems = 8;
constexpr int kScanBlock = kThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t &total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < kThreads; d <<= 1) {
    const uint32_t a = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[kThreads - 1];
  const uint32_t r = sh[t] - v;
  __syncthreads();
  return r;
}

__global__ void scan_blocks_kernel(const uint32_t *in, int64_t n, uint32_t *out, uint32_t *bsum) {
  __shared__ uint32_t sh[kThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems], s = 0;
  for (int k = 0; k < kScanItems; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, sh, total);
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void scan_top_kernel(uint32_t *bsum, int64_t nb, uint32_t *out_total) {
  __shared__ uint32_t sh[kThreads];
  uint32_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += kThreads) {
    const int64_t i = c0 + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0u;
    uint32_t total;
    const uint32_t r = block_exclusive_scan(v, sh, total);
    if (i < nb) bsum[i] = carry + r;
    carry += total;
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ void scan_add_kernel(uint32_t *out, int64_t n, const uint32_t *bsum) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += bsum[i / kScanBlock];
}

}  // namespace

hipError_t exclusive_scan(const uint32_t *in, int64_t n, uint32_t *out, uint32_t *bsum,
                          hipStream_t st) {
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(scan_blocks_kernel, dim3((unsigned)nb), dim3(kThreads), 0, st, in, n, out,
                     bsum);
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kThreads), 0, st, bsum, nb, out + n);
  hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, out, n, bsum);
  return hipGetLastError();
}

namespace {

// ---------------------------------------------------------------------------
// Neighbour lists (NeighborLists::BuildTargetNeighborLists,
// include/NeighborLists.hpp:223-251): for every target, the targets of the
// sources within the radius, the target's own entry dropped once; each list
// sorted ascending.  Two passes: counts, then (after a scan) the lists.
// ---------------------------------------------------------------------------
template <bool FILL>
__device__ __forceinline__ uint32_t visit_neighbours(const ForcingLaunch &L, int64_t i,
                                                     int32_t *out) {
  const float xi[3] = {L.x[3 * i], L.x[3 * i + 1], L.x[3 * i + 2]};
  const int cx = cell_of(xi[0], L.inv_h), cy = cell_of(xi[1], L.inv_h), cz = cell_of(xi[2], L.inv_h);
  bool self_seen = false;
  uint32_t c = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int ux = cx + dx, uy = cy + dy, uz = cz + dz;
        const uint32_t b = cell_hash(ux, uy, uz, L.tmask);
        const uint32_t e = L.bstart[b + 1];
        for (uint32_t k = L.bstart[b]; k < e; k++) {
          const int4 q = L.scell[k];
          if (q.x != ux || q.y != uy || q.z != uz) continue;
          const float4 s = L.spos[k];
          const double ddx = (double)s.x - (double)xi[0], ddy = (double)s.y - (double)xi[1],
                       ddz = (double)s.z - (double)xi[2];
          const double d = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
          if (!(d * d < L.eps2)) continue;
          if (q.w == (int32_t)i && !self_seen) {
            self_seen = true;
            continue;
          }
          if (FILL) out[c] = q.w;
          c++;
        }
      }
  return c;
}

__global__ void neighbour_count_kernel(ForcingLaunch L, uint32_t *cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  cnt[i] = visit_neighbours<false>(L, i, nullptr);
}

__global__ void neighbour_fill_kernel(ForcingLaunch L, const uint32_t *off, int32_t *idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  int32_t *o = idx + off[i];
  const uint32_t c = visit_neighbours<true>(L, i, o);
  for (uint32_t k = 1; k < c; k++) {  // insertion sort, ascending
    const int32_t v = o[k];
    uint32_t q = k;
    while (q > 0 && o[q - 1] > v) {
      o[q] = o[q - 1];
      q--;
    }
    o[q] = v;
  }
}

// ---------------------------------------------------------------------------
// The force kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load3(const float *x, int64_t i, float p[3]) {
  p[0] = x[3 * i];
  p[1] = x[3 * i + 1];
  p[2] = x[3 * i + 2];
}

// chain_tension, bond b = (q, q+1): nodal = (r / |r|) * k_t (|r| - L), r = X_q - X_{q+1}
__device__ __forceinline__ void tension_nodal(const float *a, const float *b, float kt, float rest,
                                              float nodal[3]) {
  float r[3];
  for (int j = 0; j < 3; j++) r[j] = a[j] - b[j];
  const float rn = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  const float fm = kt * (rn - rest);
  for (int j = 0; j < 3; j++) nodal[j] = (r[j] / rn) * fm;
}

__device__ __forceinline__ void cross3(const float *a, const float *b, float *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// chain_bending, triple (t, t+1, t+2): f21 and f23 (before the k_b factor)
__device__ __forceinline__ void bending_pair(const float *x0, const float *x1, const float *x2,
                                             float f21[3], float f23[3]) {
  float a1[3], a2[3];
  for (int j = 0; j < 3; j++) a1[j] = x1[j] - x0[j], a2[j] = x2[j] - x1[j];
  const float n1 = sqrtf(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  const float n2 = sqrtf(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
  for (int j = 0; j < 3; j++) a1[j] = a1[j] / n1, a2[j] = a2[j] / n2;
  const float c = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
  float tmp = 1.0f - c * c;
  if (tmp < 0.0f) tmp = 0.0f;
  const float s = sqrtf(tmp);
  float uv[3], tq[3];
  cross3(a1, a2, uv);
  for (int j = 0; j < 3; j++) tq[j] = -uv[j] * s;
  cross3(tq, a1, f21);
  cross3(tq, a2, f23);
}

// CalculatePotentialForce (PotentialForce.hpp:115-140): De = 1, a = 1, re = 2.5
__device__ __forceinline__ void morse_add(const float *r, double acc[3]) {
  const float rn = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (!((double)rn > 1e-3)) return;
  const float a = 1.0f, re = 2.5f;
  const double e1 = exp((double)(-a * (rn - re)));
  const double e2 = exp(-2.0 * (double)a * (double)(rn - re));
  const double fmag = -2.0 * (double)a * 1.0 * (e1 - e2);
  const float ratio = (float)(fmag / (double)rn);
  for (int j = 0; j < 3; j++) acc[j] += (double)(r[j] * ratio);
}

__global__ void force_kernel(ForcingLaunch L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  float xi[3];
  load3(L.x, i, xi);
  float f[3] = {L.b[0], L.b[1], L.b[2]};

  if (L.n_chain > 1 && i < L.n_fil * L.n_chain) {
    const int64_t q = i % L.n_chain, base = i - q, N = L.n_chain;
    // bending: F[0:N-2] += f21; F[1:N-1] -= f21 + f23; F[2:N] += f23
    double fb[3] = {0, 0, 0};
    float p0[3], p1[3], p2[3], f21[3], f23[3];
    if (q <= N - 3) {
      load3(L.x, base + q + 1, p1);
      load3(L.x, base + q + 2, p2);
      bending_pair(xi, p1, p2, f21, f23);
      for (int j = 0; j < 3; j++) fb[j] += (double)f21[j];
    }
    if (q >= 1 && q <= N - 2) {
      load3(L.x, base + q - 1, p0);
      load3(L.x, base + q + 1, p2);
      bending_pair(p0, xi, p2, f21, f23);
      for (int j = 0; j < 3; j++) fb[j] -= (double)(f21[j] + f23[j]);
    }
    if (q >= 2) {
      load3(L.x, base + q - 2, p0);
      load3(L.x, base + q - 1, p1);
      bending_pair(p0, p1, xi, f21, f23);
      for (int j = 0; j < 3; j++) fb[j] += (double)f23[j];
    }
    for (int j = 0; j < 3; j++) f[j] = (float)((double)f[j] + (double)L.k_b * fb[j]);
    // tension: F[0:N-1] -= nodal; F[1:N] += nodal
    double ft[3] = {0, 0, 0};
    float nodal[3];
    if (q <= N - 2) {
      load3(L.x, base + q + 1, p1);
      tension_nodal(xi, p1, L.k_t, L.rest, nodal);
      for (int j = 0; j < 3; j++) ft[j] -= (double)nodal[j];
    }
    if (q >= 1) {
      load3(L.x, base + q - 1, p0);
      tension_nodal(p0, xi, L.k_t, L.rest, nodal);
      for (int j = 0; j < 3; j++) ft[j] += (double)nodal[j];
    }
    for (int j = 0; j < 3; j++) f[j] = (float)((double)f[j] + ft[j]);
  }

  if (L.bond_chain > 1 && i < L.bond_fil * L.bond_chain) {
    // bonded Morse: the bond (k, k+1) pulls bead k along d = x_{k+1} - x_k
    // with dU/dr and bead k+1 the other way; FP64, added once
    const int64_t q = i % L.bond_chain, base = i - q;
    double acc[3] = {0, 0, 0};
    for (int side = 0; side < 2; side++) {
      if (side == 0 ? q > L.bond_chain - 2 : q < 1) continue;
      float p[3];
      load3(L.x, side == 0 ? base + q + 1 : base + q - 1, p);
      // d = x_{k+1} - x_k of the bond
      double d[3];
      for (int j = 0; j < 3; j++)
        d[j] = side == 0 ? (double)p[j] - (double)xi[j] : (double)xi[j] - (double)p[j];
      const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double e = exp(-L.bond_alpha * (r - L.bond_r0));
      const double g = 2.0 * L.bond_D * L.bond_alpha * e * (1.0 - e) / r;  // dU/dr / r
      const double sg = side == 0 ? g : -g;
      for (int j = 0; j < 3; j++) acc[j] += sg * d[j];
    }
    for (int j = 0; j < 3; j++) f[j] = (float)((double)f[j] + acc[j]);
  }

  if (L.morse) {
    double acc[3] = {0, 0, 0};
    const int cx = cell_of(xi[0], L.inv_h), cy = cell_of(xi[1], L.inv_h),
              cz = cell_of(xi[2], L.inv_h);
    bool self_seen = false;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int ux = cx + dx, uy = cy + dy, uz = cz + dz;
          const uint32_t b = cell_hash(ux, uy, uz, L.tmask);
          const uint32_t e = L.bstart[b + 1];
          for (uint32_t k = L.bstart[b]; k < e; k++) {
            const int4 c = L.scell[k];
            if (c.x != ux || c.y != uy || c.z != uz) continue;  // another cell in this bucket
            const float4 s = L.spos[k];
            const double ddx = (double)s.x - (double)xi[0], ddy = (double)s.y - (double)xi[1],
                         ddz = (double)s.z - (double)xi[2];
            const double d = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            if (!(d * d < L.eps2)) continue;
            if (c.w == (int32_t)i && !self_seen) {  // the target's own entry is skipped once
              self_seen = true;
              continue;
            }
            float r[3];
            if (L.periodic) {
              const float4 sp = L.src[c.w];  // source entry at the neighbour's target index
              r[0] = xi[0] - sp.x, r[1] = xi[1] - sp.y, r[2] = xi[2] - sp.z;
            } else {
              r[0] = xi[0] - s.x, r[1] = xi[1] - s.y, r[2] = xi[2] - s.z;
            }
            morse_add(r, acc);
          }
        }
    for (int j = 0; j < 3; j++) f[j] = (float)((double)f[j] + acc[j]);
  }
  L.f[3 * i] = f[0];
  L.f[3 * i + 1] = f[1];
  L.f[3 * i + 2] = f[2];
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

// ---------------------------------------------------------------------------
// Integrator kernels (TimeIntegrator.hpp; the host loops of the Python
// integrators perform the same float operations in the same order).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wrap_offset(float x, float lo, float hi) {
  return x > hi ? lo - hi : (x < lo ? hi - lo : 0.0f);
}

__global__ void stage_kernel(StageLaunch S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.n) return;
  for (int j = 0; j < 3; j++) {
    float x = S.x0[3 * i + j];
    for (int k = 0; k < S.nterms; k++) x = x + S.coef[S.row * 8 + k] * S.k[k][3 * i + j];
    if (S.xs) S.xs[3 * i + j] = x;
    if (S.periodic) x = x + wrap_offset(x, S.lo[j], S.hi[j]);
    S.xin[3 * i + j] = x;
  }
}

__global__ void euler_update_kernel(float *x, const float *v, float *off, int64_t n,
                                    const float *fdt, int periodic, Lim lim) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float h = *fdt;
  for (int j = 0; j < 3; j++) {
    const float y = x[3 * i + j] + h * v[3 * i + j];
    x[3 * i + j] = y;
    if (periodic) off[3 * i + j] = wrap_offset(y, lim.lo[j], lim.hi[j]);
  }
}

__global__ void add_kernel(float *x, const float *off, int64_t n3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) x[i] = x[i] + off[i];
}

// e2_i = sum_j (dt * (K1 dc1 + K3 dc3 + K4 dc4 + K5 dc5 + K6 dc6 + K7 dc7))^2
// in float, in the order of the host loop; block partial sums in double.
__global__ void dopri_err_kernel(ErrLaunch E) {
  __shared__ double sh[kThreads];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e2d = 0.0;
  if (i < E.n) {
    const float fdt = *E.fdt;
    float e2 = 0.0f;
    for (int j = 0; j < 3; j++) {
      const int64_t o = 3 * i + j;
      float v = E.k[0][o] * E.dc[0] + E.k[2][o] * E.dc[2];
      v = v + E.k[3][o] * E.dc[3];
      v = v + E.k[4][o] * E.dc[4];
      v = v + E.k[5][o] * E.dc[5];
      v = v + E.k[6][o] * E.dc[6];
      v = fdt * v;
      e2 += v * v;
    }
    e2d = (double)e2;
  }
  sh[threadIdx.x] = e2d;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) E.part[blockIdx.x] = sh[0];
}

__global__ void sum_partials_kernel(const double *part, int64_t np, double *out) {
  __shared__ double sh[kThreads];
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < np; k += kThreads) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace

hipError_t launch_cell_list(const ForcingLaunch &L, hipStream_t st) {
  if (L.n <= 0) return hipSuccess;
  if (L.periodic) {
    hipLaunchKernelGGL(image_count_kernel, grid_for(L.n), dim3(kThreads), 0, st, L);
    hipError_t e = exclusive_scan(L.cnt, L.n, L.off, L.bsum, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(image_fill_kernel, grid_for(L.n), dim3(kThreads), 0, st, L);
  } else {
    hipLaunchKernelGGL(plain_sources_kernel, grid_for(L.n), dim3(kThreads), 0, st, L);
  }
  hipLaunchKernelGGL(fill_u32_kernel, grid_for(L.tsize), dim3(kThreads), 0, st, L.bcnt,
                     (int64_t)L.tsize, 0u);
  hipLaunchKernelGGL(cell_count_kernel, grid_for(L.src_cap), dim3(kThreads), 0, st, L);
  hipError_t e = exclusive_scan(L.bcnt, (int64_t)L.tsize, L.bstart, L.bsum, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fill_u32_kernel, grid_for(L.tsize), dim3(kThreads), 0, st, L.bcnt,
                     (int64_t)L.tsize, 0u);
  hipLaunchKernelGGL(cell_scatter_kernel, grid_for(L.src_cap), dim3(kThreads), 0, st, L);
  hipLaunchKernelGGL(bucket_sort_kernel, grid_for(L.tsize), dim3(kThreads), 0, st, L);
  return hipGetLastError();
}

hipError_t launch_forcing(const ForcingLaunch &L, hipStream_t st) {
  if (L.n <= 0) return hipSuccess;
  if (L.morse) {
    hipError_t e = launch_cell_list(L, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(force_kernel, grid_for(L.n), dim3(kThreads), 0, st, L);
  return hipGetLastError();
}

hipError_t launch_neighbour_count(const ForcingLaunch &L, uint32_t *cnt, uint32_t *off,
                                  uint32_t *bsum, hipStream_t st) {
  if (L.n <= 0) return hipSuccess;
  hipError_t e = launch_cell_list(L, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(neighbour_count_kernel, grid_for(L.n), dim3(kThreads), 0, st, L, cnt);
  return exclusive_scan(cnt, L.n, off, bsum, st);
}

hipError_t launch_neighbour_fill(const ForcingLaunch &L, const uint32_t *off, int32_t *idx,
                                 hipStream_t st) {
  if (L.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(neighbour_fill_kernel, grid_for(L.n), dim3(kThreads), 0, st, L, off, idx);
  return hipGetLastError();
}

int64_t scan_partials_needed(int64_t n) { return (n + kScanBlock - 1) / kScanBlock + 1; }

hipError_t launch_stage(const StageLaunch &S, hipStream_t st) {
  if (S.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(stage_kernel, grid_for(S.n), dim3(kThreads), 0, st, S);
  return hipGetLastError();
}

hipError_t launch_euler_update(float *x, const float *v, float *off, int64_t n, const float *fdt,
                               int periodic, Lim lim, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(euler_update_kernel, grid_for(n), dim3(kThreads), 0, st, x, v, off, n, fdt,
                     periodic, lim);
  return hipGetLastError();
}

hipError_t launch_add(float *x, const float *off, int64_t n3, hipStream_t st) {
  if (n3 <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_kernel, grid_for(n3), dim3(kThreads), 0, st, x, off, n3);
  return hipGetLastError();
}

__global__ void set_f32_kernel(float *p, float v) {
  if (threadIdx.x == 0) p[0] = v;
}

hipError_t launch_set_f32(float *p, float v, hipStream_t st) {
  hipLaunchKernelGGL(set_f32_kernel, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

int64_t err_partials_needed(int64_t n) { return (n + kThreads - 1) / kThreads; }

hipError_t launch_dopri_err(const ErrLaunch &E, double *out, hipStream_t st) {
  const int64_t np = err_partials_needed(E.n);
  if (np > 0) hipLaunchKernelGGL(dopri_err_kernel, dim3((unsigned)np), dim3(kThreads), 0, st, E);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kThreads), 0, st, E.part, np, out);
  return hipGetLastError();
}

}  // namespace hh
