// Cluster tree, node boxes and close/far block lists built on the device:
// HignnModel::Update (src/HignnModel.cpp:457-461), i.e. Build
// (src/Build.cpp:17-265) with its divide step (src/HignnModel.cpp:119-198),
// the node boxes (src/CloseFarCheck.cpp:23-175) and the dual traversal
// CloseFarCheck (src/CloseFarCheck.cpp:181-454).
//
// The result is the host build's (hh_tree.cpp) bit for bit, for one or for
// several independent instances (a forest, one root per instance, all
// divided in the same level-synchronous passes).
//
// Divide step, per level of the tree, over all nodes of the level that are
// larger than a leaf:
//   * FP64 sums in the host's order: every 256-particle chunk of a node is
//     summed sequentially by one thread (mean_chunks / cov_chunks), then a
//     node's chunk totals sequentially by one thread (node_sums);
//   * the principal axis of each node's covariance on the host (the host
//     build's own routine, tree_principal_axis), a few hundred bytes per level;
//   * projections t in FP64 with the host's operation order, turned into
//     order-preserving 64-bit keys (-0 counted as +0), and the level's
//     particles reordered by (node, t, position) with two stable radix sorts:
//     by key, then by node;
//   * the split count #{t <= 0} per node from wave ballots.
// Boxes: one thread per leaf over its particles in tree order (the host's
// compare-and-select, so ties between -0 and +0 resolve as on the host), then
// internal nodes level by level from the deepest, as std::min/std::max.
// Dual traversal, per tree level: every (node, candidate) pair is one work
// item; a count pass decides far / keep / descend and how many outputs it
// makes, exclusive scans place the outputs, and a write pass emits them in the
// host's order (descendants right child first, found by descending with
// per-node subtree counts, so no stack).  A node's kept candidates are the
// candidate range of both children; leaves' candidates become the close list.
// The symmetric filter (J >= I) is a stable compaction.  With several
// instances the level-major lists are finally stably sorted by I, which gives
// the merged host order.
//
// The file is compiled with -ffp-contract=off (as hh_tree.cpp) so every
// product is rounded on its own.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <hipcub/hipcub.hpp>
#include <vector>

#include "hh_host.h"
#include "hh_internal.h"

namespace hh {

namespace {

constexpr int kChunk = 256;  // equals hh_tree.cpp's summation chunk

struct SplitNode {
  int32_t s, e;    // particle range
  int32_t off;     // first slot of the level's sort arrays
  int32_t cfirst;  // first summation chunk
};
struct SumChunk {
  int32_t node, p0, p1, pad;
};
// Traversal node table (global ids).
struct TNode {
  int32_t size, c0, c1, nleaf, nbound, pad0, pad1, pad2;
};

// ---------------------------------------------------------------------------
// divide step
// ---------------------------------------------------------------------------
__global__ void mean_chunks_kernel(const float *__restrict__ x, const int32_t *__restrict__ perm,
                                   const SumChunk *__restrict__ ch, int nch,
                                   double *__restrict__ cs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const SumChunk k = ch[c];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int p = k.p0; p < k.p1; ++p) {
    const float *q = x + 3 * (size_t)perm[p];
    a0 += (double)q[0];
    a1 += (double)q[1];
    a2 += (double)q[2];
  }
  double *o = cs + 6 * (size_t)c;
  o[0] = a0, o[1] = a1, o[2] = a2;
}

__global__ void cov_chunks_kernel(const float *__restrict__ x, const int32_t *__restrict__ perm,
                                  const SumChunk *__restrict__ ch, int nch,
                                  const double *__restrict__ stats, double *__restrict__ cs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const SumChunk k = ch[c];
  const double *mu = stats + 9 * (size_t)k.node;
  const double m0 = mu[0], m1 = mu[1], m2 = mu[2];
  double xx = 0.0, yy = 0.0, zz = 0.0, xy = 0.0, yz = 0.0, zx = 0.0;
  for (int p = k.p0; p < k.p1; ++p) {
    const float *q = x + 3 * (size_t)perm[p];
    const double c0 = (double)q[0] - m0, c1 = (double)q[1] - m1, c2 = (double)q[2] - m2;
    xx += c0 * c0;
    yy += c1 * c1;
    zz += c2 * c2;
    xy += c0 * c1;
    yz += c1 * c2;
    zx += c2 * c0;
  }
  double *o = cs + 6 * (size_t)c;
  o[0] = xx, o[1] = yy, o[2] = zz, o[3] = xy, o[4] = yz, o[5] = zx;
}

// stats[9k + 0..2] = mean (pass 0) or stats[9k + 3..8] = covariance sums (pass 1).
__global__ void node_sums_kernel(const SplitNode *__restrict__ nodes, int K, int nch,
                                 const double *__restrict__ cs, int cov, double *__restrict__ stats) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const SplitNode nd = nodes[k];
  const int c1 = k + 1 < K ? nodes[k + 1].cfirst : nch;
  double *o = stats + 9 * (size_t)k;
  if (!cov) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int c = nd.cfirst; c < c1; ++c) {
      a0 += cs[6 * (size_t)c];
      a1 += cs[6 * (size_t)c + 1];
      a2 += cs[6 * (size_t)c + 2];
    }
    const double len = (double)(nd.e - nd.s);
    o[0] = a0 / len, o[1] = a1 / len, o[2] = a2 / len;
  } else {
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = nd.cfirst; c < c1; ++c)
      for (int v = 0; v < 6; ++v) a[v] += cs[6 * (size_t)c + v];
    for (int v = 0; v < 6; ++v) o[3 + v] = a[v];
  }
}

// One workgroup of 256 per summation chunk: projection, sort key, slot and
// the node's count of t <= 0.
__global__ __launch_bounds__(256) void project_kernel(
    const float *__restrict__ x, const int32_t *__restrict__ perm, const SumChunk *__restrict__ ch,
    const SplitNode *__restrict__ nodes, const double *__restrict__ stats,
    const double *__restrict__ axes, uint64_t *__restrict__ keys, int32_t *__restrict__ vals,
    uint32_t *__restrict__ nkey, int32_t *__restrict__ nonpos) {
  const SumChunk k = ch[blockIdx.x];
  const int p = k.p0 + (int)threadIdx.x;
  bool np = false;
  if (p < k.p1) {
    const SplitNode nd = nodes[k.node];
    const double *mu = stats + 9 * (size_t)k.node;
    const double *ax = axes + 3 * (size_t)k.node;
    const float *q = x + 3 * (size_t)perm[p];
    const double c0 = (double)q[0] - mu[0], c1 = (double)q[1] - mu[1], c2 = (double)q[2] - mu[2];
    double t = (ax[0] * c0 + ax[1] * c1) + ax[2] * c2;
    np = t <= 0.0;
    if (t == 0.0) t = 0.0;  // -0 and +0 compare equal: one key
    const uint64_t b = (uint64_t)__double_as_longlong(t);
    const int j = nd.off + (p - nd.s);
    keys[j] = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    vals[j] = j;
    nkey[j] = (uint32_t)k.node;
  }
  const uint64_t bal = __ballot(np);
  if (__lane_id() == 0 && bal) atomicAdd(&nonpos[k.node], (int)__popcll(bal));
}

// The divided ranges of the reordered permutation back into perm.
__global__ void copy_back_kernel(const SplitNode *__restrict__ nodes, const uint32_t *__restrict__ nk, int M,
                                 const int32_t *__restrict__ src, int32_t *__restrict__ dst) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= M) return;
  const SplitNode nd = nodes[nk[o]];
  const int p = nd.s + (o - nd.off);
  dst[p] = src[p];
}

__global__ void gather_nkey_kernel(const int32_t *__restrict__ vals, const uint32_t *__restrict__ nkey,
                                   int M, uint32_t *__restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < M) out[o] = nkey[vals[o]];
}

__global__ void apply_order_kernel(const SplitNode *__restrict__ nodes, const uint32_t *__restrict__ nk,
                                   const int32_t *__restrict__ vals, int M,
                                   const int32_t *__restrict__ pin, int32_t *__restrict__ pout) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= M) return;
  const SplitNode nd = nodes[nk[o]];
  const int j = vals[o];
  pout[nd.s + (o - nd.off)] = pin[nd.s + (j - nd.off)];
}

// ---------------------------------------------------------------------------
// boxes
// ---------------------------------------------------------------------------
__global__ void leaf_box_kernel(const float *__restrict__ x, const int32_t *__restrict__ perm,
                                const int32_t *__restrict__ lv, const int32_t *__restrict__ range,
                                int nl, float *__restrict__ aux) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nl) return;
  float b[6];
  for (int d = 0; d < 3; ++d) b[2 * d] = FLT_MAX, b[2 * d + 1] = -FLT_MAX;
  for (int p = range[2 * k]; p < range[2 * k + 1]; ++p) {
    const float *q = x + 3 * (size_t)perm[p];
    for (int d = 0; d < 3; ++d) {
      const float v = q[d];
      b[2 * d] = (b[2 * d] > v) ? v : b[2 * d];
      b[2 * d + 1] = (b[2 * d + 1] < v) ? v : b[2 * d + 1];
    }
  }
  float *o = aux + 6 * (size_t)lv[k];
  for (int v = 0; v < 6; ++v) o[v] = b[v];
}

__global__ void inner_box_kernel(const int32_t *__restrict__ nds, const int32_t *__restrict__ kids,
                                 int K, float *__restrict__ aux) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const float *l = aux + 6 * (size_t)kids[2 * k];
  const float *r = aux + 6 * (size_t)kids[2 * k + 1];
  float *o = aux + 6 * (size_t)nds[k];
  for (int d = 0; d < 3; ++d) {
    o[2 * d] = (r[2 * d] < l[2 * d]) ? r[2 * d] : l[2 * d];                  // std::min(l, r)
    o[2 * d + 1] = (l[2 * d + 1] < r[2 * d + 1]) ? r[2 * d + 1] : l[2 * d + 1];  // std::max(l, r)
  }
}

// ---------------------------------------------------------------------------
// dual traversal
// ---------------------------------------------------------------------------
// Admissibility, src/HignnModel.cpp:482-518, float arithmetic in the
// reference's order (as hh_tree.cpp).
__device__ bool admissible_dev(const float *__restrict__ aux, int p, int q) {
  const float *a = aux + 6 * (size_t)p, *b = aux + 6 * (size_t)q;
  const bool apart = a[0] > b[1] || a[1] < b[0] || a[2] > b[3] || a[3] < b[2] ||
                     a[4] > b[5] || a[5] < b[4];
  if (!apart) return false;
  float da = 0.0f, db = 0.0f, cc = 0.0f;
  for (int j = 0; j < 3; ++j) {
    float t = a[2 * j] - a[2 * j + 1];
    da += t * t;
    t = b[2 * j] - b[2 * j + 1];
    db += t * t;
    t = a[2 * j] + a[2 * j + 1] - b[2 * j] - b[2 * j + 1];
    cc += t * t;
  }
  cc *= 0.25f;
  return cc > da && cc > db;
}

enum TraverseAction : int8_t {
  TA_FAR = 0,        // (nd, m) far
  TA_FAR_DESC = 1,   // far: the bounded descendants of m
  TA_KEEP_KIDS = 2,  // keep both children of m
  TA_KEEP_SELF = 3,  // keep m
  TA_KEEP_LEAVES = 4 // keep the leaves under m
};

__device__ inline int owner_of(const int32_t *__restrict__ wbase, int K, int w) {
  int lo = 0, hi = K;  // largest k with wbase[k] <= w (ranges may be empty)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (wbase[mid] <= w) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void cand_len_kernel(const int32_t *__restrict__ cb, const int32_t *__restrict__ ce, int K,
                                int32_t *__restrict__ len) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) len[k] = ce[k] - cb[k];
  if (k == K) len[k] = 0;
}

__global__ void traverse_count_kernel(const TNode *__restrict__ tn, const float *__restrict__ aux,
                                      const int32_t *__restrict__ lnode, const int32_t *__restrict__ cb,
                                      const int32_t *__restrict__ wbase, int K, int W,
                                      const int32_t *__restrict__ cin, int64_t R,
                                      int8_t *__restrict__ act, int32_t *__restrict__ fc,
                                      int32_t *__restrict__ kc) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > W) return;
  if (w == W) {
    fc[w] = 0, kc[w] = 0;
    return;
  }
  const int k = owner_of(wbase, K, w);
  const int nd = lnode[k];
  const int m = cin[cb[k] + (w - wbase[k])];
  const TNode a = tn[nd], b = tn[m];
  const bool adm = admissible_dev(aux, nd, m);
  int8_t t;
  int nf = 0, nk = 0;
  if (a.c0 != 0) {  // internal node
    if (adm && (int64_t)a.size < R && (int64_t)b.size < R) t = TA_FAR, nf = 1;
    else if (b.c0 != 0) t = TA_KEEP_KIDS, nk = 2;
    else t = TA_KEEP_SELF, nk = 1;
  } else {
    if (adm) {
      if ((int64_t)b.size < R) t = TA_FAR, nf = 1;
      else t = TA_FAR_DESC, nf = b.nbound;
    } else if (b.c0 == 0) {
      t = TA_KEEP_SELF, nk = 1;
    } else {
      t = TA_KEEP_LEAVES, nk = b.nleaf;
    }
  }
  act[w] = t;
  fc[w] = nf;
  kc[w] = nk;
}

// The j-th node of m's subtree that satisfies the stop rule, in the host's
// LIFO order (right child first); cnt[] counts them per subtree.
template <bool kBounded>
__device__ int descend(const TNode *__restrict__ tn, int m, int j, int64_t R) {
  int c = m;
  for (;;) {
    const TNode t = tn[c];
    if (t.c0 == 0 || (kBounded && (int64_t)t.size < R)) return c;
    const TNode r = tn[t.c1];
    const int rc = kBounded ? r.nbound : r.nleaf;
    if (j < rc) {
      c = t.c1;
    } else {
      j -= rc;
      c = t.c0;
    }
  }
}

__global__ void traverse_write_kernel(const TNode *__restrict__ tn, const int32_t *__restrict__ lnode,
                                      const int32_t *__restrict__ cb, const int32_t *__restrict__ wbase,
                                      int K, int W, const int32_t *__restrict__ cin, int64_t R,
                                      const int8_t *__restrict__ act, const int32_t *__restrict__ fo,
                                      const int32_t *__restrict__ ko, int32_t *__restrict__ farI,
                                      int32_t *__restrict__ farJ, int32_t *__restrict__ keep,
                                      int32_t *__restrict__ kowner) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const int k = owner_of(wbase, K, w);
  const int nd = lnode[k];
  const int m = cin[cb[k] + (w - wbase[k])];
  const int f0 = fo[w], k0 = ko[w];
  switch (act[w]) {
    case TA_FAR:
      farI[f0] = nd, farJ[f0] = m;
      break;
    case TA_FAR_DESC: {
      const int cnt = fo[w + 1] - f0;
      for (int j = 0; j < cnt; ++j) farI[f0 + j] = nd, farJ[f0 + j] = descend<true>(tn, m, j, R);
      break;
    }
    case TA_KEEP_KIDS:
      keep[k0] = tn[m].c0, keep[k0 + 1] = tn[m].c1;
      kowner[k0] = k, kowner[k0 + 1] = k;
      break;
    case TA_KEEP_SELF:
      keep[k0] = m, kowner[k0] = k;
      break;
    default: {
      const int cnt = ko[w + 1] - k0;
      for (int j = 0; j < cnt; ++j) keep[k0 + j] = descend<false>(tn, m, j, R), kowner[k0 + j] = k;
    }
  }
}

// Candidate ranges of the next level: both children of an internal node get
// its kept candidates.
__global__ void next_ranges_kernel(const int32_t *__restrict__ lchild, const int32_t *__restrict__ wbase,
                                   const int32_t *__restrict__ ko, int K, int32_t *__restrict__ ncb,
                                   int32_t *__restrict__ nce) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const int c = lchild[k];
  if (c < 0) return;
  const int b = ko[wbase[k]], e = ko[wbase[k + 1]];
  ncb[c] = b, nce[c] = e;
  ncb[c + 1] = b, nce[c + 1] = e;
}

// flags (with a zero sentinel) of the pairs the symmetric filter keeps:
// far pairs J >= I; close pairs from leaves' kept candidates.
__global__ void far_flag_kernel(const int32_t *__restrict__ I, const int32_t *__restrict__ J, int F,
                                int sym, int32_t *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F) flag[i] = (!sym || J[i] >= I[i]) ? 1 : 0;
  if (i == F) flag[i] = 0;
}

__global__ void close_flag_kernel(const int32_t *__restrict__ keep, const int32_t *__restrict__ kowner,
                                  const int32_t *__restrict__ lnode, const int32_t *__restrict__ lchild,
                                  int Kt, int sym, int32_t *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Kt) {
    const int k = kowner[i];
    flag[i] = (lchild[k] < 0 && (!sym || keep[i] >= lnode[k])) ? 1 : 0;
  }
  if (i == Kt) flag[i] = 0;
}

__global__ void far_scatter_kernel(const int32_t *__restrict__ I, const int32_t *__restrict__ J,
                                   const int32_t *__restrict__ flag, const int32_t *__restrict__ pos,
                                   int F, int32_t *__restrict__ oI, int32_t *__restrict__ oJ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F && flag[i]) oI[pos[i]] = I[i], oJ[pos[i]] = J[i];
}

__global__ void close_scatter_kernel(const int32_t *__restrict__ keep, const int32_t *__restrict__ kowner,
                                     const int32_t *__restrict__ lnode, const int32_t *__restrict__ flag,
                                     const int32_t *__restrict__ pos, int Kt, int32_t *__restrict__ oI,
                                     int32_t *__restrict__ oJ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Kt && flag[i]) oI[pos[i]] = lnode[kowner[i]], oJ[pos[i]] = keep[i];
}

__global__ void gather_i32_kernel(const int32_t *__restrict__ src, const int32_t *__restrict__ idx, int n,
                                  int32_t *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void iota_kernel(int32_t *__restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int bits_for(int64_t v) {  // bits to hold 0..v-1
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < v) ++b;
  return b;
}

}  // namespace

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
struct DeviceTreeBuilder {
  DBuf perm2, ch, sn, cs, stats, axes, keys[2], vals[2], nk[2], nonpos, temp;
  DBuf aux, tn, lnode, lchild, cb[2], ce[2], len, wbase, act, fc, kc, fo, ko, cand[2], kowner, ftI, ftJ,
      flag, pos, closeI, closeJ, farI, farJ, sortI, sortJ, idx, leafv, leafr, inner, kids;
  PBuf pin;
  size_t temp_bytes = 0;
  hipStream_t st = nullptr;

  void sync_read(void *dst, const void *src, size_t bytes) {
    pin.reserve(std::max<size_t>(bytes, 64));
    HIPCHK(hipMemcpyAsync(pin.p, src, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::memcpy(dst, pin.p, bytes);
  }
  // host -> device; returns once the host data has been read
  void write(DBuf &b, const void *src, size_t bytes) {
    b.reserve(std::max<size_t>(bytes, 16));
    if (!bytes) return;
    HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  void ensure_temp(size_t need) {
    if (need <= temp_bytes) return;
    HIPCHK(hipStreamSynchronize(st));
    temp.reserve(need);
    temp_bytes = need;
  }
  // exclusive sum of n + 1 int32 values (the last one a zero sentinel):
  // out[n] is the total
  void exclusive_sum(const int32_t *in, int32_t *out, int n1) {
    size_t need = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, n1, st));
    ensure_temp(need);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(temp.p, need, in, out, n1, st));
  }
  template <class K>
  void sort_pairs(const K *kin, K *kout, const int32_t *vin, int32_t *vout, int n, int bits) {
    size_t need = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, n, 0, bits, st));
    ensure_temp(need);
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(temp.p, need, kin, kout, vin, vout, n, 0, bits, st));
  }
};

DeviceTreeBuilder *device_tree_builder_new() { return new DeviceTreeBuilder(); }

void device_tree_builder_free(DeviceTreeBuilder *b) {
  if (!b) return;
  for (DBuf *d : {&b->aux, &b->perm2, &b->ch, &b->sn, &b->cs, &b->stats, &b->axes, &b->keys[0], &b->keys[1],
                  &b->vals[0], &b->vals[1], &b->nk[0], &b->nk[1], &b->nonpos, &b->temp, &b->tn,
                  &b->lnode, &b->lchild, &b->cb[0], &b->cb[1], &b->ce[0], &b->ce[1], &b->len,
                  &b->wbase, &b->act, &b->fc, &b->kc, &b->fo, &b->ko, &b->cand[0], &b->cand[1],
                  &b->kowner, &b->ftI, &b->ftJ, &b->flag, &b->pos, &b->closeI, &b->closeJ,
                  &b->farI, &b->farJ, &b->sortI, &b->sortJ, &b->idx, &b->leafv, &b->leafr,
                  &b->inner, &b->kids})
    d->release();
  b->pin.release();
  delete b;
}

namespace {

// One instance's nodes during the build (local numbering of hh_tree.cpp:
// root 0, the local stage's nodes 1.. in breadth-first order).
struct InstTree {
  int64_t pb = 0, n = 0;
  struct Node {
    int64_t s, e, lvl, c0, c1;
  };
  std::vector<Node> nodes;
  std::vector<int64_t> level;  // current local-stage level
  uint64_t r1[5] = {0, 0, 0, 0, 0}, r2[5] = {0, 0, 0, 0, 0};
};

// Divides the given node ranges (global particle positions) once: sorts each
// range by (t, position) and returns #{t <= 0} per range.
void divide_ranges(DeviceTreeBuilder &B, const float *x, int32_t *perm,
                   const std::vector<std::pair<int64_t, int64_t>> &rg, std::vector<int32_t> &nonpos) {
  const int K = (int)rg.size();
  nonpos.assign((size_t)K, 0);
  if (K == 0) return;
  std::vector<SplitNode> sn((size_t)K);
  std::vector<SumChunk> ch;
  int64_t M = 0;
  for (int k = 0; k < K; ++k) {
    const int64_t s = rg[(size_t)k].first, e = rg[(size_t)k].second;
    sn[(size_t)k] = SplitNode{(int32_t)s, (int32_t)e, (int32_t)M, (int32_t)ch.size()};
    for (int64_t p = s; p < e; p += kChunk)
      ch.push_back(SumChunk{k, (int32_t)p, (int32_t)std::min(e, p + kChunk), 0});
    M += e - s;
  }
  const int nch = (int)ch.size();
  hipStream_t st = B.st;
  B.write(B.sn, sn.data(), sn.size() * sizeof(SplitNode));
  B.write(B.ch, ch.data(), ch.size() * sizeof(SumChunk));
  B.cs.reserve(48 * (size_t)nch);
  B.stats.reserve(72 * (size_t)K);
  mean_chunks_kernel<<<blocks_for(nch, 64), 64, 0, st>>>(x, perm, B.ch.as<SumChunk>(), nch, B.cs.as<double>());
  node_sums_kernel<<<blocks_for(K, 64), 64, 0, st>>>(B.sn.as<SplitNode>(), K, nch, B.cs.as<double>(), 0,
                                                     B.stats.as<double>());
  cov_chunks_kernel<<<blocks_for(nch, 64), 64, 0, st>>>(x, perm, B.ch.as<SumChunk>(), nch,
                                                        B.stats.as<double>(), B.cs.as<double>());
  node_sums_kernel<<<blocks_for(K, 64), 64, 0, st>>>(B.sn.as<SplitNode>(), K, nch, B.cs.as<double>(), 1,
                                                     B.stats.as<double>());
  HIPCHK(hipGetLastError());
  std::vector<double> stats(9 * (size_t)K), axes(3 * (size_t)K);
  B.sync_read(stats.data(), B.stats.p, stats.size() * 8);
  for (int k = 0; k < K; ++k) tree_principal_axis(&stats[9 * (size_t)k + 3], &axes[3 * (size_t)k]);
  B.write(B.axes, axes.data(), axes.size() * 8);
  for (int i = 0; i < 2; ++i) {
    B.keys[i].reserve(8 * (size_t)M);
    B.vals[i].reserve(4 * (size_t)M);
    B.nk[i].reserve(4 * (size_t)M);
  }
  B.nonpos.reserve(4 * (size_t)K);
  HIPCHK(hipMemsetAsync(B.nonpos.p, 0, 4 * (size_t)K, st));
  project_kernel<<<nch, 256, 0, st>>>(x, perm, B.ch.as<SumChunk>(), B.sn.as<SplitNode>(),
                                      B.stats.as<double>(), B.axes.as<double>(), B.keys[0].as<uint64_t>(),
                                      B.vals[0].as<int32_t>(), B.nk[0].as<uint32_t>(),
                                      B.nonpos.as<int32_t>());
  HIPCHK(hipGetLastError());
  // stable: by key (input in slot = position order), then by node
  B.sort_pairs(B.keys[0].as<uint64_t>(), B.keys[1].as<uint64_t>(), B.vals[0].as<int32_t>(),
               B.vals[1].as<int32_t>(), (int)M, 64);
  gather_nkey_kernel<<<blocks_for(M), 256, 0, st>>>(B.vals[1].as<int32_t>(), B.nk[0].as<uint32_t>(), (int)M,
                                                   B.nk[1].as<uint32_t>());
  B.sort_pairs(B.nk[1].as<uint32_t>(), B.nk[0].as<uint32_t>(), B.vals[1].as<int32_t>(),
               B.vals[0].as<int32_t>(), (int)M, bits_for(K));
  int32_t *pout = B.perm2.as<int32_t>();
  apply_order_kernel<<<blocks_for(M), 256, 0, st>>>(B.sn.as<SplitNode>(), B.nk[0].as<uint32_t>(),
                                                    B.vals[0].as<int32_t>(), (int)M, perm, pout);
  // positions outside the divided ranges keep their entries
  copy_back_kernel<<<blocks_for(M), 256, 0, st>>>(B.sn.as<SplitNode>(), B.nk[0].as<uint32_t>(), (int)M, pout,
                                                  perm);
  HIPCHK(hipGetLastError());
  B.sync_read(nonpos.data(), B.nonpos.p, 4 * (size_t)K);
}

}  // namespace

void build_tree_lists_device(DeviceTreeBuilder *Bp, const float *x, const std::vector<int64_t> &pbase,
                             int leaf, bool sym, int64_t R, void *stream, TreeLists &out,
                             int32_t *perm_dev_out) {
  DeviceTreeBuilder &B = *Bp;
  B.st = (hipStream_t)stream;
  hipStream_t st = B.st;
  const int64_t ninst = (int64_t)pbase.size() - 1;
  const int64_t n = pbase.back();
  if (n <= 0 || n >= (int64_t)1 << 31) throw HHError(HH_ERR_ARG, "device tree build: n out of range");
  out = TreeLists();
  out.n = n;
  // perm in tree order: identity per instance, kept in perm_dev_out; the
  // reorder of each divide writes perm2 and copies the divided ranges back
  int32_t *perm = perm_dev_out;
  B.perm2.reserve(4 * (size_t)n);
  iota_kernel<<<blocks_for(n), 256, 0, st>>>(perm, (int)n);
  HIPCHK(hipGetLastError());
  // ---- tree ----
  std::vector<InstTree> it((size_t)ninst);
  for (int64_t i = 0; i < ninst; ++i) {
    it[(size_t)i].pb = pbase[(size_t)i];
    it[(size_t)i].n = pbase[(size_t)i + 1] - pbase[(size_t)i];
    it[(size_t)i].nodes.push_back({0, it[(size_t)i].n, 0, 0, 0});
    it[(size_t)i].level.push_back(0);
  }
  auto divide = [&](const std::vector<std::pair<int64_t, int64_t>> &rg, std::vector<int32_t> &np) {
    divide_ranges(B, x, perm, rg, np);
  };
  // initial stage (src/Build.cpp:38-97): every root larger than a leaf is
  // divided once; rows 1 and 2 record that split
  {
    std::vector<std::pair<int64_t, int64_t>> rg;
    std::vector<int64_t> who;
    for (int64_t i = 0; i < ninst; ++i)
      if (it[(size_t)i].n > leaf) rg.push_back({it[(size_t)i].pb, it[(size_t)i].pb + it[(size_t)i].n}), who.push_back(i);
    std::vector<int32_t> np;
    divide(rg, np);
    for (size_t k = 0; k < who.size(); ++k) {
      InstTree &t = it[(size_t)who[k]];
      const int64_t d = np[k];
      if (d != 0 && d != t.n) {
        t.r1[2] = 0, t.r1[3] = (uint64_t)d, t.r1[4] = 1;
        t.r2[2] = (uint64_t)d, t.r2[3] = (uint64_t)t.n, t.r2[4] = 1;
      }
    }
  }
  // local stage (src/Build.cpp:111-180): breadth first from the roots
  for (;;) {
    std::vector<std::pair<int64_t, int64_t>> rg;
    std::vector<std::pair<int64_t, int64_t>> who;  // (instance, local node)
    bool any = false;
    for (int64_t i = 0; i < ninst; ++i) {
      InstTree &t = it[(size_t)i];
      if (!t.level.empty()) any = true;
      for (int64_t id : t.level) {
        const auto &nd = t.nodes[(size_t)id];
        if (nd.e - nd.s > leaf) rg.push_back({t.pb + nd.s, t.pb + nd.e}), who.push_back({i, id});
      }
    }
    if (!any) break;
    std::vector<int32_t> np;
    divide(rg, np);
    std::vector<std::vector<int64_t>> split((size_t)ninst);
    for (int64_t i = 0; i < ninst; ++i) split[(size_t)i].assign(it[(size_t)i].level.size(), -1);
    {
      size_t k = 0;
      for (int64_t i = 0; i < ninst; ++i) {
        InstTree &t = it[(size_t)i];
        for (size_t q = 0; q < t.level.size(); ++q) {
          const auto &nd = t.nodes[(size_t)t.level[q]];
          if (nd.e - nd.s > leaf) {
            const int64_t sp = nd.s + np[k++];
            split[(size_t)i][q] = (sp == nd.s || sp == nd.e) ? nd.s : sp;
          } else {
            split[(size_t)i][q] = nd.s;
          }
        }
      }
    }
    for (int64_t i = 0; i < ninst; ++i) {
      InstTree &t = it[(size_t)i];
      std::vector<int64_t> next;
      for (size_t q = 0; q < t.level.size(); ++q) {
        const int64_t id = t.level[q];
        const int64_t sp = split[(size_t)i][q];
        if (sp == t.nodes[(size_t)id].s) continue;
        const InstTree::Node par = t.nodes[(size_t)id];
        const int64_t a = (int64_t)t.nodes.size(), b = a + 1;
        t.nodes[(size_t)id].c0 = a;
        t.nodes[(size_t)id].c1 = b;
        t.nodes.push_back({par.s, sp, par.lvl + 1, 0, 0});
        t.nodes.push_back({sp, par.e, par.lvl + 1, 0, 0});
        next.push_back(a);
        next.push_back(b);
      }
      t.level.swap(next);
    }
  }
  // rows (src/Build.cpp:199-247 per instance; instances merged with node and
  // particle offsets, leaves keeping child id 0)
  std::vector<int64_t> nbase((size_t)ninst + 1, 0);
  for (int64_t i = 0; i < ninst; ++i) nbase[(size_t)i + 1] = nbase[(size_t)i] + 3 + ((int64_t)it[(size_t)i].nodes.size() - 1);
  const int64_t T = nbase[(size_t)ninst];
  if (T >= (int64_t)1 << 31) throw HHError(HH_ERR_ARG, "device tree build: too many nodes");
  out.T = T;
  out.rows.assign((size_t)(5 * T), 0);
  for (int64_t i = 0; i < ninst; ++i) {
    const InstTree &t = it[(size_t)i];
    const uint64_t nb = (uint64_t)nbase[(size_t)i], pb = (uint64_t)t.pb;
    const int64_t nloc = (int64_t)t.nodes.size() - 1;
    std::vector<uint64_t> loc((size_t)(5 * (3 + nloc)), 0);
    loc[2] = 0, loc[3] = (uint64_t)t.n, loc[4] = 0;
    for (int c = 0; c < 5; ++c) loc[5 + c] = t.r1[c], loc[10 + c] = t.r2[c];
    if (nloc > 0) loc[0] = 3, loc[1] = 4;
    for (int64_t l = 1; l <= nloc; ++l) {
      const auto &nd = t.nodes[(size_t)l];
      uint64_t *row = &loc[(size_t)(5 * (l + 2))];
      if (nd.c0) row[0] = (uint64_t)(nd.c0 + 2), row[1] = (uint64_t)(nd.c1 + 2);
      row[2] = (uint64_t)nd.s, row[3] = (uint64_t)nd.e, row[4] = (uint64_t)nd.lvl;
    }
    for (int64_t l = 0; l < 3 + nloc; ++l) {
      const uint64_t *r = &loc[(size_t)(5 * l)];
      uint64_t *o = &out.rows[(size_t)(5 * (nbase[(size_t)i] + l))];
      o[0] = r[0] ? r[0] + nb : 0;
      o[1] = r[0] ? r[1] + nb : r[1];
      o[2] = r[2] + pb;
      o[3] = r[3] + pb;
      o[4] = r[4];
    }
  }
  auto isleaf = [&](int64_t g) { return out.rows[(size_t)(5 * g)] == 0; };
  auto kid = [&](int64_t g, int c) { return (int64_t)out.rows[(size_t)(5 * g + c)]; };
  // reachable nodes per instance in breadth-first (= id) order, by depth
  std::vector<std::vector<int32_t>> depth_nodes;
  for (int64_t i = 0; i < ninst; ++i) {
    std::vector<int64_t> lv{nbase[(size_t)i]};
    for (size_t dpt = 0; !lv.empty(); ++dpt) {
      if (depth_nodes.size() <= dpt) depth_nodes.emplace_back();
      std::vector<int64_t> nx;
      for (int64_t g : lv) {
        depth_nodes[dpt].push_back((int32_t)g);
        if (!isleaf(g)) nx.push_back(kid(g, 0)), nx.push_back(kid(g, 1));
      }
      lv.swap(nx);
    }
  }
  // leaves in id order (per instance level by level = id order)
  for (int64_t i = 0; i < ninst; ++i)
    for (int64_t g = nbase[(size_t)i]; g < nbase[(size_t)i + 1]; ++g) {
      const int64_t l = g - nbase[(size_t)i];
      if ((l == 1 || l == 2)) continue;  // rows of the initial stage: unreachable
      if (isleaf(g)) out.leaves.push_back(g);
    }
  // ---- boxes ----
  B.tn.reserve(sizeof(TNode) * (size_t)T);
  B.aux.reserve(24 * (size_t)T);
  float *aux = B.aux.as<float>();
  HIPCHK(hipMemsetAsync(aux, 0, 24 * (size_t)T, st));
  {
    std::vector<int32_t> lv, lr;
    for (int64_t g : out.leaves) {
      lv.push_back((int32_t)g);
      lr.push_back((int32_t)out.rows[(size_t)(5 * g + 2)]);
      lr.push_back((int32_t)out.rows[(size_t)(5 * g + 3)]);
    }
    B.write(B.leafv, lv.data(), lv.size() * 4);
    B.write(B.leafr, lr.data(), lr.size() * 4);
    leaf_box_kernel<<<blocks_for((int64_t)lv.size()), 256, 0, st>>>(x, perm, B.leafv.as<int32_t>(),
                                                                     B.leafr.as<int32_t>(), (int)lv.size(), aux);
    HIPCHK(hipGetLastError());
    // internal nodes, deepest level first, from one upload
    std::vector<int32_t> nds, kd;
    std::vector<std::pair<size_t, size_t>> grp;
    for (size_t dd = depth_nodes.size(); dd-- > 0;) {
      const size_t g0 = nds.size();
      for (int32_t g : depth_nodes[dd])
        if (!isleaf(g)) nds.push_back(g), kd.push_back((int32_t)kid(g, 0)), kd.push_back((int32_t)kid(g, 1));
      if (nds.size() > g0) grp.push_back({g0, nds.size() - g0});
    }
    B.write(B.inner, nds.data(), nds.size() * 4);
    B.write(B.kids, kd.data(), kd.size() * 4);
    for (const auto &g : grp) {
      inner_box_kernel<<<blocks_for((int64_t)g.second), 256, 0, st>>>(
          B.inner.as<int32_t>() + g.first, B.kids.as<int32_t>() + 2 * g.first, (int)g.second, aux);
      HIPCHK(hipGetLastError());
    }
  }
  // ---- traversal ----
  {
    std::vector<TNode> tn((size_t)T);
    std::vector<int32_t> nleaf((size_t)T, 0), nbound((size_t)T, 0);
    for (size_t dd = depth_nodes.size(); dd-- > 0;)
      for (int32_t g : depth_nodes[dd]) {
        const int64_t sz = (int64_t)(out.rows[(size_t)(5 * g + 3)] - out.rows[(size_t)(5 * g + 2)]);
        if (isleaf(g)) {
          nleaf[(size_t)g] = 1, nbound[(size_t)g] = 1;
        } else {
          nleaf[(size_t)g] = nleaf[(size_t)kid(g, 0)] + nleaf[(size_t)kid(g, 1)];
          nbound[(size_t)g] = sz < R ? 1 : nbound[(size_t)kid(g, 0)] + nbound[(size_t)kid(g, 1)];
        }
      }
    for (int64_t g = 0; g < T; ++g) {
      TNode &t = tn[(size_t)g];
      std::memset(&t, 0, sizeof t);
      t.size = (int32_t)(out.rows[(size_t)(5 * g + 3)] - out.rows[(size_t)(5 * g + 2)]);
      t.c0 = (int32_t)out.rows[(size_t)(5 * g)];
      t.c1 = (int32_t)out.rows[(size_t)(5 * g + 1)];
      t.nleaf = nleaf[(size_t)g];
      t.nbound = nbound[(size_t)g];
    }
    B.write(B.tn, tn.data(), tn.size() * sizeof(TNode));
    int64_t nclose = 0, nfar = 0;
    // level 0: every root is its own candidate
    int cur = 0;
    {
      std::vector<int32_t> roots, cbv, cev;
      for (int64_t i = 0; i < ninst; ++i) {
        roots.push_back((int32_t)nbase[(size_t)i]);
        cbv.push_back((int32_t)i), cev.push_back((int32_t)i + 1);
      }
      B.write(B.cand[cur], roots.data(), roots.size() * 4);
      B.write(B.cb[cur], cbv.data(), cbv.size() * 4);
      B.write(B.ce[cur], cev.data(), cev.size() * 4);
      HIPCHK(hipStreamSynchronize(st));
    }
    const TNode *dtn = B.tn.as<TNode>();
    for (size_t dd = 0; dd < depth_nodes.size(); ++dd) {
      const std::vector<int32_t> &lv = depth_nodes[dd];
      const int K = (int)lv.size();
      std::vector<int32_t> lch((size_t)K, -1);
      int nnext = 0;
      for (int k = 0; k < K; ++k)
        if (!isleaf(lv[(size_t)k])) lch[(size_t)k] = nnext, nnext += 2;
      B.write(B.lnode, lv.data(), lv.size() * 4);
      B.write(B.lchild, lch.data(), lch.size() * 4);
      B.len.reserve(4 * (size_t)(K + 1));
      B.wbase.reserve(4 * (size_t)(K + 1));
      cand_len_kernel<<<blocks_for(K + 1), 256, 0, st>>>(B.cb[cur].as<int32_t>(), B.ce[cur].as<int32_t>(), K,
                                                         B.len.as<int32_t>());
      B.exclusive_sum(B.len.as<int32_t>(), B.wbase.as<int32_t>(), K + 1);
      int32_t W = 0;
      B.sync_read(&W, B.wbase.as<int32_t>() + K, 4);
      const int nxt = cur ^ 1;
      B.cb[nxt].reserve(4 * (size_t)std::max(nnext, 1));
      B.ce[nxt].reserve(4 * (size_t)std::max(nnext, 1));
      if (nnext) HIPCHK(hipMemsetAsync(B.cb[nxt].p, 0, 4 * (size_t)nnext, st));
      if (nnext) HIPCHK(hipMemsetAsync(B.ce[nxt].p, 0, 4 * (size_t)nnext, st));
      if (W == 0) {
        cur = nxt;
        continue;
      }
      B.act.reserve((size_t)W + 1);
      B.fc.reserve(4 * (size_t)(W + 1));
      B.kc.reserve(4 * (size_t)(W + 1));
      B.fo.reserve(4 * (size_t)(W + 1));
      B.ko.reserve(4 * (size_t)(W + 1));
      traverse_count_kernel<<<blocks_for((int64_t)W + 1), 256, 0, st>>>(
          dtn, aux, B.lnode.as<int32_t>(), B.cb[cur].as<int32_t>(), B.wbase.as<int32_t>(), K, W,
          B.cand[cur].as<int32_t>(), R, B.act.as<int8_t>(), B.fc.as<int32_t>(), B.kc.as<int32_t>());
      HIPCHK(hipGetLastError());
      B.exclusive_sum(B.fc.as<int32_t>(), B.fo.as<int32_t>(), W + 1);
      B.exclusive_sum(B.kc.as<int32_t>(), B.ko.as<int32_t>(), W + 1);
      int32_t FK[2] = {0, 0};
      B.sync_read(&FK[0], B.fo.as<int32_t>() + W, 4);
      B.sync_read(&FK[1], B.ko.as<int32_t>() + W, 4);
      const int F = FK[0], Kt = FK[1];
      B.ftI.reserve(4 * (size_t)std::max(F, 1));
      B.ftJ.reserve(4 * (size_t)std::max(F, 1));
      B.cand[nxt].reserve(4 * (size_t)std::max(Kt, 1));
      B.kowner.reserve(4 * (size_t)std::max(Kt, 1));
      traverse_write_kernel<<<blocks_for(W), 256, 0, st>>>(
          dtn, B.lnode.as<int32_t>(), B.cb[cur].as<int32_t>(), B.wbase.as<int32_t>(), K, W,
          B.cand[cur].as<int32_t>(), R, B.act.as<int8_t>(), B.fo.as<int32_t>(), B.ko.as<int32_t>(),
          B.ftI.as<int32_t>(), B.ftJ.as<int32_t>(), B.cand[nxt].as<int32_t>(), B.kowner.as<int32_t>());
      HIPCHK(hipGetLastError());
      next_ranges_kernel<<<blocks_for(K), 256, 0, st>>>(B.lchild.as<int32_t>(), B.wbase.as<int32_t>(),
                                                        B.ko.as<int32_t>(), K, B.cb[nxt].as<int32_t>(),
                                                        B.ce[nxt].as<int32_t>());
      HIPCHK(hipGetLastError());
      // compactions into the global lists
      const int FKmax = std::max(F, Kt);
      B.flag.reserve(4 * (size_t)(FKmax + 1));
      B.pos.reserve(4 * (size_t)(FKmax + 1));
      int32_t cnt = 0;
      if (F > 0) {
        far_flag_kernel<<<blocks_for((int64_t)F + 1), 256, 0, st>>>(B.ftI.as<int32_t>(), B.ftJ.as<int32_t>(), F,
                                                                    sym ? 1 : 0, B.flag.as<int32_t>());
        B.exclusive_sum(B.flag.as<int32_t>(), B.pos.as<int32_t>(), F + 1);
        B.sync_read(&cnt, B.pos.as<int32_t>() + F, 4);
        if (cnt > 0) {
          const size_t need = 4 * (size_t)(nfar + cnt);
          if (need > B.farI.bytes) {
            DBuf a, b;
            a.reserve(2 * need), b.reserve(2 * need);
            if (nfar) {
              HIPCHK(hipMemcpyAsync(a.p, B.farI.p, 4 * (size_t)nfar, hipMemcpyDeviceToDevice, st));
              HIPCHK(hipMemcpyAsync(b.p, B.farJ.p, 4 * (size_t)nfar, hipMemcpyDeviceToDevice, st));
              HIPCHK(hipStreamSynchronize(st));
            }
            B.farI.release(), B.farJ.release();
            B.farI = a, B.farJ = b;
          }
          far_scatter_kernel<<<blocks_for(F), 256, 0, st>>>(B.ftI.as<int32_t>(), B.ftJ.as<int32_t>(),
                                                            B.flag.as<int32_t>(), B.pos.as<int32_t>(), F,
                                                            B.farI.as<int32_t>() + nfar,
                                                            B.farJ.as<int32_t>() + nfar);
          HIPCHK(hipGetLastError());
          nfar += cnt;
        }
      }
      if (Kt > 0) {
        close_flag_kernel<<<blocks_for((int64_t)Kt + 1), 256, 0, st>>>(
            B.cand[nxt].as<int32_t>(), B.kowner.as<int32_t>(), B.lnode.as<int32_t>(), B.lchild.as<int32_t>(), Kt,
            sym ? 1 : 0, B.flag.as<int32_t>());
        B.exclusive_sum(B.flag.as<int32_t>(), B.pos.as<int32_t>(), Kt + 1);
        B.sync_read(&cnt, B.pos.as<int32_t>() + Kt, 4);
        if (cnt > 0) {
          const size_t need = 4 * (size_t)(nclose + cnt);
          if (need > B.closeI.bytes) {
            DBuf a, b;
            a.reserve(2 * need), b.reserve(2 * need);
            if (nclose) {
              HIPCHK(hipMemcpyAsync(a.p, B.closeI.p, 4 * (size_t)nclose, hipMemcpyDeviceToDevice, st));
              HIPCHK(hipMemcpyAsync(b.p, B.closeJ.p, 4 * (size_t)nclose, hipMemcpyDeviceToDevice, st));
              HIPCHK(hipStreamSynchronize(st));
            }
            B.closeI.release(), B.closeJ.release();
            B.closeI = a, B.closeJ = b;
          }
          close_scatter_kernel<<<blocks_for(Kt), 256, 0, st>>>(
              B.cand[nxt].as<int32_t>(), B.kowner.as<int32_t>(), B.lnode.as<int32_t>(), B.flag.as<int32_t>(),
              B.pos.as<int32_t>(), Kt, B.closeI.as<int32_t>() + nclose, B.closeJ.as<int32_t>() + nclose);
          HIPCHK(hipGetLastError());
          nclose += cnt;
        }
      }
      cur = nxt;
    }
    // with several instances the level-major lists are put in id order
    auto by_owner = [&](DBuf &I, DBuf &J, int64_t cnt) {
      if (ninst <= 1 || cnt == 0) return;
      B.sortI.reserve(4 * (size_t)cnt);
      B.sortJ.reserve(4 * (size_t)cnt);
      B.idx.reserve(8 * (size_t)cnt);
      int32_t *iv = B.idx.as<int32_t>(), *iv2 = iv + cnt;
      iota_kernel<<<blocks_for(cnt), 256, 0, st>>>(iv, (int)cnt);
      B.sort_pairs((const uint32_t *)I.as<int32_t>(), (uint32_t *)B.sortI.as<int32_t>(), iv, iv2, (int)cnt,
                   bits_for(T));
      gather_i32_kernel<<<blocks_for(cnt), 256, 0, st>>>(J.as<int32_t>(), iv2, (int)cnt, B.sortJ.as<int32_t>());
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(I.p, B.sortI.p, 4 * (size_t)cnt, hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(J.p, B.sortJ.p, 4 * (size_t)cnt, hipMemcpyDeviceToDevice, st));
    };
    by_owner(B.closeI, B.closeJ, nclose);
    by_owner(B.farI, B.farJ, nfar);
    // ---- results to the host ----
    std::vector<int32_t> ci((size_t)nclose), cj((size_t)nclose), fi((size_t)nfar), fj((size_t)nfar),
        pm((size_t)n);
    out.aux.resize((size_t)(6 * T));
    if (nclose) {
      HIPCHK(hipMemcpyAsync(ci.data(), B.closeI.p, 4 * (size_t)nclose, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(cj.data(), B.closeJ.p, 4 * (size_t)nclose, hipMemcpyDeviceToHost, st));
    }
    if (nfar) {
      HIPCHK(hipMemcpyAsync(fi.data(), B.farI.p, 4 * (size_t)nfar, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(fj.data(), B.farJ.p, 4 * (size_t)nfar, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipMemcpyAsync(pm.data(), perm, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(out.aux.data(), aux, 24 * (size_t)T, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    out.closeI.assign(ci.begin(), ci.end());
    out.closeJ.assign(cj.begin(), cj.end());
    out.farI.assign(fi.begin(), fi.end());
    out.farJ.assign(fj.begin(), fj.end());
    out.perm.assign(pm.begin(), pm.end());
  }
}

}  // namespace hh
