// Internal types shared by the host side (tree, lists, orchestration) and the
// HIP kernels of libhh.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hh {

// ---------------------------------------------------------------------------
// Cluster tree and block lists (host), built by hh_tree.cpp.
// ---------------------------------------------------------------------------
struct TreeLists {
  int64_t n = 0;
  int64_t T = 0;                  // tree rows (incl. the two unreachable rows 1,2)
  std::vector<uint64_t> rows;     // T x 5: c0, c1, start, end, level
  std::vector<int64_t> perm;      // tree position -> original index
  std::vector<float> aux;         // T x 6 boxes
  std::vector<int64_t> leaves;    // leaf list in traversal order
  std::vector<int64_t> closeI, closeJ;
  std::vector<int64_t> farI, farJ;
  int64_t size(int64_t nd) const { return (int64_t)(rows[5 * nd + 3] - rows[5 * nd + 2]); }
  int64_t start(int64_t nd) const { return (int64_t)rows[5 * nd + 2]; }
  bool leaf(int64_t nd) const { return rows[5 * nd] == 0; }
};

// Builds tree + boxes + lists for coordinates x (n,3, original order).
void build_tree_lists(const float *x, int64_t n, int leaf, bool use_symmetry,
                      int64_t max_rel_coord, int nthreads, TreeLists &out);

// The divide step's principal axis (cyclic Jacobi on the FP64 covariance
// [xx, yy, zz, xy, yz, zx]), shared by the host and the device builds.
void tree_principal_axis(const double cov[6], double axis[3]);

// The same tree and lists built on the device (hh_tree_gpu.hip) for one or
// more independent instances: instance k holds the particles
// [pbase[k], pbase[k+1]) of x_dev (n,3 floats, original order, device
// memory).  Node ids, particle ranges and list order are those of
// build_tree_lists per instance, merged as hh_model_create_instances merges
// them.  Device work buffers live in the builder and are reused.
struct DeviceTreeBuilder;
DeviceTreeBuilder *device_tree_builder_new();
void device_tree_builder_free(DeviceTreeBuilder *b);
void build_tree_lists_device(DeviceTreeBuilder *b, const float *x_dev,
                             const std::vector<int64_t> &pbase, int leaf, bool use_symmetry,
                             int64_t max_rel_coord, void *stream, TreeLists &out,
                             int32_t *perm_dev_out);

// ---------------------------------------------------------------------------
// Device-side work descriptors.
// ---------------------------------------------------------------------------
enum BlockFlags : int32_t {
  BF_REUSE = 1,       // also u_j += M f_i (symmetric reuse, not transposed)
  BF_SYMM = 2,        // use 0.5 (M + M^T) (dense path)
  BF_NEG = 4,         // displacement x_i - x_j instead of x_j - x_i
  BF_EXPLICIT_I = 8,  // targets come from a list; u indexed by list position
};

// A row block: targets i in [i0, i0+m) (or tlist[i0 + t]) against sources
// j in [j0, j0+n).
struct RowBlock {
  int32_t i0, m, j0, n, flags, pad0, pad1, pad2;
};

// A near task: 4 consecutive targets of a block.
struct NearTask {
  int32_t block, ioff;
};

// A far block for the ACA kernels.
struct FarBlock {
  int32_t is, m, js, n;
  int64_t coff, qoff;  // pool offsets (3x3 entries) of C_0 and Q_0
};

enum Precision : int32_t {
  PREC_FP32 = 0,        // plain: FP32 MLP (MFMA f32), FP64 ACA pools
  PREC_BF16 = 1,        // chosen: BF16 MLP (MFMA bf16, FP32 accumulate), FP32 ACA
  PREC_FP16 = 2,        // as BF16 with FP16 operands
  PREC_FP64 = 3,        // accuracy check: FP64 MLP (MFMA f64)
};

}  // namespace hh
