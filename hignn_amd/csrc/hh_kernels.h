// Launch descriptors of the HIP kernels in hh_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hh_internal.h"

namespace hh {

// Device copies of the pair-function weights in MFMA fragment order.
struct DevWeightsF {  // FP32 or FP64 elements
  const void *w1b, *w2f, *b2, *w3f, *b3;
};
struct DevWeightsH {  // 16-bit fragments + FP32 layer 1 and biases
  const float *w1b, *b2, *b3;
  const void *w2f, *w3f;
};

// Tile shapes of the 16-bit near kernel: 16-pair tiles per MLP call and
// threads per persistent workgroup (index NearLaunch::hcfg).
struct NearShape {
  int tiles, threads;
};
constexpr NearShape kNearShapes[] = {{4, 512}, {2, 512}, {2, 1024}, {4, 256}};
constexpr int kNearShapeCount = 4;

struct NearLaunch {
  int prec;
  int hcfg;        // 16-bit kernel tile shape (kNearShapes)
  const float4 *xp;
  const float4 *fp;
  const int32_t *tlist;
  const RowBlock *blocks;
  const NearTask *tasks;
  int64_t ntasks;
  double *u;
  DevWeightsF wf;
  DevWeightsH wh;
  int max_blocks;  // grid cap for the FP32/FP64 kernels
  int cus;         // compute units (grid of the LDS-resident 16-bit kernel)
  int32_t *next;   // work-queue counter of the 16-bit kernel (zeroed by the
                   // launch, on its stream), or null: round-robin tasks
};

// Counts of one batched ACA iteration.  They live in device memory, so the
// kernels run on counts that either the host uploaded or the device plan of
// the iteration computed (no host round trip per iteration); launch grids
// come from host-side upper bounds and the kernels loop grid-stride.
struct AcaCounts {
  int32_t na;           // active blocks (slots)
  int32_t nl;           // blocks to apply (stopped in the previous iteration)
  int32_t ntc[2];       // column-pass 16-item tiles: [0] full MLP, [1] surrogate
  int32_t ntr[2];       // row-pass tiles
  int32_t ncc, ncr;     // step chunks over I / J of the active blocks
  int32_t nci, ncj;     // apply chunks over I / J of the blocks to apply
  int64_t ysum_c, ysum_r;  // items of the active blocks (rows of Y per pass)
  int64_t entries[2];   // running totals of pass entries: [0] full MLP, [1] surrogate
};

// Tile lists of one ACA pass: every active slot writes its ceil(items/16)
// tiles at toff[slot] of the list of its kind (full-MLP or surrogate
// entries).  Column pass: pivot x[js + jpiv[k]], items x[is + i]; row pass:
// pivot x[is + ipiv[k]], items x[js + j].
struct AcaExpandLaunch {
  const FarBlock *blk;
  const int32_t *active;
  const int32_t *it;
  const int32_t *ipiv;
  const int32_t *jpiv;
  const int64_t *ybase;
  const int32_t *toff;
  const int32_t *kind;    // per batch block: 1 = surrogate entries; NULL: all full MLP
  const AcaCounts *cnt;
  int64_t grid;           // slots bound (grid size)
  int max_iter;
  int row_pass;
  void *tiles_full, *tiles_sur;  // 24 B per tile
};

struct AcaMlpLaunch {
  int prec;
  const float4 *xp;
  const void *tiles;      // resolved tiles (24 B each)
  const int32_t *dntiles; // device count of the tiles
  int64_t ntiles;         // bound of the count (grid size)
  int row_pass;
  float *Y;
  DevWeightsF wf;
  DevWeightsH wh;
  int max_blocks;
  int cus;
};

// Far-field ACA steps and apply work on chunks of at most kAcaChunk items of
// a block (rows of I or columns of J); kernels find a chunk's (slot, first
// item) from per-slot prefix counts.
constexpr int kAcaChunk = 256;
constexpr size_t kAcaPartBytes = 64;  // per-chunk partials (arg-max + 6 Gram sums)

struct AcaStepLaunch {
  const FarBlock *blk;
  const int32_t *active;
  const int64_t *ybase_c, *ybase_r;
  const float *Y;
  int32_t *it;
  int32_t *ipiv;
  int32_t *jpiv;
  double *mu2;
  double *nu2;
  int32_t *stop;
  int max_iter;
  double eps2;
  const int32_t *cfirst_c, *cfirst_r;  // slot -> first column-side (I) / row-side (J) chunk
                                       // (nactive + 1 entries)
  void *part_c, *part_r;            // kAcaPartBytes per chunk
  double *x_c, *x_r;                // max_iter x 9 per chunk
  double *inv;                      // 9 per slot
  int chunk;                        // items per chunk, 1..kAcaChunk
  const AcaCounts *cnt;             // na, ncc, ncr
  int64_t nactive;                  // bounds of the counts (grid sizes)
  int64_t nchunks_c, nchunks_r;
  int64_t grid_cap;                 // at most this many workgroups per kernel
  int pool_f64;
  int pool_c16;  // BF16 C pool (16-bit modes)
  int pool_q16;  // BF16 Q pool (16-bit modes)
  void *C;
  void *Q;
};

struct AcaApplyLaunch {
  const FarBlock *blk;
  const int32_t *list;
  const int32_t *it;
  const float4 *fp;
  double *u;
  int use_symmetry;
  const AcaCounts *cnt;             // nl, nci, ncj
  int64_t nlist;                    // bounds of the counts (grid sizes)
  int64_t grid_cap;
  int pool_f64;
  int pool_c16;
  int pool_q16;
  const void *C;
  const void *Q;
  const int32_t *first_i, *first_j;  // apply slot -> first chunk over I / J (nlist + 1)
  int64_t nchunks_i, nchunks_j;
  double *w_i, *w_j;                // max_iter x 3 per chunk
  int max_iter;
  int chunk;                        // items per chunk, 1..kAcaChunk
};

struct PAcaLaunch {
  int prec;
  const float4 *xp;
  const float4 *fp;
  double *u;
  const FarBlock *blk;
  const int32_t *order;
  int32_t *rank_out;
  int64_t nblocks;
  int max_iter;
  double eps2;
  int use_symmetry;
  void *poolC, *poolQ;
  int64_t slotC, slotQ;
  float *Y;
  int64_t slotY;
  int nslots;
  DevWeightsF wf;
  DevWeightsH wh;
};

// Surrogate entries for the surrogate tile list of an ACA pass (blocks that
// passed the judge).
struct AcaSurLaunch {
  const float4 *xp;
  const void *tiles;
  const int32_t *dntiles;  // device count of the tiles
  int64_t ntiles;          // bound of the count (grid size)
  int row_pass;
  float *Y;
  const float *params;
  float inv_scale;
  int cus;
};

// Surrogate judge of nblk far blocks: samples pairs per block, the full FP32
// MLP at them (r: nblk*samples*3 floats, m9: nblk*samples*9 floats of
// workspace), flag[b] = 1 when the surrogate passes, err[b] its sampled
// relative error.
struct JudgeLaunch {
  const FarBlock *blk;
  int64_t nblk;
  const float4 *xp;
  int samples;
  float *r, *m9, *s6;  // workspace: samples' displacements, full-MLP blocks, surrogate outputs
  const float *params;
  float r_lo, r_hi;
  double tol;
  int32_t *flag;
  float *err;
  DevWeightsF wf;
  DevWeightsH wh;
  int cus;
};

hipError_t launch_aca_sur(const AcaSurLaunch &L, hipStream_t st);
hipError_t launch_judge(const JudgeLaunch &L, hipStream_t st);
// Surrogate training samples and their targets |r| S(M) (n,6) from blocks m9.
hipError_t launch_sur_samples(int64_t n, float r_lo, float r_hi, uint64_t seed, float *r,
                              hipStream_t st);
hipError_t launch_sur_targets(const float *r, const float *m9, int64_t n, float *t,
                              hipStream_t st);
hipError_t launch_persistent_aca(const PAcaLaunch &L, hipStream_t st);
hipError_t launch_near(const NearLaunch &L, hipStream_t st);
// Device plan of one batched ACA iteration: from the previous active list
// and the stop flags its finalize step wrote (or, for the first plan, all nb
// blocks of the batch), the new active list with its per-slot prefix counts
// (Y rows, tile offsets per kind, step chunks) and the list of blocks that
// stopped, with their apply chunk offsets; counts into cnt_out.  The order of
// the input list is kept.  Three launches (group sums, one scan of the
// groups, ordered write), no host involvement.
constexpr int kPlanGroups = 256;
constexpr int kPlanFields = 14;
struct AcaPlanLaunch {
  const FarBlock *blk;
  const int32_t *kind;     // per batch block: 1 = surrogate entries; NULL: all full MLP
  const int32_t *act_in;   // previous active list; NULL: blocks 0..nb-1
  const int32_t *stop_in;  // per previous slot; NULL: none stopped
  const AcaCounts *cnt_in; // previous counts; NULL: first plan (nb blocks, zero totals)
  int select_kind;         // first plan: only blocks of this kind (-1: all blocks)
  int nb;
  int chunk;
  int32_t *act_out;
  int64_t *ybase_c, *ybase_r;
  int32_t *toff_c, *toff_r;
  int32_t *cfirst_c, *cfirst_r;  // na + 1 entries
  int32_t *alist, *first_i, *first_j;  // nl (+ 1) entries
  AcaCounts *cnt_out;
  int64_t *part;           // kPlanGroups x kPlanFields workspace
};
hipError_t launch_aca_plan(const AcaPlanLaunch &L, hipStream_t st);
// Batch start: it = 0, mu2 = nu2 = 0, jpiv[b][0] = n_b / 2 for nb blocks.
hipError_t launch_aca_init(const FarBlock *blk, int nb, int max_iter, int32_t *it, int32_t *jpiv,
                           double *mu2, double *nu2, hipStream_t st);
// acc[0..1] += cnt->entries[0..1] (entries per kind of a finished batch).
hipError_t launch_aca_entries(const AcaCounts *cnt, int64_t *acc, hipStream_t st);
// Final ranks of a batch into the per-far-block rank array; with kind, only
// the blocks of kind select_kind.
hipError_t launch_copy_ranks(const int32_t *it, int nb, int32_t *rank_out, const int32_t *kind,
                             int select_kind, hipStream_t st);
hipError_t launch_aca_expand(const AcaExpandLaunch &L, hipStream_t st);
hipError_t launch_aca_mlp(const AcaMlpLaunch &L, hipStream_t st);
// which = 0: column residual, row pivot, column normalisation;
// which = 1: row residual, column pivot, norms and stop test.
hipError_t launch_aca_steps(const AcaStepLaunch &L, int which, hipStream_t st);
hipError_t launch_aca_apply(const AcaApplyLaunch &L, hipStream_t st);
hipError_t launch_pair(int prec, const float *r, int64_t P, float *out, DevWeightsF wf,
                       DevWeightsH wh, int cus, hipStream_t st);
// Copies n16 16-byte vectors with a kernel.
hipError_t launch_copy16(void *dst, const void *src, int64_t n16, hipStream_t st);
// Clears n doubles with a kernel (usable inside captured graphs).
hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t st);
hipError_t launch_gather_forces(const float *f, int64_t n, const int32_t *perm, float4 *fp,
                                hipStream_t st);
hipError_t launch_scatter_velocity(const double *u, int64_t n, const int32_t *perm, float *out,
                                   hipStream_t st);

}  // namespace hh
