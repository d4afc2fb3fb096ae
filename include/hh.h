/*
 * hh.h -- C ABI of libhh, the MI355X (gfx950) implementation of the H-HIGNN
 * mobility matvec u = M(x) f.
 *
 * Every entry point returns 0 on success or a negative status; the message
 * of the last failure on the calling thread is returned by hh_last_error().
 * Arrays are plain C pointers: coordinates, forces and velocities are
 * row-major (n,3) float32 in the caller's original particle order.  A model
 * handle is not thread-safe; one device stream per handle.
 *
 * Each declaration names the reference interface it replaces
 * (jonaspleyer/hignn; pybind11 module `hignn`, interface/HignnInterface.cpp).
 */
#ifndef HH_H
#define HH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HH_OK 0
#define HH_ERR_ARG (-1)     /* invalid argument or shape */
#define HH_ERR_STATE (-2)   /* call not valid in the current state */
#define HH_ERR_DEVICE (-3)  /* HIP runtime error */
#define HH_ERR_ALLOC (-4)   /* device or host allocation failed */

/* precision modes (hh_model_set_precision) */
#define HH_PREC_FP32 0 /* plain: FP32 MLP, FP64 ACA */
#define HH_PREC_BF16 1 /* BF16 MLP operands, FP32 accumulate; ACA arithmetic in FP32 on
                          factors stored in BF16 (FP32 with a surrogate attached) */
#define HH_PREC_FP16 2 /* FP16 MLP operands, otherwise as HH_PREC_BF16 */

typedef struct hh_model hh_model;

const char *hh_last_error(void);

/* Init() -- src/HignnModel.cpp:16-42 (MPI + Kokkos init; here: selects the
 * HIP device and queries its properties). */
int hh_init(void);
/* Finalize() -- src/HignnModel.cpp:48-52. */
int hh_finalize(void);

/* HignnModel(coord, blockSize) -- src/HignnModel.cpp:335-386. */
int hh_model_create(const float *coord, int64_t n, int block_size, hh_model **out);

/* Many independent suspensions in one model (no reference counterpart; the
 * reference runs one system per HignnModel).  Instance k holds rows
 * [o_k, o_k + counts[k]) of coord (o_k the sum of the earlier counts); each
 * instance gets its own cluster tree and close/far lists (built on host
 * threads, one instance per thread), and no block couples two instances.
 * The merged lists run as one near-field launch and one far field over all
 * instances, so u of every instance equals that of its own model.  Every
 * other call works as for a single system on the concatenated arrays
 * (tree export: instance k's node ids follow those of instances 0..k-1). */
int hh_model_create_instances(const float *coord, const int64_t *counts, int64_t ninst,
                              int block_size, hh_model **out);
void hh_model_destroy(hh_model *m);

/* load_two_body_model(path) -- src/HignnModel.cpp:402-433.  The host layer
 * resolves the file; this call takes the 71,177 raw float32 weights
 * W1(128,3) b1(128) W2(512,128) b2(512) W3(9,512) b3(9). */
int hh_model_load_two_body(hh_model *m, const float *weights, int64_t count);

/* Setters -- src/HignnModel.cpp:773-874 (interface/HignnInterface.cpp:21-30). */
int hh_model_set_epsilon(hh_model *m, double epsilon);
int hh_model_set_eta(hh_model *m, double eta); /* stored, not used (as the reference) */
int hh_model_set_max_iter(hh_model *m, int max_iter);
int hh_model_set_mat_pool_size_factor(hh_model *m, int factor);
int hh_model_set_max_far_dot_work_node_size(hh_model *m, int size);
int hh_model_set_max_relative_coord(hh_model *m, int64_t size);
int hh_model_set_post_check_flag(hh_model *m, int flag);
int hh_model_set_use_symmetry_flag(hh_model *m, int flag);
int hh_model_set_max_far_field_distance(hh_model *m, double distance); /* stored, unused */

/* update_coord(coord) -- src/HignnModel.cpp:744-762: new coordinates, then
 * tree + close/far lists (Build + CloseFarCheck). */
int hh_model_update_coord(hh_model *m, const float *coord, int64_t n);

/* Opt-in tree reuse (no counterpart in the reference, which rebuilds on every
 * update_coord): with flag != 0, update_coord after a first build keeps the
 * permutation, tree and close/far lists and only refreshes the positions --
 * for the stages of one integrator step, where particles move little.  The
 * caller decides when to rebuild (flag 0, then update_coord). */
int hh_model_set_tree_reuse(hh_model *m, int flag);

/* Where update_coord builds the tree and lists (the reference builds them on
 * the host, src/HignnModel.cpp:457-461): 0 on host threads, 1 on the device
 * (the default; see hh_tree_build_device: the same result bit for bit).
 * HH_TREE_BUILD=host in the environment makes new models build on the host. */
int hh_model_set_tree_build(hh_model *m, int device);

/* dot(u, f) -- src/HignnModel.cpp:545-634: u = M(x) f, host arrays. */
int hh_model_dot(hh_model *m, float *u, const float *f, int64_t n);

/* dense_dot(u, f) -- src/HignnModel.cpp:660-728 with src/DenseDot.cpp. */
int hh_model_dense_dot(hh_model *m, float *u, const float *f, int64_t n);

/* Device-resident dot: f_dev and u_dev are device pointers (n,3) float32 in
 * the original order; stream is a hipStream_t (NULL: the model's stream).
 * Same arithmetic as hh_model_dot, without host transfers. */
int hh_model_dot_device(hh_model *m, float *u_dev, const float *f_dev, int64_t n, void *stream);

/* Precision of dot (no counterpart in the reference, which is FP32 only). */
int hh_model_set_precision(hh_model *m, int prec);

/* Tree and block lists of the last update_coord (the reference keeps them in
 * mClusterTreePtr, mReorderedMap, mCloseMatI/J, mFarMatI/J, mLeafNodeList,
 * include/HignnModel.hpp:54-75).  counts: [tree rows, close pairs, far pairs,
 * leaves].  Any output pointer may be NULL. */
int hh_model_tree_counts(hh_model *m, int64_t counts[4]);
int hh_model_tree_export(hh_model *m, uint64_t *tree5, uint64_t *perm, float *aux6,
                         int64_t *close_ij, int64_t *far_ij, int64_t *leaves);

/* Final ACA rank of every far block in the last dot (far-list order). */
int hh_model_far_ranks(hh_model *m, int32_t *ranks);

/* PostCheckDot -- src/PostCheckDot.cpp:3-377: FP64 evaluation of u at the
 * given target particles (original indices).  convention 0: the H-matrix's
 * own block convention with exact far blocks; 1: ordered-pair dense with
 * symmetrised blocks (src/DenseDot.cpp:230-246). */
int hh_model_check_fp64(hh_model *m, const float *f, const int64_t *targets, int64_t nt,
                        int convention, double *u_out);

/* Pair function on the device: out9[p] = M(r[p]) (row-major 3x3) for P
 * displacements r (host arrays), i.e. the TorchScript Net.forward of
 * python/convert.py:11-20 as called at src/CloseDot.cpp:233.  prec:
 * HH_PREC_FP32/BF16/FP16; 3 = FP64 (the result is rounded to float). */
int hh_model_eval_pair(hh_model *m, const float *r, int64_t P, float *out9, int prec);

/* Host-only cluster tree and block lists (Build + CloseFarCheck,
 * src/Build.cpp:17-265, src/CloseFarCheck.cpp:11-465) for coordinates
 * x (n,3); no device is touched.  Export as hh_model_tree_export. */
typedef struct hh_tree hh_tree;
int hh_tree_build(const float *x, int64_t n, int block_size, int use_symmetry,
                  int64_t max_relative_coord, hh_tree **out);
int hh_tree_counts(hh_tree *t, int64_t counts[4]);
int hh_tree_export(hh_tree *t, uint64_t *tree5, uint64_t *perm, float *aux6, int64_t *close_ij,
                   int64_t *far_ij, int64_t *leaves);

/* The same tree and lists built on the device (Build + CloseFarCheck,
 * src/Build.cpp:17-265, src/CloseFarCheck.cpp:181-454, as HignnModel::Update
 * runs them, src/HignnModel.cpp:457-461) for ninst independent systems whose
 * particles are consecutive rows of x_dev ((sum counts, 3) floats, device
 * memory): instance k's node ids follow those of instances 0..k-1 and its
 * particle ranges theirs, as hh_model_create_instances merges them.  With
 * ninst = 1 the result equals hh_tree_build's bit for bit.  stream may be
 * NULL (the null stream). */
int hh_tree_build_device(const float *x_dev, const int64_t *counts, int64_t ninst, int block_size,
                         int use_symmetry, int64_t max_relative_coord, void *stream, hh_tree **out);
void hh_tree_free(hh_tree *t);

/* Device times (ms) of the last dot: [close, far, total], total from the
 * force gather to the velocity scatter (the near and far fields may run at
 * the same time on two streams, so total <= close + far). */
int hh_model_last_timings(hh_model *m, double t[3]);

/* Pair-function kernels of the last dot (device time in ms from HIP events
 * on the model's stream, pair-function evaluations, launches):
 * [0] the near-field kernel, [1] the ACA column/row passes of the far field.
 * No reference counterpart (the reference times phases with Kokkos timers,
 * src/FarDot.cpp:74-79); used for roofline accounting. */
int hh_model_last_profile(hh_model *m, double ms[2], int64_t pairs[2], int64_t launches[2]);

/* Far-field launch tuning on the caller's forces f (device array (n,3),
 * stream as hh_model_dot_device).  Candidates: the persistent far-field
 * kernel (only with at most 4 far blocks per compute unit) and the batched
 * ACA iterations with step chunks of 64, 128 and 256 items.  Each runs one
 * untimed and one timed matvec; the fastest is kept for the model's later
 * matvecs.  table receives up to
 * max_rows rows of {mode (0 persistent, 1 batched), chunk, matvec ms},
 * *rows their number.  Every candidate computes the same matvec up to
 * rounding.  No reference counterpart (the reference has fixed launch
 * shapes, src/FarDot.cpp:64-120). */
int hh_model_tune_far_field(hh_model *m, const float *f_dev, int64_t n, void *stream,
                            double *table, int max_rows, int *rows);

/* Tile shape of the 16-bit (BF16/FP16) near-field kernel: index 0..3 of
 * {16-pair tiles per MLP call, threads per persistent workgroup} =
 * {4, 512}, {2, 512}, {2, 1024}, {4, 256}; default 0 (HH_NEAR_SHAPE sets the
 * default of new models).  hh_model_tune_near times launch choices on the
 * caller's forces (device array f_dev, (n,3) float32; one untimed and one
 * timed matvec each, since the near field shares the GPU with the far field):
 * every shape on one workgroup per compute unit and, when the far field runs
 * beside the near field, shapes 0..2 on 5/8, 1/2 and 3/8 of the compute
 * units; it keeps the fastest shape and workgroup count.  table receives up
 * to max_rows rows of {shape, tiles, threads, workgroups, matvec ms} (13
 * rows at most), *rows their number (0 in the FP32 mode, whose kernel has one
 * shape).  Every shape computes the same sums up to rounding.  No reference
 * counterpart: the reference's close-field launch is fixed
 * (src/CloseDot.cpp:25-337). */
int hh_model_set_near_shape(hh_model *m, int shape);

/* Workgroups of the persistent 16-bit near kernel: 0 (the default;
 * HH_NEAR_WORKGROUPS sets the default of new models) or any count at or above
 * the number of compute units gives one per compute unit (its LDS holds one
 * per compute unit); fewer leave whole compute units to the far-field
 * kernels that run beside it.  The workgroups take their tasks from a shared
 * queue, so every count computes the same sums up to rounding.  Negative
 * counts are rejected.  No reference counterpart. */
int hh_model_set_near_workgroups(hh_model *m, int workgroups);

/* Near field beside the batched far field (1, the default; HH_NEAR_OVERLAP=0
 * in the environment makes new models start with 0): the near kernel runs on
 * a second stream beside the ACA iterations (with a surrogate: beside those
 * of the blocks on surrogate entries); 0 runs the near field first and alone,
 * then the far field, e.g. to time the near kernel by itself.  The
 * persistent far kernel always runs beside the near field.  Same result up to the
 * order of the velocity sums.  No reference counterpart: the reference runs
 * CloseDot and FarDot one after the other (src/HignnModel.cpp:545-634). */
int hh_model_set_near_overlap(hh_model *m, int on);
int hh_model_tune_near(hh_model *m, const float *f_dev, int64_t n, void *stream, double *table,
                       int max_rows, int *rows);

#ifdef __cplusplus
}
#endif

#endif /* HH_H */
