/* Distilled far-field surrogate of the two-body pair function.
 *
 * No reference counterpart: the reference evaluates every ACA column and row
 * entry with the full MLP (src/FarDot.cpp:346-458 and 667-782).  Here a
 * small net generates those entries for the far blocks it is good enough
 * for; the full MLP stays the near-field path, judges the surrogate per far
 * block on sampled entries, and is the fallback for blocks that fail.
 *
 * The net: s(r) ~ |r| S(M(r)) on the inputs z = r / r_hi, 3 -> 32 -> 32 -> 6
 * with SiLU hidden layers (silu x = x / (1 + e^-x)); the outputs are the
 * entries (xx, yy, zz, xy, yz, zx) of the symmetrised block S(M) (M with
 * its off-diagonal pairs averaged, as the ACA uses it) times |r|.
 * Parameters: W1 [32][3], b1 [32], W2 [32][32], b2 [32], W3 [6][32], b3 [6].
 * Loss: mean over samples of sum_k w_k (s_k - t_k)^2, w = (1,1,1,2,2,2),
 * i.e. the squared Frobenius error of the 3x3 block times |r|^2.
 * It is trained on the GPU by this library's own forward and backward
 * kernels (minibatch Adam), on displacements labelled by the full MLP.
 *
 * Device arrays: r (n,3) float32 displacements, targets and outputs (n,6).
 * stream is a hipStream_t (NULL: the default stream); every call returns
 * after its work on that stream is complete.
 */
#ifndef HH_SURROGATE_H
#define HH_SURROGATE_H

#include <stdint.h>

#include "hh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HH_SURROGATE_PARAMS 1382

typedef struct hh_surrogate hh_surrogate;

/* A surrogate for displacements with r_lo <= |r| <= r_hi (0 < r_lo < r_hi;
 * inputs are scaled by 1/r_hi), all parameters zero. */
int hh_surrogate_create(float r_lo, float r_hi, hh_surrogate **out);
void hh_surrogate_destroy(hh_surrogate *s);
int hh_surrogate_range(hh_surrogate *s, float *r_lo, float *r_hi);

/* Host copies of the HH_SURROGATE_PARAMS parameters; setting them also
 * resets the optimiser state. */
int hh_surrogate_set_params(hh_surrogate *s, const float *params);
int hh_surrogate_get_params(hh_surrogate *s, float *params);

/* out = s(r) for n displacements. */
int hh_surrogate_eval_device(hh_surrogate *s, const float *r_dev, int64_t n, float *out_dev,
                             void *stream);

/* Gradient of the loss on (r, targets) with respect to the parameters (host
 * array of HH_SURROGATE_PARAMS, may be NULL) and the loss (may be NULL).
 * Block partial sums are added in a fixed order: the result does not depend
 * on scheduling. */
int hh_surrogate_gradient_device(hh_surrogate *s, const float *r_dev, const float *t_dev,
                                 int64_t n, float *grad, float *loss, void *stream);

/* As hh_surrogate_gradient_device, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight training loops).  grad_dev (HH_SURROGATE_PARAMS
 * floats) and loss_dev (1 float) are caller-owned device outputs; the
 * gradient pass's per-block partial sums live in a fixed 2.8 MB device
 * workspace of the library (512 blocks x (HH_SURROGATE_PARAMS + 1) floats,
 * one per GPU, kept until the process ends).  Every call on a GPU shares
 * that workspace: calls that overlap on one GPU (different streams or
 * threads) race.  The work is enqueued on stream and not waited for; the
 * result equals hh_surrogate_gradient_device's bit for bit.  n must be in
 * 1..HH_SURROGATE_NOALLOC_MAX_SAMPLES. */
#define HH_SURROGATE_NOALLOC_MAX_SAMPLES 2147483647LL
int hh_surrogate_gradient_noalloc(hh_surrogate *s, const float *r_dev, const float *t_dev,
                                  int64_t n, float *grad_dev, float *loss_dev, void *stream);

/* steps minibatch Adam steps (beta 0.9 / 0.999, eps 1e-8) on consecutive
 * batches of `batch` samples (wrapping round; batch <= 0 or > n: all n), the
 * learning rate following a cosine from lr0 to lr1 over the call's steps.
 * *final_loss (may be NULL) is the loss of the last step's batch; with
 * steps == 0 it is the loss over all n samples at the current parameters,
 * which stay unchanged. */
int hh_surrogate_train_device(hh_surrogate *s, const float *r_dev, const float *t_dev, int64_t n,
                              int steps, int64_t batch, float lr0, float lr1, float *final_loss,
                              void *stream);

/* Training samples from the model's full pair function: n displacements with
 * |r| log-uniform in [r_lo, r_hi] and uniform directions (counter-based
 * random numbers from seed), and their targets |r| S(M(r)) from the FP32
 * MLP.  r_dev (n,3) and t_dev (n,6) are caller-owned device arrays. */
int hh_model_surrogate_samples(hh_model *m, int64_t n, float r_lo, float r_hi, uint64_t seed,
                               float *r_dev, float *t_dev, void *stream);

/* Attaches a trained surrogate to the model's far field (NULL detaches; the
 * caller keeps it alive while attached).  Before each far field the full
 * FP32 MLP judges it on `samples` entries of every far block (sample pairs
 * spread over the block by a fixed hash): a block whose sampled relative
 * Frobenius error exceeds tol, or with a sampled |r| outside the surrogate's
 * range, runs its ACA on full-MLP entries; every other block's ACA column
 * and row entries come from the surrogate.  With a surrogate attached the
 * far field always runs the batched ACA iterations. */
int hh_model_set_surrogate(hh_model *m, hh_surrogate *s, double tol, int samples);

/* Of the last matvec: [far blocks judged, blocks on surrogate entries,
 * blocks that fell back to the full MLP, surrogate entries evaluated,
 * full-MLP far entries evaluated]. */
int hh_model_surrogate_stats(hh_model *m, int64_t stats[5]);

/* Sampled relative error of the surrogate per far block (far-list order) in
 * the last matvec's judge; +inf for a block with a sampled |r| outside the
 * surrogate's range. */
int hh_model_surrogate_errors(hh_model *m, float *err);

#ifdef __cplusplus
}
#endif

#endif
