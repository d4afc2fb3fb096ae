/* Distilled far-field surrogate of the two-body pair function.
 *
 * No reference counterpart: the reference evaluates every ACA entry with the
 * full MLP (src/FarDot.cpp:330-470).  This is a small net
 * s(r) ~ |r| M(r), 4 -> 32 -> 32 -> 9 with tanh hidden layers, on the inputs
 * (r/|r|, ln(|r|/r0)), trained on the GPU by this library's own forward and
 * backward kernels on pairs labelled by the full MLP (hh_model_eval_pair).
 * Parameters: W1 [32][4], b1 [32], W2 [32][32], b2 [32], W3 [9][32], b3 [9].
 * Loss: mean squared error over the 9 outputs of n samples.
 * Device arrays: r (n,3) float32 displacements, targets and outputs (n,9).
 */
#ifndef HH_SURROGATE_H
#define HH_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HH_SURROGATE_PARAMS 1513

typedef struct hh_surrogate hh_surrogate;

/* A surrogate with feature scale r0 (> 0), all parameters zero (set them
 * with hh_surrogate_set_params before training). */
int hh_surrogate_create(float r0, hh_surrogate **out);
void hh_surrogate_destroy(hh_surrogate *s);

/* Host copies of the HH_SURROGATE_PARAMS parameters; setting them also resets
 * the optimiser state. */
int hh_surrogate_set_params(hh_surrogate *s, const float *params);
int hh_surrogate_get_params(hh_surrogate *s, float *params);

/* out = s(r) for n displacements (device arrays). */
int hh_surrogate_eval_device(hh_surrogate *s, const float *r_dev, int64_t n, float *out_dev);

/* Gradient of the loss on (r, targets) with respect to the parameters (host
 * array of HH_SURROGATE_PARAMS, may be NULL) and the loss (may be NULL). */
int hh_surrogate_gradient_device(hh_surrogate *s, const float *r_dev, const float *t_dev,
                                 int64_t n, float *grad, float *loss);

/* steps full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8) at learning rate
 * lr; *final_loss (may be NULL) is the loss at the last step's gradient
 * (steps == 0: the loss at the current parameters, which stay unchanged). */
int hh_surrogate_train_device(hh_surrogate *s, const float *r_dev, const float *t_dev, int64_t n,
                              int steps, float lr, float *final_loss);

#ifdef __cplusplus
}
#endif

#endif
