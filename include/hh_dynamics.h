/*
 * hh_dynamics.h -- device-resident forcing and time integration around the
 * matvec of hh.h: one integrator stage = forcing on the device, coordinate
 * update, u = M(x) f, stage combination, with no host round trip of the
 * particle arrays and no Python callback.
 *
 * The reference runs these steps through a Python velocity callback per
 * stage (include/TimeIntegrator.hpp:92-141 calling python/simulate.py:90-144
 * or python/morse_potential.py:11-23).  Conventions as in hh.h: status codes,
 * hh_last_error(), row-major (n,3) float32 arrays in the original order.
 */
#ifndef HH_DYNAMICS_H
#define HH_DYNAMICS_H

#include <stdint.h>

#include "hh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hh_forcing hh_forcing;

/* A forcing program; all terms are off until set.  f = b + bending + tension
 * + Morse bonds + Morse neighbours, each term as the reference computes it. */
int hh_forcing_create(hh_forcing **out);
void hh_forcing_destroy(hh_forcing *fc);

/* Constant body force added to every particle: `force += b` of
 * velocity_update, python/simulate.py:131-132. */
int hh_forcing_set_body(hh_forcing *fc, const float b[3]);

/* Filament chains: particles [0, n_filament * n_chain) form n_filament chains
 * of n_chain consecutive particles; chain_bending (python/simulate.py:14-53,
 * stiffness k_b) and chain_tension (python/simulate.py:55-88, stiffness k_t,
 * rest length) as added at python/simulate.py:134-140.  n_filament = 0
 * removes the term. */
int hh_forcing_set_filaments(hh_forcing *fc, int64_t n_filament, int64_t n_chain,
                             float rest_length, float k_t, double k_b);

/* Morse bonds between consecutive beads: particles [0, n_filament *
 * n_chain) form chains of n_chain; each bond (k, k+1) of a chain adds
 * -dU/dr along the bond to both beads, U(r) = D (1 - exp(-alpha (r - r0)))^2
 * (the Morse form of CalculatePotentialForce, include/PotentialForce.hpp:
 * 115-140, with the bond's own D, alpha, r0 and neighbours given by the
 * chain instead of a radius search).  Sums in FP64, added to the float force
 * once.  n_filament = 0 removes the term. */
int hh_forcing_set_bonds(hh_forcing *fc, int64_t n_filament, int64_t n_chain, double D,
                         double alpha, double r0);

/* Morse pair force over neighbours within epsilon: PotentialForce
 * (include/PotentialForce.hpp): set_two_body_epsilon (:229-231),
 * CalculatePotentialForce (:115-140, De = 1, a = 1, re = 2.5),
 * GetPotentialForce (:468-580).  epsilon <= 0 removes the term. */
int hh_forcing_set_morse(hh_forcing *fc, float epsilon);

/* set_periodic(flag) (:151-153) and set_domain(domain) (:166-185, which also
 * turns periodicity on): lo/hi are the rows of the (2,3) domain.  A periodic
 * evaluation without a domain is rejected. */
int hh_forcing_set_periodic(hh_forcing *fc, int flag);
int hh_forcing_set_domain(hh_forcing *fc, const float lo[3], const float hi[3]);

/* f = forcing(x) for device arrays x, f (n,3); stream is a hipStream_t
 * (NULL: the default stream). */
int hh_forcing_eval_device(hh_forcing *fc, const float *x_dev, float *f_dev, int64_t n,
                           void *stream);

/* Host arrays: get_potential_force(coord) (PotentialForce.hpp:468) when only
 * the Morse term is set. */
int hh_forcing_eval(hh_forcing *fc, const float *x, float *f, int64_t n);

/* update_coord (src/HignnModel.cpp:744-762) from device coordinates, on the
 * given stream.  With tree reuse on and a built tree this is one gather on the
 * device; otherwise the coordinates are copied to the host for the tree build. */
int hh_model_update_coord_device(hh_model *m, const float *coord_dev, int64_t n, void *stream);

/* Whether hh_model_dot_enqueue may run now (*capturable = 1): tree reuse on,
 * a built tree, the far field on the persistent kernel and its buffers set up
 * by an earlier dot at the current precision and iteration cap.  *signature
 * changes whenever anything such a matvec launches with changes (buffers,
 * tree, weights, settings); a graph captured with one signature is valid
 * while it holds.  Same role as Dot (src/HignnModel.cpp:765-836) split into a
 * state query and a host-wait-free enqueue; no reference counterpart of its
 * own. */
int hh_model_dot_capturable(hh_model *m, int *capturable, uint64_t *signature);

/* u = M(x) f for device arrays (n,3), enqueued on stream with no host wait,
 * timing or allocation, so that it can be captured into a HIP graph.  Fails
 * with HH_ERR_STATE unless hh_model_dot_capturable reports 1. */
int hh_model_dot_enqueue(hh_model *m, float *u_dev, const float *f_dev, int64_t n, void *stream);

typedef struct hh_stepper hh_stepper;

#define HH_STEP_EULER 0 /* ExplicitEuler::Run, include/TimeIntegrator.hpp:431-488 */
#define HH_STEP_DOPRI 1 /* ExplicitRk4::Run (Dormand-Prince 5(4)), TimeIntegrator.hpp:602-860 */

/* A device-resident integrator driving model m with forcing fc (both must
 * outlive it).  Velocity of a stage: f = forcing(x), update_coord(x), u = M f. */
int hh_stepper_create(hh_model *m, hh_forcing *fc, int method, hh_stepper **out);
void hh_stepper_destroy(hh_stepper *s);

/* set_time_step, set_final_time, set_threshold (DOPRI only). */
int hh_stepper_set(hh_stepper *s, double time_step, double final_time, double threshold);

/* set_x_lim / set_y_lim / set_z_lim (TimeIntegrator.hpp:330-361): periodic
 * wrapping of positions along axis 0..2. */
int hh_stepper_set_lim(hh_stepper *s, int axis, float lo, float hi);

/* Adds the three-body and self corrections of tb (include/hh_three_body.h)
 * to every stage velocity, as python/gravity_field.py:58-75 adds
 * velocity_update_without_2body to the matvec; NULL removes them.  tb must
 * outlive the stepper.  No reference counterpart as a native call. */
struct hh_three_body;
int hh_stepper_set_three_body(hh_stepper *s, struct hh_three_body *tb);

/* With the model's tree kept across stages (hh_model_set_tree_reuse), a full
 * tree and list rebuild at the first stage of every `every`-th accepted step
 * (0, the default: the first build is kept for the whole run).  No reference
 * counterpart (the reference rebuilds on every update_coord,
 * src/HignnModel.cpp:744-762). */
int hh_stepper_set_tree_rebuild(hh_stepper *s, int64_t every);

/* Per-stage HIP graphs for the stage combination and forcing (default on). */
int hh_stepper_set_graphs(hh_stepper *s, int flag);

/* initialize(position) (TimeIntegrator.hpp:376-406): host (n,3) positions;
 * time and counters restart at zero. */
int hh_stepper_initialize(hh_stepper *s, const float *x, int64_t n);

/* Output hook of set_output_step / set_output_func: called after accepted
 * step `step` when step % output_step == 0, with device pointers to the
 * positions and the last velocity (n,3), before periodic offsets are added
 * (as the reference's output). */
typedef void (*hh_output_fn)(void *user, int64_t step, double t, const float *x_dev,
                             const float *v_dev);
int hh_stepper_set_output(hh_stepper *s, int output_step, hh_output_fn fn, void *user);

/* run() (TimeIntegrator.hpp:431-488 / 602-860): advance until the final time
 * or until max_steps more accepted steps (max_steps <= 0: no limit). */
int hh_stepper_run(hh_stepper *s, int64_t max_steps);

/* Host copies of the positions and of the last velocity (either may be NULL);
 * t: current time; counts: [velocity evaluations, accepted, rejected steps,
 * evaluations run as one whole-stage graph (forcing and matvec together)]. */
int hh_stepper_state(hh_stepper *s, float *x, float *v, double *t, int64_t counts[4]);

#ifdef __cplusplus
}
#endif

#endif /* HH_DYNAMICS_H */
