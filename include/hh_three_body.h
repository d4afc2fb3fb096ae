/*
 * hh_three_body.h -- neighbour lists and the three-body / self corrections of
 * the H-HIGNN velocity, on the device.  Conventions as in hh.h.
 *
 * The reference builds the lists on the host (NeighborLists,
 * include/NeighborLists.hpp, bound at interface/HignnInterface.cpp:73-91)
 * and evaluates the corrections with torch
 * (python/HIGNN/model_structure.py:74-140, python/gravity_field.py:34-59).
 */
#ifndef HH_THREE_BODY_H
#define HH_THREE_BODY_H

#include <stdint.h>

#include "hh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hh_neighbors hh_neighbors;

/* NeighborLists() -- NeighborLists.hpp:218-223 (radii 0, not periodic). */
int hh_neighbors_create(hh_neighbors **out);
void hh_neighbors_destroy(hh_neighbors *nb);

/* set_two_body_epsilon / set_three_body_epsilon -- :281-287.  The two-body
 * radius is also the margin of the periodic images (:61). */
int hh_neighbors_set_two_body_epsilon(hh_neighbors *nb, float epsilon);
int hh_neighbors_set_three_body_epsilon(hh_neighbors *nb, float epsilon);
/* set_periodic / set_domain -- :228-250 (set_domain turns periodicity on). */
int hh_neighbors_set_periodic(hh_neighbors *nb, int flag);
int hh_neighbors_set_domain(hh_neighbors *nb, const float lo[3], const float hi[3]);

/* update_coord(coord) -- :252-279: host (n,3) positions. */
int hh_neighbors_update_coord(hh_neighbors *nb, const float *x, int64_t n);

/* build_two_body_info -- :466-510: edges (i, j) for the neighbours j of
 * every target i within the two-body radius. */
int hh_neighbors_build_two_body(hh_neighbors *nb, int64_t *num_edges);
/* build_three_body_info -- :512-588: edges (i, n, p), n a neighbour of i and
 * p a neighbour of n other than i, and self edges (i, n), within the
 * three-body radius. */
int hh_neighbors_build_three_body(hh_neighbors *nb, int64_t *num_edges, int64_t *num_self_edges);

/* get_two_body_edge_info (:289-307), get_three_body_edge_info (:342-360),
 * get_three_body_edge_self_info (:405-422): row-major (2,E) / (3,E) / (2,E)
 * index arrays.  Edges of one target are contiguous and targets ascend; inside
 * a target the order is ascending by neighbour (the reference's is its
 * kd-tree's visiting order). */
int hh_neighbors_two_body_edges(hh_neighbors *nb, int64_t *out);
int hh_neighbors_three_body_edges(hh_neighbors *nb, int64_t *out);
int hh_neighbors_three_body_self_edges(hh_neighbors *nb, int64_t *out);

typedef struct hh_three_body hh_three_body;

int hh_three_body_create(hh_three_body **out);
void hh_three_body_destroy(hh_three_body *tb);

/* Weights of Three_body_net and Two_body_self_net
 * (python/HIGNN/model_structure.py:41-68; nets of
 * python/Saved_Model/<run>/HIGNN_nn_3body.pkl and HIGNN_nn_self.pkl): flat float32
 * W1(256,d) b1(256) W2(1024,256) b2(1024) W3(512,1024) b3(512) W4(9,512) b4(9)
 * with d = 6 (three-body, 794,377 floats) and d = 3 (self, 793,609 floats). */
int hh_three_body_load(hh_three_body *tb, const float *w3, int64_t n3, const float *wself,
                       int64_t nself);

/* Three-body radius (the cutoff passed to set_three_body_epsilon,
 * python/gravity_field.py:18-22, 36). */
int hh_three_body_set_epsilon(hh_three_body *tb, float epsilon);
/* HH_PREC_FP32 (FP32 GEMMs) or HH_PREC_FP16 (FP16 operands, FP32 accumulate). */
int hh_three_body_set_precision(hh_three_body *tb, int prec);

/* velocity_update_without_2body (python/gravity_field.py:34-59): u = v_3body
 * + v_self for device arrays x, f, u (n,3); accumulate != 0 adds to u. */
int hh_three_body_velocity_device(hh_three_body *tb, const float *x_dev, const float *f_dev,
                                  int64_t n, float *u_dev, int accumulate, void *stream);
/* The same for host arrays (u is overwritten). */
int hh_three_body_velocity(hh_three_body *tb, const float *x, const float *f, int64_t n,
                           float *u);

/* Edges of the last evaluation ([three-body, self]) and the device time (ms)
 * of its MLP kernels ([three-body, self]); no reference counterpart. */
int hh_three_body_last_profile(hh_three_body *tb, int64_t edges[2], double ms[2]);

#ifdef __cplusplus
}
#endif

#endif /* HH_THREE_BODY_H */
