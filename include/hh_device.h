/*
 * hh_device.h -- device memory of libhh's own HIP runtime.
 *
 * No reference counterpart (the reference holds device arrays as Kokkos
 * Views).  For callers that have no other GPU framework in the process:
 * buffers for the device-pointer entry points (hh_model_dot_device,
 * hh_surrogate_*), filled and read back through the copies below.  Returns
 * 0 or a negative status as include/hh.h.
 */
#ifndef HH_DEVICE_H
#define HH_DEVICE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int hh_device_alloc(size_t bytes, void **p);
int hh_device_free(void *p);
/* Synchronous copies (all earlier work on the device completes first). */
int hh_memcpy_to_device(void *dst_dev, const void *src_host, size_t bytes);
int hh_memcpy_to_host(void *dst_host, const void *src_dev, size_t bytes);
int hh_device_synchronize(void);
/* Device-to-device copy of bytes (a multiple of 16) by a copy kernel on the
 * default stream, synchronous: a transfer of known size for calibrating HBM
 * byte counters. */
int hh_device_copy(void *dst_dev, const void *src_dev, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* HH_DEVICE_H */
