/*
 * allium_device.h -- device-memory plumbing of the allium MI355X library.
 *
 * No counterpart in the reference (whose vectors live in host memory): these
 * calls let callers without their own HIP runtime bindings (the Python tests
 * and benchmark) place vectors in HBM and time work on the device.
 */
#ifndef ALLIUM_DEVICE_H
#define ALLIUM_DEVICE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif
int allium_device_count(int *count);
int allium_device_name(char *buf, int len);
int allium_device_malloc(void **ptr, size_t bytes);
int allium_device_free(void *ptr);
int allium_memcpy_h2d(void *dst, const void *src, size_t bytes);
int allium_memcpy_d2h(void *dst, const void *src, size_t bytes);
int allium_memcpy_d2d(void *dst, const void *src, size_t bytes);
int allium_device_synchronize(void);
/* op 0 records a start event on the null stream; op 1 records a stop event,
 * waits for it and returns the elapsed milliseconds in *ms */
int allium_event_timer(int op, double *ms);
#ifdef __cplusplus
}
#endif
#endif
