/*
 * host_batch.h -- paged pinned host batches.
 *
 * Drop-in for nahmedraja/GASAL2 src/host_batch.h:9-17 (implementation
 * src/host_batch.cpp).  Sequences are appended to a linked list of pinned
 * pages and each is padded with N_CODE to a multiple of 8 bytes.
 */
#ifndef __HOST_BACTH_H__
#define __HOST_BACTH_H__

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gasal.h"

#ifdef __cplusplus
extern "C" {
#endif
/* src/host_batch.cpp:11-22 */
host_batch_t *gasal_host_batch_new(uint32_t batch_bytes, uint32_t offset);
/* src/host_batch.cpp:24-41 */
void gasal_host_batch_destroy(host_batch_t *res);
/* src/host_batch.cpp:43-47 */
host_batch_t *gasal_host_batch_getlast(host_batch_t *arg);
/* src/host_batch.cpp:49-75 */
void gasal_host_batch_reset(gasal_gpu_storage_t *gpu_storage);
/* src/host_batch.cpp:79-153: copies `size` bytes at batch index idx, pads with
 * N_CODE to a multiple of 8, returns the next free index. */
uint32_t gasal_host_batch_fill(gasal_gpu_storage_t *gpu_storage, uint32_t idx, const char *data, uint32_t size,
                               data_source SRC);
/* src/host_batch.cpp:162-223: raw append without padding. */
uint32_t gasal_host_batch_add(gasal_gpu_storage_t *gpu_storage, uint32_t idx, const char *data, uint32_t size,
                              data_source SRC);
/* src/host_batch.cpp:156-159 */
uint32_t gasal_host_batch_addbase(gasal_gpu_storage_t *gpu_storage, uint32_t idx, const char base, data_source SRC);
/* src/host_batch.cpp:228-232 */
void gasal_host_batch_print(host_batch_t *res);
/* src/host_batch.cpp:235-244 */
void gasal_host_batch_printall(host_batch_t *res);
#ifdef __cplusplus
}
#endif

#endif
