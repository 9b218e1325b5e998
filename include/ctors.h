/*
 * ctors.h -- storage constructors and destructors.
 *
 * Drop-in for nahmedraja/GASAL2 src/ctors.h:5-15 (implementation src/ctors.cpp).
 */
#ifndef __CTORS_H__
#define __CTORS_H__

#include "gasal.h"
#include "args_parser.h"

#ifdef __cplusplus
extern "C" {
#endif
/* src/ctors.cpp:17-23 */
gasal_gpu_storage_v gasal_init_gpu_storage_v(int n_streams);
/* src/ctors.cpp:26-128: per storage, pinned host pages and arrays, device
 * buffers sized for max_n_alns pairs of the given maximum lengths, result
 * arrays for the requested mode, and one HIP stream. */
void gasal_init_streams(gasal_gpu_storage_v *gpu_storage_vec, int max_query_len, int max_target_len, int max_n_alns,
                        Parameters *params);
/* src/ctors.cpp:202-231 (deprecated in the reference, kept for source compatibility) */
void gasal_gpu_mem_alloc(gasal_gpu_storage_t *gpu_storage, int gpu_max_query_batch_bytes,
                         int gpu_max_target_batch_bytes, int gpu_max_n_alns, Parameters *params);
/* src/ctors.cpp:234-252 (deprecated in the reference, kept for source compatibility) */
void gasal_gpu_mem_free(gasal_gpu_storage_t *gpu_storage, Parameters *params);
/* src/ctors.cpp:130-190 */
void gasal_destroy_streams(gasal_gpu_storage_v *gpu_storage_vec, Parameters *params);
/* src/ctors.cpp:193-196 */
void gasal_destroy_gpu_storage_v(gasal_gpu_storage_v *gpu_storage_vec);
#ifdef __cplusplus
}
#endif

#endif
