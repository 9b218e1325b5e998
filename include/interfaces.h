/*
 * interfaces.h -- resizing, per-pair operations and device selection.
 *
 * Drop-in for nahmedraja/GASAL2 src/interfaces.h:5-14 (src/interfaces.cpp).
 */
#ifndef __GASAL_INTERFACES_H__
#define __GASAL_INTERFACES_H__

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gasal.h"
#include "args_parser.h"

#ifdef __cplusplus
extern "C" {
#endif
/* src/interfaces.cpp:26-78: grows the host per-pair arrays and result buffers. */
void gasal_host_alns_resize(gasal_gpu_storage_t *gpu_storage, int new_max_alns, Parameters *params);
/* src/interfaces.cpp:81-96: copies per-pair reverse/complement operations. */
void gasal_op_fill(gasal_gpu_storage_t *gpu_storage_t, uint8_t *data, uint32_t nbr_seqs_in_stream, data_source SRC);
/* src/interfaces.cpp:98-128 */
#ifdef __cplusplus
void gasal_set_device(int gpu_select = 0, bool isPrintingProp = true);
#else
void gasal_set_device(int gpu_select, bool isPrintingProp);
#endif
#ifdef __cplusplus
}
#endif

#endif
