/*
 * gasal_align.h -- batch alignment entry points.
 *
 * Drop-in for nahmedraja/GASAL2 src/gasal_align.h:114-120.  C linkage so that
 * non-C++ hosts (ctypes, cgo, JNI) bind the same symbols; the signatures are
 * the reference's.
 */
#ifndef __GASAL_ALIGN_H__
#define __GASAL_ALIGN_H__

#include "gasal.h"
#include "args_parser.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces gasal_copy_subst_scores, src/gasal_align.cu:329-339 (declared
 * src/gasal_align.h:114).  Sets the process-wide match/mismatch/gap scores used
 * by every later gasal_aln_async (they are passed to the kernels by value). */
void gasal_copy_subst_scores(gasal_subst_scores *subst);

/* Replaces gasal_aln_async, src/gasal_align.cu:29-307 (declared
 * src/gasal_align.h:116).  Validates, grows device buffers, enqueues H2D copies,
 * pack, reverse/complement, the DP kernel(s), traceback and D2H copies on
 * gpu_storage->str and returns immediately. */
void gasal_aln_async(gasal_gpu_storage_t *gpu_storage, const uint32_t actual_query_batch_bytes,
                     const uint32_t actual_target_batch_bytes, const uint32_t actual_n_alns, Parameters *params);

/* Replaces gasal_is_aln_async_done, src/gasal_align.cu:310-326 (declared
 * src/gasal_align.h:120).  0 = done (and the storage is reset for the next
 * batch), -1 = still running, -2 = nothing was launched on this storage. */
int gasal_is_aln_async_done(gasal_gpu_storage_t *gpu_storage);

#ifdef __cplusplus
}
#endif

#endif
