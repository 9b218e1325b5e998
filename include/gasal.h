/*
 * gasal.h -- public types of the MI355X batched pairwise-alignment library.
 *
 * Drop-in for nahmedraja/GASAL2 src/gasal.h: same macro, enum, struct and field
 * names, same enum numbering and field order (src/gasal.h:37-165), with HIP
 * types in place of CUDA ones (hipStream_t for cudaStream_t, HIP's uint4).
 * One field is appended to gasal_gpu_storage_t (`impl`, the library's private
 * per-storage state); code written against the reference compiles unchanged.
 */
#ifndef __GASAL_H__
#define __GASAL_H__

#include <stdlib.h>
#include <stdint.h>
#include <stdio.h>

/* C++ hosts get the whole HIP runtime header; C hosts (cgo, a plain C
 * program) only need its runtime API types and compile with HIP's platform
 * flags (`hipconfig --cpp_config`: -D__HIP_PLATFORM_AMD__ -I<rocm>/include). */
#ifdef __cplusplus
#include <hip/hip_runtime.h>
#else
#include <stdbool.h>
#include <hip/hip_runtime_api.h>
#endif

#ifndef HOST_MALLOC_SAFETY_FACTOR
#define HOST_MALLOC_SAFETY_FACTOR 5
#endif

/* Compile-time knobs of the reference (Makefile:1-4).  Here N_CODE and
 * N_PENALTY only set the process defaults; gasal2_set_n_code() and
 * gasal2_set_n_penalty() change them at run time.  MAX_QUERY_LEN is kept for
 * source compatibility only: this library has no compile-time length cap. */
#ifndef N_CODE
#define N_CODE 0x4E
#endif
#ifndef MAX_QUERY_LEN
#define MAX_QUERY_LEN 1024
#endif

/* src/gasal.h:15-22.  Evaluates a HIP runtime call; on failure prints the
 * reference's message and leaves through the library's error path. */
#define CHECKCUDAERROR(error)                                                              \
    do {                                                                                   \
        err = error;                                                                       \
        if (hipSuccess != err) {                                                           \
            fprintf(stderr, "[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n", \
                    hipGetErrorString(err), err, __LINE__, __FILE__);                     \
            gasal2_fail(EXIT_FAILURE);                                                     \
        }                                                                                  \
    } while (0)

#ifdef __cplusplus
extern "C" {
#endif
/* Error exit used by every API check (the reference calls exit(EXIT_FAILURE),
 * src/gasal.h:20, src/gasal_align.cu:32-67).  Default: exit.  See
 * gasal2_set_error_mode() in gasal2_ext.h. */
void gasal2_fail(int code);
#ifdef __cplusplus
}
#endif

static inline int CudaCheckKernelLaunch(void) {
    hipError_t err = hipGetLastError();
    if (hipSuccess != err) return -1;
    return 0;
}

enum comp_start { WITHOUT_START, WITH_START, WITH_TB };

enum Bool { FALSE, TRUE };

enum data_source { NONE, QUERY, TARGET, BOTH };

enum algo_type { UNKNOWN, GLOBAL, SEMI_GLOBAL, LOCAL, MICROLOCAL, BANDED, KSW };

enum operation_on_seq { FORWARD_NATURAL, REVERSE_NATURAL, FORWARD_COMPLEMENT, REVERSE_COMPLEMENT };

#ifndef __cplusplus
/* C spells the enum types without the `enum` keyword only through typedefs. */
typedef enum comp_start comp_start;
typedef enum Bool Bool;
typedef enum data_source data_source;
typedef enum algo_type algo_type;
typedef enum operation_on_seq operation_on_seq;
#endif

/* Linked list of pinned host pages (src/gasal.h:75-83). */
struct host_batch {
    uint8_t *data;
    uint32_t page_size;
    uint32_t data_size;
    uint32_t offset;
    int is_locked;
    struct host_batch *next;
};
typedef struct host_batch host_batch_t;

/* Result arrays, host- or device-side (src/gasal.h:86-95). */
struct gasal_res {
    int32_t *aln_score;
    int32_t *query_batch_end;
    int32_t *target_batch_end;
    int32_t *query_batch_start;
    int32_t *target_batch_start;
    uint8_t *cigar;
    uint32_t *n_cigar_ops;
};
typedef struct gasal_res gasal_res_t;

/* One stream's worth of buffers (src/gasal.h:98-150). */
typedef struct {
    uint8_t *unpacked_query_batch;
    uint8_t *unpacked_target_batch;
    uint32_t *packed_query_batch;
    uint32_t *packed_target_batch;
    uint32_t *query_batch_offsets;
    uint32_t *target_batch_offsets;
    uint32_t *query_batch_lens;
    uint32_t *target_batch_lens;

    uint32_t *host_seed_scores;
    uint32_t *seed_scores;

    host_batch_t *extensible_host_unpacked_query_batch;
    host_batch_t *extensible_host_unpacked_target_batch;

    uint8_t *host_query_op;
    uint8_t *host_target_op;
    uint8_t *query_op;
    uint8_t *target_op;

    uint32_t *host_query_batch_offsets;
    uint32_t *host_target_batch_offsets;
    uint32_t *host_query_batch_lens;
    uint32_t *host_target_batch_lens;

    gasal_res_t *host_res;
    gasal_res_t *device_cpy;
    gasal_res_t *device_res;

    gasal_res_t *host_res_second;
    gasal_res_t *device_res_second;
    gasal_res_t *device_cpy_second;

    uint32_t gpu_max_query_batch_bytes;
    uint32_t gpu_max_target_batch_bytes;

    uint32_t host_max_query_batch_bytes;
    uint32_t host_max_target_batch_bytes;

    uint32_t gpu_max_n_alns;
    uint32_t host_max_n_alns;
    uint32_t current_n_alns;

    uint64_t packed_tb_matrix_size;
    uint4 *packed_tb_matrices;

    hipStream_t str;
    int is_free;
    int id;

    void *impl; /* library-private state (work queue, scratch, traceback, graph) */
} gasal_gpu_storage_t;

/* Vector of storages (src/gasal.h:153-156). */
typedef struct {
    int n;
    gasal_gpu_storage_t *a;
} gasal_gpu_storage_v;

/* Match/mismatch and gap penalties (src/gasal.h:160-165). */
typedef struct {
    int32_t match;
    int32_t mismatch;
    int32_t gap_open;
    int32_t gap_extend;
} gasal_subst_scores;

#endif
