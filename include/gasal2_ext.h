/*
 * gasal2_ext.h -- C-ABI additions of the MI355X library.
 *
 * The reference API (gasal_align.h, ctors.h, host_batch.h, interfaces.h) takes
 * a C++ `Parameters*`; these helpers let a C / ctypes / cgo / JNI host build one
 * and reach the library's run-time switches.  Each entry states the reference
 * interface it extends or that it has no counterpart.
 */
#ifndef GASAL2_EXT_H
#define GASAL2_EXT_H

#include "gasal.h"

#ifdef __cplusplus
class Parameters;
extern "C" {
#else
typedef struct Parameters Parameters;
#endif

/* Plain-data view of the algorithm fields of class Parameters
 * (src/args_parser.h:30-48); field meanings and enum values are the reference's. */
typedef struct {
    int32_t sa, sb, gapo, gape;
    int32_t start_pos;               /* comp_start  */
    int32_t print_out, n_threads;
    int32_t k_band;
    int32_t secondBest;              /* Bool        */
    int32_t isPacked;
    int32_t isReverseComplement;
    int32_t semiglobal_skipping_head;/* data_source */
    int32_t semiglobal_skipping_tail;/* data_source */
    int32_t algo;                    /* algo_type   */
} gasal2_params_t;

/* Parameters(int argc, char **argv) with argc = 0 (src/args_parser.cpp:8-39):
 * reference defaults.  No reference counterpart as a C function. */
Parameters *gasal2_parameters_new(void);
void gasal2_parameters_free(Parameters *p);
void gasal2_parameters_get(const Parameters *p, gasal2_params_t *out);
void gasal2_parameters_set(Parameters *p, const gasal2_params_t *in);

/* Run-time values of the reference's compile-time N_CODE and N_PENALTY
 * (Makefile:1-4, src/gasal_kernels.h:37-56).  has_penalty = 0 means
 * "N_PENALTY undefined".  Process-wide, like gasal_copy_subst_scores. */
void gasal2_set_n_code(int32_t n_code);
int32_t gasal2_get_n_code(void);
void gasal2_set_n_penalty(int32_t has_penalty, int32_t penalty);

/* Error path (src/gasal.h:15-22; src/gasal_align.cu:32-67 exit on misuse).
 * mode 0 (default): print and exit(EXIT_FAILURE) like the reference.
 * mode 1: print, record the message and return from the failing API call
 * without side effects, so a test harness can observe it. */
void gasal2_set_error_mode(int32_t mode);
int32_t gasal2_error_count(void);
const char *gasal2_last_error(void);

/* Split form of gasal_aln_async (src/gasal_align.cu:29-307) for callers whose
 * inputs are already resident in device memory (and for measuring the device
 * path alone):
 *   gasal2_upload   : the H2D half (src/gasal_align.cu:152-175,207-235), blocking;
 *   gasal2_run_resident : pack, reverse/complement, DP, traceback, all on
 *                     gpu_storage->str, results left in device memory;
 *   gasal2_download : the D2H half (src/gasal_align.cu:266-304), blocking.
 * gasal_aln_async == upload + run_resident + download, asynchronously. */
void gasal2_upload(gasal_gpu_storage_t *s, uint32_t query_batch_bytes, uint32_t target_batch_bytes, uint32_t n_alns,
                   Parameters *params);
void gasal2_run_resident(gasal_gpu_storage_t *s, uint32_t query_batch_bytes, uint32_t target_batch_bytes,
                         uint32_t n_alns, Parameters *params);
void gasal2_download(gasal_gpu_storage_t *s, uint32_t query_batch_bytes, uint32_t n_alns, Parameters *params);
/* Blocks until everything enqueued on gpu_storage->str has finished. */
void gasal2_synchronize(gasal_gpu_storage_t *s);

/* Device time in milliseconds of the last run_resident's stages, measured with
 * HIP events on gpu_storage->str: stage 0 = whole run, 1 = pack, 2 =
 * reverse/complement, 3 = DP kernel(s), 4 = traceback.  Events are recorded only
 * after gasal2_enable_timing(s, 1).  No reference counterpart. */
void gasal2_enable_timing(gasal_gpu_storage_t *s, int32_t on);
float gasal2_stage_ms(gasal_gpu_storage_t *s, int32_t stage);

/* Full CIGAR of pair k of the last finished WITH_TB batch (reverse order, one
 * byte per run: op | count << 2, op 0 M, 1 X, 2 D, 3 I; reference format
 * README.md:203-211).  host_res->cigar holds the same bytes at
 * host_query_batch_offsets[k], clipped to the pair's padded query length, as in
 * the reference; this accessor has no such limit.  Returns n_ops, copies at most
 * cap bytes. */
uint32_t gasal2_get_cigar(gasal_gpu_storage_t *s, uint32_t k, uint8_t *out, uint32_t cap);

/* Kernel dispatch policy (no reference counterpart; results are identical
 * under every policy):
 *   0 auto: long pairs run one wave each, and so does every pair of a
 *     batch of at most 128 pairs per compute unit (64 for WITH_TB), which
 *     the lane path would spread over a fraction of the SIMDs (unless
 *     gasal2_set_wavefront or a tuned entry gives the thresholds);
 *   1 reference-order kernels only (one lane per pair, no score tables);
 *   2 one lane per pair only (no one-wave-per-pair path);
 *   3 one wave per pair for every pair the path accepts (LOCAL without
 *     second best, score-only or WITH_TB). */
void gasal2_set_dispatch(int32_t policy);
int32_t gasal2_get_dispatch(void);

/* Tuned switch thresholds (written by the autotuner, gasal2_amd/autotune.py):
 * for batches of mode (algo, start, second best) whose mean query length lies
 * in [lq_lo, lq_hi), use these gasal2_set_wavefront values instead of the
 * global ones.  Any setting gives identical results; only speed changes.
 * No reference counterpart. */
void gasal2_tune_set(int32_t algo, int32_t start, int32_t second, uint32_t lq_lo, uint32_t lq_hi,
                     uint32_t min_target_words, uint32_t min_cells, uint32_t max_blocks);
void gasal2_tune_clear(void);
int32_t gasal2_tune_count(void);

/* Learned-band path (no reference counterpart; see gasal2_band.h).
 *
 * gasal2_prep_resident: pack and reverse-complement an uploaded batch
 * (gasal2_upload) in place; gasal2_device_views then gives the device
 * pointers of the packed words, offsets and lengths.
 *
 * gasal2_run_banded: LOCAL (WITHOUT_START or WITH_TB, no second best) inside
 * per-pair diagonal bands d = query column - target row in [dlo[k], dhi[k]]
 * (host arrays), over the 8x8 tiles the band touches.  Results stay in device
 * memory (gasal2_download fetches them); flags[k] (host) come back set:
 * bit0 best cell on a band-edge tile, bit1 traceback left the band, bit2 the
 * check band -- the band widened by check_widen diagonals on each side, run
 * first for score and ends only (0: no check) -- found another score or end,
 * bit3 not run.
 * Returns 0, or 1 when the substitution scores do not fit the 4-bit score
 * tables (every pair flagged), -1 on error.
 *
 * gasal2_run_subset: full DP for pairs idx[0..cnt) of the last run on this
 * storage, written over their results; chunked so the full-matrix direction
 * tiles stay under max_tb_bytes (0: 32 GiB). */
/* Device pointers of the last run's results on this storage (score, query
 * end, target end, n_ops, CIGAR slots of *cap bytes), for device-side
 * consumers such as gasal2_lb_path_diag. */
void gasal2_result_views(gasal_gpu_storage_t *s, void **score, void **qend, void **tend, void **nops, void **cigar,
                         uint32_t *cap);
int32_t gasal2_prep_resident(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params);
void gasal2_device_views(gasal_gpu_storage_t *s, void **pq, void **pt, void **qoff, void **toff, void **qlen, void **tlen);
int32_t gasal2_run_banded(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params,
                          const int32_t *dlo, const int32_t *dhi, uint8_t *flags, int32_t check_widen);
int32_t gasal2_run_subset(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params,
                          const uint32_t *idx, uint32_t cnt, uint64_t max_tb_bytes);

/* Direction tiles of the one-lane-per-pair kernels (WITH_TB) use one region
 * size for every 64-pair group while that totals at most `bytes` (default
 * 64 GiB); above it each group gets a region sized for its own pairs' shape
 * bucket (uneven batches).  With the checkpointed traceback the tiles take
 * two regions of that layout (the budget applies to one); checkpoints for
 * one-wave pairs (mode 2) engage only when both of theirs fit in it.
 * Results are identical either way.  No reference counterpart
 * (src/ctors.cpp:48-53 sizes for the maximum lengths). */
void gasal2_set_tile_budget(uint64_t bytes);

/* Packed 16-bit DP, where every value of the batch fits (default mode 1):
 * LOCAL score-and-ends batches run two 8-row strips of a pair per lane in the
 * halves of each register; LOCAL with second best or the checkpointed
 * traceback runs two pairs per lane, one per half; GLOBAL and SEMI_GLOBAL
 * scores and KSW rows are packed too.  Mode 2 runs every LOCAL mode two pairs
 * per lane; 0 keeps the int32 kernels.  Results are identical.  No reference
 * counterpart. */
void gasal2_set_pk16(int32_t mode);

/* KSW extension (seed extension as bwa's ksw_extend2, which the reference's
 * KSW kernel transcribes with no band and z-drop 0,
 * src/kernels/ksw_kernel_template.h:27-29,62): w >= 0 bands row i to query
 * columns [i - w, i + w] (w first clamped like ksw_extend2), zdrop > 0
 * enables the z-drop test of :164-170 and, as in ksw_extend2, ends the
 * pair's extension when it fires (the reference's loop nesting would leave
 * only the current 8-row tile; its zdrop is 0, so it never fires there).
 * Defaults -1, 0 reproduce the reference. */
void gasal2_set_ksw_extension(int32_t w, int32_t zdrop);
/* KSW end bonus: the score of reaching the query's end is reported (query end =
 * its length) unless it is <= the best score - end_bonus; with a band it also
 * enters the clamp of w, as bwa's ksw_extend2 end_bonus does.  Default 5, the
 * reference's fixed PEN_CLIP5 (src/kernels/ksw_kernel_template.h:20,188). */
void gasal2_set_ksw_end_bonus(int32_t end_bonus);
/* KSW settings for one storage (own = 1), overriding the process-wide ones
 * above for its batches; own = 0 returns it to them.  No reference
 * counterpart. */
void gasal2_storage_set_ksw(gasal_gpu_storage_t *s, int32_t own, int32_t w, int32_t zdrop, int32_t end_bonus);

/* Checkpointed traceback for LOCAL WITH_TB.  The DP stores, per 8x8 tile,
 * the left column state (H, F) and, per strip, the bottom row (H, E) -- two
 * regions of the tile layout, twice the bytes of the direction words the
 * reference stores (src/kernels/local_kernel_template.h:45-60) -- and
 * gasal_get_tb's kernel recomputes the tiles on each path from them.  Used
 * when H fits 16 bits (match * min(Lq, Lt) <= 32767); results are identical
 * either way.  mode 0: direction words everywhere; 1 (default): checkpoints
 * on the one-lane-per-pair path; 2: also for the pairs that run one wave per
 * pair, when their checkpoints (twice the bytes of direction words) fit in
 * 80% of the device memory free at the launch.  Mode 2 pays where the long
 * pairs' matrices dominate (uneven 50-5000 bp batches) and costs where their
 * traceback walks do (1-10 kbp divergent pairs). */
void gasal2_set_tb_checkpoint(int32_t mode);

/* Fused traceback for checkpointed lane-path pairs (default off; no
 * counterpart in the reference, whose gasal_get_tb runs after the DP kernel,
 * src/gasal_align.h:50-53).  With checkpoints (mode 1 or 2 above), the DP
 * kernel walks each lane-path pair's traceback right after the pair's DP, so
 * the walks' memory waits overlap the other waves' DP; the separate traceback
 * kernel then walks only the remaining pairs.  0 walks every pair in the
 * separate kernel.  Results are identical either way.  On 1M full-feature
 * pairs it saves about 1% of the step; on batches with one-wave pairs it
 * costs about as much, hence off by default. */
void gasal2_set_tb_fused(int32_t on);

/* Transposed sequence panels (default on; no counterpart in the reference).
 * After the shape sort, the fast LOCAL one-lane-per-pair kernels read their
 * packed query and target words from per-group panels laid out
 * [group][word][lane], written by one gather kernel, instead of each lane
 * reading its own pair's words.  Results are identical either way. */
void gasal2_set_panels(int32_t on);

/* Which kernel paths the storage's last batch took (no counterpart in the
 * reference; for measurement): bit 0 a packed 16-bit kernel on the
 * one-lane-per-pair path (or KSW's packed row), bit 1 one-wave-per-pair
 * blocks, bit 2 checkpointed traceback on the one-lane-per-pair path. */
int32_t gasal2_last_paths(gasal_gpu_storage_t *storage);

/* Graph mode (default on): each storage captures its device program (pack,
 * reverse/complement, shape ordering, DP, traceback) into a HIP graph the first
 * time a batch shape is seen and replays it with one launch while the shape
 * (n, byte counts, workspace, scores, mode) repeats; H2D and D2H copies stay
 * outside the graph.  Runs with per-stage timing enabled launch eagerly.
 * gasal2_graph_stats reports captures and graph launches of a storage.  No
 * reference counterpart (the reference launches each kernel,
 * src/gasal_align.cu:180-264). */
void gasal2_set_graph_mode(int32_t on);

/* The device program of gasal2_run_resident (pack, reverse/complement,
 * shape sort, DP; LOCAL, GLOBAL and SEMI_GLOBAL, WITHOUT_START or WITH_START,
 * any second best) on caller-owned device arrays, calling no allocator: its
 * scratch is one file-scope device array of 32 MiB per GPU, held until the
 * process ends.  No counterpart in the reference; for callers that cannot
 * allocate device memory (HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).
 *   query_batch / target_batch: the batch as gasal_aln_async would upload it
 *     (bytes padded to multiples of 8, or with params->isPacked the packed
 *     words; the byte counts count bases either way); offsets, lengths and,
 *     with isReverseComplement, the per-pair operations as device arrays;
 *   max_query_len / max_target_len bound every pair's lengths;
 *   outputs: device arrays of n_alns int32 (score, query_end, target_end
 *     required; starts and second-best arrays may be NULL); bad_pairs, if not
 *     NULL, receives the count of pairs outside their batch or longer than the
 *     declared maxima (those pairs' outputs are undefined).
 * Enqueued on `stream`; returns 0, or -1 with a message on stderr (never
 * exits) for WITH_TB, KSW, BANDED, or a batch whose scratch would exceed
 * 32 MiB.  The first call on a device must not be under graph capture (it
 * looks up the scratch's address).  Calls that overlap on one GPU share the
 * scratch and race; inputs are never written. */
int32_t gasal2_aln_noalloc(const void *query_batch, const void *target_batch, uint32_t query_batch_bytes,
                           uint32_t target_batch_bytes, const uint32_t *query_offsets, const uint32_t *target_offsets,
                           const uint32_t *query_lens, const uint32_t *target_lens, const uint8_t *query_ops,
                           const uint8_t *target_ops, uint32_t n_alns, uint32_t max_query_len, uint32_t max_target_len,
                           int32_t *score, int32_t *query_end, int32_t *target_end, int32_t *query_start,
                           int32_t *target_start, int32_t *second_score, int32_t *second_query_end,
                           int32_t *second_target_end, uint32_t *bad_pairs, Parameters *params, hipStream_t stream);

/* Batch order across storages.  0 (default): every storage's stream is
 * independent, as in the reference (src/ctors.cpp:118; each stream starts its
 * copies at once, src/gasal_align.cu:152-175).  1: each gasal_aln_async
 * enqueues its H2D copies after the previously submitted batch's (on any
 * storage), so batches take PCIe one after another; 2: also its device
 * program after that batch's program.  Results are identical.  No reference
 * counterpart. */
void gasal2_set_batch_order(int32_t mode);
/* An order chain of its own for a set of storages (the same modes as
 * gasal2_set_batch_order, applied among the storages attached to it instead of
 * process-wide); NULL detaches a storage back to the process-wide chain.  Free
 * a chain after its storages are destroyed or detached.  No reference
 * counterpart. */
typedef struct gasal2_order gasal2_order_t;
gasal2_order_t *gasal2_order_new(int32_t mode);
void gasal2_storage_set_order(gasal_gpu_storage_t *s, gasal2_order_t *order);
void gasal2_order_free(gasal2_order_t *order);
int32_t gasal2_get_graph_mode(void);
void gasal2_graph_stats(gasal_gpu_storage_t *s, uint64_t *captures, uint64_t *launches);

/* Thresholds of the per-pair switch between one lane per pair and one wave per
 * pair (LOCAL without start/traceback/second best): a pair runs one wave per
 * pair when its target has at least min_target_words 8-base words and
 * 64 * query words * target words >= min_cells; at most max_blocks persistent
 * one-wave blocks take such pairs from a device queue.  Defaults 48, 2^20,
 * 2048 (WITH_TB: 32, 2^18, 2048 until this is called); (0, 0, 0) restores
 * those per-mode defaults.  No reference counterpart (the reference has one
 * thread per pair, src/kernels/local_kernel_template.h:74-75). */
void gasal2_set_wavefront(uint32_t min_target_words, uint32_t min_cells, uint32_t max_blocks);

/* Properties of the current device used to state the integer-VALU peak:
 * compute units, peak engine clock (kHz) and the name.  Returns 0 on success.
 * No reference counterpart (gasal_set_device prints the name only,
 * src/interfaces.cpp:98-128). */
int32_t gasal2_device_info(int32_t *n_cu, int32_t *clock_khz, char *name, int32_t name_cap);

#ifdef __cplusplus
}
#endif
#endif
