/*
 * gasal2_stream.h -- the host input path of the batched pipeline: paired
 * FASTA / FASTQ files read into padded host batches, streamed through several
 * storages by one host thread each (fill the storage's pinned pages,
 * gasal_aln_async, wait, copy the results out), results gathered per pair.
 *
 * The reference has no library counterpart: its example program does this
 * work inline (reading test_prog/test_prog.cpp:79-137, batching and threads
 * :237-330, paging src/host_batch.cpp:79-153).  Everything here is built on the
 * public API of gasal_align.h / ctors.h / host_batch.h.
 */
#ifndef GASAL2_STREAM_H
#define GASAL2_STREAM_H

#include "gasal.h"
#include "gasal2_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Pairs in host memory, laid out as the reference's host batches: every
 * sequence padded with N_CODE to a multiple of 8 bases, offsets and lengths
 * in bases.  query_words / target_words, when not NULL, hold the same batch as
 * 4-bit packed words (what an isPacked caller hands over).  Names are
 * optional. */
typedef struct {
    uint8_t *query, *target;
    uint64_t query_bytes, target_bytes;
    uint32_t *query_words, *target_words;
    uint32_t *query_offsets, *target_offsets;
    uint32_t *query_lens, *target_lens;
    uint8_t *query_ops, *target_ops; /* operation_on_seq per pair */
    char **query_names, **target_names;
    uint32_t n;
    uint32_t *seed_scores;           /* KSW: h0 per pair (NULL otherwise; the caller's, gasal2_pairs_free leaves it) */
} gasal2_pairs_t;

/* gasal2_pairs_read flags */
#define GASAL2_READ_NAMES 1u  /* keep the header names */
#define GASAL2_READ_PACK 2u   /* also fill query_words / target_words */

/* Reads paired files of the same number of records.  FASTA ('>' headers; as
 * in the reference's example program a header starting with '<', '/' or '+'
 * instead selects REVERSE_NATURAL, FORWARD_COMPLEMENT or REVERSE_COMPLEMENT
 * for that sequence, test_prog/test_prog.cpp:79-137) or FASTQ ('@' records of
 * four lines; qualities are skipped).  Sequence lines may wrap.  Returns NULL
 * after a message on stderr if a file cannot be read or the record counts
 * differ.  Free with gasal2_pairs_free. */
gasal2_pairs_t *gasal2_pairs_read(const char *query_path, const char *target_path, uint32_t flags);
void gasal2_pairs_free(gasal2_pairs_t *pairs);

/* Host result arrays of pairs->n entries each; NULL = not wanted.  cigar:
 * per pair a slot of cigar_cap bytes holding its CIGAR bytes in the
 * reference's format (op | count << 2, reverse order), n_cigar_ops[k] of
 * them, clipped to the slot (whole when they fit, also when they exceed the
 * pair's padded query length, where host_res clips them). */
typedef struct {
    int32_t *score, *query_end, *target_end, *query_start, *target_start;
    int32_t *second_score, *second_query_end, *second_target_end;
    uint32_t *n_cigar_ops;
    uint8_t *cigar;
    uint32_t cigar_cap;
} gasal2_results_t;

typedef struct {
    double seconds;            /* first gasal_aln_async (its H2D) to the last results copied out */
    double seconds_from_fill;  /* first host fill to the last results copied out */
    uint32_t sub_batches;
    uint64_t cells;            /* sum of query length x target length */
    /* host time of the worker threads, summed over them: filling pages,
     * inside gasal_aln_async (enqueueing), waiting for the storage, copying
     * results out */
    double fill_seconds, launch_seconds, wait_seconds, copy_seconds;
} gasal2_stream_stats_t;

/* A set of storages on the caller's current device, kept across calls so
 * their device workspace is allocated once: n_storages storages (each a HIP
 * stream) sized for sub-batches of batch_pairs pairs of at most the given
 * lengths (longer batches grow them, as gasal_aln_async does).  n_storages 3
 * uses the four hardware queues a process gets without two streams sharing
 * one.  The Parameters must outlive the stream. */
typedef struct gasal2_stream gasal2_stream_t;
gasal2_stream_t *gasal2_stream_new(Parameters *params, uint32_t batch_pairs, int32_t n_storages,
                                   uint32_t max_query_len, uint32_t max_target_len);
void gasal2_stream_free(gasal2_stream_t *stream);

/* The stream's storages share an order chain (gasal2_order_new): by default
 * mode 2, so sub-batches take PCIe and run their device programs one after
 * another in submission order (each program fills the device; the next
 * sub-batch's H2D and the last one's D2H overlap it).  0 lets the storages'
 * streams run independently.  Waits for work in flight.  Results are
 * identical in every mode. */
#define GASAL2_STREAM_ORDER_DEFAULT 2
void gasal2_stream_set_order(gasal2_stream_t *stream, int32_t mode);

/* Aligns every pair: sub-batches of the stream's batch_pairs go round robin
 * to its storages, each driven by its own host thread (fill the pinned pages,
 * gasal_aln_async, wait, copy the results out).  With params->isPacked the
 * packed words are handed over (pairs->query_words / target_words must be
 * set); KSW takes pairs->seed_scores.  Returns 0, or -1 after a message on
 * stderr. */
int32_t gasal2_stream_align(gasal2_stream_t *stream, const gasal2_pairs_t *pairs, gasal2_results_t *out,
                            gasal2_stream_stats_t *stats);

/* Two-sided seed extension (the step before alignment in a seed-and-extend
 * mapper; bwa-mem's left/right ksw_extend2 calls around each seed, the use the
 * reference's KSW kernel was written for, src/kernels/ksw_kernel_template.h
 * :27-29): for pair k the seed query[query_beg, +len) ~ target[target_beg, +len)
 * of score h0 is extended to the left by KSW on the reversed prefixes
 * (query[0, query_beg), target[0, target_beg)) from h0, then to the right on
 * the suffixes from the left extension's score.  Each side runs as a
 * stream_align batch with the given band w, z-drop and end bonus (the KSW
 * settings of gasal2_storage_set_ksw); a side with an empty query or target is
 * not extended.  The stream must be made with algo = KSW. */
typedef struct {
    uint32_t query_beg, target_beg, len;
    int32_t score; /* h0 of the left extension, e.g. len * match */
} gasal2_seed_t;

typedef struct {
    int32_t w, zdrop;                   /* as gasal2_set_ksw_extension (w < 0: no band, zdrop 0: off) */
    int32_t end_bonus_left, end_bonus_right; /* as gasal2_set_ksw_end_bonus, per side */
} gasal2_extend_opts_t;

/* Output arrays of pairs->n entries (NULL = not wanted): the extended
 * alignment's score (the right extension's, which started from the left
 * one's), half-open spans [beg, end) on query and target, and each side's
 * KSW score (the seed's score where a side was not extended). */
typedef struct {
    int32_t *score, *query_beg, *query_end, *target_beg, *target_end, *left_score, *right_score;
} gasal2_extend_out_t;

int32_t gasal2_extend_seeds(gasal2_stream_t *stream, const gasal2_pairs_t *pairs, const gasal2_seed_t *seeds,
                            const gasal2_extend_opts_t *opts, gasal2_extend_out_t *out, gasal2_stream_stats_t *stats);

/* gasal2_stream_new + gasal2_stream_align + gasal2_stream_free in one call
 * (the workspace is allocated inside, so the first sub-batches pay for it). */
int32_t gasal2_pairs_align(const gasal2_pairs_t *pairs, gasal2_results_t *out, Parameters *params,
                           uint32_t batch_pairs, int32_t n_storages, gasal2_stream_stats_t *stats);

#ifdef __cplusplus
}
#endif

#endif
