/*
 * gasal_oracle.h -- CPU oracle for the batched alignment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++17/OpenMP restatement of the
 * reference kernels' recurrences (nahmedraja/GASAL2, src/kernels/ headers and the
 * stage order of src/gasal_align.cu:29-307).  Only tests/, the smoke() entry of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and only
 * as the checker / the reported CPU baseline.  The product library never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates.  The enum numbering
 * follows src/gasal.h:37-72 (algo: 1 GLOBAL, 2 SEMI_GLOBAL, 3 LOCAL, 5 BANDED,
 * 6 KSW; start: 0 WITHOUT_START, 1 WITH_START, 2 WITH_TB; data_source: 0 NONE,
 * 1 QUERY, 2 TARGET, 3 BOTH).
 */
#ifndef GASAL_ORACLE_H
#define GASAL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t match;          /* _cudaMatchScore    src/gasal_kernels.h:32 */
    int32_t mismatch;       /* _cudaMismatchScore src/gasal_kernels.h:33 */
    int32_t gap_open;       /* _cudaGapO          src/gasal_kernels.h:29 */
    int32_t gap_extend;     /* _cudaGapExtend     src/gasal_kernels.h:31 */
    int32_t n_code;         /* compile-time N_CODE (Makefile:4); padding byte  */
    int32_t has_n_penalty;  /* N_PENALTY defined? (src/gasal_kernels.h:39)     */
    int32_t n_penalty;      /* its value                                       */
} oracle_scores_t;

typedef struct {
    int32_t algo;                  /* Parameters::algo        src/args_parser.h:48 */
    int32_t start_pos;             /* Parameters::start_pos   src/args_parser.h:35 */
    int32_t second_best;           /* Parameters::secondBest  src/args_parser.h:40 */
    int32_t head;                  /* semiglobal_skipping_head src/args_parser.h:45 */
    int32_t tail;                  /* semiglobal_skipping_tail src/args_parser.h:46 */
    int32_t k_band;                /* Parameters::k_band      src/args_parser.h:38 */
    int32_t is_packed;             /* Parameters::isPacked    src/args_parser.h:42 */
    int32_t is_reverse_complement; /* Parameters::isReverseComplement :43 */
    int32_t ksw_w;                 /* KSW band half-width, < 0: none (the reference: none) */
    int32_t ksw_zdrop;             /* KSW z-drop, 0: off (the reference: 0, ksw_kernel_template.h:62) */
    int32_t ksw_end_bonus;         /* KSW end bonus: the to-end score wins unless it is <= max - end_bonus
                                      (the reference: PEN_CLIP5 = 5, ksw_kernel_template.h:20,188); with a
                                      band it also enters bwa's clamp of w */
} oracle_params_t;

typedef struct {
    int32_t *aln_score;
    int32_t *query_batch_end;
    int32_t *target_batch_end;
    int32_t *query_batch_start;
    int32_t *target_batch_start;
    int32_t *second_score;
    int32_t *second_query_end;
    int32_t *second_target_end;
    uint32_t *n_cigar_ops;
    uint8_t *cigar;                 /* pair k writes at cigar + cigar_offsets[k] */
    const uint64_t *cigar_offsets;  /* capacity of pair k: oracle_cigar_capacity */
} oracle_out_t;

/* Per-pair CIGAR capacity (bytes) that always holds a pair's CIGAR produced in
 * isolation (the walk moves at least one cell per step, plus the final byte and
 * the GLOBAL residual runs, src/kernels/get_tb.h:47-139). */
uint32_t oracle_cigar_capacity(uint32_t query_len, uint32_t target_len);

/* src/kernels/pack_rc_seqs.h:13-53: 8 bytes -> one word, first base in bits 31:28. */
void oracle_pack(const uint8_t *unpacked, uint32_t n_bytes, uint32_t *packed);

/* src/kernels/pack_rc_seqs.h:56-212, one call per pair and per side. */
void oracle_rc_one(uint32_t *packed_words, uint32_t len, uint8_t op, int32_t n_code);

/* Whole gasal_aln_async stage order (src/gasal_align.cu:190-283) on host data:
 * pack (unless is_packed) -> reverse/complement (if is_reverse_complement) ->
 * DP kernel chosen by the dispatch table (src/gasal_align.h:7-107) -> get_tb.
 * Returns 0 on success, nonzero on a rejected input. */
int oracle_align_batch(const oracle_scores_t *sc, const oracle_params_t *pr,
                       const uint8_t *query_batch, uint32_t query_batch_bytes,
                       const uint8_t *target_batch, uint32_t target_batch_bytes,
                       const uint32_t *query_offsets, const uint32_t *target_offsets,
                       const uint32_t *query_lens, const uint32_t *target_lens,
                       const uint8_t *query_ops, const uint8_t *target_ops,
                       const uint32_t *seed_scores, uint32_t n_alns,
                       oracle_out_t *out, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
