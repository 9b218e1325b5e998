/*
 * gasal2_sam.h -- the results' text forms: the CIGAR bytes of a WITH_TB
 * alignment (op | count << 2 per byte, last run first; README.md:203-211) as
 * CIGAR text, and whole SAM records.
 *
 * The reference prints only its own merged M/X/D/I text
 * (test_prog/test_prog.cpp:382-428).  A byte's op 2 moves along the target
 * only and op 3 along the query only (src/kernels/get_tb.h:107-108), so with
 * the target as the reference sequence op 2 is SAM's D and op 3 SAM's I.
 */
#ifndef GASAL2_SAM_H
#define GASAL2_SAM_H

#include <stdint.h>

#include "gasal.h"
#include "gasal2_stream.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CIGAR text styles */
#define GASAL2_CIGAR_HARNESS 0 /* the reference harness's text: equal adjacent runs merged, M X D I */
#define GASAL2_CIGAR_SAM 1     /* SAM: M (match or mismatch) I D, adjacent equal ops merged */

/* Writes the forward CIGAR text of n_ops bytes into buf (NUL-terminated,
 * at most cap bytes with the NUL) and returns its length without the NUL,
 * also when it did not fit (as snprintf does); -1 for an unknown style. */
int32_t gasal2_cigar_text(const uint8_t *cigar, uint32_t n_ops, int32_t style, char *buf, uint32_t cap);

/* gasal2_sam_record / gasal2_sam_write flags */
#define GASAL2_SAM_EQX 1u     /* =/X instead of M, decided from the bases */
#define GASAL2_SAM_NO_SEQ 2u  /* SEQ '*' */
#define GASAL2_SAM_HEADER 4u  /* gasal2_sam_write: @HD, one @SQ per distinct target name, @PG */

/* One pair's inputs and outputs.  query/target are the sequences as handed to
 * the library (bytes, before operation_on_seq); their ops are applied here as
 * the library applied them (reverse/complement of the padded 4-bit words), so
 * SEQ is the query as aligned and POS counts in the target as aligned.
 * query_end/target_end are the 0-based ends the library reported (LOCAL); for
 * GLOBAL pass -1 and the alignment is anchored at the sequences' starts.
 * has_second = 0 leaves out XS. */
typedef struct {
    const char *qname, *rname;
    const uint8_t *query, *target;
    uint32_t query_len, target_len;
    uint8_t query_op, target_op;
    int32_t score, query_end, target_end;
    int32_t has_second, second_score;
    const uint8_t *cigar; /* NULL or n_cigar_ops = 0: CIGAR '*' */
    uint32_t n_cigar_ops;
} gasal2_sam_in_t;

/* One SAM line (with its '\n') for a pair:
 *   QNAME FLAG RNAME POS MAPQ CIGAR * 0 0 SEQ * AS:i NM:i [XS:i]
 * QNAME and RNAME are the names' first words ("*" if NULL or empty).  FLAG 16
 * when the query was reverse-complemented (operation_on_seq 3 with
 * isReverseComplement), 4 (unmapped: RNAME, POS, CIGAR '*') when score <= 0 or
 * there is no CIGAR.  MAPQ 255.  The CIGAR's span is anchored at the reported
 * ends; leading and trailing query bases outside it become soft clips, so the
 * query length always equals the CIGAR's.  NM counts mismatching aligned
 * bases (N included) and gap bases.  Returns the line's length (written into
 * buf, NUL-terminated, when it fits in cap), or -1 for an invalid input.
 * rc_enabled is the Parameters' isReverseComplement, n_code the N_CODE the
 * batch was padded with. */
int32_t gasal2_sam_record(const gasal2_sam_in_t *in, int32_t rc_enabled, int32_t n_code, uint32_t flags, char *buf,
                          uint32_t cap);

/* All pairs' SAM records in pair order (header first with GASAL2_SAM_HEADER)
 * written to path ("-" = stdout) from what gasal2_stream_align produced:
 * res->cigar slots of res->cigar_cap bytes (a pair whose n_cigar_ops exceeds
 * the slot gets CIGAR '*'), ends from res->query_end / target_end (NULL for
 * GLOBAL), second scores when res->second_score is set.  Formats with up to
 * n_threads host threads (0 = all cores).  Returns 0, or -1 after a message on
 * stderr. */
int32_t gasal2_sam_write(const char *path, const gasal2_pairs_t *pairs, const gasal2_results_t *res,
                         int32_t rc_enabled, uint32_t flags, int32_t n_threads);

#ifdef __cplusplus
}
#endif

#endif
