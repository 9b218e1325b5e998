// engine.h -- internal interface between the host runtime (runtime.cpp) and the
// device pipeline (kernels.hip).  Not installed; the public API is include/.
#ifndef GASAL2_ENGINE_H
#define GASAL2_ENGINE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

// Scores as the kernels see them, passed by value (the reference keeps them in
// __constant__ memory, src/gasal_kernels.h:29-33).
struct G2Scores {
    int32_t match, mismatch, gapo, gape, gapoe;
    uint32_t n_value;   // N_CODE & 0xF (src/gasal_kernels.h:37)
    int32_t n_code;     // full N_CODE byte (reverse kernel compares with it)
    int32_t has_np;     // N_PENALTY defined
    int32_t np;         // N_PENALTY
};

// Algorithm switches (Parameters fields, src/args_parser.h:30-48).
struct G2Opts {
    int32_t algo, start, second, head, tail, k_band, is_packed, is_rc;
    int32_t banded;      // LOCAL inside per-pair diagonal bands (k_local_band)
    int32_t skip_prep;   // batch already packed and reverse-complemented on the device
    int32_t prep_only;   // pack and reverse-complement only
    int32_t ksw_w, ksw_zdrop;   // KSW band half-width (< 0: none) and z-drop (0: off), the reference's defaults
    int32_t ksw_eb;             // KSW end bonus (the reference: PEN_CLIP5 = 5)
};

// Device view of one batch plus the workspace it runs in.
struct G2Dev {
    // inputs
    const uint8_t *uq, *ut;          // unpacked bytes
    uint32_t *pq, *pt;               // packed words (alias uq/ut when is_packed)
    const uint32_t *qoff, *toff, *qlen, *tlen;
    const uint8_t *qop, *top;
    const uint32_t *seeds;
    uint32_t q_bytes, t_bytes, n;
    uint32_t max_qw, max_tw;         // max words over the batch (host-computed)
    // outputs (device)
    int32_t *score, *qend, *tend, *qstart, *tstart;
    int32_t *s_score, *s_qend, *s_tend;
    uint32_t *n_ops;
    uint8_t *cigar;                  // per-pair slots of cigar_cap bytes (pair order)
    uint32_t cigar_cap;
    uint8_t *cigar_qlayout;          // reference layout: CIGAR at query offsets
    // workspace
    uint32_t *keys, *perm;           // shape key and sorted permutation, n each
    uint32_t *hist;                  // G2_SHAPE_BUCKETS counters / bucket offsets
    uint32_t *bnd;                   // boundary rows: per group bnd_cols*64 dwords
    uint32_t bnd_cols;
    uint32_t *tb;                    // direction words (or tile left states): per group tb_tw*tb_qw8*64
    uint32_t tb_tw, tb_qw8;
    // per-group tile regions (set when one size for all groups would not fit):
    // group g starts at dword tb_gbase[g], strips of tb_gqw[g] query words
    const uint64_t *tb_gbase;
    const uint32_t *tb_gqw;
    // checkpointed traceback (LOCAL WITH_TB, fast lane path): when set, the DP
    // stores no direction words; tb (same layout) holds each 8x8 tile's left
    // column state (H, F) and tbh each strip's bottom row (H, E), from which
    // the traceback recomputes the tiles on its path
    uint32_t *tbh;
    uint32_t tb_fused;               // with tbh: the DP kernel walks each lane-path pair's traceback itself
    // transposed sequence panels of the lane path ([group][word][lane]; null: off)
    uint32_t *qpan, *tpan;
    uint32_t qpan_w, tpan_w;
    int32_t *eh;                     // KSW rows: per group eh_cols*64*2 ints
    uint32_t eh_cols;
    int32_t ksw16;                   // KSW rows as packed uint16 (H, E): values fit 16 bits
    uint32_t *aux;                   // small device counters (work queue etc.)
    uint32_t *sink;                  // 4 KiB a kernel may write discarded records to (never read)
    // score-table kernels (k_local_fast, k_global / k_semi LUT variants): 0 =
    // off, 1 = launched beside the generic kernel, the nibble census in aux[0]
    // picks which one runs.  enc_ok: the LOCAL lane-per-pair position encoding
    // fits 31 bits (otherwise k_local takes the pairs the wavefront leaves).
    int32_t fast, enc_ok;
    int32_t pk16;                    // LOCAL score and ends in packed 16-bit halves (lane_local_pk)
    int32_t pk2;                     // LOCAL, two pairs per lane in packed 16-bit halves (k_local_pk2)
    int32_t pk16g;                   // GLOBAL score in packed 16-bit halves (lane_global_pk)
    int32_t pk16s;                   // SEMI_GLOBAL (tails none / target) in packed 16-bit halves (semi_pass_pk)
    int32_t enc_s, enc_cb;           // position encoding: h << enc_s | ~strip << (3+enc_cb) | ~col << 3 | 7-row
    uint32_t enc_cm, enc_sm;         // column / strip field masks
    // intra-pair wavefront for long pairs (LOCAL score + ends): pairs whose
    // shape passes wf_min_tw / wf_min_cells are sorted first and taken from an
    // atomic queue (aux[1], count in aux[2]) by the first wf_blocks blocks.
    uint32_t wf_blocks, wf_min_tw, wf_min_cells;
    uint32_t *wf_scratch;            // per wavefront block: per pass 17 x 64 dwords of strip state
    uint32_t wf_stride;              // dwords per wavefront block
    uint32_t *wf_tb;                 // WITH_TB: tiles of the long pairs, per pair [strip][word] x 8 dwords
    const uint64_t *wf_tbo;          // first 32-byte tile of each long pair in wf_tb
    uint32_t *wf_tbh;                // checkpointed traceback of the long pairs: strip bottom rows
                                     // (wf_tb then holds the tiles' left column states); null: direction words
    // banded LOCAL (k_local_band): per-pair diagonal band [band_lo, band_hi]
    // of d = query column - target row; direction tiles of pair k start at
    // 32-byte chunk band_tbo[k], band_bw[k] chunks per strip.
    const int32_t *band_lo, *band_hi;
    const uint64_t *band_tbo;
    const uint32_t *band_bw;
    uint8_t *band_flags;             // per pair: bit0 edge contact, bit1 walk left the band, bit2 check band disagrees
    int32_t band_widen;              // check band: [lo - widen, hi + widen], score and ends only (0: none)
    int32_t *band_wres;              // check band results: score, query end, target end (3 x n)
    // pairs of a sub-batch run (gasal2_run_subset): results go to slot sub_idx[k]
    const uint32_t *sub_idx;
};

struct G2Timing {
    bool enabled;
    hipEvent_t ev[6];   // start, after pack, after rc, after dp, after tb, end
};

// Shape buckets of the pair ordering (long flag, 8-bit query and target word classes).
constexpr uint32_t G2_SHAPE_BUCKETS = 1u << 17;

// Enqueue the whole device pipeline on `stream`.  Returns hipSuccess or the
// first launch error.
hipError_t g2_run(const G2Scores &sc, const G2Opts &op, const G2Dev &d, hipStream_t stream, G2Timing *tm);

// Copies the results of a sub-batch run (pairs sub.sub_idx[k]) into the main
// batch's result arrays and CIGAR slots.
hipError_t g2_scatter_subset(const G2Dev &sub, const G2Dev &main, bool tb, hipStream_t stream);

// Kernel choices for a batch that depend only on the mode, the scores and the
// largest shape the lane path runs (runtime.cpp): score-table kernels, then
// the packed 16-bit kernels and the LOCAL position encoding.
void g2_choose_fast(G2Dev &d, const G2Opts &op, const G2Scores &sc, uint32_t qw, uint32_t tw);
void g2_choose_packed(G2Dev &d, const G2Opts &op, const G2Scores &sc, uint32_t eq, uint32_t et, bool banded);
G2Scores g2_scores();

// Dispatch policy for LOCAL (0 auto, 1 reference-order kernel only).
void g2_set_dispatch(int policy);
int g2_get_dispatch();

static inline uint32_t g2_regs(uint32_t len) { return (len >> 3) + ((len & 7) ? 1u : 0u); }

// CIGAR slot size that always holds a pair's CIGAR (see get_tb walk bounds).
static inline uint32_t g2_cigar_capacity(uint32_t max_qlen, uint32_t max_tlen) {
    uint32_t q8 = 8 * g2_regs(max_qlen), t8 = 8 * g2_regs(max_tlen);
    return (q8 + t8 + (q8 + t8) / 32 + 16 + 15) & ~15u;   // 16-byte aligned slots
}

#endif
