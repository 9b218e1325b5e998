/*
 * gasal_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY; see gasal_oracle.h).
 *
 * A literal, sequential restatement of the reference kernels, one pair at a
 * time, parallel over pairs with OpenMP.  Citations are relative to the
 * reference repository root.  Behaviour the reference leaves undefined is
 * defined here once and documented at the place it happens:
 *   - reads of direction nibbles outside the pair's computed matrix return 0
 *     (src/kernels/get_tb.h:57 would read another pair's or stale memory);
 *   - each pair's CIGAR is produced in isolation (src/kernels/get_tb.h:94 could
 *     overrun into the neighbouring pair's query slot);
 *   - the per-thread arrays the reference sizes MAX_QUERY_LEN are sized to the
 *     pair (identical whenever the reference's own size limit holds);
 *   - second-best outputs of algorithms that never write them are 0.
 */
#include "gasal_oracle.h"

#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { A_UNKNOWN = 0, A_GLOBAL = 1, A_SEMI_GLOBAL = 2, A_LOCAL = 3, A_MICROLOCAL = 4, A_BANDED = 5, A_KSW = 6 };
enum { S_WITHOUT_START = 0, S_WITH_START = 1, S_WITH_TB = 2 };
enum { D_NONE = 0, D_QUERY = 1, D_TARGET = 2, D_BOTH = 3 };

const int32_t MINUS_INF = -32768;  /* SHRT_MIN, src/gasal_kernels.h:35 */

struct Sc {
    int32_t match, mismatch, gapo, gape, gapoe;
    uint32_t n_value;      /* N_CODE & 0xF, src/gasal_kernels.h:37 */
    int32_t n_code;
    bool has_np;
    int32_t np;
};

/* short2 storage of the reference's global[] arrays: int16 truncation. */
inline int32_t t16(int32_t v) { return (int32_t)(int16_t)(uint16_t)((uint32_t)v & 0xFFFFu); }

/* DEV_GET_SUB_SCORE_LOCAL, src/gasal_kernels.h:40-42 (N_PENALTY) and :49-51. */
inline int32_t sub_local(const Sc &c, uint32_t r, uint32_t g) {
    int32_t s = (r == g) ? c.match : -c.mismatch;
    if (r == c.n_value || g == c.n_value) s = c.has_np ? -c.np : 0;
    return s;
}
/* DEV_GET_SUB_SCORE_GLOBAL, src/gasal_kernels.h:44-46 and :53-54. */
inline int32_t sub_global(const Sc &c, uint32_t r, uint32_t g) {
    int32_t s = (r == g) ? c.match : -c.mismatch;
    if (c.has_np && (r == c.n_value || g == c.n_value)) s = -c.np;
    return s;
}

/* Word reader with bounds: words past the batch read as 0 (only the KSW kernel
 * reads one word past a pair, src/kernels/ksw_kernel_template.h:56, and it never
 * uses a base from it). */
struct Words {
    const uint32_t *w;
    uint64_t n;
    uint32_t at(int64_t i) const { return (i >= 0 && (uint64_t)i < n) ? w[i] : 0u; }
};

struct PairOut {
    int32_t score = 0, qend = 0, tend = 0, qstart = 0, tstart = 0;
    int32_t s_score = 0, s_qend = 0, s_tend = 0;
};

/* Direction nibbles of one pair: one word per (strip, padded column), nibble of
 * row r at bit 28 - 4*(r&7).  Flat cell index (strip*8QW + col)*8 + (r&7) equals
 * the reference's tile addressing (src/kernels/local_kernel_template.h:257,380;
 * src/kernels/get_tb.h:50-71). */
struct TB {
    std::vector<uint32_t> w;
    uint32_t qw8 = 0;   /* 8*QW */
    void init(uint32_t tw, uint32_t qw) { qw8 = 8 * qw; w.assign((size_t)tw * qw8, 0u); }
    void put(uint32_t strip, uint32_t col, uint32_t word) { w[(size_t)strip * qw8 + col] = word; }
    uint32_t cell(int64_t flat) const {
        int64_t word = flat >> 3;
        if (flat < 0 || word >= (int64_t)w.size()) return 0u;  /* outside: defined as 0 */
        return (w[(size_t)word] >> (28 - 4 * (flat & 7))) & 15u;
    }
};

inline uint32_t regs_of(uint32_t len) { return (len >> 3) + ((len & 7) ? 1 : 0); }

/* ---------------------------------------------------------------- LOCAL ---- */
/* gasal_local_kernel<LOCAL,S,B>, src/kernels/local_kernel_template.h:71-519;
 * cores CORE_LOCAL_COMPUTE(:19-30), _START(:32-43), _TB(:45-60). */
void local_pair(const Sc &c, const Words &Q, const Words &T, uint64_t qb, uint64_t tb,
                uint32_t read_len, uint32_t ref_len, int start, bool second, PairOut &o, TB *tbm) {
    const int32_t QW = (int32_t)regs_of(read_len), TW = (int32_t)regs_of(ref_len);
    int32_t maxHH = 0, maxXY_y = 0, prev_maxHH = 0, maxXY_x = 0;
    int32_t maxHH_second = 0, prev_maxHH_second = 0, maxXY_x_second = 0, maxXY_y_second = 0;
    std::vector<int32_t> gH((size_t)8 * QW + 8, 0), gE((size_t)8 * QW + 8, 0);   /* :118-120 */
    int32_t h[9], f[9], p[9];
    if (tbm) tbm->init(TW, QW);
    for (int32_t i = 0; i < TW; i++) {                                           /* :122 */
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }            /* :123-127 */
        uint32_t gpac = T.at((int64_t)tb + i);
        int32_t gidx = i << 3, ridx = 0;
        for (int32_t j = 0; j < QW; j++) {                                        /* :133 */
            uint32_t rpac = Q.at((int64_t)qb + j);
            for (int k = 28; k >= 0; k -= 4) {                                    /* :387 */
                uint32_t rbase = (rpac >> k) & 15u;
                h[0] = gH[ridx];
                int32_t e = gE[ridx];
                uint32_t dir = 0;
                for (int l = 28, m = 1; m < 9; l -= 4, m++) {
                    uint32_t gbase = (gpac >> l) & 15u;
                    int32_t s = sub_local(c, rbase, gbase);
                    int32_t tmp_hm = p[m] + s;
                    uint32_t m_or_x = tmp_hm >= p[m] ? 0u : 1u;                  /* :49 */
                    h[m] = std::max(tmp_hm, f[m]);
                    h[m] = std::max(h[m], e);
                    h[m] = std::max(h[m], 0);
                    int sh = 28 - ((m - 1) << 2);
                    dir |= h[m] == tmp_hm ? (m_or_x << sh) : (h[m] == f[m] ? (3u << sh) : (2u << sh)); /* :53 */
                    dir |= (tmp_hm - c.gapoe) > (f[m] - c.gape) ? 0u : (1u << (31 - ((m - 1) << 2)));  /* :54 */
                    f[m] = std::max(tmp_hm - c.gapoe, f[m] - c.gape);
                    dir |= (tmp_hm - c.gapoe) > (e - c.gape) ? 0u : (1u << (30 - ((m - 1) << 2)));     /* :56 */
                    e = std::max(tmp_hm - c.gapoe, e - c.gape);
                    maxXY_y = (maxHH < h[m]) ? gidx + (m - 1) : maxXY_y;          /* :28-29 */
                    maxHH = (maxHH < h[m]) ? h[m] : maxHH;
                    if (second) {                                                 /* :145-150 */
                        bool ov = (maxHH_second < h[m]) && (maxHH > h[m]);
                        maxXY_y_second = ov ? gidx + (m - 1) : maxXY_y_second;
                        maxHH_second = ov ? h[m] : maxHH_second;
                    }
                    p[m] = h[m - 1];
                }
                gH[ridx] = t16(h[8]);                                             /* :152-154 short2 */
                gE[ridx] = t16(e);
                maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;                  /* :158 */
                if (second) {                                                     /* :160-164 */
                    maxXY_x_second = (prev_maxHH_second < maxHH) ? ridx : maxXY_x_second;
                    prev_maxHH_second = std::max(maxHH_second, prev_maxHH_second);
                }
                prev_maxHH = std::max(maxHH, prev_maxHH);
                if (tbm) tbm->put(i, ridx, dir);                                   /* :257,380 */
                ridx++;
            }
        }
    }
    o.score = maxHH; o.qend = maxXY_x; o.tend = maxXY_y;                          /* :428-430 */
    if (second) { o.s_score = maxHH_second; o.s_qend = maxXY_x_second; o.s_tend = maxXY_y_second; } /* :432-437 */

    if (start == S_WITH_START) {                                                  /* :441-511 */
        int32_t rend_pos = maxXY_x, gend_pos = maxXY_y, fwd_score = maxHH;
        int32_t rend_reg = ((rend_pos >> 3) + 1) < QW ? ((rend_pos >> 3) + 1) : QW;
        int32_t gend_reg = ((gend_pos >> 3) + 1) < TW ? ((gend_pos >> 3) + 1) : TW;
        int64_t qidx = (int64_t)qb + (rend_reg - 1), tidx = (int64_t)tb + (gend_reg - 1);
        maxHH = 0; prev_maxHH = 0; maxXY_x = 0; maxXY_y = 0;
        std::fill(gH.begin(), gH.end(), 0); std::fill(gE.begin(), gE.end(), 0);
        int32_t gidx = ((gend_reg << 3) + 8) - 1;
        for (int32_t i = 0; i < gend_reg && maxHH < fwd_score; i++) {
            for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }
            uint32_t gpac = T.at(tidx - i);
            gidx = gidx - 8;
            int32_t ridx = (rend_reg << 3) - 1;
            int32_t global_idx = 0;
            for (int32_t j = 0; j < rend_reg && maxHH < fwd_score; j++) {
                uint32_t rpac = Q.at(qidx - j);
                for (int k = 0; k <= 28 && maxHH < fwd_score; k += 4) {
                    uint32_t rbase = (rpac >> k) & 15u;
                    h[0] = gH[global_idx];
                    int32_t e = gE[global_idx];
                    for (int l = 0, m = 1; l <= 28; l += 4, m++) {               /* CORE_LOCAL_COMPUTE_START */
                        uint32_t gbase = (gpac >> l) & 15u;
                        int32_t s = sub_local(c, rbase, gbase);
                        int32_t tmp_hm = p[m] + s;
                        h[m] = std::max(tmp_hm, f[m]);
                        h[m] = std::max(h[m], e);
                        h[m] = std::max(h[m], 0);
                        f[m] = std::max(tmp_hm - c.gapoe, f[m] - c.gape);
                        e = std::max(tmp_hm - c.gapoe, e - c.gape);
                        maxXY_y = (maxHH < h[m]) ? gidx + (m - 1) : maxXY_y;      /* :41, recorded as gidx+(m-1) */
                        maxHH = (maxHH < h[m]) ? h[m] : maxHH;
                        p[m] = h[m - 1];
                    }
                    gH[global_idx] = t16(h[8]);
                    gE[global_idx] = t16(e);
                    maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;
                    prev_maxHH = std::max(maxHH, prev_maxHH);
                    ridx--; global_idx++;
                }
            }
        }
        o.qstart = maxXY_x; o.tstart = maxXY_y;                                   /* :508-509 */
    }
}

/* --------------------------------------------------------------- GLOBAL ---- */
/* gasal_global_kernel<S>, src/kernels/global.h:30-303; cores :4-26. */
void global_pair(const Sc &c, const Words &Q, const Words &T, uint64_t qb, uint64_t tb,
                 uint32_t read_len, uint32_t ref_len, PairOut &o, TB *tbm) {
    const int32_t QW = (int32_t)regs_of(read_len), TW = (int32_t)regs_of(ref_len);
    std::vector<int32_t> gH((size_t)8 * QW + 8), gE((size_t)8 * QW + 8);
    gH[0] = 0; gE[0] = MINUS_INF;                                                 /* :57-60 */
    for (size_t i = 1; i < gH.size(); i++) { gH[i] = t16(-(c.gapo + (c.gape * (int32_t)i))); gE[i] = MINUS_INF; }
    int32_t h[9], f[9], p[9], max_h[9];
    for (int m = 0; m < 9; m++) max_h[m] = 0;   /* uninitialised in the reference when read_len == 0 */
    int32_t u = 0, r = 0;
    h[u++] = 0; p[r++] = 0;                                                       /* :63-64 */
    if (tbm) tbm->init(TW, QW);
    for (int32_t i = 0; i < TW; i++) {
        int32_t ridx = 0;
        for (int m = 1; m < 9; m++, u++, r++) {                                   /* :67-71 */
            h[m] = -(c.gapo + (c.gape * u));
            f[m] = MINUS_INF;
            p[m] = r == 1 ? 0 : -(c.gapo + (c.gape * (r - 1)));
        }
        uint32_t gpac = T.at((int64_t)tb + i);
        for (int32_t j = 0; j < QW; j++) {
            uint32_t rpac = Q.at((int64_t)qb + j);
            for (int k = 28; k >= 0; k -= 4) {
                uint32_t rbase = (rpac >> k) & 15u;
                h[0] = gH[ridx];
                int32_t e = gE[ridx];
                uint32_t dir = 0;
                for (int l = 28, m = 1; m < 9; l -= 4, m++) {
                    uint32_t gbase = (gpac >> l) & 15u;
                    int32_t s = sub_global(c, rbase, gbase);
                    int32_t tmp_hm = p[m] + s;
                    uint32_t m_or_x = tmp_hm >= p[m] ? 0u : 1u;
                    h[m] = std::max(tmp_hm, f[m]);
                    h[m] = std::max(h[m], e);
                    int sh = 28 - ((m - 1) << 2);
                    dir |= h[m] == tmp_hm ? (m_or_x << sh) : (h[m] == f[m] ? (3u << sh) : (2u << sh));  /* :21 */
                    dir |= (tmp_hm - c.gapoe) > (f[m] - c.gape) ? 0u : (1u << (31 - ((m - 1) << 2)));
                    f[m] = (tmp_hm - c.gapoe) > (f[m] - c.gape) ? (tmp_hm - c.gapoe) : (f[m] - c.gape);
                    dir |= (tmp_hm - c.gapoe) > (e - c.gape) ? 0u : (1u << (30 - ((m - 1) << 2)));
                    e = (tmp_hm - c.gapoe) > (e - c.gape) ? (tmp_hm - c.gapoe) : (e - c.gape);
                    p[m] = h[m - 1];
                }
                gH[ridx] = t16(h[8]);
                gE[ridx] = t16(e);
                if (tbm) tbm->put(i, ridx, dir);
                ridx++;
                if (ridx == (int32_t)read_len)                                      /* :98-103 */
                    for (int m = 1; m < 9; m++) max_h[m] = h[m];
            }
        }
    }
    if (TW > 0) o.score = max_h[8 - ((TW << 3) - (int32_t)ref_len)];             /* :299 */
    else o.score = 0;
}

/* ----------------------------------------------------------- SEMI_GLOBAL ---- */
/* gasal_semi_global_kernel<T,S,B,HEAD,TAIL>, src/kernels/semiglobal_kernel_template.h:39-388,
 * core CORE_COMPUTE_SEMIGLOBAL :17-28. */
struct SemiState {
    std::vector<int32_t> gH, gE;
};

inline void semi_core_column(const Sc &c, uint32_t rbase, uint32_t gpac, int32_t *h, int32_t *f, int32_t *p,
                             int32_t &hd_h, int32_t &hd_e) {
    h[0] = hd_h;
    int32_t e = hd_e;
    int32_t prev_hm_diff = h[0] - c.gapoe;
    for (int l = 28, m = 1; m < 9; l -= 4, m++) {
        uint32_t gbase = (gpac >> l) & 15u;
        int32_t s = sub_local(c, rbase, gbase);                                   /* :19 LOCAL macro */
        int32_t curr_hm_diff = h[m] - c.gapoe;
        f[m] = std::max(curr_hm_diff, f[m] - c.gape);
        curr_hm_diff = p[m] + s;
        curr_hm_diff = std::max(curr_hm_diff, f[m]);
        e = std::max(prev_hm_diff, e - c.gape);
        curr_hm_diff = std::max(curr_hm_diff, e);
        h[m] = curr_hm_diff;
        p[m] = prev_hm_diff + c.gapoe;
        prev_hm_diff = curr_hm_diff - c.gapoe;
    }
    hd_h = t16(h[8]);
    hd_e = t16(e);
}

void semi_init_global(const Sc &c, int head, std::vector<int32_t> &gH, std::vector<int32_t> &gE) {
    if (head == D_QUERY || head == D_BOTH) {                                      /* :87-92 */
        std::fill(gH.begin(), gH.end(), 0); std::fill(gE.begin(), gE.end(), 0);
    } else {                                                                      /* :93-99 */
        gH[0] = 0; gE[0] = MINUS_INF;
        for (size_t i = 1; i < gH.size(); i++) { gH[i] = t16(-(c.gapo + (c.gape * (int32_t)i))); gE[i] = MINUS_INF; }
    }
}

void semi_strip_init(const Sc &c, int head, int32_t *h, int32_t *f, int32_t *p, int32_t &u, int32_t &r) {
    if (head == D_TARGET || head == D_BOTH) {                                     /* :114-121 */
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = MINUS_INF; p[m] = 0; }
    } else {                                                                      /* :122-129 */
        for (int m = 1; m < 9; m++, u++, r++) {
            h[m] = -(c.gapo + (c.gape * (u - 1)));
            f[m] = MINUS_INF;
            p[m] = r == 1 ? 0 : -(c.gapo + (c.gape * (r - 1)));
        }
    }
}

void semi_pair(const Sc &c, const Words &Q, const Words &T, uint64_t qb, uint64_t tb,
               uint32_t read_len, uint32_t ref_len, int start, bool second, int head, int tail,
               PairOut &o) {
    const int32_t QW = (int32_t)regs_of(read_len), TW = (int32_t)regs_of(ref_len);
    int32_t maxHH = MINUS_INF;                                                    /* :49 */
    int32_t maxXY_x = (int32_t)ref_len, maxXY_y = (int32_t)read_len;             /* :63-64 */
    int32_t maxHH_second = MINUS_INF;                                             /* :71 */
    int32_t maxXY_x_second = (int32_t)ref_len, maxXY_y_second = (int32_t)read_len; /* :73-74 */
    /* The reference's global[] has MAX_QUERY_LEN entries and the QUERY-tail scans
     * walk all of them (:187, :364); here the array is conceptually unbounded:
     * entries past the computed columns keep their initial values. */
    const size_t gsz = std::max<size_t>((size_t)8 * QW, std::max<size_t>(ref_len, read_len)) + 8;
    std::vector<int32_t> gH(gsz), gE(gsz);
    int32_t h[9] = {0}, f[9] = {0}, p[9] = {0};
    int32_t u = 0, r = 0;
    semi_init_global(c, head, gH, gE);
    if (head == D_QUERY || head == D_NONE) { u = 0; r = 0; h[u++] = 0; p[r++] = 0; } /* :101-107 */

    for (int32_t i = 0; i < TW; i++) {
        int32_t gidx = i << 3, ridx = 0;
        semi_strip_init(c, head, h, f, p, u, r);
        uint32_t gpac = T.at((int64_t)tb + i);
        for (int32_t j = 0; j < QW; j++) {
            uint32_t rpac = Q.at((int64_t)qb + j);
            for (int k = 28; k >= 0; k -= 4) {
                uint32_t rbase = (rpac >> k) & 15u;
                semi_core_column(c, rbase, gpac, h, f, p, gH[ridx], gE[ridx]);
                ridx++;
                if (tail == D_TARGET || tail == D_BOTH) {                         /* :160-178 */
                    if (ridx == (int32_t)read_len) {
                        for (int m = 1; m < 9; m++) {
                            bool upd = h[m] > maxHH && (gidx + m - 1) < (int32_t)ref_len;
                            maxXY_y = upd ? gidx + (m - 1) : maxXY_y;
                            maxHH = upd ? h[m] : maxHH;
                            if (second) {
                                bool ov = (h[m] > maxHH_second && h[m] < maxHH && (gidx + m - 1) < (int32_t)ref_len);
                                maxXY_y_second = ov ? gidx + (m - 1) : maxXY_y_second;
                                maxHH_second = ov ? h[m] : maxHH_second;
                            }
                        }
                    }
                }
            }
        }
    }
    if (tail == D_QUERY || tail == D_BOTH) {                                      /* :185-214 */
        for (size_t m = 0; m < gsz; m++) {
            int32_t score_tmp = gH[m];
            if (score_tmp > maxHH && (int64_t)m < (int64_t)read_len) { maxXY_x = (int32_t)m; maxHH = score_tmp; }
            if (second) {
                bool ov = (score_tmp > maxHH_second && score_tmp < maxHH && (int64_t)m < (int64_t)ref_len);
                maxXY_x_second = ov ? (int32_t)m : maxXY_x_second;
                maxHH_second = ov ? score_tmp : maxHH_second;
            }
        }
        if (maxXY_x != (int32_t)ref_len) maxXY_y = (int32_t)read_len;
        if (second && maxXY_x_second != (int32_t)ref_len) maxXY_y_second = (int32_t)read_len;
    }
    o.score = maxHH; o.tend = maxXY_y; o.qend = maxXY_x;                          /* :216-218 */
    if (second) { o.s_score = maxHH_second; o.s_tend = maxXY_y_second; o.s_qend = maxXY_x_second; }

    if (start == S_WITH_START) {                                                  /* :227-384 */
        /* reversed copies with padding nibble 0 (:232-266) */
        std::vector<uint32_t> rq((size_t)QW + 1, 0u), rt((size_t)TW + 1, 0u);
        for (int32_t ii = (int32_t)read_len - 1, kk = 0; ii >= 0; ii--, kk++) {
            uint32_t oreg = (uint32_t)ii >> 3, opos = (((oreg + 1) << 3) - (uint32_t)ii) - 1;
            uint32_t rreg = (uint32_t)kk >> 3, rpos = (((rreg + 1) << 3) - (uint32_t)kk) - 1;
            uint32_t sym = (Q.at((int64_t)qb + oreg) >> (opos << 2)) & 15u;
            rq[rreg] |= (sym << (rpos << 2));
        }
        for (int32_t ii = (int32_t)ref_len - 1, kk = 0; ii >= 0; ii--, kk++) {
            uint32_t oreg = (uint32_t)ii >> 3, opos = (((oreg + 1) << 3) - (uint32_t)ii) - 1;
            uint32_t rreg = (uint32_t)kk >> 3, rpos = (((rreg + 1) << 3) - (uint32_t)kk) - 1;
            uint32_t sym = (T.at((int64_t)tb + oreg) >> (opos << 2)) & 15u;
            rt[rreg] |= (sym << (rpos << 2));
        }
        int32_t gend_pos = maxXY_y, fwd_score = maxHH;
        int32_t d = TW - ((gend_pos >> 3) + 1);
        int32_t gend_reg = d > 0 ? d - 1 : d;                                     /* :273 */
        maxHH = MINUS_INF; maxXY_y = 0;
        semi_init_global(c, head, gH, gE);
        if (head == D_QUERY || head == D_NONE) { u = 0; r = 0; h[u++] = 0; p[r++] = 0; }
        for (int32_t i = gend_reg; i < TW && maxHH < fwd_score; i++) {
            int32_t gidx = i << 3, ridx = 0;
            semi_strip_init(c, head, h, f, p, u, r);
            uint32_t gpac = (i >= 0 && i < (int32_t)rt.size()) ? rt[i] : 0u;
            for (int32_t j = 0; j < QW && maxHH < fwd_score; j++) {
                uint32_t rpac = rq[j];
                for (int k = 28; k >= 0; k -= 4) {
                    uint32_t rbase = (rpac >> k) & 15u;
                    semi_core_column(c, rbase, gpac, h, f, p, gH[ridx], gE[ridx]);
                    ridx++;
                    if (tail == D_TARGET || tail == D_BOTH) {
                        if (ridx == (int32_t)read_len) {
                            for (int m = 1; m < 9; m++) {
                                bool upd = h[m] > maxHH && (gidx + m - 1) < (int32_t)ref_len;
                                maxXY_y = upd ? gidx + (m - 1) : maxXY_y;
                                maxHH = upd ? h[m] : maxHH;
                            }
                        }
                    }
                }
            }
        }
        if (tail == D_QUERY || tail == D_BOTH) {                                  /* :362-378 */
            for (size_t m = 0; m < gsz; m++) {
                int32_t score_tmp = gH[m];
                if (score_tmp > maxHH && (int64_t)m < (int64_t)read_len) { maxXY_x = (int32_t)m; maxHH = score_tmp; }
            }
            if (maxXY_x != (int32_t)ref_len) maxXY_y = (int32_t)read_len;
        }
        o.tstart = ((int32_t)ref_len - 1) - maxXY_y;                              /* :380-381 */
        o.qstart = ((int32_t)read_len - 1) - maxXY_x;
    }
}

/* --------------------------------------------------------------- BANDED ---- */
/* gasal_banded_tiled_kernel, src/kernels/banded.h:10-139. */
void banded_pair(const Sc &c, const Words &Q, const Words &T, uint64_t qb, uint64_t tb,
                 uint32_t read_len, uint32_t ref_len, int32_t k_band_width, PairOut &o) {
    const int32_t QW = (int32_t)regs_of(read_len), TW = (int32_t)regs_of(ref_len);
    int32_t maxHH = 0, prev_maxHH = 0, maxXY_x = 0, maxXY_y = 0;
    std::vector<int32_t> gH((size_t)8 * QW + 8, 0), gE((size_t)8 * QW + 8, 0);
    int32_t h[9], f[9], p[9];
    const int32_t k_other_band_width = (TW - (QW - k_band_width));              /* :35 */
    for (int32_t i = 0; i < TW; i++) {
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }
        uint32_t gpac = T.at((int64_t)tb + i);
        int32_t gidx = i << 3;
        int32_t ridx = std::max(0, i - k_other_band_width + 1) << 3;             /* :83 */
        int32_t last_tile = std::min(k_band_width + i, QW);                      /* :84 */
        for (int32_t j = ridx >> 3; j < last_tile; j++) {
            uint32_t rpac = Q.at((int64_t)qb + j);
            for (int k = 28; k >= 0; k -= 4) {
                uint32_t rbase = (rpac >> k) & 0x0Fu;
                h[0] = gH[ridx];
                int32_t e = gE[ridx];
                for (int l = 28, m = 1; m < 9; l -= 4, m++) {                     /* :100-115 */
                    uint32_t gbase = (gpac >> l) & 15u;
                    int32_t s = sub_local(c, rbase, gbase);
                    f[m] = std::max(h[m] - c.gapoe, f[m] - c.gape);
                    h[m] = p[m] + s;
                    h[m] = std::max(h[m], f[m]);
                    h[m] = std::max(h[m], 0);
                    e = std::max(h[m - 1] - c.gapoe, e - c.gape);
                    h[m] = std::max(h[m], e);
                    maxXY_y = (maxHH < h[m]) ? gidx + (m - 1) : maxXY_y;          /* FIND_MAX src/gasal_kernels.h:61-63 */
                    maxHH = (maxHH < h[m]) ? h[m] : maxHH;
                    p[m] = h[m - 1];
                }
                gH[ridx] = t16(h[8]);
                gE[ridx] = t16(e);
                maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;
                prev_maxHH = std::max(maxHH, prev_maxHH);
                ridx++;
            }
        }
    }
    o.score = maxHH; o.qend = maxXY_x; o.tend = maxXY_y;                          /* :133-135 */
}

/* ------------------------------------------------------------------ KSW ---- */
/* gasal_ksw_kernel<B>, src/kernels/ksw_kernel_template.h:46-199.
 * Extension (not in the reference): w >= 0 bands each row to query columns
 * [i - w, i + w] and zdrop > 0 enables the reference's own z-drop branch
 * (:164-170, hard-coded off there).  The band follows bwa's ksw_extend2
 * (ksw.c, which the reference transcribes, :27-29): w is first clamped to
 * (qlen * max_score - gap_open) / gap_extend + 1 for insertions and
 * deletions alike, then beg = max(beg, i - w), end = min(end, i + w + 1, qlen)
 * at the start of row i.  Parity of the extension is unpinned (bwa is not
 * available here). */
void ksw_pair(const Sc &c, const Words &Q, const Words &T, uint64_t qb, uint64_t tb,
              uint32_t qlen, uint32_t tlen, uint32_t h0, int32_t w, int32_t zdrop_ext, int32_t end_bonus, PairOut &o) {
    const uint32_t query_batch_regs = (qlen >> 3) + 1;                            /* :56-57 */
    const uint32_t target_batch_regs = (tlen >> 3) + 1;
    const int zdrop = zdrop_ext;                                                  /* :62 (0 there) */
    if (w >= 0) {   /* bwa's ksw_extend2 clamp (insertions and deletions cost alike here) */
        const int maxsc = std::max(c.match, 0);
        int mx = (int)((double)((int)qlen * maxsc + end_bonus - c.gapo) / c.gape + 1.0);
        mx = mx > 1 ? mx : 1;
        w = w < mx ? w : mx;
    }
    const int o_del = c.gapo, o_ins = c.gapo, e_del = c.gape, e_ins = c.gape;
    struct eh_t { int32_t h, e; };
    std::vector<eh_t> eh((size_t)qlen + 2, eh_t{0, 0});                          /* :69-75 */
    int i, j, oe_del = o_del + e_del, oe_ins = o_ins + e_ins, beg, end, max, max_i, max_j, max_ie, gscore, max_off;
    eh[0].h = (int32_t)h0;                                                        /* :78-81 */
    eh[1].h = h0 > (uint32_t)oe_ins ? (int32_t)(h0 - (uint32_t)oe_ins) : 0;
    for (j = 2; j <= (int)qlen && eh[j - 1].h > e_ins; ++j) eh[j].h = eh[j - 1].h - e_ins;
    max = (int32_t)h0, max_i = max_j = -1;                                        /* :84-87 */
    max_ie = -1, gscore = -1;
    max_off = 0;
    beg = 0, end = (int)qlen;
    /* z-drop (an extension: the reference compiles it out with zdrop 0) ends
     * the whole extension as in bwa's ksw_extend2; the m == 0 break below
     * keeps the reference's loop structure and leaves the 8-row tile only. */
    bool zstop = false;
    for (uint32_t target_tile_id = 0; target_tile_id < target_batch_regs && !zstop; target_tile_id++) {
        uint32_t gpac = T.at((int64_t)tb + target_tile_id);
        for (uint32_t target_base_id = 0; target_base_id < 8; target_base_id++) {
            i = (int)(target_tile_id * 8 + target_base_id);
            if (i >= (int)tlen) break;                                            /* :98-99 */
            uint32_t gbase = (gpac >> (32 - (target_base_id + 1) * 4)) & 0x0Fu;
            int t, f = 0, h1, m = 0, mj = -1;
            if (w >= 0) {                                                         /* band (extension) */
                if (beg < i - w) beg = i - w;
                if (end > i + w + 1) end = i + w + 1;
                if (end > (int)qlen) end = (int)qlen;
            }
            if (beg == 0) {                                                       /* :105-110 */
                h1 = (int32_t)(h0 - (uint32_t)(o_del + e_del * (i + 1)));
                if (h1 < 0) h1 = 0;
            } else h1 = 0;
            for (uint32_t query_tile_id = 0; query_tile_id < query_batch_regs; query_tile_id++) {
                uint32_t rpac = Q.at((int64_t)qb + query_tile_id);
                for (uint32_t query_base_id = 0; query_base_id < 8; query_base_id++) {
                    j = (int)(query_tile_id * 8 + query_base_id);
                    if (j < beg) continue;
                    if (j >= end) break;
                    uint32_t rbase = (rpac >> (32 - (query_base_id + 1) * 4)) & 0x0Fu;
                    eh_t *pp = &eh[j];
                    int h, M = pp->h, e = pp->e;
                    pp->h = h1;
                    int32_t s = sub_local(c, rbase, gbase);
                    M = M ? M + s : 0;
                    h = M > e ? M : e;
                    h = h > f ? h : f;
                    h1 = h;
                    mj = m > h ? mj : j;
                    m = m > h ? m : h;
                    t = M - oe_del; t = t > 0 ? t : 0;
                    e -= e_del; e = e > t ? e : t;
                    pp->e = e;
                    t = M - oe_ins; t = t > 0 ? t : 0;
                    f -= e_ins; f = f > t ? f : t;
                }
            }
            eh[end].h = h1;                                                       /* :153-154 */
            eh[end].e = 0;
            if (j == (int)qlen) {                                                 /* :155-158 */
                max_ie = gscore > h1 ? max_ie : i;
                gscore = gscore > h1 ? gscore : h1;
            }
            if (m == 0) break;                                                    /* :159-160, leaves the 8-row tile only */
            if (m > max) {
                max = m, max_i = i, max_j = mj;
                max_off = max_off > abs(mj - i) ? max_off : abs(mj - i);
            } else if (zdrop > 0) {
                if (i - max_i > mj - max_j) {
                    if (max - m - ((i - max_i) - (mj - max_j)) * e_del > zdrop) zstop = true;
                } else {
                    if (max - m - ((mj - max_j) - (i - max_i)) * e_ins > zdrop) zstop = true;
                }
                if (zstop) break;
            }
            for (j = beg; (j < end) && eh[j].h == 0 && eh[j].e == 0; ++j) ;      /* :178-183 */
            beg = j;
            for (j = end; (j >= beg) && eh[j].h == 0 && eh[j].e == 0; --j) ;
            end = j + 2 < (int)qlen ? j + 2 : (int)qlen;
        }
    }
    (void)max_off;
    if (gscore <= 0 || gscore <= max - end_bonus) {                               /* :188-197, PEN_CLIP5 */
        o.score = max; o.qend = max_j + 1; o.tend = max_i + 1;
    } else {
        o.score = gscore; o.qend = (int32_t)qlen; o.tend = max_ie + 1;
    }
}

/* --------------------------------------------------------------- get_tb ---- */
/* gasal_get_tb<T>, src/kernels/get_tb.h:4-149. */
void get_tb_pair(const Sc &c, bool is_local, const TB &tbm, uint32_t read_len, uint32_t ref_len,
                 PairOut &o, uint8_t *cigar, uint32_t &n_ops_out) {
    int i, j, total_score = 0, curr_score = 0;
    if (is_local) { i = o.tend; j = o.qend; total_score = o.score; curr_score = 0; }   /* :16-20 */
    else { i = (int)ref_len; j = (int)read_len; }                                        /* :21-24 */
    uint32_t prev_op_to_fill = 0;
    int read_len_8 = read_len % 8 ? (int)(read_len + (8 - (read_len % 8))) : (int)read_len;
    int n_ops = 0, op_select = 3, op_shift = 0, count = 0, offset = 0;
    uint32_t op_to_fill;
    while (i >= 0 && j >= 0) {                                                    /* :47 */
        int64_t cell = ((((int64_t)(i >> 3) * read_len_8) << 3) + ((int64_t)j << 3) + (i & 7));
        uint32_t cell_op = tbm.cell(cell);
        uint32_t op = (cell_op >> op_shift) & (uint32_t)op_select;                 /* :74 */
        op_to_fill = op == 0 || op_select == 3 ? op : (uint32_t)op_shift;        /* :78 */
        op_select = op == 0 || (op == 1 && op_select == 3) ? 3 : 1;               /* :80 */
        op_shift = op == 0 || (op == 1 && op_select == 3) ? 0 : ((op == 2 || op == 3) ? (int)op : op_shift); /* :82 */
        if (count < 63 && op_to_fill == prev_op_to_fill) {                        /* :87-98 */
            count++;
        } else {
            if (count > 0) { cigar[offset++] = (uint8_t)(prev_op_to_fill | (uint32_t)(count << 2)); n_ops++; }
            count = 1;
        }
        if (is_local) {                                                           /* :100-103 */
            curr_score += ((op_to_fill == 2 || op_to_fill == 3) && prev_op_to_fill != op_to_fill) ? -c.gapoe
                        : ((op_to_fill == 2 || op_to_fill == 3) ? -c.gape : (op_to_fill == 1 ? -c.mismatch : c.match));
            if (curr_score == total_score) break;
        }
        prev_op_to_fill = op_to_fill;
        i = op_to_fill == 0 || op_to_fill == 1 || op_to_fill == 2 ? i - 1 : i;  /* :107-108 */
        j = op_to_fill == 0 || op_to_fill == 1 || op_to_fill == 3 ? j - 1 : j;
    }
    cigar[offset++] = (uint8_t)(prev_op_to_fill | (uint32_t)(count << 2));       /* :113-117 */
    n_ops++;
    if (!is_local) {                                                              /* :119-139 */
        while (i >= 0) {
            uint8_t resd = (uint8_t)((i + 1) <= 63 ? (i + 1) : 63);
            cigar[offset++] = (uint8_t)(2u | (uint32_t)(resd << 2)); n_ops++; i -= 63;
        }
        while (j >= 0) {
            uint8_t resd = (uint8_t)((j + 1) <= 63 ? (j + 1) : 63);
            cigar[offset++] = (uint8_t)(3u | (uint32_t)(resd << 2)); n_ops++; j -= 63;
        }
    }
    if (is_local) { o.tstart = i; o.qstart = j; }                                 /* :142-145 */
    n_ops_out = (uint32_t)n_ops;                                                  /* :146 */
}

}  // namespace

extern "C" uint32_t oracle_cigar_capacity(uint32_t query_len, uint32_t target_len) {
    uint32_t q8 = 8 * regs_of(query_len), t8 = 8 * regs_of(target_len);
    return q8 + t8 + (q8 + t8) / 32 + 16;
}

extern "C" void oracle_pack(const uint8_t *unpacked, uint32_t n_bytes, uint32_t *packed) {
    /* src/kernels/pack_rc_seqs.h:19-34: reg1 = bytes 0..3, reg2 = bytes 4..7 (little-endian words) */
    for (uint32_t w = 0; w < n_bytes / 8; w++) {
        const uint8_t *b = unpacked + 8 * (size_t)w;
        uint32_t r = 0;
        for (int k = 0; k < 8; k++) r |= ((uint32_t)b[k] & 15u) << (28 - 4 * k);
        packed[w] = r;
    }
}

namespace {
/* PTX shl/shr clamp shift amounts >= 32 to a zero result (public PTX ISA); the
 * reference's `<< (32-nbr_N)` with nbr_N==0 relies on it
 * (src/kernels/pack_rc_seqs.h:136,153). */
inline uint32_t shl32(uint32_t x, uint32_t n) { return n >= 32 ? 0u : (x << n); }
inline uint32_t shr32(uint32_t x, uint32_t n) { return n >= 32 ? 0u : (x >> n); }
}

extern "C" void oracle_rc_one(uint32_t *w, uint32_t len, uint8_t op, int32_t n_code) {
    /* src/kernels/pack_rc_seqs.h:56-212 for one sequence (w = its first word). */
    uint32_t regs = regs_of(len);
    uint32_t to_swap = (regs >> 1) + (regs & 1);
    if (regs == 0) return;
    auto at = [&](int64_t k) -> uint32_t { return k >= 0 ? w[k] : 0u; };  /* word -1: value discarded */
    if (op & 1) {
        uint32_t nbr_N = 0;
        for (int jj = 0; jj < 32; jj += 4)
            nbr_N += (((w[regs - 1] & (0x0Fu << jj)) >> jj) == (uint32_t)n_code);  /* :115 full N_CODE */
        nbr_N = (nbr_N << 2) & 0xFF;                                                   /* uint8_t */
        for (uint32_t i = 0; i < to_swap; i++) {
            uint32_t rpac_1 = w[i];
            uint32_t rpac_2 = shl32(at((int64_t)regs - 2 - i), 32 - nbr_N) | shr32(w[regs - 1 - i], nbr_N);
            uint32_t rev1 = 0, rev2 = 0;
            for (int k = 28; k >= 0; k -= 4) {
                rev1 |= ((rpac_1 & (0x0Fu << k)) >> k) << (28 - k);
                rev2 |= ((rpac_2 & (0x0Fu << k)) >> k) << (28 - k);
            }
            uint32_t lowmask = shl32(1u, nbr_N) - 1u;
            uint32_t to_queue_1 = shl32(rev1, nbr_N) | (w[regs - 1 - i] & lowmask);
            uint32_t to_queue_2 = (at((int64_t)regs - 2 - i) & (0xFFFFFFFFu - lowmask)) | shr32(rev1, 32 - nbr_N);
            w[i] = rev2;
            w[regs - 1 - i] = to_queue_1;
            if (i != to_swap - 1) w[regs - 2 - i] = to_queue_2;
        }
    }
    if (op & 2) {                                                                /* :169-205 */
        for (uint32_t i = 0; i < regs; i++) {
            uint32_t rpac = w[i];
            for (int k = 28; k >= 0; k -= 4) {
                uint32_t nt = (rpac & (0x0Fu << k)) >> k;
                switch (nt) {
                    case 1: nt = 4; break;   /* A -> T */
                    case 3: nt = 7; break;   /* C -> G */
                    case 4: nt = 1; break;   /* T -> A */
                    case 7: nt = 3; break;   /* G -> C */
                    default: break;
                }
                rpac = (rpac & (0xFFFFFFFFu - (0x0Fu << k))) | (nt << k);
            }
            w[i] = rpac;
        }
    }
}

extern "C" int oracle_align_batch(const oracle_scores_t *scp, const oracle_params_t *pr,
                                  const uint8_t *query_batch, uint32_t query_batch_bytes,
                                  const uint8_t *target_batch, uint32_t target_batch_bytes,
                                  const uint32_t *query_offsets, const uint32_t *target_offsets,
                                  const uint32_t *query_lens, const uint32_t *target_lens,
                                  const uint8_t *query_ops, const uint8_t *target_ops,
                                  const uint32_t *seed_scores, uint32_t n_alns,
                                  oracle_out_t *out, int n_threads) {
    if (n_alns == 0 || query_batch_bytes == 0 || target_batch_bytes == 0) return 1;   /* src/gasal_align.cu:32-43 */
    if ((query_batch_bytes % 8) || (target_batch_bytes % 8)) return 2;                /* :45-52 */
    Sc c;
    c.match = scp->match; c.mismatch = scp->mismatch; c.gapo = scp->gap_open; c.gape = scp->gap_extend;
    c.gapoe = scp->gap_open + scp->gap_extend;                                       /* :334 */
    c.n_code = scp->n_code; c.n_value = (uint32_t)scp->n_code & 0xFu;
    c.has_np = scp->has_n_penalty != 0; c.np = scp->n_penalty;

    /* pack or alias (src/gasal_align.cu:190-201; src/ctors.cpp:64-67) */
    std::vector<uint32_t> pq, pt;
    if (pr->is_packed) {
        pq.assign(query_batch_bytes / 4, 0u); pt.assign(target_batch_bytes / 4, 0u);
        memcpy(pq.data(), query_batch, (size_t)query_batch_bytes);
        memcpy(pt.data(), target_batch, (size_t)target_batch_bytes);
    } else {
        pq.assign(query_batch_bytes / 8, 0u); pt.assign(target_batch_bytes / 8, 0u);
        oracle_pack(query_batch, query_batch_bytes, pq.data());
        oracle_pack(target_batch, target_batch_bytes, pt.data());
    }
    /* reverse/complement (src/gasal_align.cu:232-246), pair order */
    if (pr->is_reverse_complement) {
        for (uint32_t k = 0; k < n_alns; k++) {
            if (query_ops && query_ops[k]) oracle_rc_one(pq.data() + (query_offsets[k] >> 3), query_lens[k], query_ops[k], c.n_code);
            if (target_ops && target_ops[k]) oracle_rc_one(pt.data() + (target_offsets[k] >> 3), target_lens[k], target_ops[k], c.n_code);
        }
    }
    Words Q{pq.data(), pq.size()}, T{pt.data(), pt.size()};
    const int algo = pr->algo, start = pr->start_pos;
    const bool second = pr->second_best != 0;
    const bool tb_alg = (algo == A_LOCAL || algo == A_GLOBAL) && start == S_WITH_TB;

#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t kk = 0; kk < (int64_t)n_alns; kk++) {
        uint32_t k = (uint32_t)kk;
        PairOut o;
        uint64_t qb = query_offsets[k] >> 3, tb = target_offsets[k] >> 3;
        uint32_t ql = query_lens[k], tl = target_lens[k];
        TB tbm;
        uint32_t n_ops = 0;
        switch (algo) {
            case A_LOCAL:
                local_pair(c, Q, T, qb, tb, ql, tl, start, second, o, tb_alg ? &tbm : nullptr);
                break;
            case A_GLOBAL:
                global_pair(c, Q, T, qb, tb, ql, tl, o, tb_alg ? &tbm : nullptr);
                break;
            case A_SEMI_GLOBAL:
                semi_pair(c, Q, T, qb, tb, ql, tl, start, second, pr->head, pr->tail, o);
                break;
            case A_BANDED:
                banded_pair(c, Q, T, qb, tb, ql, tl, pr->k_band >> 3, o);         /* src/gasal_align.h:80 */
                break;
            case A_KSW:
                ksw_pair(c, Q, T, qb, tb, ql, tl, seed_scores ? seed_scores[k] : 0u, pr->ksw_w, pr->ksw_zdrop, pr->ksw_end_bonus, o);
                break;
            default:
                break;   /* UNKNOWN / MICROLOCAL: nothing launched (src/gasal_align.cu:20-21) */
        }
        if (start == S_WITH_TB) {
            uint8_t *cg = out->cigar ? out->cigar + out->cigar_offsets[k] : nullptr;
            if (tb_alg) {
                std::vector<uint8_t> tmp(oracle_cigar_capacity(ql, tl));
                get_tb_pair(c, algo == A_LOCAL, tbm, ql, tl, o, tmp.data(), n_ops);
                if (cg) memcpy(cg, tmp.data(), n_ops);
            } else {
                /* No traceback kernel for this algorithm: the host still copies the
                 * query batch bytes as "CIGAR" and the untouched query lengths as
                 * n_cigar_ops (src/gasal_align.cu:280-283). */
                n_ops = ql;
                if (cg) {
                    uint64_t off = query_offsets[k];
                    for (uint32_t b = 0; b < ql; b++)
                        cg[b] = (off + b < query_batch_bytes) ? query_batch[off + b] : 0;
                }
            }
        }
        if (out->aln_score) out->aln_score[k] = o.score;
        if (out->query_batch_end) out->query_batch_end[k] = o.qend;
        if (out->target_batch_end) out->target_batch_end[k] = o.tend;
        if (out->query_batch_start) out->query_batch_start[k] = o.qstart;
        if (out->target_batch_start) out->target_batch_start[k] = o.tstart;
        if (out->second_score) out->second_score[k] = o.s_score;
        if (out->second_query_end) out->second_query_end[k] = o.s_qend;
        if (out->second_target_end) out->second_target_end[k] = o.s_tend;
        if (out->n_cigar_ops) out->n_cigar_ops[k] = n_ops;
    }
    return 0;
}
