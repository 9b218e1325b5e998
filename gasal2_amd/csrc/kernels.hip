// kernels.hip -- device pipeline of the batched alignment path for gfx950.
//
// Stage order follows gasal_aln_async (reference src/gasal_align.cu:190-283):
//   pack -> reverse/complement -> sort pairs by shape -> DP -> traceback.
//
// DP kernels here give ONE LANE PER PAIR.  Pairs are sorted by (query words,
// target words) so the 64 lanes of a wave walk the same loop bounds, and each
// wave owns a contiguous slice of device scratch laid out [column][lane] so the
// per-column boundary row (H,E as two int16 halves of one dword -- the
// reference's short2, which also reproduces its int16 truncation between 8-row
// strips) is read and written with fully coalesced 256-byte accesses.  The
// direction nibbles of WITH_TB are stored the same way, [strip][column][lane].
// Each lane walks its pair in the reference's order (8-row strips over the
// target, columns left to right, rows top to bottom), so every order-dependent
// output (end tie-breaks, second best, early exits) is the reference's.
#include "engine.h"

#include <string.h>
#include <type_traits>

namespace {

constexpr int32_t MINUS_INF = -32768;   // SHRT_MIN, src/gasal_kernels.h:35
constexpr int WAVE = 64;

enum { A_GLOBAL = 1, A_SEMI_GLOBAL = 2, A_LOCAL = 3, A_BANDED = 5, A_KSW = 6 };
enum { S_WITHOUT_START = 0, S_WITH_START = 1, S_WITH_TB = 2 };
enum { D_NONE = 0, D_QUERY = 1, D_TARGET = 2, D_BOTH = 3 };

__device__ __forceinline__ int32_t sx16(uint32_t v) { return (int32_t)(int16_t)(uint16_t)(v & 0xFFFFu); }
// (h & 0xFFFF) | (e << 16): one byte select (bytes 0-1 of h, then bytes 0-1 of e)
__device__ __forceinline__ uint32_t pk16(int32_t h, int32_t e) {
    return __builtin_amdgcn_perm((uint32_t)e, (uint32_t)h, 0x05040100u);
}
__device__ __forceinline__ int32_t t16(int32_t v) { return (int32_t)(int16_t)(uint16_t)((uint32_t)v & 0xFFFFu); }
__device__ __forceinline__ int32_t dregs(uint32_t len) { return (int32_t)((len >> 3) + ((len & 7) ? 1 : 0)); }

// DEV_GET_SUB_SCORE_LOCAL (src/gasal_kernels.h:40-42, :49-51)
__device__ __forceinline__ int32_t sub_local(const G2Scores &c, int32_t nval, uint32_t r, uint32_t g) {
    int32_t s = (r == g) ? c.match : -c.mismatch;
    return (r == c.n_value || g == c.n_value) ? nval : s;
}
// DEV_GET_SUB_SCORE_GLOBAL (src/gasal_kernels.h:44-46, :53-54)
__device__ __forceinline__ int32_t sub_global(const G2Scores &c, uint32_t r, uint32_t g) {
    int32_t s = (r == g) ? c.match : -c.mismatch;
    return (c.has_np && (r == c.n_value || g == c.n_value)) ? -c.np : s;
}

struct LaneCtx {
    uint32_t grp, lane, slot, pair;
    uint32_t qlen, tlen;
    int32_t QW, TW;
    const uint32_t *Q, *T;
    // sequence words as the fast LOCAL kernels read them: word j at Qp[j * ps]
    // (the group's transposed panels, ps = 64, or the pair's words, ps = 1)
    const uint32_t *Qp, *Tp;
    uint32_t ps;
    uint32_t *bd;     // boundary row: element c at bd[c * WAVE]
};

// Direction words of a wave slot are laid out per 8x8 tile, [strip][query
// word], 32 bytes per lane (8 columns), and a traceback walk stays inside one
// tile for eight columns.  A lane's tiles: its first chunk and the strip
// stride in query words, taken once per kernel (one region size for all
// groups, or per-group regions).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// A tile's eight words are 32 contiguous bytes per lane ([lane][8] within the
// 2 KiB chunk of a wave), so a traceback lane, whose walk has left its
// neighbours' tiles, fetches one line per tile record instead of two; the
// DP's two 16-byte stores per lane fill each line within one pair of
// instructions.  (Halves of all 64 lanes laid out contiguously, [half][lane][4],
// made every DP store a whole line but cost the traceback 4.78 against 3.72 ms
// on 1M pairs and the DP 0.2-0.7 ms less.)  Stores and the DP's loads are
// non-temporal: the tiles stream through L2 (written once, read once much
// later) and would otherwise evict the sequence words the lanes re-read.
constexpr uint32_t TB_HALF = 4u;   // word offset of a lane's second half
constexpr uint32_t TB_LANE = 8u;   // word stride between lanes
struct TbView {
    uint32_t *base;   // the group's chunk (0, 0) plus TB_LANE * lane
    uint32_t qw;
    __device__ __forceinline__ uint32_t *chunk(int32_t strip, int32_t word) const {
        return base + ((size_t)strip * qw + (size_t)word) * (WAVE * 8);
    }
    __device__ __forceinline__ void store(int32_t strip, int32_t word, const uint32_t *w) const {
        uint32_t *p = chunk(strip, word);
        const u32x4 a = {w[0], w[1], w[2], w[3]}, b = {w[4], w[5], w[6], w[7]};
        __builtin_nontemporal_store(a, reinterpret_cast<u32x4 *>(p));
        __builtin_nontemporal_store(b, reinterpret_cast<u32x4 *>(p + TB_HALF));
    }
    __device__ __forceinline__ void load_nt(int32_t strip, int32_t word, uint32_t *w) const {
        const uint32_t *p = chunk(strip, word);
        const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
        const u32x4 b = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p + TB_HALF));
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
        w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
    }
    __device__ __forceinline__ void load(int32_t strip, int32_t word, uint32_t *w) const {
        const uint32_t *p = chunk(strip, word);
        const uint4 a = *reinterpret_cast<const uint4 *>(p), b = *reinterpret_cast<const uint4 *>(p + TB_HALF);
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
        w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
    }
    __device__ __forceinline__ uint32_t at(int32_t strip, int32_t word, int k) const {
        return chunk(strip, word)[k < 4 ? k : TB_HALF + k - 4];
    }
};
__device__ __forceinline__ TbView tb_view(const G2Dev &d, uint32_t grp, uint32_t lane, uint32_t *region = nullptr) {
    TbView v;
    uint32_t *const r = region ? region : d.tb;
    if (d.tb_gbase) {
        v.base = r + d.tb_gbase[grp] + (size_t)lane * TB_LANE;
        v.qw = d.tb_gqw[grp];
    } else {
        v.qw = d.tb_qw8 >> 3;
        v.base = r + (size_t)grp * d.tb_tw * v.qw * WAVE * 8 + (size_t)lane * TB_LANE;
    }
    return v;
}
// a lane's whole record as ordinary stores: with two lane groups' records
// in flight per wave, non-temporal stores left lines half-written at
// eviction (1.27x the record bytes written to HBM against 1.04x, measured)
__device__ __forceinline__ void tb_store_rec(uint32_t *p, const uint32_t *w) {
    reinterpret_cast<u32x4 *>(p)[0] = u32x4{w[0], w[1], w[2], w[3]};
    reinterpret_cast<u32x4 *>(p + TB_HALF)[0] = u32x4{w[4], w[5], w[6], w[7]};
}
__device__ __forceinline__ void tb_load(const uint32_t *p, uint32_t *w) {
    const uint4 a = reinterpret_cast<const uint4 *>(p)[0], b = reinterpret_cast<const uint4 *>(p)[1];
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}
__device__ __forceinline__ void tb_store(uint32_t *p, const uint32_t *w) {
    reinterpret_cast<uint4 *>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4 *>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// goff: blocks before this kernel's lane-per-pair groups; slot0: sorted slots
// taken by the wavefront path.
__device__ __forceinline__ bool lane_setup(const G2Dev &d, LaneCtx &x, uint32_t goff = 0, uint32_t slot0 = 0) {
    x.lane = threadIdx.x & (WAVE - 1);
    x.grp = (blockIdx.x - goff) * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    x.slot = slot0 + x.grp * WAVE + x.lane;
    if (x.slot >= d.n) return false;
    x.pair = d.perm[x.slot];
    x.qlen = d.qlen[x.pair];
    x.tlen = d.tlen[x.pair];
    x.QW = dregs(x.qlen);
    x.TW = dregs(x.tlen);
    x.Q = d.pq + (d.qoff[x.pair] >> 3);
    x.T = d.pt + (d.toff[x.pair] >> 3);
    if (d.qpan) {
        x.Qp = d.qpan + (size_t)x.grp * d.qpan_w * WAVE + x.lane;
        x.Tp = d.tpan + (size_t)x.grp * d.tpan_w * WAVE + x.lane;
        x.ps = WAVE;
    } else {
        x.Qp = x.Q;
        x.Tp = x.T;
        x.ps = 1;
    }
    x.bd = d.bnd + (size_t)x.grp * d.bnd_cols * WAVE + x.lane;
    return true;
}

// ------------------------------------------------------------------ pack ----
// gasal_pack_kernel (src/kernels/pack_rc_seqs.h:13-53): byte b -> nibble b&15,
// 8 bases per word, base 0 in bits 31:28.  16 bytes in, 8 bytes out per thread.
__device__ __forceinline__ uint32_t pack8(uint32_t lo, uint32_t hi) {
    uint32_t r = ((lo & 15u) << 28) | (((lo >> 8) & 15u) << 24) | (((lo >> 16) & 15u) << 20) | (((lo >> 24) & 15u) << 16);
    r |= ((hi & 15u) << 12) | (((hi >> 8) & 15u) << 8) | (((hi >> 16) & 15u) << 4) | ((hi >> 24) & 15u);
    return r;
}

__global__ __launch_bounds__(256) void k_pack(const uint4 *__restrict__ in, uint2 *__restrict__ out, uint32_t n16,
                                              const uint2 *__restrict__ in_tail, uint32_t *__restrict__ out_tail,
                                              uint32_t has_tail) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        uint4 v = in[i];
        uint2 o;
        o.x = pack8(v.x, v.y);
        o.y = pack8(v.z, v.w);
        out[i] = o;
    }
    if (has_tail && blockIdx.x == 0 && threadIdx.x == 0) {
        uint2 v = *in_tail;
        *out_tail = pack8(v.x, v.y);
    }
}

// -------------------------------------------------------- reverse / comp ----
// gasal_reversecomplement_kernel (src/kernels/pack_rc_seqs.h:56-212).  Shift
// amounts of 32 give 0, the result of PTX shl/shr the reference relies on.
__device__ __forceinline__ uint32_t shl32(uint32_t x, uint32_t n) { return n >= 32 ? 0u : (x << n); }
__device__ __forceinline__ uint32_t shr32(uint32_t x, uint32_t n) { return n >= 32 ? 0u : (x >> n); }
__device__ __forceinline__ uint32_t nib_reverse(uint32_t x) {
    x = ((x & 0x0F0F0F0Fu) << 4) | ((x >> 4) & 0x0F0F0F0Fu);
    return __builtin_bswap32(x);
}
// Complement of the packed bases, all eight nibbles at once: A (1) <-> T (4)
// toggles bits 0 and 2, C (3) <-> G (7) bit 2; every other code is kept
// (src/kernels/pack_rc_seqs.h:56-212 maps nibble by nibble).  Bit 4k of b0..b3
// is bit 0..3 of nibble k.
__device__ __forceinline__ uint32_t nib_complement(uint32_t x) {
    const uint32_t M = 0x11111111u;
    const uint32_t b0 = x & M, b1 = (x >> 1) & M, b2 = (x >> 2) & M, b3 = (x >> 3) & M;
    const uint32_t at = (b0 ^ b2) & ~(b1 | b3) & M;   // nibble 1 or 4
    const uint32_t cg = b0 & b1 & ~b3;                // nibble 3 or 7
    return x ^ at ^ ((at | cg) << 2);
}

__device__ void rc_one(uint32_t *w, uint32_t len, uint32_t op, int32_t n_code) {
    uint32_t regs = (len >> 3) + ((len & 7) ? 1 : 0);
    if (regs == 0) return;
    uint32_t to_swap = (regs >> 1) + (regs & 1);
    if (op & 1u) {
        uint32_t nbr_N = 0;
        uint32_t last = w[regs - 1];
        for (int jj = 0; jj < 32; jj += 4) nbr_N += (((last & (0x0Fu << jj)) >> jj) == (uint32_t)n_code);
        nbr_N = (nbr_N << 2) & 0xFFu;
        for (uint32_t i = 0; i < to_swap; i++) {
            uint32_t rpac_1 = w[i];
            uint32_t wm2 = (regs >= 2 + i) ? w[regs - 2 - i] : 0u;   // word -1 is read but discarded
            uint32_t wm1 = w[regs - 1 - i];
            uint32_t rpac_2 = shl32(wm2, 32 - nbr_N) | shr32(wm1, nbr_N);
            uint32_t rev1 = nib_reverse(rpac_1), rev2 = nib_reverse(rpac_2);
            uint32_t lowmask = shl32(1u, nbr_N) - 1u;
            uint32_t to_queue_1 = shl32(rev1, nbr_N) | (wm1 & lowmask);
            uint32_t to_queue_2 = (wm2 & (0xFFFFFFFFu - lowmask)) | shr32(rev1, 32 - nbr_N);
            w[i] = rev2;
            w[regs - 1 - i] = to_queue_1;
            if (i != to_swap - 1) w[regs - 2 - i] = to_queue_2;
        }
    }
    if (op & 2u)
        for (uint32_t i = 0; i < regs; i++) w[i] = nib_complement(w[i]);
}

// One sequence (a query or target whose op is nonzero) per RC_T lanes.  When
// the last word holds no nibble equal to N_CODE (always so for a byte-sized
// N_CODE such as 'N'), the reference's reverse is a plain reversal of the
// padded word sequence: lane t swaps words (i, regs-1-i) for i = t, t+RC_T,
// ..., so each round's loads and stores cover RC_T consecutive words at both
// ends of the sequence (two 64-byte runs); otherwise the first lane runs the
// reference's sequential loop.  The last word and the first round's words are
// loaded together, so a sequence of up to 2 * RC_T words costs one memory round
// trip.  Every round's loads precede its stores, and rounds touch disjoint words.
constexpr uint32_t RC_T = 16;
__device__ __forceinline__ void rc_side(uint32_t *w, uint32_t len, uint32_t op, int32_t n_code, uint32_t t) {
    const uint32_t regs = (len >> 3) + ((len & 7) ? 1 : 0);
    if (regs == 0 || op == 0) return;
    const uint32_t to_swap = (regs >> 1) + (regs & 1);
    uint32_t i = t;
    uint32_t a = 0, b = 0;
    if (i < to_swap) {
        a = w[i];
        b = w[regs - 1 - i];
    }
    if (op & 1u) {   // decide between the plain reversal and the reference's loop
        const uint32_t last = w[regs - 1];
        bool plain = true;
        for (int jj = 0; jj < 32; jj += 4) plain &= (((last >> jj) & 15u) != (uint32_t)n_code);
        if (!plain) {
            if (t == 0) rc_one(w, len, op, n_code);
            return;
        }
    }
    while (i < to_swap) {
        const uint32_t j = regs - 1 - i;
        uint32_t na = (op & 1u) ? nib_reverse(b) : a;
        uint32_t nb = (op & 1u) ? nib_reverse(a) : b;
        if (op & 2u) {
            na = nib_complement(na);
            nb = nib_complement(nb);
        }
        w[i] = na;
        if (j != i) w[j] = nb;
        i += RC_T;
        if (i < to_swap) {
            a = w[i];
            b = w[regs - 1 - i];
        }
    }
}

// One pair per lane for the metadata (coalesced loads of 64 consecutive
// pairs), then the wave's sequences with a nonzero op -- queries, then targets
// -- compacted into LDS and taken four at a time, RC_T lanes each.  Pairs with
// both ops zero cost no memory beyond their metadata.
constexpr uint32_t RC_WAVES = 4;
struct RcItem {
    uint32_t word, len, op;
};
__global__ __launch_bounds__(RC_WAVES * WAVE) void k_rc(G2Dev d, int32_t n_code) {
    __shared__ RcItem items[RC_WAVES][2 * WAVE];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const uint32_t k = (blockIdx.x * RC_WAVES + wv) * WAVE + lane;
    uint32_t qo = 0, to = 0, qf = 0, ql = 0, tf = 0, tl = 0;
    if (k < d.n) {
        qo = d.qop[k];
        to = d.top[k];
        qf = d.qoff[k];
        ql = d.qlen[k];
        tf = d.toff[k];
        tl = d.tlen[k];
    }
    const uint64_t mq = __ballot(qo != 0), mt = __ballot(to != 0);
    const uint32_t nq = (uint32_t)__popcll(mq), nt = (uint32_t)__popcll(mt);
    const uint64_t below = lane ? (~0ull >> (WAVE - lane)) : 0ull;   // lanes under this one
    if (qo) items[wv][__popcll(mq & below)] = RcItem{qf >> 3, ql, qo};
    if (to) items[wv][nq + __popcll(mt & below)] = RcItem{tf >> 3, tl, to};
    __syncthreads();
    const uint32_t t = lane % RC_T;
    for (uint32_t g = lane / RC_T; g < nq + nt; g += WAVE / RC_T) {
        const RcItem it = items[wv][g];
        rc_side((g < nq ? d.pq : d.pt) + it.word, it.len, it.op, n_code, t);
    }
}

// ------------------------------------------------------------------ sort ----
__device__ __forceinline__ bool census_fast(uint32_t m);

// Pairs are ordered by a shape bucket, largest first, so that the 64 lanes of
// a wave share loop bounds: a counting sort over G2_SHAPE_BUCKETS keys
// (histogram, descending scan, scatter).  It has no memset or library call,
// so the whole device program can be captured in a graph.  Results never
// depend on the order (every output is per pair), only the load balance does.
//
// Key: long-pair flag (one wave per pair, taken first), then query words,
// then target words, each as an 8-bit bucket: exact below 128 words, then 16
// steps per octave (at most 1/16 apart) up to 2^15 words.
__device__ __forceinline__ uint32_t shape8(uint32_t x) {
    if (x < 128u) return x;
    const uint32_t e = 31u - __builtin_clz(x);   // >= 7
    const uint32_t v = 128u + (e - 7u) * 16u + ((x >> (e - 4u)) & 15u);
    return v > 255u ? 255u : v;
}

__device__ __forceinline__ uint32_t pair_key(const G2Dev &d, uint32_t k, bool &lng) {
    const uint32_t qw = (uint32_t)dregs(d.qlen[k]), tw = (uint32_t)dregs(d.tlen[k]);
    lng = d.wf_blocks > 0 && tw >= d.wf_min_tw && (uint64_t)qw * tw * 64u >= (uint64_t)d.wf_min_cells;
    // banded runs: the band's words per strip stand in for the query words
    uint32_t cw = qw;
    if (d.band_lo) {
        const int64_t span = (int64_t)d.band_hi[k] - d.band_lo[k];
        const int64_t bw = span < 0 ? 1 : ((span + 14) >> 3) + 1;
        cw = (uint32_t)(bw < (int64_t)qw ? bw : (int64_t)qw);
    }
    return (lng ? 0x10000u : 0u) | (shape8(cw) << 8) | shape8(tw);
}

// A batch holds a few hundred distinct keys, so one global atomic per pair
// serialises on the same counters.  Each block takes KH_CHUNK pairs and counts
// them in an LDS hash table (open addressing, at most KH_PROBE probes), then
// adds each distinct key's count to the global histogram once.  A key that
// finds no slot within KH_PROBE probes goes to the global counter directly.
constexpr uint32_t KH_SLOTS = 2048, KH_PROBE = 32, KH_CHUNK = 8192, KH_THREADS = 1024;
constexpr uint32_t KH_EMPTY = 0xFFFFFFFFu;

__device__ __forceinline__ int kh_slot(uint32_t *tkey, uint32_t key) {
    uint32_t h = (key * 2654435761u) >> 21;   // 11-bit multiplicative hash
    for (uint32_t p = 0; p < KH_PROBE; p++, h = (h + 1u) & (KH_SLOTS - 1u)) {
        const uint32_t prev = atomicCAS(&tkey[h], KH_EMPTY, key);
        if (prev == KH_EMPTY || prev == key) return (int)h;
    }
    return -1;
}

// aux[2] counts the long pairs.
__global__ __launch_bounds__(KH_THREADS) void k_keys(G2Dev d) {
    static_assert(KH_SLOTS == 1u << 11, "kh_slot hashes to 11 bits");
    __shared__ uint32_t tkey[KH_SLOTS], tcnt[KH_SLOTS];
    __shared__ uint32_t nlong;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < KH_SLOTS; i += KH_THREADS) { tkey[i] = KH_EMPTY; tcnt[i] = 0; }
    if (t == 0) nlong = 0;
    __syncthreads();
    const uint32_t k0 = blockIdx.x * KH_CHUNK, k1 = d.n - k0 < KH_CHUNK ? d.n : k0 + KH_CHUNK;
    for (uint32_t k = k0 + t; k < k1; k += KH_THREADS) {
        bool lng;
        const uint32_t key = pair_key(d, k, lng);
        d.keys[k] = key;
        const int h = kh_slot(tkey, key);
        if (h >= 0) atomicAdd(&tcnt[h], 1u);
        else atomicAdd(&d.hist[key], 1u);
        if (lng) atomicAdd(&nlong, 1u);
    }
    __syncthreads();
    for (uint32_t i = t; i < KH_SLOTS; i += KH_THREADS)
        if (tcnt[i]) atomicAdd(&d.hist[tkey[i]], tcnt[i]);
    if (t == 0 && nlong) atomicAdd(&d.aux[2], nlong);
}

// Exclusive scan of the histogram in descending key order, in place: hist[b]
// becomes the first sorted slot of bucket b.  One block of 1024 threads, each
// owning G2_SHAPE_BUCKETS / 1024 consecutive buckets.
__global__ __launch_bounds__(1024) void k_scan_desc(uint32_t *hist) {
    constexpr uint32_t PER = G2_SHAPE_BUCKETS / 1024;
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t hi = G2_SHAPE_BUCKETS - 1 - t * PER;   // this thread's buckets: hi, hi-1, ..., hi-PER+1
    uint32_t sum = 0;
    for (uint32_t i = 0; i < PER; i++) sum += hist[hi - i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {   // Hillis-Steele inclusive scan
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t i = 0; i < PER; i++) {
        const uint32_t c = hist[hi - i];
        hist[hi - i] = run;
        run += c;
    }
}

// Places each pair at its bucket's next free slot.  As in k_keys, a block
// ranks its KH_CHUNK pairs within their keys in LDS, claims each key's range
// with one global atomic, then writes base + rank.  The order inside a bucket
// is whatever the atomics give; no output depends on it.
__global__ __launch_bounds__(KH_THREADS) void k_scatter(G2Dev d) {
    __shared__ uint32_t tkey[KH_SLOTS], tcnt[KH_SLOTS];
    __shared__ uint32_t rank[KH_CHUNK];   // slot << 16 | rank inside the slot, or KH_EMPTY
    static_assert(KH_CHUNK <= 65536u && KH_SLOTS <= 65536u, "slot and rank share one word");
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < KH_SLOTS; i += KH_THREADS) { tkey[i] = KH_EMPTY; tcnt[i] = 0; }
    __syncthreads();
    const uint32_t k0 = blockIdx.x * KH_CHUNK, k1 = d.n - k0 < KH_CHUNK ? d.n : k0 + KH_CHUNK;
    for (uint32_t k = k0 + t; k < k1; k += KH_THREADS) {
        const uint32_t key = d.keys[k];
        const int h = kh_slot(tkey, key);
        if (h >= 0) {
            rank[k - k0] = ((uint32_t)h << 16) | atomicAdd(&tcnt[h], 1u);
        } else {
            rank[k - k0] = KH_EMPTY;
            d.perm[atomicAdd(&d.hist[key], 1u)] = k;
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < KH_SLOTS; i += KH_THREADS)   // count -> first global slot
        if (tcnt[i]) tcnt[i] = atomicAdd(&d.hist[tkey[i]], tcnt[i]);
    __syncthreads();
    for (uint32_t k = k0 + t; k < k1; k += KH_THREADS) {
        const uint32_t v = rank[k - k0];
        if (v != KH_EMPTY) d.perm[tcnt[v >> 16] + (v & 0xFFFFu)] = k;
    }
}

// Transposed sequence panels of the lane path: word j of the pair in lane l of
// group g at qpan[(g * qpan_w + j) * 64 + l] (0 past the pair's words), the
// same for targets.  A wave's read of one word is then 256 contiguous bytes
// instead of 64 scattered lines.  One block per lane group: 32 words of the
// group's 64 pairs at a time are read pair-contiguously into LDS and written
// out word-contiguously.
__global__ __launch_bounds__(256) void k_panel(G2Dev d) {
    __shared__ uint32_t tile[32][WAVE + 1];
    __shared__ uint32_t s_qo[WAVE], s_to[WAVE], s_qw[WAVE], s_tw[WAVE];
    const uint32_t slot0 = d.wf_blocks ? d.aux[2] : 0u;
    const uint32_t g = blockIdx.x, t = threadIdx.x;
    if ((uint64_t)g * WAVE + slot0 >= d.n) return;   // the whole block: past the lane groups
    if (t < WAVE) {
        const uint32_t slot = slot0 + g * WAVE + t;
        const bool ok = slot < d.n;
        const uint32_t pr = ok ? d.perm[slot] : 0u;
        s_qo[t] = ok ? d.qoff[pr] >> 3 : 0u;
        s_to[t] = ok ? d.toff[pr] >> 3 : 0u;
        s_qw[t] = ok ? (uint32_t)dregs(d.qlen[pr]) : 0u;
        s_tw[t] = ok ? (uint32_t)dregs(d.tlen[pr]) : 0u;
    }
    __syncthreads();
    for (int side = 0; side < 2; side++) {
        const uint32_t W = side ? d.tpan_w : d.qpan_w;
        const uint32_t *src = side ? d.pt : d.pq;
        uint32_t *out = (side ? d.tpan : d.qpan) + (size_t)g * W * WAVE;
        for (uint32_t w0 = 0; w0 < W; w0 += 32) {
            for (uint32_t idx = t; idx < 32 * WAVE; idx += 256) {   // pair-contiguous reads
                const uint32_t l = idx >> 5, w = idx & 31u;
                const uint32_t cnt = side ? s_tw[l] : s_qw[l];
                tile[w][l] = w0 + w < cnt ? src[(side ? s_to[l] : s_qo[l]) + w0 + w] : 0u;
            }
            __syncthreads();
            for (uint32_t idx = t; idx < 32 * WAVE; idx += 256) {   // word-contiguous writes
                const uint32_t w = idx >> 6, l = idx & (WAVE - 1);
                if (w0 + w < W) out[(size_t)(w0 + w) * WAVE + l] = tile[w][l];
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- LOCAL -----
// gasal_local_kernel<LOCAL,S,B> (src/kernels/local_kernel_template.h:71-519).
__device__ __forceinline__ bool census_fast(uint32_t m);

template <int START, bool SECOND>
__global__ __launch_bounds__(64) void k_local(G2Dev d, G2Scores c) {
    if (d.fast && (d.enc_ok || d.pk2 || ((START == S_WITHOUT_START || START == S_WITH_START) && !SECOND && d.pk16)) &&
        census_fast(d.aux[0]))
        return;   // k_local_fast handles this batch
    LaneCtx x;
    // long pairs of a score-only batch sort first and run one wave per pair
    if (!lane_setup(d, x, 0, d.wf_blocks ? d.aux[2] : 0u)) return;
    constexpr bool TB = START == S_WITH_TB;
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t gapoe = c.gapoe, gape = c.gape;
    const int32_t QW = x.QW, TW = x.TW;
    uint32_t *__restrict__ bd = x.bd;
    const TbView tbv = TB ? tb_view(d, x.grp, x.lane) : TbView{nullptr, 0};

    for (int32_t cc = 0; cc < 8 * QW; cc++) bd[cc * WAVE] = 0u;   // initHD (:118-120)

    int32_t maxHH = 0, maxXY_y = 0, prev_maxHH = 0, maxXY_x = 0;
    int32_t sec = 0, prev_sec = 0, secx = 0, secy = 0;
    for (int32_t i = 0; i < TW; i++) {
        int32_t h[9], f[9], p[9];
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }
        const uint32_t gpac = x.T[i];
        uint32_t gb[9];
#pragma unroll
        for (int m = 1; m < 9; m++) gb[m] = (gpac >> (32 - 4 * m)) & 15u;
        const int32_t gidx = i << 3;
        uint32_t nxt[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[kk * WAVE];
        for (int32_t j = 0; j < QW; j++) {
            const uint32_t rpac = x.Q[j];
            uint32_t cur[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) cur[kk] = nxt[kk];
            if (j + 1 < QW) {
#pragma unroll
                for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[((j + 1) * 8 + kk) * WAVE];
            }
            uint32_t dirw[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int32_t ridx = j * 8 + kk;
                const uint32_t rbase = (rpac >> (28 - 4 * kk)) & 15u;
                h[0] = sx16(cur[kk]);
                int32_t e = sx16(cur[kk] >> 16);
                uint32_t dir = 0;
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const int32_t s = sub_local(c, nval, rbase, gb[m]);
                    const int32_t tmp = p[m] + s;
                    int32_t hm = max(max(tmp, f[m]), max(e, 0));
                    if (TB) {
                        const int sh = 28 - ((m - 1) << 2);
                        const uint32_t m_or_x = tmp >= p[m] ? 0u : 1u;                     // :49
                        dir |= hm == tmp ? (m_or_x << sh) : (hm == f[m] ? (3u << sh) : (2u << sh)); // :53
                        dir |= (tmp - gapoe) > (f[m] - gape) ? 0u : (1u << (sh + 3));      // :54
                        dir |= (tmp - gapoe) > (e - gape) ? 0u : (1u << (sh + 2));         // :56
                    }
                    f[m] = max(tmp - gapoe, f[m] - gape);
                    e = max(tmp - gapoe, e - gape);
                    maxXY_y = (maxHH < hm) ? gidx + (m - 1) : maxXY_y;
                    maxHH = max(maxHH, hm);
                    if (SECOND) {                                                          // :145-150
                        bool ov = (sec < hm) && (maxHH > hm);
                        secy = ov ? gidx + (m - 1) : secy;
                        sec = ov ? hm : sec;
                    }
                    p[m] = h[m - 1];
                    h[m] = hm;
                }
                bd[ridx * WAVE] = pk16(h[8], e);
                maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;
                if (SECOND) {                                                              // :160-164
                    secx = (prev_sec < maxHH) ? ridx : secx;
                    prev_sec = max(sec, prev_sec);
                }
                prev_maxHH = max(maxHH, prev_maxHH);
                if (TB) dirw[kk] = dir;
            }
            if (TB) tbv.store(i, j, dirw);
        }
    }
    const uint32_t pr = x.pair;
    d.score[pr] = maxHH;
    d.qend[pr] = maxXY_x;
    d.tend[pr] = maxXY_y;
    if (SECOND) { d.s_score[pr] = sec; d.s_qend[pr] = secx; d.s_tend[pr] = secy; }

    if (START == S_WITH_START) {                                                           // :441-511
        const int32_t fwd_score = maxHH;
        const int32_t rend_reg = ((maxXY_x >> 3) + 1) < QW ? ((maxXY_x >> 3) + 1) : QW;
        const int32_t gend_reg = ((maxXY_y >> 3) + 1) < TW ? ((maxXY_y >> 3) + 1) : TW;
        const uint32_t *Qe = x.Q + (rend_reg - 1), *Te = x.T + (gend_reg - 1);
        maxHH = 0; prev_maxHH = 0; maxXY_x = 0; maxXY_y = 0;
        for (int32_t cc = 0; cc < 8 * rend_reg; cc++) bd[cc * WAVE] = 0u;
        int32_t gidx = ((gend_reg << 3) + 8) - 1;
        bool go = maxHH < fwd_score;
        for (int32_t i = 0; i < gend_reg && go; i++) {
            int32_t h[9], f[9], p[9];
#pragma unroll
            for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }
            const uint32_t gpac = *(Te - i);
            uint32_t gb[9];
#pragma unroll
            for (int m = 1; m < 9; m++) gb[m] = (gpac >> (4 * (m - 1))) & 15u;
            gidx -= 8;
            int32_t ridx = (rend_reg << 3) - 1;
            int32_t gi = 0;
            for (int32_t j = 0; j < rend_reg && go; j++) {
                const uint32_t rpac = *(Qe - j);
                for (int kk = 0; kk < 8 && go; kk++) {
                    const uint32_t rbase = (rpac >> (4 * kk)) & 15u;
                    uint32_t hd = bd[gi * WAVE];
                    h[0] = sx16(hd);
                    int32_t e = sx16(hd >> 16);
#pragma unroll
                    for (int m = 1; m < 9; m++) {                                          // CORE_LOCAL_COMPUTE_START
                        const int32_t s = sub_local(c, nval, rbase, gb[m]);
                        const int32_t tmp = p[m] + s;
                        int32_t hm = max(max(tmp, f[m]), max(e, 0));
                        f[m] = max(tmp - gapoe, f[m] - gape);
                        e = max(tmp - gapoe, e - gape);
                        maxXY_y = (maxHH < hm) ? gidx + (m - 1) : maxXY_y;                 // :41
                        maxHH = max(maxHH, hm);
                        p[m] = h[m - 1];
                        h[m] = hm;
                    }
                    bd[gi * WAVE] = pk16(h[8], e);
                    maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;
                    prev_maxHH = max(maxHH, prev_maxHH);
                    ridx--; gi++;
                    go = maxHH < fwd_score;
                }
            }
        }
        d.qstart[pr] = maxXY_x;
        d.tstart[pr] = maxXY_y;
    }
}

// ------------------------------------------------------- LOCAL, fast path ----
// Same recurrence, order and outputs as k_local, with three changes that cut
// the integer-VALU work per cell:
//  * substitution: per 8-row strip each row gets a 32-bit table of eight
//    signed 4-bit scores indexed by the query base's class (nibble & 7), so a
//    cell's score is one v_bfe_i32 (valid when every base nibble present in the
//    batch has a distinct class and all scores fit in 4 bits; the nibble census
//    in aux[0] and the host check decide, otherwise k_local runs);
//  * max / second best: cells are ranked by h << S | ~strip | ~column | 7-row,
//    so "first maximal cell in (strip, column, row) order" becomes a plain
//    integer max (no per-cell position selects);
//  * direction bits (WITH_TB): the four comparisons of the reference's nibble
//    are taken as sign bits of differences and shifted into the column word
//    with v_alignbit (no SGPR masks); bit order per row from the top:
//    h!=M, h!=F_in, F restarted, E restarted.  k_get_tb decodes them.
__device__ __forceinline__ bool census_fast(uint32_t m) { return ((m & (m >> 8)) & 0xFFu) == 0u; }

// GLOBAL_SEM: DEV_GET_SUB_SCORE_GLOBAL (N is an ordinary base unless N_PENALTY).
template <bool GLOBAL_SEM = false>
__device__ __forceinline__ uint32_t score_lut(const G2Scores &c, int32_t nval, uint32_t census, uint32_t tnib) {
    uint32_t lut = 0;
#pragma unroll
    for (uint32_t cls = 0; cls < 8; cls++) {
        const uint32_t q = ((census >> cls) & 1u) ? cls : cls + 8u;   // the nibble this class stands for
        int32_t s = GLOBAL_SEM ? sub_global(c, q, tnib) : sub_local(c, nval, q, tnib);
        lut |= ((uint32_t)s & 15u) << (4 * cls);
    }
    return lut;
}

__device__ __forceinline__ uint32_t nib_set(uint32_t w) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 32; k += 4) m |= 1u << ((w >> k) & 15u);
    return m;
}

// Bit v of aux[0] is set when some packed nibble equals v.  The packed arrays
// are 256-byte aligned, so their whole 16-byte groups are read as uint4 (more
// bytes in flight per wave); the last nq % 4 and nt % 4 words one by one.
__global__ __launch_bounds__(256) void k_census(const uint32_t *__restrict__ pq, uint32_t nq, const uint32_t *__restrict__ pt,
                                               uint32_t nt, uint32_t *__restrict__ out) {
    uint32_t m = 0;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    const uint32_t nq4 = nq >> 2, n4 = nq4 + (nt >> 2);
    const uint4 *q4 = reinterpret_cast<const uint4 *>(pq), *t4 = reinterpret_cast<const uint4 *>(pt);
    constexpr uint32_t U = 2;   // independent 16-byte loads in flight per thread
    for (uint32_t i = gtid; i < n4; i += U * stride) {
        uint4 v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t j = i + u * stride;
            v[u] = j < nq4 ? q4[j] : (j < n4 ? t4[j - nq4] : uint4{0u, 0u, 0u, 0u});
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t mw = nib_set(v[u].x) | nib_set(v[u].y) | nib_set(v[u].z) | nib_set(v[u].w);
            m |= i + u * stride < n4 ? mw : 0u;
        }
    }
    if (gtid < (nq & 3u)) m |= nib_set(pq[(nq & ~3u) + gtid]);
    if (gtid < (nt & 3u)) m |= nib_set(pt[(nt & ~3u) + gtid]);
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off);
    if ((threadIdx.x & 63) == 0 && m) atomicOr(out, m);
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// ------------------------------------------------ LOCAL, one wave per pair ----
// Intra-pair wavefront for long pairs (score and ends).  Strip s (8 target
// rows) lives on lane s % 64 in pass s / 64; lane l works on column c at step
// c + l, so the (H, E) row leaving lane l's strip at column c is handed to lane
// l + 1 one step later with DPP wave_shr:1, and the query class travels the
// same way.  Lane 63's bottom row is parked in an LDS ring for lane 0 of the
// next pass.  Queries longer than WF_W columns are cut into column blocks;
// between blocks each lane saves its strip state (h, f per row and the
// diagonal corner) in per-block scratch.  The recurrences and the int16
// boundary rows are those of k_local_fast.
//
// End position: the reference keeps the first strict maximum in (strip,
// column, row) order (src/kernels/local_kernel_template.h:121-135, 347-371).
// Within one run (one strip, one column block) columns arrive in order, so a
// strict compare keeps the first; runs are merged per lane and then over the
// wave on the key (h, -strip, -column, -row).
constexpr int WF_W = 1024;

struct WfBest {
    int32_t v;      // h << 3 | (7 - row in strip); 0 = nothing positive yet
    int32_t strip, col;
};

// true when a precedes b in the reference's choice: higher h, then lower
// strip, then lower column (one (strip, column) cell group is never split
// over two runs, so the row bits never decide)
__device__ __forceinline__ bool wf_before(int32_t av, int32_t as, int32_t ac, int32_t bv, int32_t bs, int32_t bc) {
    if ((av >> 3) != (bv >> 3)) return (av >> 3) > (bv >> 3);
    if (as != bs) return as < bs;
    return ac < bc;
}

// TB: direction words as in k_local_fast; each lane keeps its last eight
// columns in a window and stores the 32-byte tile when a column closes an 8x8
// tile, to the pair's own region wf_tb + wf_tbo[pair] laid out [strip][word].
// LUT: score tables (the batch's nibble classes allow them), else the
// substitution rule per cell with the query base travelling instead of its class.
// CK (TB only): checkpointed traceback, as k_local_fast<.., CK>: per tile the
// left column state (at wf_tb) and per strip the bottom row (at wf_tbh), per
// pair [strip][word] x 8 dwords, instead of direction words.
template <bool TB, bool LUT, bool CK = false>
__device__ void wf_local(const G2Dev &d, const G2Scores &c, uint32_t census) {
    __shared__ uint32_t ring[WF_W];
    const int32_t l = (int32_t)threadIdx.x;
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t gapoe = c.gapoe, gape = c.gape;
    uint32_t *vs = d.wf_scratch + (size_t)blockIdx.x * d.wf_stride;
    // Every value that steers a loop below is wave-uniform, and no branch on the
    // lane index feeds one: with `if (lane == 0) item = atomicAdd(...)` the
    // compiler splits the queue loop per lane (lane 0 re-enters the atomic,
    // lanes 1..63 keep iterating with item 0), which never terminates.  All
    // lanes take part in the atomic; only lane 0 adds, so lane 0 reads the
    // counter before its own increment.
    const uint32_t nlong = __builtin_amdgcn_readfirstlane(d.aux[2]);
    constexpr bool use_lut = LUT;
    for (;;) {
        const uint32_t item = __builtin_amdgcn_readfirstlane(atomicAdd(&d.aux[1], l == 0 ? 1u : 0u));
        if (item >= nlong) break;
        const uint32_t pr = __builtin_amdgcn_readfirstlane(d.perm[item]);
        const uint32_t qlen = __builtin_amdgcn_readfirstlane(d.qlen[pr]);
        const uint32_t tlen = __builtin_amdgcn_readfirstlane(d.tlen[pr]);
        const int32_t QW = dregs(qlen), TW = dregs(tlen);
        const uint32_t *Q = d.pq + (d.qoff[pr] >> 3);
        const uint32_t *T = d.pt + (d.toff[pr] >> 3);
        const int32_t C = 8 * QW;
        const int32_t npass = (TW + 63) / 64;
        const int32_t nblk = (C + WF_W - 1) / WF_W;
        // tiles per pair [strip][word] x 8 dwords
        uint32_t *tbp = TB ? d.wf_tb + d.wf_tbo[pr] * 8 : nullptr;
        uint32_t *thp = CK ? d.wf_tbh + d.wf_tbo[pr] * 8 : nullptr;
        WfBest lb = {0, 0x7FFFFFFF, 0x7FFFFFFF};
        for (int32_t b = 0; b < nblk; b++) {
            const int32_t c0 = b * WF_W;
            const int32_t wb = min(WF_W, C - c0);
            for (int32_t k = 0; k < npass; k++) {
                const int32_t s = l + 64 * k;
                const bool act = s < TW;
                uint32_t *st = vs + (size_t)k * 17 * 64 + l;   // this lane's strip state, [17][64]
                int32_t h[9], f[9];
                int32_t ptop = 0;   // H(strip top - 1, column - 1): diagonal of the strip's first row
#pragma unroll
                for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; }
                if (b > 0 && act) {
#pragma unroll
                    for (int m = 1; m < 9; m++) { h[m] = (int32_t)st[(m - 1) * 64]; f[m] = (int32_t)st[(m + 7) * 64]; }
                    ptop = (int32_t)st[16 * 64];
                }
                uint32_t lut[9];          // score table per row, or the row's base when !use_lut
                const uint32_t gpac = act ? T[s] : 0u;
#pragma unroll
                for (int m = 1; m < 9; m++)
                    lut[m] = use_lut ? score_lut(c, nval, census, (gpac >> (32 - 4 * m)) & 15u) : (gpac >> (32 - 4 * m)) & 15u;
                int32_t rv = 0, rt = 0;   // run best: encoded value, step
                uint32_t win[8];          // direction words of this lane's last eight columns
#pragma unroll
                for (int m = 0; m < 8; m++) win[m] = 0;
                uint32_t out = 0, qcls = 0;
                uint32_t qcur = 0, qnext = Q[c0 >> 3];   // query word of lane 0's column, next one prefetched
                // the last active lane (strip) reaches column wb - 1 at step
                // wb - 1 + (nact - 1); lanes past the pair's strips do nothing
                const int32_t nact = min(64, TW - 64 * k);
                for (int32_t tau = 0; tau < wb + nact - 1; tau++) {
                    // inputs of lane 0 (column c0 + tau): top boundary and query class
                    uint32_t top0 = 0, q0 = 0;
                    if (tau < wb) {
                        if ((tau & 7) == 0) {
                            qcur = qnext;
                            if (tau + 8 < wb) qnext = Q[(c0 + tau + 8) >> 3];
                        }
                        if (k > 0) top0 = ring[tau];
                        const uint32_t qn = (qcur >> (28 - 4 * (tau & 7))) & 15u;
                        q0 = use_lut ? (qn & 7u) << 2 : qn;
                    }
                    const uint32_t tin = (uint32_t)__builtin_amdgcn_update_dpp((int)top0, (int)out, 0x138, 0xF, 0xF, false);
                    qcls = (uint32_t)__builtin_amdgcn_update_dpp((int)q0, (int)qcls, 0x138, 0xF, 0xF, false);
                    const int32_t rel = tau - l;
                    if (act && rel >= 0 && rel < wb) {
                        const int32_t col = c0 + rel;
                        if (CK && (col & 7) == 0) {   // the tile's left column state
                            uint32_t lw[8];
#pragma unroll
                            for (int m = 1; m < 9; m++) lw[m - 1] = pk16(h[m], f[m]);
                            tb_store(tbp + ((size_t)s * QW + (size_t)(col >> 3)) * 8, lw);
                        }
                        int32_t e = (int32_t)tin >> 16;
                        int32_t p = ptop;
                        ptop = sx16(tin);
                        int32_t colbest = 0;
                        uint32_t dir = 0;
#pragma unroll
                        for (int m = 1; m < 9; m++) {
                            const int32_t sc = use_lut ? __builtin_amdgcn_sbfe(lut[m], qcls, 4) : sub_local(c, nval, qcls, lut[m]);
                            const int32_t tmp = p + sc;
                            int32_t hm;
                            const int32_t t = tmp - gapoe;
                            if (TB) {   // a traceback reads the reference's E and F, negative ones too
                                hm = max(max(max(tmp, f[m]), e), 0);
                                const int32_t fx = f[m] - gape, ex = e - gape;
                                if (!CK) {
                                    dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(tmp - hm), 31);
                                    dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(f[m] - hm), 31);
                                    dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(fx - t), 31);
                                    dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(ex - t), 31);
                                }
                                f[m] = max(t, fx);
                                e = max(t, ex);
                            } else {   // score and ends: E and F held as max(E, 0) and max(F, 0)
                                       // (an extension that saturates at 0, then a max with t),
                                       // so H = max(tmp, F, E, 0) needs no max with 0
                                hm = max(max(tmp, f[m]), e);
                                f[m] = max(t, (int32_t)__builtin_elementwise_sub_sat((uint32_t)f[m], (uint32_t)gape));
                                e = max(t, (int32_t)__builtin_elementwise_sub_sat((uint32_t)e, (uint32_t)gape));
                            }
                            const int32_t en = (hm << 3) | (8 - m);
                            colbest = m == 1 ? en : max(colbest, en);
                            p = h[m];
                            h[m] = hm;
                        }
                        // strictly higher h only: an equal h in a later column never wins
                        if (colbest > (rv | 7)) { rv = colbest; rt = tau; }
                        out = pk16(h[8], e);
                        if (TB) {   // direction words, or with CK the strip's bottom row
#pragma unroll
                            for (int m = 0; m < 7; m++) win[m] = win[m + 1];
                            win[7] = CK ? out : dir;
                            if ((col & 7) == 7) tb_store((CK ? thp : tbp) + ((size_t)s * QW + (size_t)(col >> 3)) * 8, win);
                        }
                        if (l == 63 && k + 1 < npass) ring[rel] = out;
                    }
                }
                if (rv >= 8) {   // h > 0 in this run
                    const int32_t rc = c0 + rt - l;
                    if (wf_before(rv, s, rc, lb.v, lb.strip, lb.col)) lb = {rv, s, rc};
                }
                if (b + 1 < nblk && act) {
#pragma unroll
                    for (int m = 1; m < 9; m++) { st[(m - 1) * 64] = (uint32_t)h[m]; st[(m + 7) * 64] = (uint32_t)f[m]; }
                    st[16 * 64] = (uint32_t)ptop;
                }
                __syncthreads();   // ring written by lane 63 is read by lane 0 in the next pass
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int32_t ov = __shfl_xor(lb.v, off), os = __shfl_xor(lb.strip, off), oc = __shfl_xor(lb.col, off);
            if (wf_before(ov, os, oc, lb.v, lb.strip, lb.col)) lb = {ov, os, oc};
        }
        if (l == 0) {
            const int32_t v = lb.v >> 3;
            d.score[pr] = v;
            d.qend[pr] = v > 0 ? lb.col : 0;
            d.tend[pr] = v > 0 ? lb.strip * 8 + (7 - (lb.v & 7)) : 0;
        }
    }
}

// The WITH_START reverse pass of the LOCAL kernels (src/kernels/local_kernel_template.h
// :441-511): from the words holding the forward end, the DP over the reversed
// prefix rectangle until a cell reaches the forward score; the start is the
// reverse pass's first maximal cell.  Scores from the reversed rows' tables
// (the batch census allows them for the kernels that call it).
__device__ void local_reverse_start(const G2Dev &d, const G2Scores &c, const LaneCtx &x, uint32_t census, int32_t fwd_score,
                                    int32_t maxXY_x, int32_t maxXY_y) {
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t gapoe = c.gapoe, gape = c.gape;
    const int32_t QW = x.QW, TW = x.TW;
    uint32_t *__restrict__ bd = x.bd;
    const uint32_t pr = x.pair;
    {
        int32_t maxHH = 0, prev_maxHH = 0;
        const int32_t rend_reg = ((maxXY_x >> 3) + 1) < QW ? ((maxXY_x >> 3) + 1) : QW;
        const int32_t gend_reg = ((maxXY_y >> 3) + 1) < TW ? ((maxXY_y >> 3) + 1) : TW;
        const uint32_t *Qe = x.Q + (rend_reg - 1), *Te = x.T + (gend_reg - 1);
        maxXY_x = 0; maxXY_y = 0;
        for (int32_t cc = 0; cc < 8 * rend_reg; cc++) bd[cc * WAVE] = 0u;
        int32_t gidx = ((gend_reg << 3) + 8) - 1;
        bool go = maxHH < fwd_score;
        for (int32_t i = 0; i < gend_reg && go; i++) {
            int32_t h[9], f[9], p[9];
#pragma unroll
            for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }
            const uint32_t gpac = *(Te - i);
            uint32_t gb[9];   // score tables of the reversed rows (the batch census allows them here)
#pragma unroll
            for (int m = 1; m < 9; m++) gb[m] = score_lut(c, nval, census, (gpac >> (4 * (m - 1))) & 15u);
            gidx -= 8;
            int32_t ridx = (rend_reg << 3) - 1;
            int32_t gi = 0;
            // A word at a time, its boundary-row entries loaded one word
            // ahead; the exit is tested after each word.  The cells a word
            // runs past the reference's exit change nothing it reports: the
            // maxima move only on a strictly higher h, and no h here exceeds
            // the forward score.
            uint32_t nxt[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[kk * WAVE];
            for (int32_t j = 0; j < rend_reg && go; j++) {
                const uint32_t rpac = *(Qe - j);
                uint32_t cur[8];
#pragma unroll
                for (int kk = 0; kk < 8; kk++) cur[kk] = nxt[kk];
                if (j + 1 < rend_reg) {
#pragma unroll
                    for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[((j + 1) * 8 + kk) * WAVE];
                }
#pragma unroll
                for (int kk = 0; kk < 8; kk++) {
                    const uint32_t qsh = ((rpac >> (4 * kk)) & 7u) << 2;
                    const uint32_t hd = cur[kk];
                    h[0] = sx16(hd);
                    int32_t e = sx16(hd >> 16);
#pragma unroll
                    for (int m = 1; m < 9; m++) {
                        const int32_t s = __builtin_amdgcn_sbfe(gb[m], qsh, 4);
                        const int32_t tmp = p[m] + s;
                        // E and F held as max(E, 0) and max(F, 0) (saturating
                        // extensions): H = max(tmp, F, E, 0) with no max with 0;
                        // only H reaches this pass's outputs
                        const int32_t hm = max(max(tmp, f[m]), e);
                        f[m] = max(tmp - gapoe, (int32_t)__builtin_elementwise_sub_sat((uint32_t)f[m], (uint32_t)gape));
                        e = max(tmp - gapoe, (int32_t)__builtin_elementwise_sub_sat((uint32_t)e, (uint32_t)gape));
                        maxXY_y = (maxHH < hm) ? gidx + (m - 1) : maxXY_y;
                        maxHH = max(maxHH, hm);
                        p[m] = h[m - 1];
                        h[m] = hm;
                    }
                    bd[gi * WAVE] = pk16(h[8], e);
                    maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;
                    prev_maxHH = max(maxHH, prev_maxHH);
                    ridx--; gi++;
                }
                go = maxHH < fwd_score;
            }
        }
        d.qstart[pr] = maxXY_x;
        d.tstart[pr] = maxXY_y;
    }
}

// ------------------------------------------ LOCAL, packed 16-bit, one lane ----
// Score and ends with two 8-row strips per lane, one in each 16-bit half of
// every register (v_pk_add_u16 / v_pk_sub_i16 / v_pk_max_i16): the low half
// runs strip 2k at column t, the high half strip 2k+1 at column t-1, whose top
// row is the low half's bottom row of the previous step, kept in a register.
// The high strip's bottom row goes to the boundary row for the next pair of
// strips.  Used when every value fits int16 (H <= match * min(8QW, 8TW) <=
// 32767; E and F are held as max(E, 0) and max(F, 0)), so the H values are
// those of the int32 kernels.  Ends: per strip the first column
// whose maximum beats the strip's running best, its first row; strips merged
// in order (the reference's first strict maximum in (strip, column, row)).
typedef short pk16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ pk16x2 as_pk(uint32_t u) { return __builtin_bit_cast(pk16x2, u); }
__device__ __forceinline__ uint32_t as_u32(pk16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ pk16x2 pmax(pk16x2 a, pk16x2 b) { return __builtin_elementwise_max(a, b); }
// a - b per half, clamped at 0 (unsigned saturation; a, b >= 0)
__device__ __forceinline__ pk16x2 psubsat(pk16x2 a, pk16x2 b) {
    typedef unsigned short pu16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(pk16x2, __builtin_elementwise_sub_sat(__builtin_bit_cast(pu16x2, a),
                                                                    __builtin_bit_cast(pu16x2, b)));
}
// {lo16(lo), lo16(hi)} and {hi16(lo), hi16(hi)}
__device__ __forceinline__ uint32_t pack_lo(uint32_t lo, uint32_t hi) { return __builtin_amdgcn_perm(hi, lo, 0x05040100u); }
__device__ __forceinline__ uint32_t pack_hi(uint32_t lo, uint32_t hi) { return __builtin_amdgcn_perm(hi, lo, 0x07060302u); }

__device__ void lane_local_pk(const G2Dev &d, const G2Scores &c, const LaneCtx &x, uint32_t census, bool with_start) {
    const int32_t nval = c.has_np ? -c.np : 0;
    const pk16x2 gapoe2 = {(short)c.gapoe, (short)c.gapoe}, gape2 = {(short)c.gape, (short)c.gape};
    const pk16x2 zero2 = {0, 0};
    const int32_t QW = x.QW, TW = x.TW, C = 8 * QW;
    uint32_t *__restrict__ bd = x.bd;
    for (int32_t cc = 0; cc < C; cc++) bd[cc * WAVE] = 0u;
    int32_t bv = 0, bs = 0, bc = 0, br = 0;   // best: h, strip, column, row in strip
    for (int32_t s0 = 0; s0 < TW; s0 += 2) {
        const bool hi_ok = s0 + 1 < TW;
        uint32_t lutL[9], lutH[9];
        const uint32_t gL = x.Tp[(size_t)s0 * x.ps], gH = hi_ok ? x.Tp[(size_t)(s0 + 1) * x.ps] : 0u;
#pragma unroll
        for (int m = 1; m < 9; m++) {
            lutL[m] = score_lut(c, nval, census, (gL >> (32 - 4 * m)) & 15u);
            lutH[m] = score_lut(c, nval, census, (gH >> (32 - 4 * m)) & 15u);
        }
        pk16x2 h[9], f[9];
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = zero2; f[m] = zero2; }
        uint32_t carry = 0;            // low strip's bottom (H | E << 16) of the previous step
        uint32_t qshH = 0;
        int32_t lv = 0, lc = 0, lr = 0, hv = 0, hc = 0, hr = 0;
        // one step: the low strip at column t (query code qshL, row above
        // topL), the high strip at column t - 1
        auto step = [&](int32_t t, uint32_t qshL, uint32_t topL) {
            pk16x2 e = as_pk(pack_hi(topL, carry));
            pk16x2 p = h[0];
            h[0] = as_pk(pack_lo(topL, carry));
            pk16x2 colmax = zero2;
#pragma unroll
            for (int m = 1; m < 9; m++) {
                const uint32_t sl = (uint32_t)__builtin_amdgcn_sbfe(lutL[m], qshL, 4);
                const uint32_t sh = (uint32_t)__builtin_amdgcn_sbfe(lutH[m], qshH, 4);
                const pk16x2 tmp = p + as_pk(pack_lo(sl, sh));
                // E and F kept as max(E, 0) and max(F, 0) (an extension that
                // saturates at 0, then a max with tt): H = max(tmp, F, E, 0)
                // then needs no separate max with 0
                const pk16x2 hm = pmax(pmax(tmp, f[m]), e);
                const pk16x2 tt = tmp - gapoe2;
                f[m] = pmax(tt, psubsat(f[m], gape2));
                e = pmax(tt, psubsat(e, gape2));
                p = h[m];
                h[m] = hm;
                colmax = pmax(colmax, hm);
            }
            qshH = qshL;
            const uint32_t h8 = as_u32(h[8]), e8 = as_u32(e);
            if (t >= 1) bd[(t - 1) * WAVE] = pack_hi(h8, e8);
            carry = pack_lo(h8, e8);
            if (t == 0) {   // the high strip starts at the next step: undo its column -1
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    h[m] = as_pk(as_u32(h[m]) & 0xFFFFu);
                    f[m] = as_pk(as_u32(f[m]) & 0xFFFFu);
                }
            }
            const int32_t cl = (int32_t)(as_u32(colmax) & 0xFFFFu), ch = (int32_t)(as_u32(colmax) >> 16);
            if (t < C && cl > lv) {
                lv = cl; lc = t;
#pragma unroll
                for (int m = 8; m >= 1; m--) if ((int32_t)(as_u32(h[m]) & 0xFFFFu) == cl) lr = m - 1;
            }
            if (t >= 1 && hi_ok && ch > hv) {
                hv = ch; hc = t - 1;
#pragma unroll
                for (int m = 8; m >= 1; m--) if ((int32_t)(as_u32(h[m]) >> 16) == ch) hr = m - 1;
            }
        };
        // a word of eight columns per iteration, its row-above entries loaded
        // one word ahead
        uint32_t top[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) top[kk] = bd[kk * WAVE];
        for (int32_t jw = 0; jw < QW; jw++) {
            const uint32_t rpac = x.Qp[(size_t)jw * x.ps];
            uint32_t cur[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) cur[kk] = top[kk];
            if (jw + 1 < QW) {
#pragma unroll
                for (int kk = 0; kk < 8; kk++) top[kk] = bd[(8 * (jw + 1) + kk) * WAVE];
            }
#pragma unroll
            for (int kk = 0; kk < 8; kk++)
                step(8 * jw + kk, kk < 7 ? (rpac >> (26 - 4 * kk)) & 0x1Cu : (rpac << 2) & 0x1Cu, cur[kk]);
        }
        step(C, 0, 0u);   // the high strip's last column
        if (lv > bv) { bv = lv; bs = s0; bc = lc; br = lr; }
        if (hv > bv) { bv = hv; bs = s0 + 1; bc = hc; br = hr; }
    }
    d.score[x.pair] = bv;
    d.qend[x.pair] = bv > 0 ? bc : 0;
    d.tend[x.pair] = bv > 0 ? bs * 8 + br : 0;
    if (with_start) local_reverse_start(d, c, x, census, bv, bv > 0 ? bc : 0, bv > 0 ? bs * 8 + br : 0);
}

template <bool LOCALT>
__device__ void tb_walk(const G2Dev &d, const G2Scores &c, uint32_t slot, const uint32_t *lut16, bool fused);

// CK (WITH_TB only): checkpointed traceback.  No direction words: per 8x8
// tile the left column state (H, F of its eight rows, 16 bits each) goes to
// the tile region and each strip's bottom row (H, E) to a second region of
// the same layout, which also replaces the boundary row; k_get_tb recomputes
// the tiles on the traceback path from them.  Exact when H fits 16 bits (the
// host enables it only then; F is at least -(gap open + extend + penalty)).
// WF: the launch carries one-wave-per-pair blocks (d.wf_blocks > 0); without
// them the kernel holds only the lane path and keeps its smaller register
// footprint.
template <int START, bool SECOND, bool CK = false, bool WF = true>
__global__ __launch_bounds__(64) void k_local_fast(G2Dev d, G2Scores c) {
    const uint32_t census = d.aux[0];
    // CK with tb_fused (opt-in): each lane walks its pair's traceback after its
    // DP (the walk's memory waits overlap other waves' DP); score tables of
    // the 16 target nibbles for the walk's tile recompute
    __shared__ uint32_t lut16[16];
    if (CK && d.tb_fused) {
        if (threadIdx.x < 16) lut16[threadIdx.x] = score_lut(c, c.has_np ? -c.np : 0, census, threadIdx.x);
        __syncthreads();
    }
    if (WF && blockIdx.x < d.wf_blocks) {
        if (START == S_WITHOUT_START && !SECOND) {
            if (census_fast(census)) wf_local<false, true>(d, c, census);
            else wf_local<false, false>(d, c, census);
        }
        if (START == S_WITH_TB && !SECOND) {
            // checkpoints need the score tables for the recompute; without them
            // the pairs store direction words in the same region
            if (census_fast(census)) {
                if (d.wf_tbh) wf_local<true, true, true>(d, c, census);
                else wf_local<true, true>(d, c, census);
            } else {
                wf_local<true, false>(d, c, census);
            }
        }
        return;
    }
    if (!census_fast(census)) return;
    if ((START == S_WITHOUT_START || START == S_WITH_START) && !SECOND && d.pk16) {
        LaneCtx x;
        if (lane_setup(d, x, d.wf_blocks, d.wf_blocks ? d.aux[2] : 0u)) lane_local_pk(d, c, x, census, START == S_WITH_START);
        return;
    }
    if (!d.enc_ok) return;   // position encoding does not fit: k_local takes the short pairs
    LaneCtx x;
    if (!lane_setup(d, x, d.wf_blocks, d.wf_blocks ? d.aux[2] : 0u)) return;
    constexpr bool TB = START == S_WITH_TB && !CK;
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t gapoe = c.gapoe, gape = c.gape;
    const int32_t QW = x.QW, TW = x.TW;
    const int32_t S = d.enc_s;
    uint32_t *__restrict__ bd = x.bd;
    const TbView tbv = START == S_WITH_TB ? tb_view(d, x.grp, x.lane) : TbView{nullptr, 0};
    const TbView hv = CK ? tb_view(d, x.grp, x.lane, d.tbh) : TbView{nullptr, 0};

    if (!CK)
        for (int32_t cc = 0; cc < 8 * QW; cc++) bd[cc * WAVE] = 0u;

    int32_t best = 0;       // encoded first maximal cell
    int32_t sbest = 0;      // encoded first maximal second-best candidate
    int32_t pm = 0;         // running max of h in processing order (the reference's maxHH)
    for (int32_t i = 0; i < TW; i++) {
        int32_t h[9], f[9];
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; }
        const uint32_t gpac = x.Tp[(size_t)i * x.ps];
        uint32_t lut[9];
#pragma unroll
        for (int m = 1; m < 9; m++) lut[m] = score_lut(c, nval, census, (gpac >> (32 - 4 * m)) & 15u);
        // column tag: ~strip and ~column fields, pre-shifted by 3
        int32_t tagc = (int32_t)((((~(uint32_t)i) & d.enc_sm) << d.enc_cb | d.enc_cm) << 3);
        uint32_t nxt[8];
        int32_t hb = 0;   // CK: H of the bottom row one column back (0 left of column 0)
        if (CK) {   // the previous strip's bottom row; zero above strip 0
            if (i > 0) hv.load_nt(i - 1, 0, nxt);
            else {
#pragma unroll
                for (int kk = 0; kk < 8; kk++) nxt[kk] = 0u;
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[kk * WAVE];
        }
        for (int32_t j = 0; j < QW; j++) {
            const uint32_t rpac = x.Qp[(size_t)j * x.ps];
            uint32_t cur[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) cur[kk] = nxt[kk];
            if (j + 1 < QW) {
                if (CK) {
                    if (i > 0) hv.load_nt(i - 1, j + 1, nxt);
                } else {
#pragma unroll
                    for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[((j + 1) * 8 + kk) * WAVE];
                }
            }
            uint32_t dirw[8];
            if (CK) {   // the tile's left column state
#pragma unroll
                for (int m = 1; m < 9; m++) dirw[m - 1] = pk16(h[m], f[m]);
                tbv.store(i, j, dirw);
            }
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int32_t ridx = j * 8 + kk;
                // 4 * (query nibble & 7): the nibble of column kk sits at bits 31-4kk .. 28-4kk
                const uint32_t qsh = kk < 7 ? (rpac >> (26 - 4 * kk)) & 0x1Cu : (rpac << 2) & 0x1Cu;
                int32_t e = (int32_t)cur[kk] >> 16;
                // p walks down the previous column: H(row-1, col-1) for each row.
                // Checkpoint rows hold H one column behind E, so entry kk
                // already is H(row-1, col-1).
                int32_t p;
                if (CK) {
                    p = sx16(cur[kk]);
                } else {
                    p = h[0];
                    h[0] = sx16(cur[kk]);
                }
                uint32_t dir = 0;
                int32_t colbest = 0, colsec = 0;
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const int32_t s = __builtin_amdgcn_sbfe(lut[m], qsh, 4);
                    const int32_t tmp = p + s;
                    const int32_t hm = max(max(max(tmp, f[m]), e), 0);
                    const int32_t t = tmp - gapoe;
                    const int32_t fx = f[m] - gape, ex = e - gape;
                    if (TB) {
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(tmp - hm), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(f[m] - hm), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(fx - t), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(ex - t), 31);
                    }
                    f[m] = max(t, fx);
                    e = max(t, ex);
                    const int32_t en = (hm << S) | (8 - m);
                    colbest = m == 1 ? en : max(colbest, en);
                    if (SECOND) {
                        const int32_t cand = hm < pm ? en : 0;
                        colsec = m == 1 ? cand : max(colsec, cand);
                        pm = max(pm, hm);
                    }
                    p = h[m];
                    h[m] = hm;
                }
                if (CK) {
                    dirw[kk] = pk16(hb, e);   // H(bottom, col - 1), E(bottom, col)
                    hb = h[8];
                } else {
                    bd[ridx * WAVE] = pk16(h[8], e);
                }
                best = max(best, colbest | tagc);
                if (SECOND) sbest = max(sbest, colsec | tagc);
                tagc -= 8;
                if (TB) dirw[kk] = dir;
            }
            if (TB) tbv.store(i, j, dirw);
            if (CK) hv.store(i, j, dirw);
        }
    }
    const uint32_t pr = x.pair;
    const int32_t vbest = best >> S;
    const uint32_t cm = d.enc_cm, sm = d.enc_sm;
    auto decode = [&](int32_t en, int32_t &row, int32_t &col) {
        const uint32_t u = (uint32_t)en;
        const uint32_t strip = (~(u >> (3 + d.enc_cb))) & sm;
        col = (int32_t)((~(u >> 3)) & cm);
        row = (int32_t)(strip * 8 + (7 - (u & 7u)));
    };
    int32_t maxXY_x = 0, maxXY_y = 0;
    if (vbest > 0) decode(best, maxXY_y, maxXY_x);
    d.score[pr] = vbest;
    d.qend[pr] = maxXY_x;
    d.tend[pr] = maxXY_y;
    if (SECOND) {
        const int32_t vsec = sbest >> S;
        int32_t sy = 0, sx = 0;
        if (vsec > 0) decode(sbest, sy, sx);
        d.s_score[pr] = vsec;
        d.s_tend[pr] = sy;
        d.s_qend[pr] = vbest > 0 ? 8 * QW - 1 : 0;   // compares with the best (src/kernels/local_kernel_template.h:162)
    }
    if (START == S_WITH_START) local_reverse_start(d, c, x, census, vbest, maxXY_x, maxXY_y);
    if (CK && d.tb_fused) {
        __threadfence();   // this lane's checkpoints and results, seen by its own loads
        tb_walk<true>(d, c, x.slot, lut16, true);
    }
}

// ------------------------------- LOCAL, two pairs per lane, packed 16-bit ----
// Lane l of pair-wave w runs the pairs of sorted slots A = 128w + l and
// B = 128w + 64 + l (lane groups 2w and 2w + 1), A in the low and B in the
// high 16-bit half of every register (v_pk_add_u16 / v_pk_sub_u16 /
// v_pk_max_i16 / v_pk_ashrrev_i16 / v_perm_b32).  Each half keeps the int32
// kernel's visiting order -- strip by strip, column by column, row by row --
// so the end (first strict maximum in (strip, column, row) order,
// src/kernels/local_kernel_template.h:145-150,414-418), the second best (the
// largest h below the running maximum when visited, :160-164,432-437) and the
// checkpoints are exactly k_local_fast's.  A cell is ranked within its column
// as h * 8 + (7 - row) (h <= 4095); after each column a half's best takes the
// column's top rank when its h beats the best's h, together with the column's
// index i * C + t (< 2^16).  Loop bounds are the larger pair's; a half's cells
// past its own matrix run on zero words and reach neither its results nor its
// checkpoints (column validity is per query word, so the masks are per word).
// The host enables it when H <= 4095, E and F stay above -32000 and the column
// index fits 16 bits.
#ifndef PK2_WAVES
#define PK2_WAVES 2
#endif
// 0xFFFF in each 16-bit half whose sign bit is set (v_perm sign-byte selects)
__device__ __forceinline__ uint32_t half_sign(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x09090808u); }
// h * 8 + r per half in one packed multiply-add (h >= 0, r < 8 in both halves
// of r2): left to itself the compiler turns the product into a shift and an or
__device__ __forceinline__ pk16x2 prank(pk16x2 h, uint32_t r2) {
    uint32_t v;
    asm("v_pk_mad_u16 %0, %1, 8, %2 op_sel_hi:[1,0,1]" : "=v"(v) : "v"(as_u32(h)), "s"(r2));
    return as_pk(v);
}

template <int START, bool SECOND>
__device__ void lane2_local_pk(const G2Dev &d, const G2Scores &c, uint32_t census, uint32_t wave, uint32_t slot0,
                               const uint32_t *lut16) {
    constexpr bool CK = START == S_WITH_TB;   // WITH_TB only with the checkpointed traceback
    LaneCtx xa, xb;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    auto setup = [&](LaneCtx &x, uint32_t grp) -> bool {
        x.lane = lane;
        x.grp = grp;
        x.slot = slot0 + grp * WAVE + lane;
        if (x.slot >= d.n) return false;
        x.pair = d.perm[x.slot];
        x.qlen = d.qlen[x.pair];
        x.tlen = d.tlen[x.pair];
        x.QW = dregs(x.qlen);
        x.TW = dregs(x.tlen);
        x.Q = d.pq + (d.qoff[x.pair] >> 3);
        x.T = d.pt + (d.toff[x.pair] >> 3);
        if (d.qpan) {
            x.Qp = d.qpan + (size_t)grp * d.qpan_w * WAVE + lane;
            x.Tp = d.tpan + (size_t)grp * d.tpan_w * WAVE + lane;
            x.ps = WAVE;
        } else {
            x.Qp = x.Q;
            x.Tp = x.T;
            x.ps = 1;
        }
        x.bd = d.bnd + (size_t)grp * d.bnd_cols * WAVE + lane;
        return true;
    };
    if (!setup(xa, 2 * wave)) return;
    const bool okB = setup(xb, 2 * wave + 1);
    if (!okB) xb = xa;   // no pair B: the high half repeats A and is never stored
    const int32_t QWa = xa.QW, TWa = xa.TW;
    const int32_t QWb = okB ? xb.QW : 0, TWb = okB ? xb.TW : 0;
    const int32_t QWm = QWa > QWb ? QWa : QWb, TWm = TWa > TWb ? TWa : TWb;
    const int32_t C = 8 * QWm;
    const pk16x2 gapoe2 = {(short)c.gapoe, (short)c.gapoe}, gape2 = {(short)c.gape, (short)c.gape};
    const pk16x2 zero2 = {0, 0};
    const TbView tvA = CK ? tb_view(d, xa.grp, lane) : TbView{nullptr, 0};
    const TbView hvA = CK ? tb_view(d, xa.grp, lane, d.tbh) : TbView{nullptr, 0};
    const TbView tvB = CK ? tb_view(d, xb.grp, lane) : TbView{nullptr, 0};
    const TbView hvB = CK ? tb_view(d, xb.grp, lane, d.tbh) : TbView{nullptr, 0};
    if (!CK) {   // boundary rows (H, E) = 0 above strip 0
        for (int32_t cc = 0; cc < 8 * QWa; cc++) xa.bd[cc * WAVE] = 0u;
        if (okB)
            for (int32_t cc = 0; cc < 8 * QWb; cc++) xb.bd[cc * WAVE] = 0u;
    }
    // A half's words, strips and checkpoint records past its own matrix are
    // read at the nearest index inside it (the values only reach that half's
    // discarded cells); the boundary rows span the batch's widest pair.
    const int32_t qa1 = max(QWa - 1, 0), qb1 = max((okB ? QWb : QWa) - 1, 0);
    const int32_t ta1 = max(TWa - 1, 0), tb1 = max((okB ? TWb : TWa) - 1, 0);
    const bool one_size = d.tb_gbase == nullptr;
    uint32_t best2 = 0, bpos2 = 0;   // per half: ranked best, its column index
    uint32_t sec2 = 0, spos2 = 0;    // per half: ranked second best, its column index
    pk16x2 pm2 = zero2;              // per half: running maximum of h (the reference's maxHH order)
    uint32_t colid2 = 0;             // column index i * C + t in both halves
    for (int32_t i = 0; i < TWm; i++) {
        const bool siA = i < TWa, siB = i < TWb;
        const uint32_t gA = xa.Tp[(uint32_t)min(i, ta1) * xa.ps], gB = xb.Tp[(uint32_t)min(i, tb1) * xb.ps];
        uint32_t lutA[9], lutB[9];
#pragma unroll
        for (int m = 1; m < 9; m++) {
            lutA[m] = lut16[(gA >> (32 - 4 * m)) & 15u];
            lutB[m] = lut16[(gB >> (32 - 4 * m)) & 15u];
        }
        pk16x2 h[9], f[9];
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = zero2; f[m] = zero2; }
        // the row above this strip at a word's columns, one pair each: CK: the
        // previous strip's bottom-row checkpoint (H one column behind E; strip
        // 0 reads strip 0's and masks it to zero); else the boundary row (H, E)
        uint32_t nA[8], nB[8];
        // pack_hi / pack_lo byte selectors; above strip 0 (CK) every byte 0
        const uint32_t sel_hi = (!CK || i > 0) ? 0x07060302u : 0x0C0C0C0Cu;
        const uint32_t sel_lo = (!CK || i > 0) ? 0x05040100u : 0x0C0C0C0Cu;
        auto fetch = [&](int32_t j) {
            if (CK) {
                hvA.load_nt(max(min(i - 1, ta1), 0), min(j, qa1), nA);
                hvB.load_nt(max(min(i - 1, tb1), 0), min(j, qb1), nB);
            } else {
#pragma unroll
                for (int kk = 0; kk < 8; kk++) {
                    nA[kk] = xa.bd[(j * 8 + kk) * WAVE];
                    nB[kk] = xb.bd[(j * 8 + kk) * WAVE];
                }
            }
        };
        fetch(0);
        uint32_t nrA = xa.Qp[0], nrB = xb.Qp[0];   // query words, one word ahead
        pk16x2 hb2 = zero2;   // CK: bottom-row H one column back (0 left of column 0)
        for (int32_t j = 0; j < QWm; j++) {
            const bool vA = siA && j < QWa, vB = siB && j < QWb;
            const uint32_t vm2 = (vA ? 0xFFFFu : 0u) | (vB ? 0xFFFF0000u : 0u);
            const uint32_t rA = nrA, rB = nrB;
            nrA = xa.Qp[(uint32_t)min(j + 1, qa1) * xa.ps];
            nrB = xb.Qp[(uint32_t)min(j + 1, qb1) * xb.ps];
            // this word's row above, packed per column: E and the diagonal H
            uint32_t pe[8], ph[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {   // CK, strip 0: the selectors give zero bytes
                pe[kk] = __builtin_amdgcn_perm(nB[kk], nA[kk], sel_hi);
                ph[kk] = __builtin_amdgcn_perm(nB[kk], nA[kk], sel_lo);
            }
            fetch(j + 1);   // past the last word: an index clamped inside the pair (unused)
            // CK: this tile's records, one per pair, stored without branches.
            // A half past its matrix writes its own region's unused tiles when
            // every group region spans the largest shape, else (per-group
            // regions, or no pair B) the sink.
            uint32_t *const sinkp = d.sink + lane * TB_LANE;
            const bool okA_st = vA || one_size, okB_st = vB || (one_size && okB);
            uint32_t *const lsA = okA_st ? tvA.chunk(i, j) : sinkp, *const lsB = okB_st ? tvB.chunk(i, j) : sinkp;
            uint32_t *const bsA = okA_st ? hvA.chunk(i, j) : sinkp, *const bsB = okB_st ? hvB.chunk(i, j) : sinkp;
            if (CK) {   // the tile's left column state (H, F)
                uint32_t w[8];
#pragma unroll
                for (int m = 1; m < 9; m++) w[m - 1] = pack_lo(as_u32(h[m]), as_u32(f[m]));
                tb_store_rec(lsA, w);
#pragma unroll
                for (int m = 1; m < 9; m++) w[m - 1] = pack_hi(as_u32(h[m]), as_u32(f[m]));
                tb_store_rec(lsB, w);
            }
            const pk16x2 pm_word = pm2;
            uint32_t wA[8], wB[8];   // CK: bottom-row records of this tile (stored whole: a record's
                                     // two halves written apart leave each line half-written twice)
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const uint32_t qshA = kk < 7 ? (rA >> (26 - 4 * kk)) & 0x1Cu : (rA << 2) & 0x1Cu;
                const uint32_t qshB = kk < 7 ? (rB >> (26 - 4 * kk)) & 0x1Cu : (rB << 2) & 0x1Cu;
                pk16x2 e = as_pk(pe[kk]);
                pk16x2 p;
                if (CK) {
                    p = as_pk(ph[kk]);
                } else {
                    p = h[0];
                    h[0] = as_pk(ph[kk]);
                }
                pk16x2 colmax = zero2, colsec = zero2;
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const uint32_t sa = (uint32_t)__builtin_amdgcn_sbfe(lutA[m], qshA, 4);
                    const uint32_t sb = (uint32_t)__builtin_amdgcn_sbfe(lutB[m], qshB, 4);
                    const pk16x2 tmp = p + as_pk(pack_lo(sa, sb));
                    pk16x2 hm;
                    const pk16x2 tt = tmp - gapoe2;
                    if (CK) {   // checkpoints hold the reference's E and F: a
                                // traceback can walk through cells where they are negative
                        hm = pmax(pmax(pmax(tmp, f[m]), e), zero2);
                        f[m] = pmax(tt, f[m] - gape2);
                        e = pmax(tt, e - gape2);
                    } else {    // E and F kept as max(E, 0) and max(F, 0) (an
                                // extension that saturates at 0, then a max with
                                // tt): H = max(tmp, F, E, 0) needs no max with 0
                        hm = pmax(pmax(tmp, f[m]), e);
                        f[m] = pmax(tt, psubsat(f[m], gape2));
                        e = pmax(tt, psubsat(e, gape2));
                    }
                    colmax = pmax(colmax, prank(hm, (uint32_t)(8 - m) * 0x00010001u));
                    p = h[m];
                    h[m] = hm;
                }
                if (SECOND) {
                    // Second best: the ranked h of cells whose h is below the
                    // running maximum at that cell, in row order.  When every
                    // valid half's column maximum is below the running maximum
                    // (most columns of most waves), all the column's cells
                    // count and the running maximum stays: the column's ranked
                    // maximum is the answer.  Otherwise the cells are taken in
                    // order.  A valid half's best h is its running maximum, so
                    // the best can move only in such a column.
                    const uint32_t rec = ~half_sign(as_u32((colmax >> (pk16x2){3, 3}) - pm2)) & vm2;
                    if (__any(rec != 0)) {
#pragma unroll
                        for (int m = 1; m < 9; m++) {
                            const pk16x2 rk = prank(h[m], (uint32_t)(8 - m) * 0x00010001u);
                            colsec = pmax(colsec, as_pk(as_u32(rk) & half_sign(as_u32(h[m] - pm2))));
                            pm2 = pmax(pm2, h[m]);
                        }
                        const uint32_t up = half_sign(as_u32(as_pk(best2 | 0x00070007u) - colmax)) & vm2;
                        best2 = (as_u32(colmax) & up) | (best2 & ~up);
                        bpos2 = (colid2 & up) | (bpos2 & ~up);
                    } else {
                        colsec = colmax;
                    }
                }
                if (CK) {
                    wA[kk] = pack_lo(as_u32(hb2), as_u32(e));   // H(bottom, col - 1), E(bottom, col)
                    wB[kk] = pack_hi(as_u32(hb2), as_u32(e));
                    hb2 = h[8];
                    if (kk == 7) {
                        tb_store_rec(bsA, wA);
                        tb_store_rec(bsB, wB);
                    }
                } else {   // a half past its matrix writes its own (unused) boundary columns
                    xa.bd[(j * 8 + kk) * WAVE] = pack_lo(as_u32(h[8]), as_u32(e));
                    xb.bd[(j * 8 + kk) * WAVE] = pack_hi(as_u32(h[8]), as_u32(e));
                }
                // a half's best moves to this column when its h beats the best's h
                if (!SECOND) {
                    const uint32_t up = half_sign(as_u32(as_pk(best2 | 0x00070007u) - colmax)) & vm2;
                    best2 = (as_u32(colmax) & up) | (best2 & ~up);
                    bpos2 = (colid2 & up) | (bpos2 & ~up);
                } else {
                    const uint32_t us = half_sign(as_u32(as_pk(sec2 | 0x00070007u) - colsec)) & vm2;
                    sec2 = (as_u32(colsec) & us) | (sec2 & ~us);
                    spos2 = (colid2 & us) | (spos2 & ~us);
                }
                colid2 += 0x00010001u;
                // keep each column's ranking with its DP: left to itself the
                // compiler defers the ranking of several columns and holds every
                // h of those columns live (an empty asm that reads and writes the
                // running results orders the columns)
                if (SECOND) asm volatile("" : "+v"(best2), "+v"(bpos2), "+v"(sec2), "+v"(spos2), "+v"(pm2));
                else asm volatile("" : "+v"(best2), "+v"(bpos2));
                __builtin_amdgcn_sched_barrier(0);
            }
            if (SECOND) pm2 = as_pk((as_u32(pm2) & vm2) | (as_u32(pm_word) & ~vm2));   // a half past its matrix keeps its maximum
        }
    }
    auto finish = [&](const LaneCtx &x, uint32_t b, uint32_t bp, uint32_t sc2, uint32_t sp) {
        const uint32_t pr = x.pair;
        const int32_t vbest = (int32_t)(b >> 3);
        int32_t qe = 0, te = 0;
        if (vbest > 0) {
            qe = (int32_t)(bp % (uint32_t)C);
            te = (int32_t)(bp / (uint32_t)C) * 8 + (7 - (int32_t)(b & 7u));
        }
        d.score[pr] = vbest;
        d.qend[pr] = qe;
        d.tend[pr] = te;
        if (SECOND) {
            const int32_t vsec = (int32_t)(sc2 >> 3);
            d.s_score[pr] = vsec;
            d.s_tend[pr] = vsec > 0 ? (int32_t)(sp / (uint32_t)C) * 8 + (7 - (int32_t)(sc2 & 7u)) : 0;
            d.s_qend[pr] = vbest > 0 ? 8 * x.QW - 1 : 0;   // compares with the best (src/kernels/local_kernel_template.h:162)
        }
        if (START == S_WITH_START) local_reverse_start(d, c, x, census, vbest, qe, te);
    };
    finish(xa, best2 & 0xFFFFu, bpos2 & 0xFFFFu, sec2 & 0xFFFFu, spos2 & 0xFFFFu);
    if (okB) finish(xb, best2 >> 16, bpos2 >> 16, sec2 >> 16, spos2 >> 16);
}

// Launch: d.wf_blocks one-wave blocks first (as in k_local_fast), then one
// block per two lane groups.
template <int START, bool SECOND, bool WF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PK2_WAVES, 8))) void k_local_pk2(G2Dev d, G2Scores c) {
    const uint32_t census = d.aux[0];
    if (WF && blockIdx.x < d.wf_blocks) {
        if (START == S_WITHOUT_START && !SECOND) {
            if (census_fast(census)) wf_local<false, true>(d, c, census);
            else wf_local<false, false>(d, c, census);
        }
        if (START == S_WITH_TB && !SECOND) {
            if (census_fast(census)) {
                if (d.wf_tbh) wf_local<true, true, true>(d, c, census);
                else wf_local<true, true>(d, c, census);
            } else {
                wf_local<true, false>(d, c, census);
            }
        }
        return;
    }
    if (!census_fast(census)) return;
    __shared__ uint32_t lut16[16];   // score tables of the 16 target nibbles
    if (threadIdx.x < 16) lut16[threadIdx.x] = score_lut(c, c.has_np ? -c.np : 0, census, threadIdx.x);
    __syncthreads();
    lane2_local_pk<START, SECOND>(d, c, census, blockIdx.x - (WF ? d.wf_blocks : 0u), d.wf_blocks ? d.aux[2] : 0u, lut16);
}

// ----------------------------------------------------- LOCAL, diagonal band ----
// Learned-band DP: k_local_fast's recurrences restricted, per pair, to the 8x8
// tiles that meet the diagonal band d = column - row in [band_lo, band_hi]:
// strip s covers query words [(8s + lo) >> 3, ((8s + 7 + hi) >> 3] clipped to
// the matrix.  A cell outside the computed tiles counts as H = 0 with no gap
// (E, F = BAND_NEG): every banded value is then the best over paths inside the
// tiles, a lower bound of the full-matrix value.  The matrix's own top row and
// first column keep the reference boundary (H = 0, E = F = 0,
// src/kernels/local_kernel_template.h:103,118-127).  Best cell: first strict
// maximum in (strip, column, row) order, as in k_local.
// Flags (band_flags): bit0 the best cell is on an edge tile (a tile with a
// neighbour outside the band), bit2 the check band (the band widened by
// band_widen diagonals each side, run first with CHECK = true, score and ends
// only) found a different score or end, bit3 not run (score tables refused).
constexpr int32_t BAND_NEG = -16000;

__device__ __forceinline__ void band_words(int32_t s, int32_t lo, int32_t hi, int32_t QW, int32_t &wlo, int32_t &whi) {
    wlo = (8 * s + lo) >> 3;            // arithmetic shift: floor
    whi = ((8 * s + 7 + hi) >> 3) + 1;
    wlo = wlo < 0 ? 0 : wlo;
    whi = whi > QW ? QW : whi;
}

template <bool TB, bool CHECK>
__global__ __launch_bounds__(64) void k_local_band(G2Dev d, G2Scores c) {
    const uint32_t census = d.aux[0];
    LaneCtx x;
    if (!lane_setup(d, x)) return;
    const uint32_t pr = x.pair;
    if (!census_fast(census)) { if (!CHECK) d.band_flags[pr] = 8; return; }
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t gapoe = c.gapoe, gape = c.gape;
    const int32_t QW = x.QW, TW = x.TW;
    const int32_t lo = d.band_lo[pr] - (CHECK ? d.band_widen : 0), hi = d.band_hi[pr] + (CHECK ? d.band_widen : 0);
    uint32_t *__restrict__ bd = x.bd;
    uint32_t *tbp = TB ? d.tb + d.band_tbo[pr] * 8 : nullptr;
    const int32_t bw = TB ? (int32_t)d.band_bw[pr] : 0;
    int32_t rv = 0, rs = 0, rc = 0;   // best: h << 3 | (8 - row in strip), strip, column
    bool best_edge = false, overflow = false;
    int32_t plo = 0, phi = 0;         // word range of the previous strip
    for (int32_t st = 0; st < TW; st++) {
        int32_t wlo, whi;
        band_words(st, lo, hi, QW, wlo, whi);
        if (wlo >= whi) { plo = phi = 0; continue; }
        int32_t h[9], f[9];
        const uint32_t gpac = x.T[st];
        uint32_t lut[9];
#pragma unroll
        for (int m = 1; m < 9; m++) lut[m] = score_lut(c, nval, census, (gpac >> (32 - 4 * m)) & 15u);
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = wlo == 0 ? 0 : BAND_NEG; }
        // The boundary row is kept band-relative: strip st writes column col at
        // bd[col - 8 wlo], so the lanes of a wave (different bands) touch at
        // most two rows per access; the previous strip's value of col sits at
        // col - 8 plo (>= the write index, read before it is overwritten).
        // H(strip top - 1, first column - 1), if the previous strip computed it
        if (st > 0 && wlo > 0 && wlo - 1 >= plo && wlo - 1 < phi) h[0] = sx16(bd[(8 * (wlo - plo) - 1) * WAVE]);
        for (int32_t w = wlo; w < whi; w++) {
            const uint32_t rpac = x.Q[w];
            const bool top_in = st == 0 || (w >= plo && w < phi);
            const bool edge = (w == wlo && wlo > 0) || (w == whi - 1 && whi < QW) || !top_in;
            uint32_t dirw[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int32_t col = 8 * w + kk;
                const uint32_t qsh = kk < 7 ? (rpac >> (26 - 4 * kk)) & 0x1Cu : (rpac << 2) & 0x1Cu;
                const uint32_t cur = st == 0 ? 0u : (top_in ? bd[(8 * (w - plo) + kk) * WAVE] : pk16(0, BAND_NEG));
                int32_t e = (int32_t)cur >> 16;
                int32_t p = h[0];
                h[0] = sx16(cur);
                uint32_t dir = 0;
                int32_t colbest = 0;
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const int32_t sc = __builtin_amdgcn_sbfe(lut[m], qsh, 4);
                    const int32_t tmp = p + sc;
                    const int32_t hm = max(max(max(tmp, f[m]), e), 0);
                    const int32_t t = tmp - gapoe;
                    const int32_t fx = f[m] - gape, ex = e - gape;
                    if (TB) {
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(tmp - hm), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(f[m] - hm), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(fx - t), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(ex - t), 31);
                    }
                    f[m] = max(t, fx);
                    e = max(t, ex);
                    const int32_t en = (hm << 3) | (8 - m);
                    colbest = m == 1 ? en : max(colbest, en);
                    p = h[m];
                    h[m] = hm;
                }
                bd[(8 * (w - wlo) + kk) * WAVE] = pk16(h[8], e);
                if (colbest > (rv | 7)) { rv = colbest; rs = st; rc = col; best_edge = edge; }
                // keep each column's bookkeeping with its DP (left alone the
                // compiler holds several columns' values live: 256 VGPRs with
                // the direction words)
                if (TB) {
                    dirw[kk] = dir;
                    asm volatile("" : "+v"(rv), "+v"(rs), "+v"(rc), "+v"(dirw[kk]));
                } else {
                    asm volatile("" : "+v"(rv), "+v"(rs), "+v"(rc));
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            if (TB) {
                if (w - wlo < bw) tb_store(tbp + ((size_t)st * bw + (size_t)(w - wlo)) * 8, dirw);
                else overflow = true;
            }
        }
        plo = wlo;
        phi = whi;
    }
    const int32_t v = rv >> 3;
    const int32_t qe = v > 0 ? rc : 0, te = v > 0 ? rs * 8 + (7 - (rv & 7)) : 0;
    if (CHECK) {
        d.band_wres[pr] = v;
        d.band_wres[d.n + pr] = qe;
        d.band_wres[2 * d.n + pr] = te;
        return;
    }
    d.score[pr] = v;
    d.qend[pr] = qe;
    d.tend[pr] = te;
    const bool differs = d.band_wres &&
                         (d.band_wres[pr] != v || d.band_wres[d.n + pr] != qe || d.band_wres[2 * d.n + pr] != te);
    d.band_flags[pr] = (uint8_t)((best_edge && v > 0 ? 1u : 0u) | (differs ? 4u : 0u) | (overflow ? 8u : 0u));
}

// --------------------------------------------------------------- GLOBAL -----
// gasal_global_kernel<S> (src/kernels/global.h:30-303).
// LUT: substitution from a per-strip 4-bit score table (see k_local_fast),
// launched beside the generic variant; the nibble census picks one.
// ------------------------------------------ GLOBAL, packed 16-bit, one lane ----
// k_global's score (WITHOUT_START / WITH_START) with two 8-row strips per lane
// in the 16-bit halves of each register, skewed by one column as in
// lane_local_pk: the low half runs strip 2k at column t, the high half strip
// 2k+1 at column t-1 with the low half's bottom row of the previous step as
// its top row.  Same boundaries as k_global (top row -(open + extend * c),
// column 0 left values per row, src/kernels/global.h:57-71), no clamp at 0.
// Used when every value fits int16 with room for the packed minus infinity
// (NEG16 stands for MINUS_INF: it never wins a maximum against a real value
// in that range, so the scores are those of the int32 kernel).
constexpr int32_t NEG16 = -30000;
__device__ void lane_global_pk(const G2Dev &d, const G2Scores &c, const LaneCtx &x, uint32_t census) {
    const int32_t gapo = c.gapo, gape = c.gape;
    const pk16x2 gapoe2 = {(short)c.gapoe, (short)c.gapoe}, gape2 = {(short)c.gape, (short)c.gape};
    const int32_t QW = x.QW, TW = x.TW, C = 8 * QW;
    uint32_t *__restrict__ bd = x.bd;
    bd[0] = pk16(0, NEG16);                                                                // :57-60
    for (int32_t cc = 1; cc < C; cc++) bd[cc * WAVE] = pk16(-(gapo + gape * cc), NEG16);
    // Diagonal entering row R + 1 (0-based R) at column 0: the reference's
    // p[m] for 1-based row R + 1, 0 for the first row, else -(open + extend * R)
    // (src/kernels/global.h:67-71)
    auto hl = [&](int32_t R) -> int32_t { return R == 0 ? 0 : -(gapo + gape * R); };
    const int32_t rowsel = 8 - ((TW << 3) - (int32_t)x.tlen);                              // :299
    const int32_t lastcol = (int32_t)x.qlen - 1;
    int32_t score = 0;
    for (int32_t s0 = 0; s0 < TW; s0 += 2) {
        const bool hi_ok = s0 + 1 < TW;
        uint32_t lutL[9], lutH[9];
        const uint32_t gL = x.Tp[(size_t)s0 * x.ps], gH = hi_ok ? x.Tp[(size_t)(s0 + 1) * x.ps] : 0u;
#pragma unroll
        for (int m = 1; m < 9; m++) {
            lutL[m] = score_lut<true>(c, 0, census, (gL >> (32 - 4 * m)) & 15u);
            lutH[m] = score_lut<true>(c, 0, census, (gH >> (32 - 4 * m)) & 15u);
        }
        // running diagonal: h[m] holds H(row m, previous column); column -1 is the left boundary
        pk16x2 h[9], f[9];
        const pk16x2 neg2 = {(short)NEG16, (short)NEG16};
#pragma unroll
        for (int m = 0; m < 9; m++) {
            h[m] = as_pk(pack_lo((uint32_t)hl(8 * s0 + m), (uint32_t)hl(8 * (s0 + 1) + m)));
            f[m] = neg2;
        }
        uint32_t carry = 0;            // low strip's bottom (H | E << 16) of the previous step
        uint32_t qshH = 0;
        const bool last_lo = s0 == TW - 1, last_hi = hi_ok && s0 + 1 == TW - 1;
        // one step: the low strip at column t (query code qshL, row above topL),
        // the high strip at column t - 1
        auto step = [&](int32_t t, uint32_t qshL, uint32_t topL, bool ends) {
            pk16x2 e = as_pk(pack_hi(topL, carry));
            pk16x2 p = h[0];
            h[0] = as_pk(pack_lo(topL, carry));
#pragma unroll
            for (int m = 1; m < 9; m++) {
                const uint32_t sl = (uint32_t)__builtin_amdgcn_sbfe(lutL[m], qshL, 4);
                const uint32_t sh = (uint32_t)__builtin_amdgcn_sbfe(lutH[m], qshH, 4);
                const pk16x2 tmp = p + as_pk(pack_lo(sl, sh));
                const pk16x2 hm = pmax(pmax(tmp, f[m]), e);
                const pk16x2 tt = tmp - gapoe2;
                f[m] = pmax(tt, f[m] - gape2);
                e = pmax(tt, e - gape2);
                p = h[m];
                h[m] = hm;
            }
            qshH = qshL;
            const uint32_t h8 = as_u32(h[8]), e8 = as_u32(e);
            if (t >= 1) bd[(t - 1) * WAVE] = pack_hi(h8, e8);
            carry = pack_lo(h8, e8);
            if (ends) {
                if (last_lo && t == lastcol) {                                             // :98-103
#pragma unroll
                    for (int m = 1; m < 9; m++) score = (m == rowsel) ? (int32_t)(short)(as_u32(h[m]) & 0xFFFFu) : score;
                }
                if (last_hi && t - 1 == lastcol) {
#pragma unroll
                    for (int m = 1; m < 9; m++) score = (m == rowsel) ? (int32_t)(short)(as_u32(h[m]) >> 16) : score;
                }
            }
            if (t == 0) {   // the high strip starts at the next step: its column -1 again
#pragma unroll
                for (int m = 0; m < 9; m++) {
                    h[m] = as_pk((as_u32(h[m]) & 0xFFFFu) | ((uint32_t)hl(8 * (s0 + 1) + m) << 16));
                    f[m] = as_pk((as_u32(f[m]) & 0xFFFFu) | ((uint32_t)NEG16 << 16));
                }
            }
        };
        // a word of eight columns per iteration, its row-above entries loaded
        // one word ahead; the score checks run only while some lane of the
        // wave is in its last pair of strips
        auto words = [&](auto ENDS) {
            uint32_t top[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) top[kk] = bd[kk * WAVE];
            for (int32_t jw = 0; jw < QW; jw++) {
                const uint32_t rpac = x.Qp[(size_t)jw * x.ps];
                uint32_t cur[8];
#pragma unroll
                for (int kk = 0; kk < 8; kk++) cur[kk] = top[kk];
                if (jw + 1 < QW) {
#pragma unroll
                    for (int kk = 0; kk < 8; kk++) top[kk] = bd[(8 * (jw + 1) + kk) * WAVE];
                }
#pragma unroll
                for (int kk = 0; kk < 8; kk++) {
                    const uint32_t qsh = kk < 7 ? (rpac >> (26 - 4 * kk)) & 0x1Cu : (rpac << 2) & 0x1Cu;
                    step(8 * jw + kk, qsh, cur[kk], decltype(ENDS)::value);
                }
            }
            step(C, 0, 0u, decltype(ENDS)::value);   // the high strip's last column
        };
        if (__any(last_lo || last_hi)) words(std::true_type{});
        else words(std::false_type{});
    }
    d.score[x.pair] = score;
}

template <bool TB, bool LUT>
__global__ __launch_bounds__(64) void k_global(G2Dev d, G2Scores c) {
    const uint32_t census = d.aux[0];
    if (LUT ? !census_fast(census) : (d.fast && census_fast(census))) return;
    LaneCtx x;
    if (!lane_setup(d, x)) return;
    if (LUT && !TB && d.pk16g) {
        lane_global_pk(d, c, x, census);
        return;
    }
    const int32_t gapo = c.gapo, gapoe = c.gapoe, gape = c.gape;
    const int32_t QW = x.QW, TW = x.TW;
    uint32_t *__restrict__ bd = x.bd;
    const TbView tbv = TB ? tb_view(d, x.grp, x.lane) : TbView{nullptr, 0};
    bd[0] = pk16(0, MINUS_INF);                                                            // :57-60
    for (int32_t cc = 1; cc < 8 * QW; cc++) bd[cc * WAVE] = pk16(-(gapo + gape * cc), MINUS_INF);
    const int32_t rowsel = 8 - ((TW << 3) - (int32_t)x.tlen);                              // :299
    const int32_t lastcol = (int32_t)x.qlen - 1;
    int32_t score = 0;
    int32_t r = 1;
    for (int32_t i = 0; i < TW; i++) {
        int32_t h[9], f[9], p[9];
        h[0] = 0; p[0] = 0; f[0] = 0;
#pragma unroll
        for (int m = 1; m < 9; m++) {                                                      // :67-71
            h[m] = 0;
            f[m] = MINUS_INF;
            p[m] = (r + m - 1) == 1 ? 0 : -(gapo + gape * (r + m - 2));
        }
        r += 8;
        const uint32_t gpac = x.T[i];
        uint32_t gb[9];
#pragma unroll
        for (int m = 1; m < 9; m++) {
            gb[m] = (gpac >> (32 - 4 * m)) & 15u;
            if (LUT) gb[m] = score_lut<true>(c, 0, census, gb[m]);
        }
        for (int32_t j = 0; j < QW; j++) {
            const uint32_t rpac = x.Q[j];
            uint32_t cur[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) cur[kk] = bd[(j * 8 + kk) * WAVE];
            uint32_t dirw[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int32_t ridx = j * 8 + kk;
                const uint32_t rbase = (rpac >> (28 - 4 * kk)) & 15u;
                h[0] = sx16(cur[kk]);
                int32_t e = sx16(cur[kk] >> 16);
                uint32_t dir = 0;
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const int32_t s = LUT ? __builtin_amdgcn_sbfe(gb[m], (rbase & 7u) << 2, 4) : sub_global(c, rbase, gb[m]);
                    const int32_t tmp = p[m] + s;
                    int32_t hm = max(max(tmp, f[m]), e);
                    if (TB) {   // sign-bit direction format of k_local_fast (k_get_tb decodes it)
                        const int32_t t = tmp - gapoe;
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(tmp - hm), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(f[m] - hm), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(f[m] - gape - t), 31);
                        dir = __builtin_amdgcn_alignbit(dir, (uint32_t)(e - gape - t), 31);
                    }
                    f[m] = max(tmp - gapoe, f[m] - gape);
                    e = max(tmp - gapoe, e - gape);
                    p[m] = h[m - 1];
                    h[m] = hm;
                }
                bd[ridx * WAVE] = pk16(h[8], e);
                if (ridx == lastcol && i == TW - 1) {                                      // :98-103
#pragma unroll
                    for (int m = 1; m < 9; m++) score = (m == rowsel) ? h[m] : score;
                }
                if (TB) {   // one column at a time (else the compiler holds several: 256 VGPRs)
                    dirw[kk] = dir;
                    asm volatile("" : "+v"(dirw[kk]), "+v"(score));
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if (TB) tbv.store(i, j, dirw);
        }
    }
    d.score[x.pair] = score;
}

// ---------------------------------------------------------- SEMI_GLOBAL -----
// gasal_semi_global_kernel<T,S,B,HEAD,TAIL> (src/kernels/semiglobal_kernel_template.h:39-388).
template <int HEAD>
__device__ __forceinline__ uint32_t semi_init_word(int32_t cc, const G2Scores &c) {
    if (HEAD == D_QUERY || HEAD == D_BOTH) return 0u;                                      // :87-92
    return cc == 0 ? pk16(0, MINUS_INF) : pk16(-(c.gapo + c.gape * cc), MINUS_INF);       // :93-99
}
template <int HEAD>
__device__ __forceinline__ int32_t semi_init_h(int32_t cc, const G2Scores &c) {
    if (HEAD == D_QUERY || HEAD == D_BOTH) return 0;
    return cc == 0 ? 0 : t16(-(c.gapo + c.gape * cc));
}
template <int HEAD>
__device__ __forceinline__ void semi_strip_init(const G2Scores &c, int32_t *h, int32_t *f, int32_t *p, int32_t &u) {
    if (HEAD == D_TARGET || HEAD == D_BOTH) {                                              // :114-121
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = MINUS_INF; p[m] = 0; }
    } else {                                                                               // :122-129 (r == u)
#pragma unroll
        for (int m = 1; m < 9; m++) {
            int32_t uu = u + m - 1;
            h[m] = -(c.gapo + (c.gape * (uu - 1)));
            f[m] = MINUS_INF;
            p[m] = uu == 1 ? 0 : -(c.gapo + (c.gape * (uu - 1)));
        }
        u += 8;
    }
}
// gb[m]: the row's target nibble, or (LUT) its 4-bit score table.
template <bool LUT>
__device__ __forceinline__ void semi_column(const G2Scores &c, int32_t nval, uint32_t rbase, const uint32_t *gb,
                                            int32_t *h, int32_t *f, int32_t *p, uint32_t &hd) {
    h[0] = sx16(hd);
    int32_t e = sx16(hd >> 16);
    int32_t prev_hm_diff = h[0] - c.gapoe;
    const uint32_t qsh = (rbase & 7u) << 2;
#pragma unroll
    for (int m = 1; m < 9; m++) {                                                          // :17-28
        const int32_t s = LUT ? __builtin_amdgcn_sbfe(gb[m], qsh, 4) : sub_local(c, nval, rbase, gb[m]);
        int32_t cur = h[m] - c.gapoe;
        f[m] = max(cur, f[m] - c.gape);
        cur = p[m] + s;
        cur = max(cur, f[m]);
        e = max(prev_hm_diff, e - c.gape);
        cur = max(cur, e);
        h[m] = cur;
        p[m] = prev_hm_diff + c.gapoe;
        prev_hm_diff = cur - c.gapoe;
    }
    hd = pk16(h[8], e);
}
// Reversed copy word k of a sequence (:244-266): base 8k+b at bits 28-4b is the
// original base len-1-(8k+b); positions past the start are nibble 0.
// It is the nibble-reversed 8-base window of the forward sequence that starts
// at base len-8(k+1): one funnel shift of two forward words, then a nibble
// reverse.
__device__ __forceinline__ uint32_t rev_word(const uint32_t *W, int32_t len, int32_t k) {
    const int32_t st = len - 8 * (k + 1);
    uint32_t win;
    if (st >= 0) {
        const int32_t w0 = st >> 3, off = st & 7;
        win = off == 0 ? W[w0] : __builtin_amdgcn_alignbit(W[w0], W[w0 + 1], 32 - 4 * off);
    } else {
        win = W[0] >> (4 * (-st));   // bases before the start are nibble 0
    }
    return nib_reverse(win);
}

// One pass of k_semi (forward, or with REV the WITH_START pass over reversed
// copies) with two strips per lane in packed 16-bit halves, skewed by one
// column as in lane_local_pk; tails that end on the target only (TT) or none.
// Same recurrences (semi_column), strip initialisation (semi_strip_init) and
// end bookkeeping per row in strip order (src/kernels/semiglobal_kernel_template.h
// :114-178, :269-360); NEG16 stands for MINUS_INF.  REV: strips from i0 while
// the best is below fwd_score; the strip after the one that reaches it is not
// run (the reference's loop exit).
template <bool SECOND, int HEAD, bool TT, bool REV>
__device__ void semi_pass_pk(const G2Dev &d, const G2Scores &c, const LaneCtx &x, uint32_t census, int32_t nval, int32_t i0,
                             int32_t fwd_score, int32_t &maxHH, int32_t &maxY, int32_t &sec, int32_t &secy) {
    const int32_t QW = x.QW, TW = x.TW, C = 8 * QW;
    const int32_t read_len = (int32_t)x.qlen, ref_len = (int32_t)x.tlen;
    const int32_t gapo = c.gapo, gape = c.gape;
    const pk16x2 gapoe2 = {(short)c.gapoe, (short)c.gapoe}, gape2 = {(short)c.gape, (short)c.gape};
    const pk16x2 neg2 = {(short)NEG16, (short)NEG16};
    constexpr bool HT = HEAD == D_TARGET || HEAD == D_BOTH;
    uint32_t *__restrict__ bd = x.bd;
    for (int32_t cc = 0; cc < C; cc++)                                                     // :87-99
        bd[cc * WAVE] = (HEAD == D_QUERY || HEAD == D_BOTH) ? 0u : (cc == 0 ? pk16(0, NEG16) : pk16(-(gapo + gape * cc), NEG16));
    // semi_strip_init for the strip whose row counter is u (HEAD NONE / QUERY)
    auto h_init = [&](int32_t u, int m) -> int32_t { return HT ? 0 : -(gapo + gape * (u + m - 2)); };
    auto p_init = [&](int32_t u, int m) -> int32_t { return (HT || u + m - 1 == 1) ? 0 : -(gapo + gape * (u + m - 2)); };
    const int32_t tend_col = read_len - 1;
    int32_t u = 1;
    bool stop = false;
    for (int32_t s0 = i0; s0 < TW && !stop; s0 += 2) {
        if (REV && !(maxHH < fwd_score)) break;
        const bool hi_ok = s0 + 1 < TW;
        auto strip_word = [&](int32_t i) -> uint32_t { return REV ? (i >= 0 ? rev_word(x.T, ref_len, i) : 0u) : x.T[i]; };
        const uint32_t gL = strip_word(s0), gH = hi_ok ? strip_word(s0 + 1) : 0u;
        uint32_t lutL[9], lutH[9];
#pragma unroll
        for (int m = 1; m < 9; m++) {
            lutL[m] = score_lut(c, nval, census, (gL >> (32 - 4 * m)) & 15u);
            lutH[m] = score_lut(c, nval, census, (gH >> (32 - 4 * m)) & 15u);
        }
        const int32_t uL = u, uH = u + 8;
        u += 16;
        pk16x2 h[9], f[9], p[9];
#pragma unroll
        for (int m = 1; m < 9; m++) {
            h[m] = as_pk(pack_lo((uint32_t)h_init(uL, m), (uint32_t)h_init(uH, m)));
            p[m] = as_pk(pack_lo((uint32_t)p_init(uL, m), (uint32_t)p_init(uH, m)));
            f[m] = neg2;
        }
        const int32_t gL8 = s0 << 3, gH8 = (s0 + 1) << 3;
        bool hi_live = hi_ok;
        uint32_t carry = 0, qshH = 0;
        // one step: the low strip at column t (query code qshL, row above
        // topL), the high strip at column t - 1
        auto step = [&](int32_t t, uint32_t qshL, uint32_t topL) {
            pk16x2 e = as_pk(pack_hi(topL, carry));
            pk16x2 ph = as_pk(pack_lo(topL, carry));   // H of the row above, this column
            pk16x2 prev = ph - gapoe2;
#pragma unroll
            for (int m = 1; m < 9; m++) {                                                  // :17-28
                const uint32_t sl = (uint32_t)__builtin_amdgcn_sbfe(lutL[m], qshL, 4);
                const uint32_t sh = (uint32_t)__builtin_amdgcn_sbfe(lutH[m], qshH, 4);
                f[m] = pmax(h[m] - gapoe2, f[m] - gape2);
                pk16x2 cur = pmax(p[m] + as_pk(pack_lo(sl, sh)), f[m]);
                e = pmax(prev, e - gape2);
                cur = pmax(cur, e);
                h[m] = cur;
                p[m] = ph;
                ph = cur;
                prev = cur - gapoe2;
            }
            const uint32_t h8 = as_u32(h[8]), e8 = as_u32(e);
            if (t >= 1 && hi_ok) bd[(t - 1) * WAVE] = pack_hi(h8, e8);
            carry = pack_lo(h8, e8);
            if (TT && t == tend_col) {                                                     // :160-178 (low strip)
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const int32_t hv = (int32_t)(short)(as_u32(h[m]) & 0xFFFFu);
                    const bool in = (gL8 + m - 1) < ref_len;
                    const bool upd = hv > maxHH && in;
                    maxY = upd ? gL8 + (m - 1) : maxY;
                    maxHH = upd ? hv : maxHH;
                    if (SECOND) {
                        const bool ov = hv > sec && hv < maxHH && in;
                        secy = ov ? gL8 + (m - 1) : secy;
                        sec = ov ? hv : sec;
                    }
                }
                if (REV && !(maxHH < fwd_score)) { hi_live = false; stop = true; }
            }
            if (TT && hi_live && t - 1 == tend_col) {                                      // (high strip)
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    const int32_t hv = (int32_t)(short)(as_u32(h[m]) >> 16);
                    const bool in = (gH8 + m - 1) < ref_len;
                    const bool upd = hv > maxHH && in;
                    maxY = upd ? gH8 + (m - 1) : maxY;
                    maxHH = upd ? hv : maxHH;
                    if (SECOND) {
                        const bool ov = hv > sec && hv < maxHH && in;
                        secy = ov ? gH8 + (m - 1) : secy;
                        sec = ov ? hv : sec;
                    }
                }
            }
            if (t == 0) {   // the high strip starts at the next step: its column -1 again
#pragma unroll
                for (int m = 1; m < 9; m++) {
                    h[m] = as_pk((as_u32(h[m]) & 0xFFFFu) | ((uint32_t)h_init(uH, m) << 16));
                    p[m] = as_pk((as_u32(p[m]) & 0xFFFFu) | ((uint32_t)p_init(uH, m) << 16));
                    f[m] = as_pk((as_u32(f[m]) & 0xFFFFu) | ((uint32_t)NEG16 << 16));
                }
            }
            qshH = qshL;
        };
        // a word of eight columns per iteration, its row-above entries loaded
        // one word ahead
        uint32_t top[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) top[kk] = bd[kk * WAVE];
        for (int32_t jw = 0; jw < QW; jw++) {
            const uint32_t rpac = REV ? rev_word(x.Q, read_len, jw) : x.Q[jw];
            uint32_t cur[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) cur[kk] = top[kk];
            if (jw + 1 < QW) {
#pragma unroll
                for (int kk = 0; kk < 8; kk++) top[kk] = bd[(8 * (jw + 1) + kk) * WAVE];
            }
#pragma unroll
            for (int kk = 0; kk < 8; kk++)
                step(8 * jw + kk, kk < 7 ? (rpac >> (26 - 4 * kk)) & 0x1Cu : (rpac << 2) & 0x1Cu, cur[kk]);
        }
        step(C, 0, 0u);   // the high strip's last column
    }
}

template <int START, bool SECOND, int HEAD, int TAIL, bool LUT>
__global__ __launch_bounds__(64) void k_semi(G2Dev d, G2Scores c) {
    // the WITH_START pass reads reversed copies padded with nibble 0
    const uint32_t census = d.aux[0] | (START == S_WITH_START ? 1u : 0u);
    if (LUT ? !census_fast(census) : (d.fast && census_fast(census))) return;
    LaneCtx x;
    if (!lane_setup(d, x)) return;
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t QW = x.QW, TW = x.TW;
    const int32_t read_len = (int32_t)x.qlen, ref_len = (int32_t)x.tlen;
    uint32_t *__restrict__ bd = x.bd;
    constexpr bool TT = (TAIL == D_TARGET || TAIL == D_BOTH);
    constexpr bool TQ = (TAIL == D_QUERY || TAIL == D_BOTH);
    if (LUT && !TQ && d.pk16s) {   // packed 16-bit passes (values fit int16, tails on the target only)
        int32_t maxHH = MINUS_INF, maxY = read_len, sec = MINUS_INF, secy = read_len;
        semi_pass_pk<SECOND, HEAD, TT, false>(d, c, x, census, nval, 0, 0, maxHH, maxY, sec, secy);
        const uint32_t pr = x.pair;
        d.score[pr] = maxHH;
        d.tend[pr] = maxY;
        d.qend[pr] = ref_len;   // maxXY_x keeps its initial value without a query tail (:63-64)
        if (SECOND) { d.s_score[pr] = sec; d.s_tend[pr] = secy; d.s_qend[pr] = ref_len; }
        if (START == S_WITH_START) {                                                       // :227-384
            const int32_t fwd_score = maxHH;
            const int32_t dd = TW - ((maxY >> 3) + 1);
            const int32_t gend_reg = dd > 0 ? dd - 1 : dd;                                  // :273
            int32_t rmax = MINUS_INF, ry = 0, rs = 0, rsy = 0;
            semi_pass_pk<false, HEAD, TT, true>(d, c, x, census, nval, gend_reg, fwd_score, rmax, ry, rs, rsy);
            d.tstart[pr] = (ref_len - 1) - ry;                                             // :380-381
            d.qstart[pr] = (read_len - 1) - ref_len;
        }
        return;
    }

    int32_t maxHH = MINUS_INF, maxXY_x = ref_len, maxXY_y = read_len;                      // :49,63-64
    int32_t sec = MINUS_INF, secx = ref_len, secy = read_len;                              // :71-74
    for (int32_t cc = 0; cc < 8 * QW; cc++) bd[cc * WAVE] = semi_init_word<HEAD>(cc, c);
    int32_t h[9], f[9], p[9];
    int32_t u = 1;
    h[0] = 0; p[0] = 0; f[0] = 0;
    for (int32_t i = 0; i < TW; i++) {
        const int32_t gidx = i << 3;
        semi_strip_init<HEAD>(c, h, f, p, u);
        const uint32_t gpac = x.T[i];
        uint32_t gb[9];
#pragma unroll
        for (int m = 1; m < 9; m++) {
            gb[m] = (gpac >> (32 - 4 * m)) & 15u;
            if (LUT) gb[m] = score_lut(c, nval, census, gb[m]);
        }
        for (int32_t j = 0; j < QW; j++) {
            const uint32_t rpac = x.Q[j];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int32_t ridx = j * 8 + kk;
                uint32_t hd = bd[ridx * WAVE];
                semi_column<LUT>(c, nval, (rpac >> (28 - 4 * kk)) & 15u, gb, h, f, p, hd);
                bd[ridx * WAVE] = hd;
                if (TT && ridx + 1 == read_len) {                                          // :160-178
#pragma unroll
                    for (int m = 1; m < 9; m++) {
                        bool in = (gidx + m - 1) < ref_len;
                        bool upd = h[m] > maxHH && in;
                        maxXY_y = upd ? gidx + (m - 1) : maxXY_y;
                        maxHH = upd ? h[m] : maxHH;
                        if (SECOND) {
                            bool ov = h[m] > sec && h[m] < maxHH && in;
                            secy = ov ? gidx + (m - 1) : secy;
                            sec = ov ? h[m] : sec;
                        }
                    }
                }
            }
        }
    }
    if (TQ) {                                                                              // :185-214
        const int32_t lim = max(max(8 * QW, read_len), ref_len);
        for (int32_t m = 0; m < lim; m++) {
            int32_t st = m < 8 * QW ? sx16(bd[m * WAVE]) : semi_init_h<HEAD>(m, c);
            if (st > maxHH && m < read_len) { maxXY_x = m; maxHH = st; }
            if (SECOND) {
                bool ov = st > sec && st < maxHH && m < ref_len;
                secx = ov ? m : secx;
                sec = ov ? st : sec;
            }
        }
        if (maxXY_x != ref_len) maxXY_y = read_len;
        if (SECOND && secx != ref_len) secy = read_len;
    }
    const uint32_t pr = x.pair;
    d.score[pr] = maxHH;
    d.tend[pr] = maxXY_y;
    d.qend[pr] = maxXY_x;
    if (SECOND) { d.s_score[pr] = sec; d.s_tend[pr] = secy; d.s_qend[pr] = secx; }

    if (START == S_WITH_START) {                                                           // :227-384
        const int32_t fwd_score = maxHH;
        const int32_t dd = TW - ((maxXY_y >> 3) + 1);
        const int32_t gend_reg = dd > 0 ? dd - 1 : dd;                                      // :273
        maxHH = MINUS_INF;
        maxXY_y = 0;
        for (int32_t cc = 0; cc < 8 * QW; cc++) bd[cc * WAVE] = semi_init_word<HEAD>(cc, c);
        u = 1;
        h[0] = 0; p[0] = 0;
        for (int32_t i = gend_reg; i < TW && maxHH < fwd_score; i++) {
            const int32_t gidx = i << 3;
            semi_strip_init<HEAD>(c, h, f, p, u);
            const uint32_t gpac = i >= 0 ? rev_word(x.T, ref_len, i) : 0u;
            uint32_t gb[9];
#pragma unroll
            for (int m = 1; m < 9; m++) {
            gb[m] = (gpac >> (32 - 4 * m)) & 15u;
            if (LUT) gb[m] = score_lut(c, nval, census, gb[m]);
        }
            for (int32_t j = 0; j < QW && maxHH < fwd_score; j++) {
                const uint32_t rpac = rev_word(x.Q, read_len, j);
#pragma unroll
                for (int kk = 0; kk < 8; kk++) {
                    const int32_t ridx = j * 8 + kk;
                    uint32_t hd = bd[ridx * WAVE];
                    semi_column<LUT>(c, nval, (rpac >> (28 - 4 * kk)) & 15u, gb, h, f, p, hd);
                    bd[ridx * WAVE] = hd;
                    if (TT && ridx + 1 == read_len) {
#pragma unroll
                        for (int m = 1; m < 9; m++) {
                            bool upd = h[m] > maxHH && (gidx + m - 1) < ref_len;
                            maxXY_y = upd ? gidx + (m - 1) : maxXY_y;
                            maxHH = upd ? h[m] : maxHH;
                        }
                    }
                }
            }
        }
        if (TQ) {                                                                          // :362-378
            const int32_t lim = max(8 * QW, read_len);
            for (int32_t m = 0; m < lim; m++) {
                int32_t st = m < 8 * QW ? sx16(bd[m * WAVE]) : semi_init_h<HEAD>(m, c);
                if (st > maxHH && m < read_len) { maxXY_x = m; maxHH = st; }
            }
            if (maxXY_x != ref_len) maxXY_y = read_len;
        }
        d.tstart[pr] = (ref_len - 1) - maxXY_y;                                           // :380-381
        d.qstart[pr] = (read_len - 1) - maxXY_x;
    }
}

// --------------------------------------------------------------- BANDED -----
// gasal_banded_tiled_kernel (src/kernels/banded.h:10-139).
// LUT: substitution scores from a per-row table of the batch's nibble classes
// (the census allows it), else per cell.  A word's boundary-row entries are
// loaded one word ahead.
template <bool LUT>
__global__ __launch_bounds__(64) void k_banded(G2Dev d, G2Scores c, int32_t k_band_width) {
    const uint32_t census = d.aux[0];
    if (LUT ? !(d.fast && census_fast(census)) : (d.fast && census_fast(census))) return;
    LaneCtx x;
    if (!lane_setup(d, x)) return;
    const int32_t nval = c.has_np ? -c.np : 0;
    const int32_t gapoe = c.gapoe, gape = c.gape;
    const int32_t QW = x.QW, TW = x.TW;
    uint32_t *__restrict__ bd = x.bd;
    for (int32_t cc = 0; cc < 8 * QW; cc++) bd[cc * WAVE] = 0u;
    int32_t maxHH = 0, prev_maxHH = 0, maxXY_x = 0, maxXY_y = 0;
    const int32_t other = TW - (QW - k_band_width);                                        // :35
    for (int32_t i = 0; i < TW; i++) {
        int32_t h[9], f[9], p[9];
#pragma unroll
        for (int m = 0; m < 9; m++) { h[m] = 0; f[m] = 0; p[m] = 0; }
        const uint32_t gpac = x.T[i];
        uint32_t gb[9];
#pragma unroll
        for (int m = 1; m < 9; m++) {
            gb[m] = (gpac >> (32 - 4 * m)) & 15u;
            if (LUT) gb[m] = score_lut(c, nval, census, gb[m]);
        }
        const int32_t gidx = i << 3;
        const int32_t j0 = max(0, i - other + 1);                                          // :83-85
        const int32_t last_tile = min(k_band_width + i, QW);
        uint32_t nxt[8];
        if (j0 < last_tile) {
#pragma unroll
            for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[(j0 * 8 + kk) * WAVE];
        }
        for (int32_t j = j0; j < last_tile; j++) {
            const uint32_t rpac = x.Q[j];
            uint32_t cur[8];
#pragma unroll
            for (int kk = 0; kk < 8; kk++) cur[kk] = nxt[kk];
            if (j + 1 < last_tile) {
#pragma unroll
                for (int kk = 0; kk < 8; kk++) nxt[kk] = bd[((j + 1) * 8 + kk) * WAVE];
            }
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int32_t ridx = j * 8 + kk;
                const uint32_t rbase = (rpac >> (28 - 4 * kk)) & 15u;
                const uint32_t qsh = (rbase & 7u) << 2;
                const uint32_t hd = cur[kk];
                h[0] = sx16(hd);
                int32_t e = sx16(hd >> 16);
#pragma unroll
                for (int m = 1; m < 9; m++) {                                              // :100-115
                    // E and F held as max(E, 0) and max(F, 0) (saturating
                    // extensions; the opens come from H >= 0): H = max(diagonal,
                    // F, E, 0) needs no max with 0, and only H is reported
                    const int32_t s = LUT ? __builtin_amdgcn_sbfe(gb[m], qsh, 4) : sub_local(c, nval, rbase, gb[m]);
                    f[m] = max(h[m] - gapoe, (int32_t)__builtin_elementwise_sub_sat((uint32_t)f[m], (uint32_t)gape));
                    int32_t hm = p[m] + s;
                    hm = max(hm, f[m]);
                    e = max(h[m - 1] - gapoe, (int32_t)__builtin_elementwise_sub_sat((uint32_t)e, (uint32_t)gape));
                    hm = max(hm, e);
                    maxXY_y = (maxHH < hm) ? gidx + (m - 1) : maxXY_y;
                    maxHH = max(maxHH, hm);
                    p[m] = h[m - 1];
                    h[m] = hm;
                }
                bd[ridx * WAVE] = pk16(h[8], e);
                maxXY_x = (prev_maxHH < maxHH) ? ridx : maxXY_x;
                prev_maxHH = max(maxHH, prev_maxHH);
            }
        }
    }
    d.score[x.pair] = maxHH;
    d.qend[x.pair] = maxXY_x;
    d.tend[x.pair] = maxXY_y;
}

// ------------------------------------------------------------------ KSW -----
// gasal_ksw_kernel<B> (src/kernels/ksw_kernel_template.h:46-199).  The per-pair
// eh[] row lives in device scratch laid out [column][lane] (int2 per entry).
// EH16: H and E of the eh row packed as two uint16 (every KSW value is >= 0 and
// the host checked h0 + match * min(Lq, Lt) <= 65535), halving the row's traffic.
// LUT: scores from the row's table (batch census allows it), else per cell.
// qlds: query words per lane held in LDS ([word][lane], the launch gives
// qlds * 256 bytes; 0: read from global memory).  Every lane re-reads its
// query on each target row; from global memory those scattered 4-byte reads
// miss in an L2 that the rows stream through.
extern __shared__ uint32_t ksw_q[];
// QL: the query words are read from LDS (qlds words per lane), else from
// global memory; separate code so that an LDS read waits on LDS only, not on
// the next block's row records in flight.
template <bool EH16, bool LUT, bool QL>
__global__ __launch_bounds__(64) void k_ksw(G2Dev d, G2Scores c, int32_t w, int32_t zdrop, int32_t end_bonus,
                                            uint32_t qlds) {
    const uint32_t census = d.aux[0];
    if (LUT != (d.fast && census_fast(census))) return;
    LaneCtx x;
    if (!lane_setup(d, x)) return;
    if (QL)
        for (int32_t q = 0; q < x.QW; q++) ksw_q[q * WAVE + x.lane] = x.Q[q];   // QW <= qlds: qlen <= the batch's longest
    const int32_t nval = c.has_np ? -c.np : 0;
    const uint32_t qlen = x.qlen, tlen = x.tlen;
    const uint32_t query_batch_regs = (qlen >> 3) + 1, target_batch_regs = (tlen >> 3) + 1;
    const uint32_t h0 = d.seeds[x.pair];
    const int o_del = c.gapo, o_ins = c.gapo, e_del = c.gape, e_ins = c.gape;
    const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    // The row: [block of 8 columns][lane][8] records, 32 bytes per lane and
    // block with EH16 (64 without), so a wave reads and writes a block of
    // columns in one or two 16-byte accesses per lane (2 or 4 KiB contiguous
    // per wave) instead of one access per column; eh_cols is a multiple of
    // 8.  The row streams through L2 with non-temporal hints: the rows of the
    // resident waves exceed it.
    constexpr int RW = EH16 ? 8 : 16;   // 32-bit words per lane record
    uint32_t *__restrict__ rec = reinterpret_cast<uint32_t *>(d.eh) + ((size_t)x.grp * d.eh_cols * WAVE) * (EH16 ? 1 : 2) +
                                 (size_t)x.lane * RW;
    auto at = [&](int k) -> uint32_t * { return rec + (size_t)(k >> 3) * WAVE * RW + (size_t)(k & 7) * (EH16 ? 1 : 2); };
    auto ld = [&](int k) -> int2 {
        if (EH16) {
            const uint32_t u = __builtin_nontemporal_load(at(k));
            return make_int2((int)(u & 0xFFFFu), (int)(u >> 16));
        }
        const uint64_t u = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(at(k)));
        return make_int2((int)(uint32_t)u, (int)(uint32_t)(u >> 32));
    };
    auto st = [&](int k, int2 v) {
        if (EH16) __builtin_nontemporal_store((uint32_t)v.x | ((uint32_t)v.y << 16), at(k));
        else __builtin_nontemporal_store((uint64_t)(uint32_t)v.x | ((uint64_t)(uint32_t)v.y << 32), reinterpret_cast<uint64_t *>(at(k)));
    };
    auto load_rec = [&](int b, uint32_t *r) {
        const u32x4 *p = reinterpret_cast<const u32x4 *>(rec + (size_t)b * WAVE * RW);
#pragma unroll
        for (int q = 0; q < RW / 4; q++) {
            const u32x4 v = __builtin_nontemporal_load(p + q);
            r[4 * q] = v.x; r[4 * q + 1] = v.y; r[4 * q + 2] = v.z; r[4 * q + 3] = v.w;
        }
    };
    auto store_rec = [&](int b, const uint32_t *r) {
        u32x4 *p = reinterpret_cast<u32x4 *>(rec + (size_t)b * WAVE * RW);
#pragma unroll
        for (int q = 0; q < RW / 4; q++) {
            u32x4 v;
            v.x = r[4 * q]; v.y = r[4 * q + 1]; v.z = r[4 * q + 2]; v.w = r[4 * q + 3];
            __builtin_nontemporal_store(v, p + q);
        }
    };
    for (uint32_t jj = 0; jj <= qlen + 1; jj++) st((int)jj, make_int2(0, 0));
    int j;
    st(0, make_int2((int32_t)h0, 0));                                                      // :78-81
    st(1, make_int2(h0 > (uint32_t)oe_ins ? (int32_t)(h0 - (uint32_t)oe_ins) : 0, 0));
    for (j = 2; j <= (int)qlen && ld(j - 1).x > e_ins; ++j) st(j, make_int2(ld(j - 1).x - e_ins, 0));
    int max = (int32_t)h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1;
    int beg = 0, end = (int)qlen;
    if (w >= 0) {   // band extension: w clamped as bwa's ksw_extend2 does (the reference has no band)
        int mx = (int)((double)((int)qlen * (c.match > 0 ? c.match : 0) + end_bonus - c.gapo) / c.gape + 1.0);
        mx = mx > 1 ? mx : 1;
        w = w < mx ? w : mx;
    }
    // z-drop (extension; 0 in the reference) ends the whole extension as in
    // bwa's ksw_extend2; the m == 0 break keeps the reference's tile-local exit
    bool zstop = false;
    int jlo = 0;   // wave-uniform lower bound of the active lanes' beg, a multiple of 8
    for (uint32_t tt = 0; tt < target_batch_regs && !zstop; tt++) {
        // Every lane still extending is active here (lanes that left the
        // previous tile on m == 0 rejoin), and beg never decreases: move the
        // wave-wide column start up by 8 while no lane's beg lies below the
        // next word.  Bounded: every beg is at most max(Lq, Lt).
        while (__ballot(beg < jlo + 8) == 0) jlo += 8;
        const uint32_t gpac = x.T[tt];
        for (uint32_t tb = 0; tb < 8; tb++) {
            const int i = (int)(tt * 8 + tb);
            if (i >= (int)tlen) break;                                                     // :98-99
            const uint32_t gbase = (gpac >> (32 - (tb + 1) * 4)) & 0x0Fu;
            const uint32_t lut = LUT ? score_lut(c, nval, census, gbase) : 0u;
            int t, f = 0, h1, m = 0, mj = -1;
            if (w >= 0) {
                if (beg < i - w) beg = i - w;
                if (end > i + w + 1) end = i + w + 1;
                if (end > (int)qlen) end = (int)qlen;
            }
            if (beg == 0) {
                h1 = (int32_t)(h0 - (uint32_t)(o_del + e_del * (i + 1)));
                if (h1 < 0) h1 = 0;
            } else
                h1 = 0;
            // j after the tile loops (only the last tile's values matter: :113-152)
            const int T = (int)(qlen >> 3);
            // Columns run from one wave-wide start for every lane, cells before
            // the lane's own beg skipped, so the wave's eh accesses stay on one
            // row (coalesced).
            uint32_t rpac = 0;
            // the row entry of the next column is loaded one step ahead (it was
            // written by the previous row; this row writes it only at its
            // step); a lane loads only its own columns [beg, end)
            // first and last column of [beg, end) whose new entry is nonzero
            // (what :178-183 rescan in memory after the row)
            int fnz = -1, lnz = -1;
            // a block's record is loaded while the previous block computes
            // (an entry is written only at its own column of this row, so an
            // early load reads the previous row's value, as the reference
            // does); columns of a block outside [beg, end) go back unchanged
            uint32_t rc_[RW], rn_[RW];
            if (jlo < end) load_rec(jlo >> 3, rc_);
            // one cell of this row at column jj, entry k of the block's record
            auto cell = [&](int k, int jj) {
                int2 pv = EH16 ? make_int2((int)(rc_[k] & 0xFFFFu), (int)(rc_[k] >> 16))
                               : make_int2((int)rc_[2 * k], (int)rc_[2 * k + 1]);
                const uint32_t rbase = (rpac >> (28 - 4 * k)) & 0x0Fu;
                int h, M = pv.x, e = pv.y;
                pv.x = h1;
                const int32_t s = LUT ? __builtin_amdgcn_sbfe(lut, (rbase & 7u) << 2, 4) : sub_local(c, nval, rbase, gbase);
                M = M ? M + s : 0;
                h = M > e ? M : e;
                h = h > f ? h : f;
                h1 = h;
                mj = m > h ? mj : jj;
                m = m > h ? m : h;
                t = M - oe_del; t = t > 0 ? t : 0;
                e -= e_del; e = e > t ? e : t;
                pv.y = e;
                if (EH16) {
                    rc_[k] = (uint32_t)pv.x | ((uint32_t)pv.y << 16);
                } else {
                    rc_[2 * k] = (uint32_t)pv.x;
                    rc_[2 * k + 1] = (uint32_t)pv.y;
                }
                const bool nz = (pv.x | pv.y) != 0;
                fnz = (nz && fnz < 0) ? jj : fnz;
                lnz = nz ? jj : lnz;
                t = M - oe_ins; t = t > 0 ? t : 0;
                f -= e_ins; f = f > t ? f : t;
            };
            for (int jb = jlo; jb < end; jb += 8) {
                if (jb + 8 < end) load_rec((jb >> 3) + 1, rn_);
                rpac = QL ? ksw_q[(jb >> 3) * WAVE + x.lane] : x.Q[jb >> 3];
                // a block inside every active lane's [beg, end) runs its
                // eight cells without per-cell lane masks
                if (__all(jb >= beg && jb + 8 <= end)) {
#pragma unroll
                    for (int k = 0; k < 8; k++) cell(k, jb + k);
                } else {
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const int jj = jb + k;
                        if (jj >= beg && jj < end) cell(k, jj);
                    }
                }
                store_rec(jb >> 3, rc_);
#pragma unroll
                for (int q = 0; q < RW; q++) rc_[q] = rn_[q];
            }
            {
                int be = beg > end ? beg : end;
                j = be < 8 * T ? 8 * T : be;
                j = j > 8 * T + 7 ? 8 * T + 7 : j;
            }
            st(end, make_int2(h1, 0));                                                     // :153-154
            if (j == (int)qlen) {
                max_ie = gscore > h1 ? max_ie : i;
                gscore = gscore > h1 ? gscore : h1;
            }
            if (m == 0) break;                                                             // :159-160
            if (m > max) {
                max = m; max_i = i; max_j = mj;
            } else if (zdrop > 0) {                                                        // :164-170
                if (i - max_i > mj - max_j) {
                    if (max - m - ((i - max_i) - (mj - max_j)) * e_del > zdrop) zstop = true;
                } else {
                    if (max - m - ((mj - max_j) - (i - max_i)) * e_ins > zdrop) zstop = true;
                }
                if (zstop) break;
            }
            // :178-183 from the row's own values: the new beg is the first
            // nonzero entry of [beg, end) (else end, or beg when beg > end);
            // the backward scan stops at once when end < beg, else at end
            // itself when h1 != 0 (stored there as (h1, 0)), else at the last
            // nonzero entry of the row (else new beg - 1)
            beg = fnz >= 0 ? fnz : (beg < end ? end : beg);
            j = (end < beg || h1 != 0) ? end : (lnz >= beg ? lnz : beg - 1);
            end = j + 2 < (int)qlen ? j + 2 : (int)qlen;
        }
    }
    const uint32_t pr = x.pair;
    if (gscore <= 0 || gscore <= max - end_bonus) {                                        // :188-197 (PEN_CLIP5 there)
        d.score[pr] = max; d.qend[pr] = max_j + 1; d.tend[pr] = max_i + 1;
    } else {
        d.score[pr] = gscore; d.qend[pr] = (int32_t)qlen; d.tend[pr] = max_ie + 1;
    }
}

// --------------------------------------------------------------- get_tb -----
// Recomputes one 8x8 tile's direction words (k_local_fast's sign-bit format)
// from the checkpoints k_local_fast<.., CK> stored: the previous strip's
// bottom row above the tile and the tile's left column state.  Same
// recurrences and score tables (lut: the strip's eight rows, as k_local_fast
// builds them), so the same words.
// top: the row above, entry kk = (H at column kk - 1, E at column kk), so the
// entry left of the tile's first column (the corner) comes with it; zero above
// strip 0.  left: the left column state; lut: the strip's score tables (the
// kernels store checkpoints only when the batch's nibble census allows tables).
// The words hold the reference's nibbles (src/kernels/local_kernel_template.h
// :45-60): [3] F continues (t <= f - extend), [2] E continues, [1:0] source:
// diagonal 0 / 1 by the score's sign, F 3, E 2, ties in that order.
__device__ __forceinline__ void ck_tile(const uint32_t *top, const uint32_t *left, uint32_t qword,
                                        const uint32_t *lut, const G2Scores &c, uint32_t *ch) {
    int32_t h[9], f[9];
#pragma unroll
    for (int m = 1; m < 9; m++) { h[m] = sx16(left[m - 1]); f[m] = (int32_t)left[m - 1] >> 16; }
    const int32_t gapoe = c.gapoe, gape = c.gape;
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
        const uint32_t qsh = kk < 7 ? (qword >> (26 - 4 * kk)) & 0x1Cu : (qword << 2) & 0x1Cu;
        int32_t e = (int32_t)top[kk] >> 16;
        int32_t p = sx16(top[kk]);
        uint32_t dir = 0;
#pragma unroll
        for (int m = 1; m < 9; m++) {
            const int32_t sc = __builtin_amdgcn_sbfe(lut[m - 1], qsh, 4);
            const int32_t tmp = p + sc;
            const int32_t hm = max(max(max(tmp, f[m]), e), 0);
            const int32_t t = tmp - gapoe;
            const int32_t fx = f[m] - gape, ex = e - gape;
            const uint32_t na = (uint32_t)(tmp - hm);                       // sign: not from the diagonal
            dir = __builtin_amdgcn_alignbit(dir, (uint32_t)t + ~(uint32_t)fx, 31);   // sign: fx >= t
            dir = __builtin_amdgcn_alignbit(dir, (uint32_t)t + ~(uint32_t)ex, 31);   // sign: ex >= t
            dir = __builtin_amdgcn_alignbit(dir, na, 31);
            // low bit: from F (hm <= f) when not diagonal, else the mismatch flag (score < 0)
            dir = __builtin_amdgcn_alignbit(dir, (na & ((uint32_t)hm + ~(uint32_t)f[m])) | (~na & (uint32_t)sc), 31);
            f[m] = max(t, fx);
            e = max(t, ex);
            p = h[m];
            h[m] = hm;
        }
        ch[kk] = dir;
    }
}

// Word a of a packed sequence array through a 4-word register cache: the
// aligned 16 bytes holding it are fetched when the walk first needs one of
// them.  The arrays are 256-byte aligned with 4 words of slack past the last
// sequence word, so the aligned read stays inside the allocation.
__device__ __forceinline__ uint32_t seq_word4(const uint32_t *arr, uint32_t a, uint4 &cache, uint32_t &key) {
    if ((a >> 2) != key) {
        key = a >> 2;
        cache = *reinterpret_cast<const uint4 *>(arr + (a & ~3u));
    }
    const uint32_t k = a & 3u;
    return k < 2 ? (k == 0 ? cache.x : cache.y) : (k == 2 ? cache.z : cache.w);
}

// gasal_get_tb<T> (src/kernels/get_tb.h:4-149).  Reads the direction words the
// DP kernel wrote for this pair's wave slot; cells outside the pair's computed
// matrix read as 0.  Writes the pair's CIGAR into its own slot.
// One pair's walk (the body of k_get_tb), for the pair in sorted slot `slot`.
// fused: called by k_local_fast<.., CK> right after the pair's DP, which walks
// exactly the pairs whose tiles it checkpointed; k_get_tb then skips those.
template <bool LOCALT>
__device__ __forceinline__ void tb_walk(const G2Dev &d, const G2Scores &c, uint32_t slot, const uint32_t *lut16,
                                        bool fused) {
    if (slot >= d.n) return;
    const uint32_t pr = d.perm[slot];
    // long pairs (sorted first) ran one wave per pair with per-pair tiles; the
    // lane path's groups count from the first slot after them
    const uint32_t nlong = (LOCALT && d.wf_tb) ? d.aux[2] : 0u;
    const bool wfp = slot < nlong;
    const uint32_t rslot = slot - nlong;
    const uint32_t lane = rslot & (WAVE - 1), grp = rslot / WAVE;
    const uint32_t read_len = d.qlen[pr], ref_len = d.tlen[pr];
    const int32_t QW = dregs(read_len), TW = dregs(ref_len);
    // LOCAL words written by k_local_fast carry the four comparisons as
    // "not" bits; the match/mismatch label comes from the bases.
    // banded runs (k_local_band) store band-compressed tiles per pair
    const bool band = LOCALT && d.band_lo != nullptr;
    const int32_t blo = band ? d.band_lo[pr] : 0, bhi = band ? d.band_hi[pr] : 0;
    const int32_t bbw = band ? (int32_t)d.band_bw[pr] : 0;
    uint32_t bflag = 0;
    if (band && (d.band_flags[pr] & 8u)) { d.n_ops[pr] = 0; return; }
    // GLOBAL words are always in the sign-bit format (k_global)
    const uint32_t census = d.aux[0];
    const bool lanefast = d.fast && d.enc_ok && census_fast(census);
    const bool fastfmt = !LOCALT || band || wfp || lanefast;
    // lane-path tiles of k_local_fast<.., CK> are recomputed from checkpoints
    const bool ck = LOCALT && !band && !wfp && lanefast && d.tbh != nullptr;
    // long pairs of k_local_fast's one-wave blocks with checkpoints (wf_local<.., CK>)
    const bool wck = LOCALT && wfp && d.wf_tbh != nullptr && census_fast(census);
    if (fused ? !ck : (ck && d.tb_fused)) return;   // walked by k_local_fast, or not fused there
    // recomputed tiles hold the reference's nibbles; stored sign-bit words are decoded
    const bool decode = fastfmt && !ck && !wck;
    const uint32_t qw0 = d.qoff[pr] >> 3;
    const uint32_t *Qw = d.pq + qw0, *Tw = d.pt + (d.toff[pr] >> 3);
    uint4 q4 = {0u, 0u, 0u, 0u};   // recompute walks: cached query words
    uint32_t q4_key = 0xFFFFFFFFu;
    const int32_t nval = c.has_np ? -c.np : 0;
    uint32_t ch[8];                 // cached 32-byte chunk: 8 columns of one 8x8 tile
    int32_t ch_key = -1;
    uint32_t lut[8];                // score tables of the strip being recomputed (ck)
    int32_t lut_key = -1;
    uint32_t qword = 0, tword = 0;  // cached sequence words (fast format)
    int32_t q_key = -1, t_key = -1;
    uint8_t *__restrict__ cg = d.cigar + (size_t)pr * d.cigar_cap;
    int i, j, total_score = 0, curr_score = 0;
    if (LOCALT) { i = d.tend[pr]; j = d.qend[pr]; total_score = d.score[pr]; }
    else { i = (int)ref_len; j = (int)read_len; }
    uint32_t prev_op_to_fill = 0;
    const int read_len_8 = 8 * QW;
    int n_ops = 0, op_select = 3, op_shift = 0, count = 0, offset = 0;
    uint32_t op_to_fill;
    // Tile loop (outer) and cell loop (inner): the lanes of a wave fetch or
    // recompute their next tile together, then each walks its cells until the
    // path leaves that tile, so tile work is not repeated once per lane.
    // Cell (i, j) in flat (strip, padded column, row) order; a column index of
    // 8QW (GLOBAL start at Lq == 8QW) is column 0 of the next strip
    // (src/kernels/get_tb.h:50).  Cells outside the pair's strips read as 0.
    auto tile_key = [&](int32_t ii, int32_t jj) -> int32_t {
        const int32_t strip = (ii >> 3) + (jj >= read_len_8 ? 1 : 0);
        const int32_t col = jj >= read_len_8 ? jj - read_len_8 : jj;
        return strip < TW ? strip * QW + (col >> 3) : -2;
    };
    bool done = !(i >= 0 && j >= 0);
    while (!done) {
        const int32_t key = tile_key(i, j);
        if (key >= 0 && key != ch_key) {
            const int32_t strip = key / QW, wv = key - strip * QW;
            const uint32_t *chunk = nullptr;
            if (band) {
                int32_t wlo, whi;
                band_words(strip, blo, bhi, QW, wlo, whi);
                if (wv < wlo || wv >= whi || wv - wlo >= bbw) { bflag |= 2u; break; }   // walk left the band
                if ((wv == wlo && wlo > 0) || (wv == whi - 1 && whi < QW)) bflag |= 1u;
                chunk = d.tb + (d.band_tbo[pr] + (size_t)strip * bbw + (size_t)(wv - wlo)) * 8;
            } else if (wfp && !wck) {
                chunk = d.wf_tb + (d.wf_tbo[pr] + (size_t)strip * QW + (size_t)wv) * 8;
            }
            if (ck || wck) {
                if (strip != lut_key) {   // the walk stays in one strip for several tiles
                    lut_key = strip;
                    const uint32_t gw = Tw[strip];
#pragma unroll
                    for (int m = 0; m < 8; m++) lut[m] = lut16[(gw >> (28 - 4 * m)) & 15u];
                }
                uint32_t top[8], left[8];
                if (ck) {
                    const TbView lv = tb_view(d, grp, lane), hv = tb_view(d, grp, lane, d.tbh);
                    if (strip > 0) hv.load(strip - 1, wv, top);
                    lv.load(strip, wv, left);
                } else {   // per pair [strip][word] x 8, H and E of the same column
                    const uint32_t *lp = d.wf_tb + d.wf_tbo[pr] * 8, *hp = d.wf_tbh + d.wf_tbo[pr] * 8;
                    if (strip > 0) {
                        uint32_t raw[8];
                        tb_load(hp + ((size_t)(strip - 1) * QW + (size_t)wv) * 8, raw);
                        const uint32_t corner = wv > 0 ? hp[((size_t)(strip - 1) * QW + (size_t)(wv - 1)) * 8 + 7] : 0u;
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            top[k] = (raw[k] & 0xFFFF0000u) | ((k == 0 ? corner : raw[k - 1]) & 0xFFFFu);
                    }
                    tb_load(lp + ((size_t)strip * QW + (size_t)wv) * 8, left);
                }
                if (strip == 0) {
#pragma unroll
                    for (int k = 0; k < 8; k++) top[k] = 0u;
                }
                ck_tile(top, left, seq_word4(d.pq, qw0 + (uint32_t)wv, q4, q4_key), lut, c, ch);
            } else if (chunk) {
                tb_load(chunk, ch);
            } else {
                tb_view(d, grp, lane).load(strip, wv, ch);
            }
            ch_key = key;
        }
        if (LOCALT && (ck || wck)) {
            // recomputed tiles: reference nibbles, and a LOCAL walk stays inside
            // the matrix, so a step reads the column's word (re-selected only on
            // a column move) and the tile changes when i >> 3 or j >> 3 does
            auto colword = [&](int32_t jj) -> uint32_t {
                const int c7 = jj & 7;
                return c7 < 4 ? (c7 < 2 ? (c7 == 0 ? ch[0] : ch[1]) : (c7 == 2 ? ch[2] : ch[3]))
                              : (c7 < 6 ? (c7 == 4 ? ch[4] : ch[5]) : (c7 == 6 ? ch[6] : ch[7]));
            };
            uint32_t wcol = colword(j);
            while (true) {
                const uint32_t cell_op = (wcol >> (28 - 4 * (i & 7))) & 15u;
                const uint32_t op = (cell_op >> op_shift) & (uint32_t)op_select;
                op_to_fill = op == 0 || op_select == 3 ? op : (uint32_t)op_shift;
                op_select = op == 0 || (op == 1 && op_select == 3) ? 3 : 1;
                op_shift = op == 0 || (op == 1 && op_select == 3) ? 0 : ((op == 2 || op == 3) ? (int)op : op_shift);
                if (count < 63 && op_to_fill == prev_op_to_fill) {
                    count++;
                } else {
                    if (count > 0) { cg[offset++] = (uint8_t)(prev_op_to_fill | (uint32_t)(count << 2)); n_ops++; }
                    count = 1;
                }
                curr_score += ((op_to_fill == 2 || op_to_fill == 3) && prev_op_to_fill != op_to_fill) ? -c.gapoe
                            : ((op_to_fill == 2 || op_to_fill == 3) ? -c.gape : (op_to_fill == 1 ? -c.mismatch : c.match));
                if (curr_score == total_score) { done = true; break; }
                prev_op_to_fill = op_to_fill;
                const int32_t i0 = i, j0 = j;
                i = op_to_fill != 3 ? i - 1 : i;   // ops 0, 1, 2 move up
                j = op_to_fill != 2 ? j - 1 : j;   // ops 0, 1, 3 move left
                if (!(i >= 0 && j >= 0)) { done = true; break; }
                if (((i0 ^ i) | (j0 ^ j)) & ~7) break;
                if (j != j0) wcol = colword(j);
            }
            continue;
        }
        while (true) {
            const int32_t strip = (i >> 3) + (j >= read_len_8 ? 1 : 0);
            const int32_t col = j >= read_len_8 ? j - read_len_8 : j;
            uint32_t cell_op = 0;
            if (strip < TW) {
                const int c7 = col & 7;
                const uint32_t word = c7 < 4 ? (c7 < 2 ? (c7 == 0 ? ch[0] : ch[1]) : (c7 == 2 ? ch[2] : ch[3]))
                                             : (c7 < 6 ? (c7 == 4 ? ch[4] : ch[5]) : (c7 == 6 ? ch[6] : ch[7]));
                cell_op = (word >> (28 - 4 * (i & 7))) & 15u;
                if (decode) {
                    // the cell the tile word belongs to: column col, row 8 strip + (i & 7)
                    const uint32_t nA = (cell_op >> 3) & 1u, nB = (cell_op >> 2) & 1u;
                    const uint32_t nC3 = (cell_op >> 1) & 1u, nC2 = cell_op & 1u;
                    if ((col >> 3) != q_key) { q_key = col >> 3; qword = Qw[q_key]; }
                    if (strip != t_key) { t_key = strip; tword = Tw[t_key]; }
                    const uint32_t qb = (qword >> (28 - 4 * (col & 7))) & 15u;
                    const uint32_t tb = (tword >> (28 - 4 * (i & 7))) & 15u;
                    const int32_t sub = LOCALT ? sub_local(c, nval, qb, tb) : sub_global(c, qb, tb);
                    const uint32_t mx = sub < 0 ? 1u : 0u;
                    const uint32_t src = nA == 0 ? mx : (nB == 0 ? 3u : 2u);
                    cell_op = src | ((nC3 ^ 1u) << 3) | ((nC2 ^ 1u) << 2);
                }
            }
            const uint32_t op = (cell_op >> op_shift) & (uint32_t)op_select;
            op_to_fill = op == 0 || op_select == 3 ? op : (uint32_t)op_shift;
            op_select = op == 0 || (op == 1 && op_select == 3) ? 3 : 1;
            op_shift = op == 0 || (op == 1 && op_select == 3) ? 0 : ((op == 2 || op == 3) ? (int)op : op_shift);
            if (count < 63 && op_to_fill == prev_op_to_fill) {
                count++;
            } else {
                if (count > 0) { cg[offset++] = (uint8_t)(prev_op_to_fill | (uint32_t)(count << 2)); n_ops++; }
                count = 1;
            }
            if (LOCALT) {
                curr_score += ((op_to_fill == 2 || op_to_fill == 3) && prev_op_to_fill != op_to_fill) ? -c.gapoe
                            : ((op_to_fill == 2 || op_to_fill == 3) ? -c.gape : (op_to_fill == 1 ? -c.mismatch : c.match));
                if (curr_score == total_score) { done = true; break; }
            }
            prev_op_to_fill = op_to_fill;
            i = op_to_fill == 0 || op_to_fill == 1 || op_to_fill == 2 ? i - 1 : i;
            j = op_to_fill == 0 || op_to_fill == 1 || op_to_fill == 3 ? j - 1 : j;
            if (!(i >= 0 && j >= 0)) { done = true; break; }
            if (tile_key(i, j) != key) break;
        }
    }
    cg[offset++] = (uint8_t)(prev_op_to_fill | (uint32_t)(count << 2));
    n_ops++;
    if (!LOCALT) {
        while (i >= 0) { uint32_t rc = (i + 1) <= 63 ? (uint32_t)(i + 1) : 63u; cg[offset++] = (uint8_t)(2u | (rc << 2)); n_ops++; i -= 63; }
        while (j >= 0) { uint32_t rc = (j + 1) <= 63 ? (uint32_t)(j + 1) : 63u; cg[offset++] = (uint8_t)(3u | (rc << 2)); n_ops++; j -= 63; }
    }
    if (LOCALT) { d.tstart[pr] = i; d.qstart[pr] = j; }
    d.n_ops[pr] = (uint32_t)n_ops;
    if (band && bflag) d.band_flags[pr] |= (uint8_t)bflag;
}

template <bool LOCALT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_get_tb(G2Dev d, G2Scores c) {
    // score table of each target nibble for the tile recompute (one per block,
    // built before any thread leaves)
    __shared__ uint32_t lut16[16];
    if (LOCALT) {
        if (threadIdx.x < 16) lut16[threadIdx.x] = score_lut(c, c.has_np ? -c.np : 0, d.aux[0], threadIdx.x);
        __syncthreads();
    }
    tb_walk<LOCALT>(d, c, blockIdx.x * blockDim.x + threadIdx.x, lut16, false);
}

// Copies each pair's CIGAR to the reference's layout (bytes at the pair's query
// offset, src/gasal_align.h:52), clipped to its padded query length.
// 16 threads per pair, 8-byte copies (slots are 8-byte aligned and the copy
// stays inside the pair's padded query slot).
__global__ __launch_bounds__(256) void k_cigar_qlayout(G2Dev d) {
    const uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 4, t = threadIdx.x & 15u;
    if (k >= d.n) return;
    const uint32_t nops = d.n_ops[k];
    const uint32_t slot = 8u * (uint32_t)dregs(d.qlen[k]);
    const uint32_t nb = nops < slot ? nops : slot;
    const uint2 *src = reinterpret_cast<const uint2 *>(d.cigar + (size_t)k * d.cigar_cap);
    uint2 *dst = reinterpret_cast<uint2 *>(d.cigar_qlayout + d.qoff[k]);
    for (uint32_t w = t; w < (nb + 7) / 8; w += 16) dst[w] = src[w];
}

// WITH_TB for algorithms without a traceback kernel: the reference copies the
// query bytes as "CIGAR" and the query lengths as n_cigar_ops
// (src/gasal_align.cu:280-283).
__global__ __launch_bounds__(256) void k_no_tb(G2Dev d) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n) return;
    uint32_t ql = d.qlen[k];
    d.n_ops[k] = ql;
    uint8_t *dst = d.cigar + (size_t)k * d.cigar_cap;
    uint64_t off = d.qoff[k];
    for (uint32_t b = 0; b < ql && b < d.cigar_cap; b++) dst[b] = (off + b < d.q_bytes) ? d.uq[off + b] : 0;
}

__global__ __launch_bounds__(256) void k_zero_i32(int32_t *p, uint32_t n) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// One thread per pair of a sub-batch; CIGAR slots copied as 4-byte words
// (slot sizes are multiples of 16 bytes).
__global__ __launch_bounds__(256) void k_sub_scatter(G2Dev sb, G2Dev mn, int tb) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= sb.n) return;
    const uint32_t dst = sb.sub_idx[k];
    mn.score[dst] = sb.score[k];
    mn.qend[dst] = sb.qend[k];
    mn.tend[dst] = sb.tend[k];
    mn.qstart[dst] = sb.qstart[k];
    mn.tstart[dst] = sb.tstart[k];
    if (tb) {
        const uint32_t nops = sb.n_ops[k];
        mn.n_ops[dst] = nops;
        const uint32_t nb = nops < sb.cigar_cap ? nops : sb.cigar_cap;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(sb.cigar + (size_t)k * sb.cigar_cap);
        uint32_t *out = reinterpret_cast<uint32_t *>(mn.cigar + (size_t)dst * mn.cigar_cap);
        const uint32_t nw = (nb + 3) / 4 < mn.cigar_cap / 4 ? (nb + 3) / 4 : mn.cigar_cap / 4;
        for (uint32_t w = 0; w < nw; w++) out[w] = src[w];
    }
}

int g_dispatch = 0;

}  // namespace

hipError_t g2_scatter_subset(const G2Dev &sub, const G2Dev &main, bool tb, hipStream_t stream) {
    if (sub.n == 0) return hipSuccess;
    k_sub_scatter<<<(sub.n + 255) / 256, 256, 0, stream>>>(sub, main, tb ? 1 : 0);
    return hipGetLastError();
}

void g2_set_dispatch(int policy) { g_dispatch = policy; }
int g2_get_dispatch() { return g_dispatch; }

#define G2_LAUNCH_CHECK()                        \
    do {                                         \
        hipError_t e_ = hipGetLastError();       \
        if (e_ != hipSuccess) return e_;         \
    } while (0)

template <int START, bool SECOND, bool LUT>
static void launch_semi_ht(int head, int tail, dim3 g, dim3 b, hipStream_t s, const G2Dev &d, const G2Scores &c) {
#define G2_SEMI_T(H)                                                                                         \
    switch (tail) {                                                                                          \
        case D_NONE: k_semi<START, SECOND, H, D_NONE, LUT><<<g, b, 0, s>>>(d, c); break;                     \
        case D_QUERY: k_semi<START, SECOND, H, D_QUERY, LUT><<<g, b, 0, s>>>(d, c); break;                   \
        case D_TARGET: k_semi<START, SECOND, H, D_TARGET, LUT><<<g, b, 0, s>>>(d, c); break;                 \
        default: k_semi<START, SECOND, H, D_BOTH, LUT><<<g, b, 0, s>>>(d, c); break;                         \
    }
    switch (head) {
        case D_NONE: G2_SEMI_T(D_NONE); break;
        case D_QUERY: G2_SEMI_T(D_QUERY); break;
        case D_TARGET: G2_SEMI_T(D_TARGET); break;
        default: G2_SEMI_T(D_BOTH); break;
    }
#undef G2_SEMI_T
}

hipError_t g2_run(const G2Scores &sc, const G2Opts &op, const G2Dev &d, hipStream_t s, G2Timing *tm) {
    const bool timing = tm && tm->enabled;
    if (timing) (void)hipEventRecord(tm->ev[0], s);
    // 1. pack (src/gasal_align.cu:190-201)
    if (!op.is_packed && !op.skip_prep) {
        auto packside = [&](const uint8_t *u, uint32_t *p, uint32_t bytes) {
            uint32_t n16 = bytes / 16;
            uint32_t tail = (bytes % 16) ? 1u : 0u;
            uint32_t blocks = (n16 + 255) / 256;
            if (blocks > 8192) blocks = 8192;
            if (blocks == 0) blocks = 1;
            k_pack<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint4 *>(u), reinterpret_cast<uint2 *>(p), n16,
                                          reinterpret_cast<const uint2 *>(u + (size_t)n16 * 16), p + (size_t)n16 * 2, tail);
        };
        packside(d.uq, d.pq, d.q_bytes);
        packside(d.ut, d.pt, d.t_bytes);
        G2_LAUNCH_CHECK();
    }
    if (timing) (void)hipEventRecord(tm->ev[1], s);
    const uint32_t nb256 = (d.n + 255) / 256;
    // 2. reverse / complement (src/gasal_align.cu:232-246)
    if (op.is_rc && !op.skip_prep) {
        k_rc<<<(d.n + RC_WAVES * WAVE - 1) / (RC_WAVES * WAVE), RC_WAVES * WAVE, 0, s>>>(d, sc.n_code);
        G2_LAUNCH_CHECK();
    }
    if (op.prep_only) return hipSuccess;
    if (timing) (void)hipEventRecord(tm->ev[2], s);
    // aux[0] nibble census over the packed batch for the score-table kernels;
    // aux[1..2] long-pair queue head and count
    k_zero_u32<<<1, 64, 0, s>>>(d.aux, 4);
    if (d.fast) {
        // packed words of the batch: the byte counts count bases (8 per word) with or without isPacked
        const uint32_t qw = d.q_bytes / 8;
        const uint32_t tw = d.t_bytes / 8;
        uint32_t cb = (qw + tw + 255) / 256;
        k_census<<<cb > 2048 ? 2048 : (cb ? cb : 1), 256, 0, s>>>(d.pq, qw, d.pt, tw, d.aux);   // threads 0..2 take the tail words
    }
    G2_LAUNCH_CHECK();
    // 3. order pairs by shape (largest first) so a wave's lanes share loop bounds;
    // long pairs for the one-wave-per-pair path sort ahead of all others
    k_zero_u32<<<G2_SHAPE_BUCKETS / 256, 256, 0, s>>>(d.hist, G2_SHAPE_BUCKETS);
    const uint32_t nkh = (d.n + KH_CHUNK - 1) / KH_CHUNK;
    k_keys<<<nkh, KH_THREADS, 0, s>>>(d);
    k_scan_desc<<<1, 1024, 0, s>>>(d.hist);
    k_scatter<<<nkh, KH_THREADS, 0, s>>>(d);
    if (d.qpan) k_panel<<<(d.n + WAVE - 1) / WAVE, 256, 0, s>>>(d);
    G2_LAUNCH_CHECK();
    // 4. DP (dispatch table src/gasal_align.h:7-107)
    const uint32_t groups = (d.n + WAVE - 1) / WAVE;
    const dim3 grid(groups), blk(WAVE);
    const bool tb = op.start == S_WITH_TB;
    const bool second = op.second != 0;
    switch (op.banded ? -1 : op.algo) {
        case -1:   // per-pair diagonal bands (score tables only; the caller checks d.fast)
            if (d.band_wres) k_local_band<false, true><<<grid, blk, 0, s>>>(d, sc);
            if (tb) k_local_band<true, false><<<grid, blk, 0, s>>>(d, sc);
            else k_local_band<false, false><<<grid, blk, 0, s>>>(d, sc);
            break;
        case A_LOCAL:
            if (d.fast && d.pk2) {   // two pairs per lane: half the lane-path blocks
                const dim3 g2((groups + 1) / 2), g2w((groups + 1) / 2 + d.wf_blocks);
                if (op.start == S_WITH_TB) {
                    if (second) k_local_pk2<S_WITH_TB, true, false><<<g2, blk, 0, s>>>(d, sc);
                    else if (d.wf_blocks) k_local_pk2<S_WITH_TB, false, true><<<g2w, blk, 0, s>>>(d, sc);
                    else k_local_pk2<S_WITH_TB, false, false><<<g2, blk, 0, s>>>(d, sc);
                } else if (op.start == S_WITH_START) {
                    if (second) k_local_pk2<S_WITH_START, true, false><<<g2, blk, 0, s>>>(d, sc);
                    else k_local_pk2<S_WITH_START, false, false><<<g2, blk, 0, s>>>(d, sc);
                } else {
                    if (second) k_local_pk2<S_WITHOUT_START, true, false><<<g2, blk, 0, s>>>(d, sc);
                    else if (d.wf_blocks) k_local_pk2<S_WITHOUT_START, false, true><<<g2w, blk, 0, s>>>(d, sc);
                    else k_local_pk2<S_WITHOUT_START, false, false><<<g2, blk, 0, s>>>(d, sc);
                }
            } else if (d.fast) {
                if (op.start == S_WITH_TB) {
                    if (d.tbh) {
                        if (second) k_local_fast<S_WITH_TB, true, true, false><<<grid, blk, 0, s>>>(d, sc);
                        else if (d.wf_blocks) k_local_fast<S_WITH_TB, false, true><<<dim3(groups + d.wf_blocks), blk, 0, s>>>(d, sc);
                        else k_local_fast<S_WITH_TB, false, true, false><<<grid, blk, 0, s>>>(d, sc);
                    } else if (second) k_local_fast<S_WITH_TB, true, false, false><<<grid, blk, 0, s>>>(d, sc);
                    else if (d.wf_blocks) k_local_fast<S_WITH_TB, false><<<dim3(groups + d.wf_blocks), blk, 0, s>>>(d, sc);
                    else k_local_fast<S_WITH_TB, false, false, false><<<grid, blk, 0, s>>>(d, sc);
                } else if (op.start == S_WITH_START) {
                    if (second) k_local_fast<S_WITH_START, true, false, false><<<grid, blk, 0, s>>>(d, sc);
                    else k_local_fast<S_WITH_START, false, false, false><<<grid, blk, 0, s>>>(d, sc);
                } else {
                    if (second) k_local_fast<S_WITHOUT_START, true, false, false><<<grid, blk, 0, s>>>(d, sc);
                    else if (d.wf_blocks) k_local_fast<S_WITHOUT_START, false><<<dim3(groups + d.wf_blocks), blk, 0, s>>>(d, sc);
                    else k_local_fast<S_WITHOUT_START, false, false, false><<<grid, blk, 0, s>>>(d, sc);
                }
            }
            if (op.start == S_WITH_TB) {
                if (second) k_local<S_WITH_TB, true><<<grid, blk, 0, s>>>(d, sc);
                else k_local<S_WITH_TB, false><<<grid, blk, 0, s>>>(d, sc);
            } else if (op.start == S_WITH_START) {
                if (second) k_local<S_WITH_START, true><<<grid, blk, 0, s>>>(d, sc);
                else k_local<S_WITH_START, false><<<grid, blk, 0, s>>>(d, sc);
            } else {
                if (second) k_local<S_WITHOUT_START, true><<<grid, blk, 0, s>>>(d, sc);
                else k_local<S_WITHOUT_START, false><<<grid, blk, 0, s>>>(d, sc);
            }
            break;
        case A_GLOBAL:
            if (d.fast) {
                if (tb) k_global<true, true><<<grid, blk, 0, s>>>(d, sc);
                else k_global<false, true><<<grid, blk, 0, s>>>(d, sc);
            }
            if (tb) k_global<true, false><<<grid, blk, 0, s>>>(d, sc);
            else k_global<false, false><<<grid, blk, 0, s>>>(d, sc);
            break;
        case A_SEMI_GLOBAL:
            if (op.start == S_WITH_START) {
                if (d.fast) {
                    if (second) launch_semi_ht<S_WITH_START, true, true>(op.head, op.tail, grid, blk, s, d, sc);
                    else launch_semi_ht<S_WITH_START, false, true>(op.head, op.tail, grid, blk, s, d, sc);
                }
                if (second) launch_semi_ht<S_WITH_START, true, false>(op.head, op.tail, grid, blk, s, d, sc);
                else launch_semi_ht<S_WITH_START, false, false>(op.head, op.tail, grid, blk, s, d, sc);
            } else {
                if (d.fast) {
                    if (second) launch_semi_ht<S_WITHOUT_START, true, true>(op.head, op.tail, grid, blk, s, d, sc);
                    else launch_semi_ht<S_WITHOUT_START, false, true>(op.head, op.tail, grid, blk, s, d, sc);
                }
                if (second) launch_semi_ht<S_WITHOUT_START, true, false>(op.head, op.tail, grid, blk, s, d, sc);
                else launch_semi_ht<S_WITHOUT_START, false, false>(op.head, op.tail, grid, blk, s, d, sc);
            }
            break;
        case A_BANDED:
            if (d.fast) k_banded<true><<<grid, blk, 0, s>>>(d, sc, op.k_band >> 3);          // src/gasal_align.h:80
            k_banded<false><<<grid, blk, 0, s>>>(d, sc, op.k_band >> 3);
            break;
        case A_KSW:
        {
            // query words in LDS when the batch's longest query is at most
            // 1024 bases (32 KiB per one-wave block)
            const uint32_t qw = (d.eh_cols - 2 + 7) / 8;
            const uint32_t qlds = qw <= 128 ? qw : 0;
            const size_t lds = (size_t)qlds * WAVE * 4;
            if (d.ksw16) {
                if (qlds) {
                    k_ksw<true, true, true><<<grid, blk, lds, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                    k_ksw<true, false, true><<<grid, blk, lds, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                } else {
                    k_ksw<true, true, false><<<grid, blk, 0, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                    k_ksw<true, false, false><<<grid, blk, 0, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                }
            } else {
                if (qlds) {
                    k_ksw<false, true, true><<<grid, blk, lds, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                    k_ksw<false, false, true><<<grid, blk, lds, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                } else {
                    k_ksw<false, true, false><<<grid, blk, 0, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                    k_ksw<false, false, false><<<grid, blk, 0, s>>>(d, sc, op.ksw_w, op.ksw_zdrop, op.ksw_eb, qlds);
                }
            }
        }
            break;
        default:
            break;  // UNKNOWN / MICROLOCAL: nothing launched (src/gasal_align.cu:20-21)
    }
    G2_LAUNCH_CHECK();
    // second-best arrays of algorithms that never write them are defined as 0
    if (second && !(op.algo == A_LOCAL || op.algo == A_SEMI_GLOBAL)) {
        k_zero_i32<<<nb256, 256, 0, s>>>(d.s_score, d.n);
        k_zero_i32<<<nb256, 256, 0, s>>>(d.s_qend, d.n);
        k_zero_i32<<<nb256, 256, 0, s>>>(d.s_tend, d.n);
        G2_LAUNCH_CHECK();
    }
    if (timing) (void)hipEventRecord(tm->ev[3], s);
    // 5. traceback (src/gasal_align.h:45-53,60-68)
    if (tb) {
        if (op.algo == A_LOCAL) k_get_tb<true><<<grid, blk, 0, s>>>(d, sc);
        else if (op.algo == A_GLOBAL) k_get_tb<false><<<grid, blk, 0, s>>>(d, sc);
        else k_no_tb<<<nb256, 256, 0, s>>>(d);
        G2_LAUNCH_CHECK();
        if (d.cigar_qlayout) {
            k_cigar_qlayout<<<(d.n + 15) / 16, 256, 0, s>>>(d);
            G2_LAUNCH_CHECK();
        }
    }
    if (timing) {
        (void)hipEventRecord(tm->ev[4], s);
        (void)hipEventRecord(tm->ev[5], s);
    }
    return hipSuccess;
}
