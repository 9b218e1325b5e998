// noalloc.hip -- gasal2_aln_noalloc: the device program of gasal2_run_resident
// (pack, reverse/complement, shape sort, DP) on caller-owned device inputs and
// outputs, with every scratch buffer carved from one file-scope device array.
// It calls no allocator, so it can run where hipMalloc cannot (under HIP graph
// capture, in serving loops, in harnesses that forbid allocation).
#include "gasal_header.h"
#include "engine.h"

#include <mutex>
#include <stdarg.h>

namespace {

// The scratch of every call on a device: packed words, sorted order, shape
// histogram, boundary rows and the outputs the caller does not ask for.
// Each device holds its own copy until the process ends; calls that overlap
// on one device race on it.
constexpr size_t NOALLOC_SCRATCH_BYTES = 32u << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t *g_scratch_ptr[MAX_DEVICES] = {};   // device address of `scratch` on each device

constexpr uint32_t AUX_BAD = 16;   // aux word counting rejected pairs (g2_run uses words 0..3)

// Copies offsets, lengths and operations into the scratch; a pair whose
// sequence does not lie inside its batch at a multiple of 8 bases, or is
// longer than the declared maximum, is replaced by a one-base pair on the
// spare word past the batch (no operation; its outputs are undefined) and
// counted in bad[0].  The kernels then never read past the batch or the
// boundary rows, and no other pair's words are touched.
__global__ __launch_bounds__(256) void k_noalloc_pairs(const uint32_t *qoff, const uint32_t *toff, const uint32_t *qlen,
                                                       const uint32_t *tlen, const uint8_t *qop, const uint8_t *top,
                                                       uint32_t n, uint32_t q_bases, uint32_t t_bases, uint32_t max_q,
                                                       uint32_t max_t, uint32_t *oqoff, uint32_t *otoff, uint32_t *oqlen,
                                                       uint32_t *otlen, uint8_t *oqop, uint8_t *otop, uint32_t *bad) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t qo = qoff[k], to = toff[k], ql = qlen[k], tl = tlen[k];
    const uint64_t qe = (uint64_t)qo + (((uint64_t)ql + 7) & ~7ull), te = (uint64_t)to + (((uint64_t)tl + 7) & ~7ull);
    const bool ok = !(qo & 7u) && !(to & 7u) && qe <= q_bases && te <= t_bases && ql <= max_q && tl <= max_t;
    oqoff[k] = ok ? qo : q_bases;
    otoff[k] = ok ? to : t_bases;
    oqlen[k] = ok ? ql : 1u;
    otlen[k] = ok ? tl : 1u;
    if (qop) {
        oqop[k] = ok ? qop[k] : (uint8_t)0;
        otop[k] = ok ? top[k] : (uint8_t)0;
    }
    if (!ok && bad) atomicAdd(bad, 1u);
}

__global__ __launch_bounds__(256) void k_noalloc_copy(const uint32_t *in, uint32_t *out, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = in[k];
}

__global__ void k_noalloc_zero(uint32_t *p, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0u;
}

int32_t fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    return -1;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// Returns 0 when the program was enqueued on `stream`, -1 (with a message on
// stderr) when the call is rejected: a mode it does not run (WITH_TB, KSW,
// BANDED), a batch whose scratch would exceed NOALLOC_SCRATCH_BYTES, or a
// HIP error.  Results are those of gasal2_run_resident on the same batch.
extern "C" int32_t gasal2_aln_noalloc(const void *query_batch, const void *target_batch, uint32_t query_batch_bytes,
                                      uint32_t target_batch_bytes, const uint32_t *query_offsets,
                                      const uint32_t *target_offsets, const uint32_t *query_lens,
                                      const uint32_t *target_lens, const uint8_t *query_ops, const uint8_t *target_ops,
                                      uint32_t n_alns, uint32_t max_query_len, uint32_t max_target_len, int32_t *score,
                                      int32_t *query_end, int32_t *target_end, int32_t *query_start,
                                      int32_t *target_start, int32_t *second_score, int32_t *second_query_end,
                                      int32_t *second_target_end, uint32_t *bad_pairs, Parameters *params,
                                      hipStream_t stream) {
    if (!params) return fail("[GASAL ERROR:] gasal2_aln_noalloc: no Parameters\n");
    if (params->algo != LOCAL && params->algo != GLOBAL && params->algo != SEMI_GLOBAL)
        return fail("[GASAL ERROR:] gasal2_aln_noalloc runs LOCAL, GLOBAL and SEMI_GLOBAL only\n");
    if (params->start_pos == WITH_TB) return fail("[GASAL ERROR:] gasal2_aln_noalloc does not run WITH_TB\n");
    if (n_alns == 0 || query_batch_bytes == 0 || target_batch_bytes == 0 || (query_batch_bytes % 8) ||
        (target_batch_bytes % 8))
        return fail("[GASAL ERROR:] gasal2_aln_noalloc: empty batch or byte counts not multiples of 8\n");
    if (!query_batch || !target_batch || !query_offsets || !target_offsets || !query_lens || !target_lens || !score ||
        !query_end || !target_end || (params->isReverseComplement && (!query_ops || !target_ops)))
        return fail("[GASAL ERROR:] gasal2_aln_noalloc: a required device array is NULL\n");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return fail("[GASAL ERROR:] gasal2_aln_noalloc: no usable current device\n");
    uint8_t *base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch_ptr[dev]) {   // once per device, with that device current
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail("[GASAL CUDA ERROR:] %s(CUDA error no.=%d): scratch address\n", hipGetErrorString(e), (int)e);
            g_scratch_ptr[dev] = (uint8_t *)p;
        }
        base = g_scratch_ptr[dev];
    }
    const uint32_t n = n_alns;
    const uint32_t qw = g2_regs(max_query_len), tw = g2_regs(max_target_len);
    if (qw == 0 || tw == 0) return fail("[GASAL ERROR:] gasal2_aln_noalloc: zero maximum length\n");
    const uint32_t groups = (n + 63) / 64;
    const uint32_t bnd_cols = 8 * qw + 8;
    // layout (bytes, 256-aligned pieces)
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += up256(bytes); return o; };
    const size_t o_pq = take(((size_t)query_batch_bytes / 8 + 4) * 4);
    const size_t o_pt = take(((size_t)target_batch_bytes / 8 + 4) * 4);
    const size_t o_qoff = take((size_t)n * 4), o_toff = take((size_t)n * 4);
    const size_t o_qlen = take((size_t)n * 4), o_tlen = take((size_t)n * 4);
    const size_t o_qop = take(n), o_top = take(n);
    const size_t o_keys = take((size_t)n * 4), o_perm = take((size_t)n * 4);
    const size_t o_hist = take((size_t)G2_SHAPE_BUCKETS * 4);
    const size_t o_aux = take(256 + 4096);
    const size_t o_bnd = take((size_t)groups * bnd_cols * 64 * 4);
    const size_t o_dummy = take((size_t)n * 4 * 5);
    if (off > NOALLOC_SCRATCH_BYTES)
        return fail("[GASAL ERROR:] gasal2_aln_noalloc: the batch needs %zu bytes of scratch, more than %zu\n", off,
                    (size_t)NOALLOC_SCRATCH_BYTES);

    G2Dev d;
    memset(&d, 0, sizeof(d));
    uint32_t *const aux = (uint32_t *)(base + o_aux);
    d.pq = (uint32_t *)(base + o_pq);
    d.pt = (uint32_t *)(base + o_pt);
    d.uq = (const uint8_t *)query_batch;
    d.ut = (const uint8_t *)target_batch;
    d.qoff = (const uint32_t *)(base + o_qoff);
    d.toff = (const uint32_t *)(base + o_toff);
    d.qlen = (const uint32_t *)(base + o_qlen);
    d.tlen = (const uint32_t *)(base + o_tlen);
    d.qop = params->isReverseComplement ? (const uint8_t *)(base + o_qop) : nullptr;
    d.top = params->isReverseComplement ? (const uint8_t *)(base + o_top) : nullptr;
    d.q_bytes = query_batch_bytes;
    d.t_bytes = target_batch_bytes;
    d.n = n;
    d.max_qw = qw;
    d.max_tw = tw;
    int32_t *dummy = (int32_t *)(base + o_dummy);
    d.score = score;
    d.qend = query_end;
    d.tend = target_end;
    d.qstart = query_start ? query_start : dummy;
    d.tstart = target_start ? target_start : dummy + n;
    d.s_score = second_score ? second_score : dummy + 2 * (size_t)n;
    d.s_qend = second_query_end ? second_query_end : dummy + 3 * (size_t)n;
    d.s_tend = second_target_end ? second_target_end : dummy + 4 * (size_t)n;
    d.keys = (uint32_t *)(base + o_keys);
    d.perm = (uint32_t *)(base + o_perm);
    d.hist = (uint32_t *)(base + o_hist);
    d.bnd = (uint32_t *)(base + o_bnd);
    d.bnd_cols = bnd_cols;
    d.aux = aux;
    d.sink = aux + 64;

    G2Opts op;
    memset(&op, 0, sizeof(op));
    op.algo = params->algo;
    op.start = params->start_pos;
    op.second = params->secondBest ? 1 : 0;
    op.head = params->semiglobal_skipping_head;
    op.tail = params->semiglobal_skipping_tail;
    op.k_band = params->k_band;
    op.is_rc = params->isReverseComplement ? 1 : 0;
    op.ksw_w = -1;
    op.ksw_eb = 5;
    // packed inputs are copied into the scratch, so reverse/complement never
    // rewrites the caller's words
    op.is_packed = 1;
    const G2Scores sc = g2_scores();
    g2_choose_fast(d, op, sc, qw, tw);
    g2_choose_packed(d, op, sc, qw, tw, false);

    const uint32_t nb = (n + 255) / 256;
    k_noalloc_zero<<<1, 64, 0, stream>>>(aux + AUX_BAD, 1);   // rejected-pair counter
    k_noalloc_pairs<<<nb, 256, 0, stream>>>(query_offsets, target_offsets, query_lens, target_lens,
                                            params->isReverseComplement ? query_ops : nullptr, target_ops, n,
                                            query_batch_bytes, target_batch_bytes, max_query_len, max_target_len,
                                            (uint32_t *)d.qoff, (uint32_t *)d.toff, (uint32_t *)d.qlen,
                                            (uint32_t *)d.tlen, (uint8_t *)d.qop, (uint8_t *)d.top, aux + AUX_BAD);
    if (params->isPacked) {   // the caller's packed words: 4 bytes per 8 bases
        k_noalloc_copy<<<1024, 256, 0, stream>>>((const uint32_t *)query_batch, d.pq, query_batch_bytes / 8);
        k_noalloc_copy<<<1024, 256, 0, stream>>>((const uint32_t *)target_batch, d.pt, target_batch_bytes / 8);
    } else {
        op.is_packed = 0;   // k_pack writes the scratch words from the caller's bytes
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = g2_run(sc, op, d, stream, nullptr);
    if (e == hipSuccess && bad_pairs) k_noalloc_copy<<<1, 64, 0, stream>>>(aux + AUX_BAD, bad_pairs, 1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail("[GASAL CUDA ERROR:] %s(CUDA error no.=%d)\n", hipGetErrorString(e), (int)e);
    return 0;
}
