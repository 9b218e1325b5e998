// stream.cpp -- host input path (include/gasal2_stream.h): paired FASTA/FASTQ
// files read into padded host batches, then streamed through several storages
// by one host thread each.  Built only on the public API; the reference does
// the same work inline in its example program (test_prog/test_prog.cpp:79-137
// reading, :237-330 threads and batches; src/host_batch.cpp:79-153 paging).
#include "gasal_header.h"
#include "gasal2_stream.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

int fail(const char *msg, const char *arg = "") {
    fprintf(stderr, "[GASAL ERROR:] %s%s\n", msg, arg);
    return -1;
}

// One file's records, padded with N_CODE to multiples of 8 bases.
struct Side {
    std::vector<uint8_t> bytes;
    std::vector<uint32_t> off, len;
    std::vector<uint8_t> op;
    std::vector<std::string> names;
    bool ok = false;
};

// Buffered line reader over a FILE (lines without their '\n' / "\r\n").
struct Lines {
    FILE *f;
    std::vector<char> buf;
    size_t pos = 0, end = 0;
    bool eof = false;
    explicit Lines(FILE *fp) : f(fp), buf(1 << 22) {}
    bool next(std::string &line) {
        line.clear();
        for (;;) {
            if (pos == end) {
                if (eof) return !line.empty();
                end = fread(buf.data(), 1, buf.size(), f);
                pos = 0;
                if (end == 0) {
                    eof = true;
                    return !line.empty();
                }
            }
            const char *s = buf.data() + pos;
            const char *nl = (const char *)memchr(s, '\n', end - pos);
            if (nl) {
                line.append(s, nl - s);
                pos += (nl - s) + 1;
                if (!line.empty() && line.back() == '\r') line.pop_back();
                return true;
            }
            line.append(s, end - pos);
            pos = end;
        }
    }
};

void close_record(Side &sd, uint8_t n_code) {
    const uint32_t start = sd.off.back();
    const uint32_t l = (uint32_t)(sd.bytes.size() - start);
    sd.len.push_back(l);
    while (sd.bytes.size() % 8) sd.bytes.push_back(n_code);   // src/host_batch.cpp:100-102
}

void read_side(const char *path, bool names, uint8_t n_code, Side &sd) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        fail("cannot open ", path);
        return;
    }
    Lines in(f);
    std::string line;
    static const char marks[] = "></+";   // test_prog/test_prog.cpp:79-137
    bool fastq = false, first = true, open = false;
    while (in.next(line)) {
        if (line.empty()) continue;
        if (first) {
            fastq = line[0] == '@';
            first = false;
        }
        if (fastq) {
            if (line[0] != '@') {
                fail("malformed FASTQ record in ", path);
                fclose(f);
                return;
            }
            std::string seq, plus, qual;
            if (!in.next(seq) || !in.next(plus) || !in.next(qual) || plus.empty() || plus[0] != '+') {
                fail("truncated FASTQ record in ", path);
                fclose(f);
                return;
            }
            sd.off.push_back((uint32_t)sd.bytes.size());
            sd.op.push_back(0);
            if (names) sd.names.push_back(line.substr(1));
            sd.bytes.insert(sd.bytes.end(), seq.begin(), seq.end());
            close_record(sd, n_code);
            continue;
        }
        const char *m = strchr(marks, line[0]);
        if (m && line[0]) {
            if (open) close_record(sd, n_code);
            sd.off.push_back((uint32_t)sd.bytes.size());
            sd.op.push_back((uint8_t)(m - marks));
            if (names) sd.names.push_back(line.substr(1));
            open = true;
        } else if (open) {
            sd.bytes.insert(sd.bytes.end(), line.begin(), line.end());
        } else {
            fail("FASTA sequence before any header in ", path);
            fclose(f);
            return;
        }
        if (sd.bytes.size() > 0xFFFFFFF0ull) {
            fail("batch of more than 4 G bases in ", path);
            fclose(f);
            return;
        }
    }
    if (open) close_record(sd, n_code);
    fclose(f);
    sd.ok = true;
}

// 4-bit packed words of padded bytes (src/kernels/pack_rc_seqs.h:24-31).
void pack_words(const uint8_t *b, uint64_t bytes, uint32_t *w) {
    for (uint64_t i = 0; i < bytes / 8; i++) {
        const uint8_t *p = b + 8 * i;
        uint32_t v = 0;
        for (int k = 0; k < 8; k++) v |= (uint32_t)(p[k] & 15u) << (28 - 4 * k);
        w[i] = v;
    }
}

// memcpy split over a few threads for large spans (a sub-batch's bases are
// hundreds of MB; one thread copies them at a fraction of the host's
// bandwidth)
void copy_wide(void *dst, const void *src, size_t bytes) {
    const size_t piece = 32u << 20;
    const unsigned parts = (unsigned)std::min<size_t>(4, (bytes + piece - 1) / piece);
    if (parts <= 1) {
        memcpy(dst, src, bytes);
        return;
    }
    const size_t step = (bytes / parts + 4095) & ~(size_t)4095;
    std::vector<std::thread> th;
    for (unsigned k = 1; k < parts; k++) {
        const size_t off = k * step;
        if (off >= bytes) break;
        th.emplace_back([=] { memcpy((char *)dst + off, (const char *)src + off, std::min(step, bytes - off)); });
    }
    memcpy(dst, src, std::min(step, bytes));
    for (auto &t : th) t.join();
}

template <typename T>
T *copy_out(const std::vector<T> &v) {
    T *p = (T *)malloc(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (p && !v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" gasal2_pairs_t *gasal2_pairs_read(const char *query_path, const char *target_path, uint32_t flags) {
    if (!query_path || !target_path) {
        fail("gasal2_pairs_read: no file name");
        return nullptr;
    }
    const bool names = (flags & GASAL2_READ_NAMES) != 0;
    const uint8_t n_code = (uint8_t)gasal2_get_n_code();
    Side q, t;
    std::thread tq([&] { read_side(query_path, names, n_code, q); });
    read_side(target_path, names, n_code, t);
    tq.join();
    if (!q.ok || !t.ok) return nullptr;
    if (q.off.size() != t.off.size()) {
        fail("query and target files hold different numbers of sequences");
        return nullptr;
    }
    gasal2_pairs_t *p = (gasal2_pairs_t *)calloc(1, sizeof(gasal2_pairs_t));
    if (!p) return nullptr;
    p->n = (uint32_t)q.off.size();
    p->query_bytes = q.bytes.size();
    p->target_bytes = t.bytes.size();
    p->query = copy_out(q.bytes);
    p->target = copy_out(t.bytes);
    p->query_offsets = copy_out(q.off);
    p->target_offsets = copy_out(t.off);
    p->query_lens = copy_out(q.len);
    p->target_lens = copy_out(t.len);
    p->query_ops = copy_out(q.op);
    p->target_ops = copy_out(t.op);
    if (flags & GASAL2_READ_PACK) {
        p->query_words = (uint32_t *)malloc(std::max<uint64_t>(p->query_bytes / 2, 4));
        p->target_words = (uint32_t *)malloc(std::max<uint64_t>(p->target_bytes / 2, 4));
        if (p->query_words && p->target_words) {
            std::thread pq([&] { pack_words(p->query, p->query_bytes, p->query_words); });
            pack_words(p->target, p->target_bytes, p->target_words);
            pq.join();
        }
    }
    if (names) {
        p->query_names = (char **)calloc(p->n ? p->n : 1, sizeof(char *));
        p->target_names = (char **)calloc(p->n ? p->n : 1, sizeof(char *));
        for (uint32_t k = 0; k < p->n && p->query_names && p->target_names; k++) {
            p->query_names[k] = strdup(q.names[k].c_str());
            p->target_names[k] = strdup(t.names[k].c_str());
        }
    }
    if (!p->query || !p->target || !p->query_offsets || !p->target_offsets || !p->query_lens || !p->target_lens ||
        !p->query_ops || !p->target_ops || ((flags & GASAL2_READ_PACK) && (!p->query_words || !p->target_words))) {
        gasal2_pairs_free(p);
        fail("gasal2_pairs_read: out of host memory");
        return nullptr;
    }
    return p;
}

extern "C" void gasal2_pairs_free(gasal2_pairs_t *p) {
    if (!p) return;
    if (p->query_names)
        for (uint32_t k = 0; k < p->n; k++) free(p->query_names[k]);
    if (p->target_names)
        for (uint32_t k = 0; k < p->n; k++) free(p->target_names[k]);
    void *ptrs[] = {p->query, p->target, p->query_words, p->target_words, p->query_offsets, p->target_offsets,
                    p->query_lens, p->target_lens, p->query_ops, p->target_ops, p->query_names, p->target_names};
    for (void *q : ptrs) free(q);
    free(p);
}

struct gasal2_stream {
    Parameters *params;
    uint32_t batch_pairs;
    int dev;
    std::vector<gasal_gpu_storage_v> storages;   // one storage each
    gasal2_order_t *order = nullptr;             // the storages' order chain
};

extern "C" gasal2_stream_t *gasal2_stream_new(Parameters *params, uint32_t batch_pairs, int32_t n_storages,
                                              uint32_t max_query_len, uint32_t max_target_len) {
    if (!params || batch_pairs == 0 || n_storages < 1 || n_storages > 16) {
        fail("gasal2_stream_new: no Parameters, zero batch size or n_storages outside 1..16");
        return nullptr;
    }
    gasal2_stream_t *st = new gasal2_stream_t();
    st->params = params;
    st->batch_pairs = batch_pairs;
    if (hipGetDevice(&st->dev) != hipSuccess) {
        delete st;
        fail("gasal2_stream_new: no current device");
        return nullptr;
    }
    st->order = gasal2_order_new(GASAL2_STREAM_ORDER_DEFAULT);
    if (!st->order) {
        delete st;
        fail("gasal2_stream_new: out of host memory");
        return nullptr;
    }
    const uint32_t per_q = (max_query_len + 7) & ~7u, per_t = (max_target_len + 7) & ~7u;
    for (int k = 0; k < n_storages; k++) {
        gasal_gpu_storage_v v = gasal_init_gpu_storage_v(1);
        gasal_init_streams(&v, (int)per_q + 8, (int)per_t + 8, (int)batch_pairs, params);
        gasal2_storage_set_order(&v.a[0], st->order);
        st->storages.push_back(v);
    }
    return st;
}

extern "C" void gasal2_stream_set_order(gasal2_stream_t *st, int32_t mode) {
    if (!st) return;
    // a fresh chain: the old one's last events may belong to batches in flight
    gasal2_order_t *o = gasal2_order_new(mode);
    if (!o) return;
    for (gasal_gpu_storage_v &v : st->storages) {
        (void)hipStreamSynchronize(v.a[0].str);
        gasal2_storage_set_order(&v.a[0], o);
    }
    gasal2_order_free(st->order);
    st->order = o;
}

extern "C" void gasal2_stream_free(gasal2_stream_t *st) {
    if (!st) return;
    for (gasal_gpu_storage_v &v : st->storages) {
        gasal_destroy_streams(&v, st->params);
        gasal_destroy_gpu_storage_v(&v);
    }
    gasal2_order_free(st->order);
    delete st;
}

extern "C" int32_t gasal2_stream_align(gasal2_stream_t *stp, const gasal2_pairs_t *pairs, gasal2_results_t *out,
                                       gasal2_stream_stats_t *stats) {
    if (!stp || !pairs || !out) return fail("gasal2_stream_align: NULL argument");
    Parameters *params = stp->params;
    const uint32_t batch_pairs = stp->batch_pairs;
    const int32_t n_storages = (int32_t)stp->storages.size();
    if (pairs->n == 0) return fail("gasal2_stream_align: no pairs");
    if (params->algo == KSW && !pairs->seed_scores) return fail("gasal2_stream_align: KSW needs pairs->seed_scores");
    const bool packed = params->isPacked;
    if (packed && (!pairs->query_words || !pairs->target_words))
        return fail("gasal2_stream_align: isPacked needs the packed words (GASAL2_READ_PACK)");
    const uint32_t n = pairs->n;
    uint64_t cells = 0;
    for (uint32_t k = 0; k < n; k++) {
        cells += (uint64_t)pairs->query_lens[k] * pairs->target_lens[k];
        const uint64_t qe = (uint64_t)pairs->query_offsets[k] + ((pairs->query_lens[k] + 7ull) & ~7ull);
        const uint64_t te = (uint64_t)pairs->target_offsets[k] + ((pairs->target_lens[k] + 7ull) & ~7ull);
        if ((pairs->query_offsets[k] & 7u) || (pairs->target_offsets[k] & 7u) || qe > pairs->query_bytes ||
            te > pairs->target_bytes)
            return fail("gasal2_stream_align: a pair lies outside its batch or off a multiple of 8");
    }
    const uint32_t nb = (n + batch_pairs - 1) / batch_pairs;
    const int dev = stp->dev;

    std::mutex mu;
    std::condition_variable cv;
    int ready = 0;
    bool go = false;
    std::atomic<int> errors{0};
    std::chrono::steady_clock::time_point t0, t_end;
    std::vector<std::chrono::steady_clock::time_point> ends(n_storages), firsts(n_storages);
    std::vector<double> t_fill(n_storages, 0.0), t_launch(n_storages, 0.0), t_wait(n_storages, 0.0),
        t_copy(n_storages, 0.0);
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };

    auto worker = [&](int st) {
        (void)hipSetDevice(dev);
        gasal_gpu_storage_t *s = &stp->storages[st].a[0];
        bool first = true;
        {
            std::unique_lock<std::mutex> lk(mu);
            ready++;
            cv.notify_all();
            cv.wait(lk, [&] { return go; });
        }
        for (uint32_t b = (uint32_t)st; b < nb && !errors.load(); b += (uint32_t)n_storages) {
            const uint32_t k0 = b * batch_pairs, k1 = std::min(n, k0 + batch_pairs), m = k1 - k0;
            const clk::time_point c0 = clk::now();
            // one contiguous span per side when the pairs are in order
            // (gasal2_pairs_read lays them out so), else per-sequence fills
            bool contiguous = true;
            for (uint32_t k = k0 + 1; k < k1 && contiguous; k++)
                contiguous = pairs->query_offsets[k] == pairs->query_offsets[k - 1] + ((pairs->query_lens[k - 1] + 7) & ~7u) &&
                             pairs->target_offsets[k] == pairs->target_offsets[k - 1] + ((pairs->target_lens[k - 1] + 7) & ~7u);
            gasal_host_batch_reset(s);
            const uint32_t q0 = pairs->query_offsets[k0], t0w = pairs->target_offsets[k0];
            uint32_t qb = 0, tb = 0;
            if (contiguous) {
                qb = pairs->query_offsets[k1 - 1] + ((pairs->query_lens[k1 - 1] + 7) & ~7u) - q0;
                tb = pairs->target_offsets[k1 - 1] + ((pairs->target_lens[k1 - 1] + 7) & ~7u) - t0w;
                host_batch_t *pq = s->extensible_host_unpacked_query_batch, *pt = s->extensible_host_unpacked_target_batch;
                const uint32_t qcopy = packed ? qb / 2 : qb, tcopy = packed ? tb / 2 : tb;
                if (qcopy > pq->page_size || tcopy > pt->page_size) contiguous = false;
                else {
                    if (packed) {
                        copy_wide(pq->data, pairs->query_words + q0 / 8, qcopy);
                        copy_wide(pt->data, pairs->target_words + t0w / 8, tcopy);
                    } else {
                        copy_wide(pq->data, pairs->query + q0, qcopy);
                        copy_wide(pt->data, pairs->target + t0w, tcopy);
                    }
                    pq->data_size = qcopy;
                    pt->data_size = tcopy;
                    for (uint32_t k = 0; k < m; k++) {
                        s->host_query_batch_offsets[k] = pairs->query_offsets[k0 + k] - q0;
                        s->host_target_batch_offsets[k] = pairs->target_offsets[k0 + k] - t0w;
                    }
                }
            }
            if (!contiguous) {
                gasal_host_batch_reset(s);
                uint32_t qi = 0, ti = 0;
                for (uint32_t k = 0; k < m; k++) {
                    const uint32_t p = k0 + k;
                    const uint32_t lq = (pairs->query_lens[p] + 7) & ~7u, lt = (pairs->target_lens[p] + 7) & ~7u;
                    s->host_query_batch_offsets[k] = packed ? 2 * qi : qi;
                    s->host_target_batch_offsets[k] = packed ? 2 * ti : ti;
                    if (packed) {   // whole words, padded to 8 bytes by the fill as the offsets expect
                        qi = gasal_host_batch_fill(s, qi, (const char *)(pairs->query_words + pairs->query_offsets[p] / 8), lq / 2, QUERY);
                        ti = gasal_host_batch_fill(s, ti, (const char *)(pairs->target_words + pairs->target_offsets[p] / 8), lt / 2, TARGET);
                    } else {
                        qi = gasal_host_batch_fill(s, qi, (const char *)(pairs->query + pairs->query_offsets[p]), pairs->query_lens[p], QUERY);
                        ti = gasal_host_batch_fill(s, ti, (const char *)(pairs->target + pairs->target_offsets[p]), pairs->target_lens[p], TARGET);
                    }
                }
                qb = packed ? 2 * qi : qi;
                tb = packed ? 2 * ti : ti;
            }
            memcpy(s->host_query_batch_lens, pairs->query_lens + k0, (size_t)m * 4);
            memcpy(s->host_target_batch_lens, pairs->target_lens + k0, (size_t)m * 4);
            if (params->isReverseComplement) {
                gasal_op_fill(s, pairs->query_ops + k0, m, QUERY);
                gasal_op_fill(s, pairs->target_ops + k0, m, TARGET);
            }
            if (params->algo == KSW) memcpy(s->host_seed_scores, pairs->seed_scores + k0, (size_t)m * 4);
            s->current_n_alns = m;
            const clk::time_point c1 = clk::now();
            if (first) {
                firsts[st] = c1;
                first = false;
            }
            gasal_aln_async(s, qb, tb, m, params);
            const clk::time_point c2 = clk::now();
            int r;
            while ((r = gasal_is_aln_async_done(s)) == -1) std::this_thread::yield();
            const clk::time_point c3 = clk::now();
            t_fill[st] += secs(c0, c1);
            t_launch[st] += secs(c1, c2);
            t_wait[st] += secs(c2, c3);
            if (r != 0) {
                errors++;
                break;
            }
            const gasal_res_t *h = s->host_res, *h2 = s->host_res_second;
            auto put = [&](int32_t *dst, const int32_t *src) {
                if (dst && src) memcpy(dst + k0, src, (size_t)m * 4);
            };
            put(out->score, h->aln_score);
            put(out->query_end, h->query_batch_end);
            put(out->target_end, h->target_batch_end);
            put(out->query_start, h->query_batch_start);
            put(out->target_start, h->target_batch_start);
            if (h2) {
                put(out->second_score, h2->aln_score);
                put(out->second_query_end, h2->query_batch_end);
                put(out->second_target_end, h2->target_batch_end);
            }
            if (params->start_pos == WITH_TB && h->n_cigar_ops) {
                if (out->n_cigar_ops) memcpy(out->n_cigar_ops + k0, h->n_cigar_ops, (size_t)m * 4);
                if (out->cigar && out->cigar_cap)
                    for (uint32_t k = 0; k < m; k++) {
                        const uint32_t slot = (pairs->query_lens[k0 + k] + 7) & ~7u;
                        uint8_t *dst = out->cigar + (size_t)(k0 + k) * out->cigar_cap;
                        if (h->n_cigar_ops[k] > slot && out->cigar_cap > slot)   // clipped in host_res: the whole CIGAR
                            gasal2_get_cigar(s, k, dst, out->cigar_cap);
                        else
                            memcpy(dst, h->cigar + s->host_query_batch_offsets[k], std::min(std::min(h->n_cigar_ops[k], slot), out->cigar_cap));
                    }
            }
            t_copy[st] += secs(c3, clk::now());
        }
        ends[st] = std::chrono::steady_clock::now();
        if (first) firsts[st] = ends[st];
    };

    std::vector<std::thread> th;
    for (int st = 0; st < n_storages; st++) th.emplace_back(worker, st);
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return ready == n_storages; });
        t0 = std::chrono::steady_clock::now();
        go = true;
        cv.notify_all();
    }
    for (auto &t : th) t.join();
    t_end = *std::max_element(ends.begin(), ends.end());
    const auto t_h2d = *std::min_element(firsts.begin(), firsts.end());
    if (stats) {
        stats->seconds = std::chrono::duration<double>(t_end - t_h2d).count();
        stats->seconds_from_fill = std::chrono::duration<double>(t_end - t0).count();
        stats->sub_batches = nb;
        stats->cells = cells;
        stats->fill_seconds = stats->launch_seconds = stats->wait_seconds = stats->copy_seconds = 0.0;
        for (int k = 0; k < n_storages; k++) {
            stats->fill_seconds += t_fill[k];
            stats->launch_seconds += t_launch[k];
            stats->wait_seconds += t_wait[k];
            stats->copy_seconds += t_copy[k];
        }
    }
    return errors.load() ? fail("gasal2_stream_align: a sub-batch failed") : 0;
}

namespace {

// A derived pair set (one side of the seed extension) owning its arrays.
struct OwnedPairs {
    std::vector<uint8_t> q, t, qop, top;
    std::vector<uint32_t> qw, tw, qoff, toff, qlen, tlen, seeds;
    std::vector<uint32_t> index;   // pair of the original set
    gasal2_pairs_t view() {
        gasal2_pairs_t p;
        memset(&p, 0, sizeof p);
        p.query = q.data();
        p.target = t.data();
        p.query_bytes = q.size();
        p.target_bytes = t.size();
        p.query_words = qw.empty() ? nullptr : qw.data();
        p.target_words = tw.empty() ? nullptr : tw.data();
        p.query_offsets = qoff.data();
        p.target_offsets = toff.data();
        p.query_lens = qlen.data();
        p.target_lens = tlen.data();
        p.query_ops = qop.data();
        p.target_ops = top.data();
        p.n = (uint32_t)index.size();
        p.seed_scores = seeds.data();
        return p;
    }
    // appends src[0, len) (reversed when rev) padded with N_CODE to 8 bases
    static uint32_t put(std::vector<uint8_t> &dst, const uint8_t *src, uint32_t len, bool rev, uint8_t n_code) {
        const uint32_t off = (uint32_t)dst.size();
        if (rev)
            for (uint32_t i = len; i-- > 0;) dst.push_back(src[i]);
        else
            dst.insert(dst.end(), src, src + len);
        while (dst.size() % 8) dst.push_back(n_code);
        return off;
    }
    void add(uint32_t k, const uint8_t *qs, uint32_t ql, const uint8_t *ts, uint32_t tl, bool rev, uint32_t h0,
             uint8_t n_code) {
        index.push_back(k);
        qoff.push_back(put(q, qs, ql, rev, n_code));
        toff.push_back(put(t, ts, tl, rev, n_code));
        qlen.push_back(ql);
        tlen.push_back(tl);
        seeds.push_back(h0);
    }
    void finish(bool packed) {
        qop.assign(index.size() + 1, 0);
        top.assign(index.size() + 1, 0);
        if (q.empty()) q.assign(8, 0);
        if (t.empty()) t.assign(8, 0);
        if (packed) {
            qw.resize(q.size() / 8 + 1);
            tw.resize(t.size() / 8 + 1);
            pack_words(q.data(), q.size(), qw.data());
            pack_words(t.data(), t.size(), tw.data());
        }
    }
};

}  // namespace

extern "C" int32_t gasal2_extend_seeds(gasal2_stream_t *st, const gasal2_pairs_t *pairs, const gasal2_seed_t *seeds,
                                       const gasal2_extend_opts_t *opts, gasal2_extend_out_t *out,
                                       gasal2_stream_stats_t *stats) {
    if (!st || !pairs || !seeds || !opts || !out) return fail("gasal2_extend_seeds: NULL argument");
    if (st->params->algo != KSW) return fail("gasal2_extend_seeds: the stream must be made with algo = KSW");
    const uint32_t n = pairs->n;
    for (uint32_t k = 0; k < n; k++) {
        const gasal2_seed_t &sd = seeds[k];
        if ((uint64_t)sd.query_beg + sd.len > pairs->query_lens[k] || (uint64_t)sd.target_beg + sd.len > pairs->target_lens[k] ||
            sd.score < 0)
            return fail("gasal2_extend_seeds: a seed lies outside its pair or has a negative score");
    }
    const uint8_t n_code = (uint8_t)gasal2_get_n_code();
    const bool packed = st->params->isPacked;
    std::vector<int32_t> lscore(n), lq(n, 0), lt(n, 0), rscore(n), rq(n, 0), rt(n, 0);
    gasal2_stream_stats_t tot;
    memset(&tot, 0, sizeof tot);
    auto run_side = [&](OwnedPairs &side, int32_t eb, std::vector<int32_t> &score, std::vector<int32_t> &qe,
                        std::vector<int32_t> &te) -> int32_t {
        if (side.index.empty()) return 0;
        side.finish(packed);
        for (gasal_gpu_storage_v &v : st->storages) gasal2_storage_set_ksw(&v.a[0], 1, opts->w, opts->zdrop, eb);
        std::vector<int32_t> sc(side.index.size()), qend(side.index.size()), tend(side.index.size());
        gasal2_results_t res;
        memset(&res, 0, sizeof res);
        res.score = sc.data();
        res.query_end = qend.data();
        res.target_end = tend.data();
        gasal2_pairs_t view = side.view();
        gasal2_stream_stats_t ss;
        const int32_t r = gasal2_stream_align(st, &view, &res, &ss);
        for (gasal_gpu_storage_v &v : st->storages) gasal2_storage_set_ksw(&v.a[0], 0, -1, 0, 5);
        if (r != 0) return r;
        for (size_t i = 0; i < side.index.size(); i++) {
            const uint32_t k = side.index[i];
            score[k] = sc[i];
            qe[k] = qend[i];
            te[k] = tend[i];
        }
        tot.seconds += ss.seconds;
        tot.seconds_from_fill += ss.seconds_from_fill;
        tot.sub_batches += ss.sub_batches;
        tot.cells += ss.cells;
        tot.fill_seconds += ss.fill_seconds;
        tot.launch_seconds += ss.launch_seconds;
        tot.wait_seconds += ss.wait_seconds;
        tot.copy_seconds += ss.copy_seconds;
        return 0;
    };
    // left: reversed prefixes from the seed's score
    {
        OwnedPairs left;
        for (uint32_t k = 0; k < n; k++) {
            lscore[k] = seeds[k].score;
            if (seeds[k].query_beg && seeds[k].target_beg)
                left.add(k, pairs->query + pairs->query_offsets[k], seeds[k].query_beg, pairs->target + pairs->target_offsets[k],
                         seeds[k].target_beg, true, (uint32_t)seeds[k].score, n_code);
        }
        if (run_side(left, opts->end_bonus_left, lscore, lq, lt) != 0) return -1;
    }
    // right: suffixes from the left extension's score
    {
        OwnedPairs right;
        for (uint32_t k = 0; k < n; k++) {
            rscore[k] = lscore[k];
            const uint32_t qs = seeds[k].query_beg + seeds[k].len, ts = seeds[k].target_beg + seeds[k].len;
            if (qs < pairs->query_lens[k] && ts < pairs->target_lens[k])
                right.add(k, pairs->query + pairs->query_offsets[k] + qs, pairs->query_lens[k] - qs,
                          pairs->target + pairs->target_offsets[k] + ts, pairs->target_lens[k] - ts, false,
                          (uint32_t)std::max(lscore[k], 0), n_code);
        }
        if (run_side(right, opts->end_bonus_right, rscore, rq, rt) != 0) return -1;
    }
    for (uint32_t k = 0; k < n; k++) {
        const gasal2_seed_t &sd = seeds[k];
        if (out->score) out->score[k] = rscore[k];
        if (out->left_score) out->left_score[k] = lscore[k];
        if (out->right_score) out->right_score[k] = rscore[k];
        if (out->query_beg) out->query_beg[k] = (int32_t)sd.query_beg - lq[k];
        if (out->target_beg) out->target_beg[k] = (int32_t)sd.target_beg - lt[k];
        if (out->query_end) out->query_end[k] = (int32_t)(sd.query_beg + sd.len) + rq[k];
        if (out->target_end) out->target_end[k] = (int32_t)(sd.target_beg + sd.len) + rt[k];
    }
    if (stats) *stats = tot;
    return 0;
}

extern "C" int32_t gasal2_pairs_align(const gasal2_pairs_t *pairs, gasal2_results_t *out, Parameters *params,
                                      uint32_t batch_pairs, int32_t n_storages, gasal2_stream_stats_t *stats) {
    if (!pairs || !out || !params) return fail("gasal2_pairs_align: NULL argument");
    uint32_t maxq = 0, maxt = 0;
    for (uint32_t k = 0; k < pairs->n; k++) {
        maxq = std::max(maxq, pairs->query_lens[k]);
        maxt = std::max(maxt, pairs->target_lens[k]);
    }
    gasal2_stream_t *st = gasal2_stream_new(params, batch_pairs, n_storages, maxq, maxt);
    if (!st) return -1;
    const int32_t r = gasal2_stream_align(st, pairs, out, stats);
    gasal2_stream_free(st);
    return r;
}
