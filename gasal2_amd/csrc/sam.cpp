// sam.cpp -- CIGAR text and SAM records from the library's results
// (include/gasal2_sam.h).  Host code only.  The CIGAR bytes follow
// README.md:203-211 (op | count << 2, last run first); the reference's own
// printer is test_prog/test_prog.cpp:382-428.
#include "gasal2_sam.h"
#include "gasal2_ext.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

inline uint32_t shl32(uint32_t x, uint32_t s) { return s >= 32 ? 0u : x << s; }
inline uint32_t shr32(uint32_t x, uint32_t s) { return s >= 32 ? 0u : x >> s; }

uint32_t nib_reverse(uint32_t x) {
    uint32_t r = 0;
    for (int k = 0; k < 8; k++) r |= ((x >> (4 * k)) & 15u) << (28 - 4 * k);
    return r;
}

uint32_t nib_complement(uint32_t w) {
    uint32_t r = 0;
    for (int k = 0; k < 32; k += 4) {
        uint32_t nt = (w >> k) & 15u;
        nt = nt == 1u ? 4u : nt == 4u ? 1u : nt == 3u ? 7u : nt == 7u ? 3u : nt;
        r |= nt << k;
    }
    return r;
}

// The device's reverse/complement of one pair's padded words (the same steps
// as the k_rc kernel: src/kernels/pack_rc_seqs.h:56-212 with shifts of 32
// giving 0), so the text shows the sequence the DP saw.
void rc_words(uint32_t *w, uint32_t len, uint32_t op, int32_t n_code) {
    const uint32_t regs = (len + 7) / 8;
    if (regs == 0) return;
    const uint32_t to_swap = (regs >> 1) + (regs & 1);
    if (op & 1u) {
        uint32_t nbr_N = 0;
        const uint32_t last = w[regs - 1];
        for (int jj = 0; jj < 32; jj += 4) nbr_N += ((last >> jj) & 15u) == (uint32_t)n_code;
        nbr_N = (nbr_N << 2) & 0xFFu;
        for (uint32_t i = 0; i < to_swap; i++) {
            const uint32_t rpac_1 = w[i];
            const uint32_t wm2 = regs >= 2 + i ? w[regs - 2 - i] : 0u;
            const uint32_t wm1 = w[regs - 1 - i];
            const uint32_t rpac_2 = shl32(wm2, 32 - nbr_N) | shr32(wm1, nbr_N);
            const uint32_t rev1 = nib_reverse(rpac_1), rev2 = nib_reverse(rpac_2);
            const uint32_t lowmask = shl32(1u, nbr_N) - 1u;
            const uint32_t q1 = shl32(rev1, nbr_N) | (wm1 & lowmask);
            const uint32_t q2 = (wm2 & ~lowmask) | shr32(rev1, 32 - nbr_N);
            w[i] = rev2;
            w[regs - 1 - i] = q1;
            if (i != to_swap - 1) w[regs - 2 - i] = q2;
        }
    }
    if (op & 2u)
        for (uint32_t i = 0; i < regs; i++) w[i] = nib_complement(w[i]);
}

inline char letter(uint32_t nib) {
    switch (nib) {
        case 1: return 'A';
        case 3: return 'C';
        case 7: return 'G';
        case 4: return 'T';
        default: return 'N';
    }
}

inline uint32_t nib_comp(uint32_t nt) { return nt == 1u ? 4u : nt == 4u ? 1u : nt == 3u ? 7u : nt == 7u ? 3u : nt; }

// The first len bases the DP saw, as letters A C G T N.
void aligned_letters(const uint8_t *seq, uint32_t len, uint8_t op, bool rc, int32_t n_code, std::vector<uint32_t> &w,
                     std::string &out) {
    out.resize(len);
    if (!rc || !op) {   // the bytes as given
        for (uint32_t i = 0; i < len; i++) out[i] = letter(seq[i] & 15u);
        return;
    }
    if ((uint32_t)n_code > 15u) {
        // no nibble equals such an N_CODE, so the reverse is a plain reversal
        // of the padded sequence (padding first), as rc_words computes it
        const uint32_t padded = 8 * ((len + 7) / 8), pad = (uint32_t)n_code & 15u;
        for (uint32_t i = 0; i < len; i++) {
            const uint32_t src = (op & 1u) ? padded - 1 - i : i;
            uint32_t nt = src < len ? seq[src] & 15u : pad;
            if (op & 2u) nt = nib_comp(nt);
            out[i] = letter(nt);
        }
        return;
    }
    const uint32_t regs = (len + 7) / 8;
    w.assign(regs, 0u);
    for (uint32_t i = 0; i < 8 * regs; i++) {
        const uint32_t b = i < len ? seq[i] : (uint32_t)(uint8_t)n_code;
        w[i / 8] |= (b & 15u) << (28 - 4 * (i % 8));
    }
    rc_words(w.data(), len, op, n_code);
    for (uint32_t i = 0; i < len; i++) out[i] = letter((w[i / 8] >> (28 - 4 * (i % 8))) & 15u);
}

void put_u(std::string &s, uint64_t v) {
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) s.push_back(tmp[--n]);
}

void put_i(std::string &s, int64_t v) {
    if (v < 0) {
        s.push_back('-');
        put_u(s, (uint64_t)(-v));
    } else {
        put_u(s, (uint64_t)v);
    }
}

// First word of a name, "*" when there is none.
void put_name(std::string &s, const char *name) {
    size_t n = 0;
    if (name)
        while (name[n] && name[n] != ' ' && name[n] != '\t' && name[n] != '\n' && name[n] != '\r') n++;
    if (n) s.append(name, n);
    else s.push_back('*');
}

struct Run {
    uint8_t op;
    uint32_t count;
};

inline bool eats_query(uint8_t op) { return op != 2; }
inline bool eats_target(uint8_t op) { return op != 3; }

// Forward runs of the CIGAR bytes (last byte first).
void forward_runs(const uint8_t *cg, uint32_t n, std::vector<Run> &runs) {
    runs.clear();
    for (uint32_t u = n; u-- > 0;) {
        const Run r{(uint8_t)(cg[u] & 3u), (uint32_t)(cg[u] >> 2)};
        if (r.count) runs.push_back(r);
    }
}

struct Scratch {
    std::vector<uint32_t> w;
    std::string q, t;
    std::vector<Run> runs;
};

void sam_line(const gasal2_sam_in_t &in, bool rc, int32_t n_code, uint32_t flags, Scratch &x, std::string &s) {
    aligned_letters(in.query, in.query_len, in.query_op, rc, n_code, x.w, x.q);
    const bool want_t = in.cigar && in.n_cigar_ops && in.score > 0;
    if (want_t) aligned_letters(in.target, in.target_len, in.target_op, rc, n_code, x.w, x.t);

    // the span: anchored at the reported ends (or at the starts without
    // ends), trimmed to the sequences
    int64_t qs = 0, ts = 0, qc = 0, tc = 0;
    bool mapped = want_t;
    if (mapped) {
        forward_runs(in.cigar, in.n_cigar_ops, x.runs);
        for (const Run &r : x.runs) {
            if (eats_query(r.op)) qc += r.count;
            if (eats_target(r.op)) tc += r.count;
        }
        if (in.query_end >= 0 && in.target_end >= 0) {
            if ((uint32_t)in.query_end >= in.query_len || (uint32_t)in.target_end >= in.target_len) {
                mapped = false;
            } else {
                qs = (int64_t)in.query_end + 1 - qc;
                ts = (int64_t)in.target_end + 1 - tc;
                size_t f = 0;
                while ((qs < 0 || ts < 0) && f < x.runs.size()) {
                    Run &r = x.runs[f];
                    const int64_t need = std::max(eats_query(r.op) && qs < 0 ? -qs : 0, eats_target(r.op) && ts < 0 ? -ts : 0);
                    const int64_t take = std::max<int64_t>(1, std::min<int64_t>(need, r.count));
                    r.count -= (uint32_t)take;
                    if (eats_query(r.op)) { qs += take; qc -= take; }
                    if (eats_target(r.op)) { ts += take; tc -= take; }
                    if (!r.count) f++;
                }
                x.runs.erase(x.runs.begin(), x.runs.begin() + f);
                if (qs < 0 || ts < 0) mapped = false;
            }
        } else {
            while (!x.runs.empty() && (qc > in.query_len || tc > in.target_len)) {
                Run &r = x.runs.back();
                r.count--;
                if (eats_query(r.op)) qc--;
                if (eats_target(r.op)) tc--;
                if (!r.count) x.runs.pop_back();
            }
        }
        if (x.runs.empty()) mapped = false;
    }

    uint32_t flag = 0;
    if (rc && in.query_op == 3) flag |= 16u;
    if (!mapped) flag |= 4u;
    put_name(s, in.qname);
    s.push_back('\t');
    put_u(s, flag);
    s.push_back('\t');
    if (mapped) put_name(s, in.rname);
    else s.push_back('*');
    s.push_back('\t');
    put_u(s, mapped ? (uint64_t)ts + 1 : 0);
    s.append(mapped ? "\t255\t" : "\t0\t");
    uint64_t nm = 0;
    if (mapped) {
        if (qs > 0) {
            put_u(s, (uint64_t)qs);
            s.push_back('S');
        }
        const bool eqx = (flags & GASAL2_SAM_EQX) != 0;
        char cur = 0;
        uint64_t cnt = 0;
        auto emit = [&](char c, uint64_t k) {
            if (c == cur) {
                cnt += k;
                return;
            }
            if (cnt) {
                put_u(s, cnt);
                s.push_back(cur);
            }
            cur = c;
            cnt = k;
        };
        uint64_t qi = (uint64_t)qs, ti = (uint64_t)ts;
        for (const Run &r : x.runs) {
            if (r.op == 2) {
                emit('D', r.count);
                ti += r.count;
                nm += r.count;
            } else if (r.op == 3) {
                emit('I', r.count);
                qi += r.count;
                nm += r.count;
            } else {
                for (uint32_t k = 0; k < r.count; k++, qi++, ti++) {
                    const char a = x.q[qi], b = x.t[ti];
                    const bool diff = a != b || a == 'N';
                    nm += diff;
                    emit(eqx ? (diff ? 'X' : '=') : 'M', 1);
                }
            }
        }
        emit(0, 0);
        const uint64_t tail = (uint64_t)in.query_len - (uint64_t)qs - (uint64_t)qc;
        if (tail) {
            put_u(s, tail);
            s.push_back('S');
        }
    } else {
        s.push_back('*');
    }
    s.append("\t*\t0\t0\t");
    if ((flags & GASAL2_SAM_NO_SEQ) || in.query_len == 0) s.push_back('*');
    else s.append(x.q);
    s.append("\t*\tAS:i:");
    put_i(s, in.score);
    if (mapped) {
        s.append("\tNM:i:");
        put_u(s, nm);
    }
    if (in.has_second) {
        s.append("\tXS:i:");
        put_i(s, in.second_score);
    }
    s.push_back('\n');
}

int fail(const char *msg, const char *arg = "") {
    fprintf(stderr, "[GASAL ERROR:] %s%s\n", msg, arg);
    return -1;
}

}  // namespace

extern "C" int32_t gasal2_cigar_text(const uint8_t *cigar, uint32_t n_ops, int32_t style, char *buf, uint32_t cap) {
    if (style != GASAL2_CIGAR_HARNESS && style != GASAL2_CIGAR_SAM) return -1;
    std::string s;
    if (cigar && n_ops) {
        // test_prog/test_prog.cpp:387-425: walk from the last byte, merging
        // runs whose op (SAM: whose letter) repeats
        static const char harness[4] = {'M', 'X', 'D', 'I'}, sam[4] = {'M', 'M', 'D', 'I'};
        const char *letters = style == GASAL2_CIGAR_SAM ? sam : harness;
        char last = letters[cigar[n_ops - 1] & 3u];
        uint64_t count = cigar[n_ops - 1] >> 2;
        for (uint32_t u = n_ops - 1; u-- > 0;) {
            const char c = letters[cigar[u] & 3u];
            if (c == last) {
                count += cigar[u] >> 2;
            } else {
                put_u(s, count);
                s.push_back(last);
                count = cigar[u] >> 2;
                last = c;
            }
        }
        put_u(s, count);
        s.push_back(last);
    }
    if (buf && cap) {
        const size_t n = std::min<size_t>(s.size(), cap - 1);
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int32_t)s.size();
}

extern "C" int32_t gasal2_sam_record(const gasal2_sam_in_t *in, int32_t rc_enabled, int32_t n_code, uint32_t flags,
                                     char *buf, uint32_t cap) {
    if (!in || (in->query_len && !in->query) || (in->cigar && in->n_cigar_ops && in->target_len && !in->target))
        return -1;
    Scratch x;
    std::string s;
    sam_line(*in, rc_enabled != 0, n_code, flags, x, s);
    if (buf && cap) {
        const size_t n = std::min<size_t>(s.size(), cap - 1);
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int32_t)s.size();
}

extern "C" int32_t gasal2_sam_write(const char *path, const gasal2_pairs_t *pairs, const gasal2_results_t *res,
                                    int32_t rc_enabled, uint32_t flags, int32_t n_threads) {
    if (!path || !pairs || !res || !res->score) return fail("gasal2_sam_write: NULL argument (path, pairs, results or scores)");
    FILE *f = strcmp(path, "-") == 0 ? stdout : fopen(path, "wb");
    if (!f) return fail("gasal2_sam_write: cannot open ", path);
    const int32_t n_code = gasal2_get_n_code();
    const uint32_t n = pairs->n;
    auto qname = [&](uint32_t k, std::string &tmp) -> const char * {
        if (pairs->query_names) return pairs->query_names[k];
        tmp = "q";
        put_u(tmp, k);
        return tmp.c_str();
    };
    auto rname = [&](uint32_t k, std::string &tmp) -> const char * {
        if (pairs->target_names) return pairs->target_names[k];
        tmp = "t";
        put_u(tmp, k);
        return tmp.c_str();
    };
    bool ok = true;
    if (flags & GASAL2_SAM_HEADER) {
        std::string h = "@HD\tVN:1.6\tSO:unsorted\n";
        std::unordered_map<std::string, uint32_t> seen;   // distinct names only (generated names are unique)
        if (pairs->target_names) seen.reserve((size_t)n * 2);
        std::string tmp, key;
        for (uint32_t k = 0; k < n; k++) {
            key.clear();
            put_name(key, rname(k, tmp));
            if (pairs->target_names && !seen.emplace(key, k).second) continue;
            h.append("@SQ\tSN:");
            h.append(key);
            h.append("\tLN:");
            put_u(h, pairs->target_lens[k]);
            h.push_back('\n');
            if (h.size() > (1u << 24)) {
                ok = ok && fwrite(h.data(), 1, h.size(), f) == h.size();
                h.clear();
            }
        }
        h.append("@PG\tID:gasal2_amd\tPN:gasal2_amd\n");
        ok = ok && fwrite(h.data(), 1, h.size(), f) == h.size();
    }
    unsigned T = n_threads > 0 ? (unsigned)n_threads : std::max(1u, std::thread::hardware_concurrency());
    T = std::min(T, 64u);
    const uint32_t block = 16384;
    std::vector<std::string> out(T);
    for (uint64_t base = 0; base < n && ok; base += (uint64_t)block * T) {
        auto work = [&](unsigned t) {
            Scratch x;
            std::string tq, tr;
            std::string &s = out[t];
            s.clear();
            const uint64_t k0 = base + (uint64_t)t * block, k1 = std::min<uint64_t>(n, k0 + block);
            for (uint64_t kk = k0; kk < k1; kk++) {
                const uint32_t k = (uint32_t)kk;
                gasal2_sam_in_t in;
                memset(&in, 0, sizeof in);
                in.qname = qname(k, tq);
                in.rname = rname(k, tr);
                in.query = pairs->query + pairs->query_offsets[k];
                in.target = pairs->target + pairs->target_offsets[k];
                in.query_len = pairs->query_lens[k];
                in.target_len = pairs->target_lens[k];
                in.query_op = pairs->query_ops ? pairs->query_ops[k] : 0;
                in.target_op = pairs->target_ops ? pairs->target_ops[k] : 0;
                in.score = res->score[k];
                in.query_end = res->query_end ? res->query_end[k] : -1;
                in.target_end = res->target_end ? res->target_end[k] : -1;
                in.has_second = res->second_score != nullptr;
                in.second_score = res->second_score ? res->second_score[k] : 0;
                if (res->cigar && res->n_cigar_ops && res->n_cigar_ops[k] <= res->cigar_cap) {
                    in.cigar = res->cigar + (size_t)k * res->cigar_cap;
                    in.n_cigar_ops = res->n_cigar_ops[k];
                }
                sam_line(in, rc_enabled != 0, n_code, flags, x, s);
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < T; t++) th.emplace_back(work, t);
        work(0);
        for (auto &t : th) t.join();
        for (unsigned t = 0; t < T && ok; t++) ok = fwrite(out[t].data(), 1, out[t].size(), f) == out[t].size();
    }
    if (f != stdout) ok = (fclose(f) == 0) && ok;
    else fflush(f);
    return ok ? 0 : fail("gasal2_sam_write: write failed on ", path);
}
