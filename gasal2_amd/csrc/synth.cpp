// synth.cpp -- seeded synthetic pair generator for the five benchmark workloads.
//
// Pure host C++ (no HIP): every pair draws from its own xoshiro256** stream
// seeded by splitmix64(seed, pair index), so generation is deterministic and
// independent of the thread count.  Sequences are ASCII (A,C,G,T,N) and are
// laid out like a GASAL2 host batch: each sequence padded with the N_CODE byte
// to a multiple of 8 (reference src/host_batch.cpp:100-102,142-145), offsets in
// bytes (== bases).
//
// Workloads (ids 1..5):
//   1  10,000-style: q = 150 iid, t = 50 | q | 50, 2% subst, 0.5% indels U{1..3}
//   2  Lq~U{100..300}, flanks F~U{0..100} split U{0..F}, 4% subst + 1% indels
//      U{1..3}, 0.5% N in q and t, query_op=3 on a Bernoulli(0.5) subset with the
//      stored query = reverse complement of q
//   3  Lq = clip(round(exp(N(ln 400, 0.7^2))), 50, 5000), Lt = Lq*U[0.8,1.2],
//      6% subst, 4% indels with geometric length (mean 2)
//   4  seed extension: Lq,Lt ~ U{100..250}, t = q mutated 3% then cut/extended
//      to Lt; h0 ~ U{10..60}
//   5  long divergent: Lq~U{1000..10000}, divergence U[0.10,0.15] split 60:40
//      substitutions:indels (geometric mean 2), plus Poisson(1/kbp) indel bursts
//      of length U{5..50}
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <string>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = seed * 0xD1B54A32D192ED03ULL + stream * 0x9E3779B97F4A7C15ULL + 0x1234567ULL;
        for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
    int64_t range(int64_t lo, int64_t hi) {                                          // U{lo..hi}
        return lo + (int64_t)(uni() * (double)(hi - lo + 1));
    }
    double normal() {  // Box-Muller
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
    int64_t geometric_mean2() {  // P(L=k) = 2^-k, k >= 1
        int64_t k = 1;
        while (uni() < 0.5) k++;
        return k;
    }
    int64_t poisson(double lambda) {  // Knuth; lambda <= ~10 here
        double L = exp(-lambda), p = 1.0;
        int64_t k = 0;
        do { k++; p *= uni(); } while (p > L);
        return k - 1;
    }
};

const char BASES[4] = {'A', 'C', 'G', 'T'};

inline char rbase(Rng &r) { return BASES[r.next() & 3]; }
inline char other_base(Rng &r, char b) {
    char c;
    do { c = rbase(r); } while (c == b);
    return c;
}

std::string random_seq(Rng &r, int64_t n) {
    std::string s((size_t)n, 'A');
    for (int64_t i = 0; i < n; i++) s[(size_t)i] = rbase(r);
    return s;
}

// Substitutions at rate ps; indel events at rate pi per base (insertion or
// deletion 50/50) with length drawn by `len`.
template <class LenFn>
std::string mutate(Rng &r, const std::string &src, double ps, double pi, LenFn len) {
    std::string out;
    out.reserve(src.size() + src.size() / 8 + 8);
    size_t i = 0;
    while (i < src.size()) {
        double u = r.uni();
        if (u < pi) {
            int64_t L = len();
            if (r.uni() < 0.5) {                       // insertion before src[i]
                for (int64_t k = 0; k < L; k++) out.push_back(rbase(r));
                out.push_back(src[i]);
                i++;
            } else {                                   // deletion of L bases
                i += (size_t)L;
            }
        } else if (u < pi + ps) {
            out.push_back(other_base(r, src[i]));
            i++;
        } else {
            out.push_back(src[i]);
            i++;
        }
    }
    return out;
}

void sprinkle_n(Rng &r, std::string &s, double pn) {
    for (auto &ch : s)
        if (r.uni() < pn) ch = 'N';
}

std::string revcomp(const std::string &s) {
    std::string o(s.rbegin(), s.rend());
    for (auto &ch : o) {
        switch (ch) {
            case 'A': ch = 'T'; break;
            case 'T': ch = 'A'; break;
            case 'C': ch = 'G'; break;
            case 'G': ch = 'C'; break;
            default: break;
        }
    }
    return o;
}

struct Pair {
    std::string q, t;
    uint8_t qop = 0, top = 0;
    uint32_t seed = 0;
};

void gen_pair(int cfg, uint64_t seed, uint64_t idx, Pair &P) {
    Rng r(seed, idx);
    auto u13 = [&r]() { return r.range(1, 3); };
    auto geo = [&r]() { return r.geometric_mean2(); };
    switch (cfg) {
        case 1: {
            std::string q = random_seq(r, 150);
            std::string t = random_seq(r, 50) + q + random_seq(r, 50);
            P.q = q;
            P.t = mutate(r, t, 0.02, 0.005, u13);
            break;
        }
        case 2: {
            int64_t lq = r.range(100, 300);
            std::string q = random_seq(r, lq);
            int64_t F = r.range(0, 100);
            int64_t left = r.range(0, F);
            std::string t = random_seq(r, left) + q + random_seq(r, F - left);
            t = mutate(r, t, 0.04, 0.01, u13);
            sprinkle_n(r, q, 0.005);
            sprinkle_n(r, t, 0.005);
            if (r.uni() < 0.5) { P.q = revcomp(q); P.qop = 3; }
            else P.q = q;
            P.t = t;
            break;
        }
        case 3: {
            double z = r.normal();
            double lqd = exp(log(400.0) + 0.7 * z);
            int64_t lq = (int64_t)llround(lqd);
            lq = std::min<int64_t>(5000, std::max<int64_t>(50, lq));
            std::string q = random_seq(r, lq);
            double f = 0.8 + 0.4 * r.uni();
            int64_t lt = std::max<int64_t>(1, (int64_t)llround((double)lq * f));
            std::string t = mutate(r, q, 0.06, 0.04 / 2.0, geo);   // 4% indel bases, mean length 2
            if ((int64_t)t.size() > lt) {
                int64_t cut = (int64_t)t.size() - lt;
                int64_t left = r.range(0, cut);
                t = t.substr((size_t)left, (size_t)lt);
            } else if ((int64_t)t.size() < lt) {
                int64_t add = lt - (int64_t)t.size();
                int64_t left = r.range(0, add);
                t = random_seq(r, left) + t + random_seq(r, add - left);
            }
            P.q = q;
            P.t = t;
            break;
        }
        case 4: {
            int64_t lq = r.range(100, 250), lt = r.range(100, 250);
            std::string q = random_seq(r, lq);
            std::string t = mutate(r, q, 0.02, 0.01, u13);
            if ((int64_t)t.size() > lt) t.resize((size_t)lt);
            else t += random_seq(r, lt - (int64_t)t.size());
            P.q = q;
            P.t = t;
            P.seed = (uint32_t)r.range(10, 60);
            break;
        }
        case 5:
        default: {
            int64_t lq = r.range(1000, 10000);
            std::string q = random_seq(r, lq);
            double d = 0.10 + 0.05 * r.uni();
            std::string t = mutate(r, q, 0.6 * d, 0.4 * d / 2.0, geo);
            int64_t nb = r.poisson((double)lq / 1000.0);
            for (int64_t b = 0; b < nb && !t.empty(); b++) {
                int64_t L = r.range(5, 50);
                int64_t pos = r.range(0, (int64_t)t.size() - 1);
                if (r.uni() < 0.5) t.insert((size_t)pos, random_seq(r, L));
                else t.erase((size_t)pos, (size_t)std::min<int64_t>(L, (int64_t)t.size() - pos));
            }
            if (t.empty()) t = "A";
            P.q = q;
            P.t = t;
            break;
        }
    }
}

struct Batch {
    std::vector<Pair> pairs;
    std::vector<uint64_t> qoff, toff;
    uint64_t qbytes = 0, tbytes = 0;
};

inline uint64_t pad8(uint64_t n) { return (n + 7) & ~7ULL; }

}  // namespace

extern "C" {

// Generates n pairs of workload `cfg`; returns an opaque handle (free with synth_free).
void *synth_generate(int cfg, uint64_t seed, uint32_t n) {
    Batch *B = new Batch();
    B->pairs.resize(n);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++) gen_pair(cfg, seed, (uint64_t)i, B->pairs[(size_t)i]);
    B->qoff.resize(n);
    B->toff.resize(n);
    uint64_t qo = 0, to = 0;
    for (uint32_t i = 0; i < n; i++) {
        B->qoff[i] = qo; B->toff[i] = to;
        qo += pad8(B->pairs[i].q.size());
        to += pad8(B->pairs[i].t.size());
    }
    B->qbytes = qo; B->tbytes = to;
    return B;
}

uint64_t synth_query_bytes(void *h) { return ((Batch *)h)->qbytes; }
uint64_t synth_target_bytes(void *h) { return ((Batch *)h)->tbytes; }

// Writes the padded host batches and per-pair arrays.  Any output pointer may be NULL.
void synth_export(void *h, uint8_t n_code, uint8_t *qbuf, uint8_t *tbuf, uint32_t *qoff, uint32_t *toff,
                  uint32_t *qlen, uint32_t *tlen, uint8_t *qop, uint8_t *top, uint32_t *seeds) {
    Batch *B = (Batch *)h;
    int64_t n = (int64_t)B->pairs.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) {
        const Pair &P = B->pairs[(size_t)i];
        if (qbuf) {
            memcpy(qbuf + B->qoff[(size_t)i], P.q.data(), P.q.size());
            memset(qbuf + B->qoff[(size_t)i] + P.q.size(), n_code, pad8(P.q.size()) - P.q.size());
        }
        if (tbuf) {
            memcpy(tbuf + B->toff[(size_t)i], P.t.data(), P.t.size());
            memset(tbuf + B->toff[(size_t)i] + P.t.size(), n_code, pad8(P.t.size()) - P.t.size());
        }
        if (qoff) qoff[i] = (uint32_t)B->qoff[(size_t)i];
        if (toff) toff[i] = (uint32_t)B->toff[(size_t)i];
        if (qlen) qlen[i] = (uint32_t)P.q.size();
        if (tlen) tlen[i] = (uint32_t)P.t.size();
        if (qop) qop[i] = P.qop;
        if (top) top[i] = P.top;
        if (seeds) seeds[i] = P.seed;
    }
}

void synth_free(void *h) { delete (Batch *)h; }

}  // extern "C"
