// test_prog.cpp -- command-line driver over the drop-in API.
//
// Same command line, batching and stdout format as the reference's example
// program (test_prog/test_prog.cpp:28-461): paired FASTA files whose header's
// first character selects the per-sequence operation ('>' 0, '<' 1, '/' 2,
// '+' 3), OpenMP threads each owning two storages (streams), batches of 5000
// pairs, one result line per pair, total wall time on stderr.  It uses only the
// public headers, so it also shows that code written against the reference
// builds unchanged.
#include "gasal_header.h"

#include <omp.h>
#include <chrono>
#include <string>
#include <vector>
#include <cstring>

#define NB_STREAMS 2
#define STREAM_BATCH_SIZE 5000

namespace {

struct Input {
    std::vector<std::string> qseq, tseq, qhdr, thdr;
    std::vector<uint8_t> qmod, tmod;
};

bool read_pairs(Parameters *args, Input &in) {
    static const char line_starts[] = "></+";
    std::string ql, tl;
    int state = 0;  // 0: expect header, 1: header read, 2: in sequence
    while (std::getline(args->query_batch_fasta, ql) && std::getline(args->target_batch_fasta, tl)) {
        const char *q = ql.empty() ? nullptr : strchr(line_starts, ql[0]);
        const char *t = tl.empty() ? nullptr : strchr(line_starts, tl[0]);
        if (q && t && ql[0] && tl[0]) {
            in.qmod.push_back((uint8_t)(q - line_starts));
            in.tmod.push_back((uint8_t)(t - line_starts));
            in.qhdr.push_back(ql.substr(1));
            in.thdr.push_back(tl.substr(1));
            state = 1;
        } else if (state == 1) {
            in.qseq.push_back(ql);
            in.tseq.push_back(tl);
            state = 2;
        } else if (state == 2) {
            in.qseq.back() += ql;
            in.tseq.back() += tl;
        } else {
            std::cerr << "Batch1 and target_batch files should be fasta having same number of sequences" << std::endl;
            return false;
        }
    }
    return true;
}

void print_result(const Parameters *args, const Input &in, gasal_gpu_storage_t *s, int i, int j) {
    std::cout << "query_name=" << in.qhdr[i];
    std::cout << "\ttarget_name=" << in.thdr[i];
    std::cout << "\tscore=" << s->host_res->aln_score[j];
    if ((args->start_pos == WITH_START || args->start_pos == WITH_TB) &&
        ((args->algo == SEMI_GLOBAL && args->semiglobal_skipping_head != NONE) || args->algo > SEMI_GLOBAL)) {
        std::cout << "\tquery_batch_start=" << s->host_res->query_batch_start[j];
        std::cout << "\ttarget_batch_start=" << s->host_res->target_batch_start[j];
    }
    if (args->algo != GLOBAL) {
        std::cout << "\tquery_batch_end=" << s->host_res->query_batch_end[j];
        std::cout << "\ttarget_batch_end=" << s->host_res->target_batch_end[j];
    }
    if (args->secondBest) {
        std::cout << "\t2nd_score=" << s->host_res_second->aln_score[j];
        std::cout << "\t2nd_query_batch_end=" << s->host_res_second->query_batch_end[j];
        std::cout << "\t2nd_target_batch_end=" << s->host_res_second->target_batch_end[j];
    }
    if (args->start_pos == WITH_TB) {
        std::cout << "\tCIGAR=";
        const uint8_t *cg = s->host_res->cigar + s->host_query_batch_offsets[j];
        int n = (int)s->host_res->n_cigar_ops[j];
        static const char opc[4] = {'M', 'X', 'D', 'I'};
        if (n > 0) {
            int last = cg[n - 1] & 3, count = cg[n - 1] >> 2;
            for (int u = n - 2; u >= 0; u--) {
                int op = cg[u] & 3;
                if (op == last) {
                    count += cg[u] >> 2;
                } else {
                    std::cout << count << opc[last];
                    count = cg[u] >> 2;
                }
                last = op;
            }
            std::cout << count << opc[last];
        }
    }
    std::cout << std::endl;
}

}  // namespace

int main(int argc, char **argv) {
    Parameters *args = new Parameters(argc, argv);
    args->parse();
    args->print();

    gasal_subst_scores sub;
    sub.match = args->sa;
    sub.mismatch = args->sb;
    sub.gap_open = args->gapo;
    sub.gap_extend = args->gape;
    gasal_copy_subst_scores(&sub);

    std::cerr << "Loading files...." << std::endl;
    Input in;
    if (!read_pairs(args, in)) return EXIT_FAILURE;
    const int total = (int)in.qseq.size();
    size_t maxlen = 0;
    for (int i = 0; i < total; i++) maxlen = std::max(maxlen, std::max(in.qseq[i].size(), in.tseq[i].size()));
    // the reference sizes the query side from the last query (test_prog.cpp:146,227)
    const int max_q = total ? (int)in.qseq.back().size() : 0;

    const int n_threads = std::max(1, args->n_threads);
    std::vector<int> first(n_threads), count(n_threads);
    const int per = (total + n_threads - 1) / n_threads;
    for (int t = 0, a = 0; t < n_threads; t++) {
        first[t] = a;
        count[t] = std::max(0, std::min(per, total - a));
        a += count[t];
    }
    std::cerr << "Processing..." << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    omp_set_num_threads(n_threads);
    std::vector<gasal_gpu_storage_v> vecs(n_threads);
    for (int t = 0; t < n_threads; t++) {
        vecs[t] = gasal_init_gpu_storage_v(NB_STREAMS);
        gasal_init_streams(&vecs[t], max_q + 7, (int)maxlen + 7, STREAM_BATCH_SIZE, args);
    }

#pragma omp parallel
    {
        const int tid = omp_get_thread_num();
        gasal_gpu_storage_v &v = vecs[tid];
        const int n = count[tid];
        int next = first[tid];
        const int end = first[tid] + n;
        const int n_batches = (n + STREAM_BATCH_SIZE - 1) / STREAM_BATCH_SIZE;
        int done_batches = 0;
        std::vector<int> batch_start(v.n, 0), batch_n(v.n, 0);
        while (done_batches < n_batches) {
            int si = 0;
            while (si < v.n && v.a[si].is_free != 1) si++;
            if (next < end && si < v.n) {
                gasal_gpu_storage_t *s = &v.a[si];
                uint32_t qi = 0, ti = 0;
                int j = 0;
                for (; next < end && j < STREAM_BATCH_SIZE; next++, j++) {
                    s->current_n_alns++;
                    if (s->current_n_alns > s->host_max_n_alns) gasal_host_alns_resize(s, s->host_max_n_alns * 2, args);
                    s->host_query_batch_offsets[j] = qi;
                    s->host_target_batch_offsets[j] = ti;
                    qi = gasal_host_batch_fill(s, qi, in.qseq[next].c_str(), (uint32_t)in.qseq[next].size(), QUERY);
                    ti = gasal_host_batch_fill(s, ti, in.tseq[next].c_str(), (uint32_t)in.tseq[next].size(), TARGET);
                    s->host_query_batch_lens[j] = (uint32_t)in.qseq[next].size();
                    s->host_target_batch_lens[j] = (uint32_t)in.tseq[next].size();
                }
                gasal_op_fill(s, in.qmod.data() + next - j, (uint32_t)j, QUERY);
                gasal_op_fill(s, in.tmod.data() + next - j, (uint32_t)j, TARGET);
                batch_start[si] = next - j;
                batch_n[si] = j;
                gasal_aln_async(s, qi, ti, (uint32_t)j, args);
                s->current_n_alns = 0;
            }
            for (int k = 0; k < v.n; k++) {
                if (gasal_is_aln_async_done(&v.a[k]) == 0) {
                    if (args->print_out) {
#pragma omp critical
                        for (int j = 0; j < batch_n[k]; j++) print_result(args, in, &v.a[k], batch_start[k] + j, j);
                    }
                    done_batches++;
                }
            }
        }
    }
    for (int t = 0; t < n_threads; t++) {
        gasal_destroy_streams(&vecs[t], args);
        gasal_destroy_gpu_storage_v(&vecs[t]);
    }
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cerr << std::endl << "Done" << std::endl;
    fprintf(stderr, "Total execution time (in milliseconds): %.3f\n", ms);
    delete args;
    return 0;
}
