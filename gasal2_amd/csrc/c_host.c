/*
 * c_host.c -- a plain C host of the library: the reference's call sequence
 * (test_prog/test_prog.cpp:237-330 and README.md:99-213) written in C99 against
 * the public headers, so the C view of the API is compiled and run, not only
 * declared.
 *
 *   c_host [-p] [-n N_PENALTY] [-r OPS] ALGO START SECOND QUERY.fa TARGET.fa
 *
 * ALGO, START and SECOND are the enum values of algo_type, comp_start and
 * Bool.  -p hands the library 4-bit packed words (isPacked = true): each
 * sequence is packed here, the byte counts and offsets stay in bases.  -r
 * reads one query operation (operation_on_seq) per line and turns on
 * isReverseComplement.  Prints one line per pair:
 *
 *   k score query_end target_end query_start target_start second_score
 *   second_query_end second_target_end n_cigar_ops cigar_hex
 *
 * with "-" for outputs the mode does not produce; the CIGAR bytes are read
 * from host_res->cigar at the pair's query offset (reverse order, clipped to
 * its padded query length).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gasal_header.h"
#include "res.h"
#include "gasal2_band.h"

typedef struct {
    char **seq;
    uint32_t *len;
    uint32_t n, cap;
} seqs_t;

static void die(const char *msg) {
    fprintf(stderr, "c_host: %s\n", msg);
    exit(EXIT_FAILURE);
}

static void push(seqs_t *s, char *buf, uint32_t len) {
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->seq = (char **)realloc(s->seq, s->cap * sizeof(char *));
        s->len = (uint32_t *)realloc(s->len, s->cap * sizeof(uint32_t));
        if (!s->seq || !s->len) die("out of memory");
    }
    s->seq[s->n] = buf;
    s->len[s->n] = len;
    s->n++;
}

/* FASTA: '>' header lines, sequence lines concatenated. */
static void read_fasta(const char *path, seqs_t *out) {
    FILE *f = fopen(path, "rb");
    if (!f) die("cannot open a FASTA file");
    char line[1 << 16];
    char *cur = NULL;
    uint32_t len = 0, cap = 0;
    int have = 0;
    while (fgets(line, sizeof line, f)) {
        size_t l = strcspn(line, "\r\n");
        if (line[0] == '>') {
            if (have) push(out, cur, len);
            cur = NULL;
            len = cap = 0;
            have = 1;
            continue;
        }
        if (len + l + 1 > cap) {
            cap = 2 * (len + (uint32_t)l + 1);
            cur = (char *)realloc(cur, cap);
            if (!cur) die("out of memory");
        }
        memcpy(cur + len, line, l);
        len += (uint32_t)l;
        cur[len] = 0;
    }
    if (have) push(out, cur ? cur : (char *)calloc(1, 1), len);
    fclose(f);
}

/* One sequence as packed words (src/kernels/pack_rc_seqs.h:24-31): nibble =
 * byte & 0xF, first base in bits 31:28, padded with N_CODE to whole words. */
static uint32_t pack_seq(const char *s, uint32_t len, uint8_t n_code, uint32_t *words) {
    const uint32_t nw = (len + 7) / 8;
    for (uint32_t w = 0; w < nw; w++) {
        uint32_t v = 0;
        for (uint32_t i = 0; i < 8; i++) {
            const uint32_t b = 8 * w + i;
            const uint8_t c = b < len ? (uint8_t)s[b] : n_code;
            v |= (uint32_t)(c & 0xF) << (28 - 4 * i);
        }
        words[w] = v;
    }
    return nw;
}

int main(int argc, char **argv) {
    int packed = 0, has_np = 0, np = 0;
    const char *ops_path = NULL;
    int a = 1;
    for (; a < argc && argv[a][0] == '-'; a++) {
        if (!strcmp(argv[a], "-p")) packed = 1;
        else if (!strcmp(argv[a], "-n") && a + 1 < argc) { has_np = 1; np = atoi(argv[++a]); }
        else if (!strcmp(argv[a], "-r") && a + 1 < argc) ops_path = argv[++a];
        else die("usage: c_host [-p] [-n N_PENALTY] [-r OPS] ALGO START SECOND QUERY.fa TARGET.fa");
    }
    if (argc - a != 5) die("usage: c_host [-p] [-n N_PENALTY] [-r OPS] ALGO START SECOND QUERY.fa TARGET.fa");
    const int algo = atoi(argv[a]), start = atoi(argv[a + 1]), second = atoi(argv[a + 2]);
    seqs_t q = {0}, t = {0};
    read_fasta(argv[a + 3], &q);
    read_fasta(argv[a + 4], &t);
    if (q.n != t.n || q.n == 0) die("query and target files must hold the same, non-zero number of sequences");
    const uint32_t n = q.n;
    uint8_t *qops = (uint8_t *)calloc(n, 1), *tops = (uint8_t *)calloc(n, 1);
    if (ops_path) {
        FILE *f = fopen(ops_path, "r");
        if (!f) die("cannot open the operations file");
        for (uint32_t k = 0; k < n; k++) {
            int v = 0;
            if (fscanf(f, "%d", &v) != 1) die("short operations file");
            qops[k] = (uint8_t)v;
        }
        fclose(f);
    }

    /* Parameters through the C helpers (no C++ object layout needed) */
    Parameters *params = gasal2_parameters_new();
    gasal2_params_t pod;
    gasal2_parameters_get(params, &pod);
    pod.algo = algo;
    pod.start_pos = start;
    pod.secondBest = second;
    pod.isPacked = packed;
    pod.isReverseComplement = ops_path != NULL;
    gasal2_parameters_set(params, &pod);
    if (has_np) gasal2_set_n_penalty(1, np);
    const uint8_t n_code = (uint8_t)gasal2_get_n_code();

    gasal_subst_scores sc;
    sc.match = pod.sa;
    sc.mismatch = pod.sb;
    sc.gap_open = pod.gapo;
    sc.gap_extend = pod.gape;
    gasal_copy_subst_scores(&sc);

    uint32_t maxq = 0, maxt = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (q.len[k] > maxq) maxq = q.len[k];
        if (t.len[k] > maxt) maxt = t.len[k];
    }
    gasal_gpu_storage_v v = gasal_init_gpu_storage_v(1);
    gasal_init_streams(&v, (int)maxq + 15, (int)maxt + 15, (int)n, params);
    gasal_gpu_storage_t *s = &v.a[0];

    /* host batch: gasal_host_batch_fill per sequence (src/host_batch.cpp:79-153) */
    uint32_t *words = (uint32_t *)malloc(((maxq > maxt ? maxq : maxt) / 8 + 2) * sizeof(uint32_t));
    uint32_t qi = 0, ti = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (packed) {
            s->host_query_batch_offsets[k] = 2 * qi;
            s->host_target_batch_offsets[k] = 2 * ti;
            uint32_t nw = pack_seq(q.seq[k], q.len[k], n_code, words);
            qi = gasal_host_batch_fill(s, qi, (const char *)words, 4 * nw, QUERY);
            nw = pack_seq(t.seq[k], t.len[k], n_code, words);
            ti = gasal_host_batch_fill(s, ti, (const char *)words, 4 * nw, TARGET);
        } else {
            s->host_query_batch_offsets[k] = qi;
            s->host_target_batch_offsets[k] = ti;
            qi = gasal_host_batch_fill(s, qi, q.seq[k], q.len[k], QUERY);
            ti = gasal_host_batch_fill(s, ti, t.seq[k], t.len[k], TARGET);
        }
        s->host_query_batch_lens[k] = q.len[k];
        s->host_target_batch_lens[k] = t.len[k];
    }
    const uint32_t qb = packed ? 2 * qi : qi, tb = packed ? 2 * ti : ti;
    if (ops_path) {
        gasal_op_fill(s, qops, n, QUERY);
        gasal_op_fill(s, tops, n, TARGET);
    }
    s->current_n_alns = n;

    gasal_aln_async(s, qb, tb, n, params);
    for (;;) {
        int r = gasal_is_aln_async_done(s);
        if (r == 0) break;
        if (r != -1) die("gasal_is_aln_async_done failed");
    }

    const gasal_res_t *h = s->host_res, *h2 = s->host_res_second;
    for (uint32_t k = 0; k < n; k++) {
        printf("%u", k);
        const int32_t *cols[8] = {h->aln_score, h->query_batch_end, h->target_batch_end, h->query_batch_start,
                                  h->target_batch_start, h2 ? h2->aln_score : NULL, h2 ? h2->query_batch_end : NULL,
                                  h2 ? h2->target_batch_end : NULL};
        for (int c = 0; c < 8; c++) {
            if (cols[c]) printf(" %d", cols[c][k]);
            else printf(" -");
        }
        if (start == WITH_TB && h->n_cigar_ops && h->cigar) {
            const uint32_t nops = h->n_cigar_ops[k];
            const uint32_t slot = 8 * ((q.len[k] + 7) / 8);
            const uint32_t nb = nops < slot ? nops : slot;
            printf(" %u ", nops);
            const uint8_t *cg = h->cigar + s->host_query_batch_offsets[k];
            for (uint32_t b = 0; b < nb; b++) printf("%02x", cg[b]);
        } else {
            printf(" - -");
        }
        printf("\n");
    }

    gasal_destroy_streams(&v, params);
    gasal_destroy_gpu_storage_v(&v);
    gasal2_parameters_free(params);
    for (uint32_t k = 0; k < n; k++) {
        free(q.seq[k]);
        free(t.seq[k]);
    }
    free(q.seq); free(q.len); free(t.seq); free(t.len);
    free(words); free(qops); free(tops);
    return 0;
}
