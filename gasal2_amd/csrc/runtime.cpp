// runtime.cpp -- host side of the drop-in API (include/*.h).
//
// Behaviour follows the reference's host runtime: storage construction
// (src/ctors.cpp), paged pinned host batches (src/host_batch.cpp), result
// buffers (src/res.cpp), resizing/op fill/device selection (src/interfaces.cpp),
// Parameters (src/args_parser.cpp) and the asynchronous launcher
// (src/gasal_align.cu).  The device work itself is g2_run() (kernels.hip).
#include "gasal_header.h"
#include "res.h"
#include "engine.h"

#include <string.h>
#include <stdarg.h>
#include <math.h>
#include <string>
#include <vector>
#include <atomic>
#include <mutex>
#include <algorithm>

// ------------------------------------------------------------ globals ------
// An order chain (the batch order below).
struct gasal2_order {
    std::mutex mu;
    std::atomic<int> mode{0};
    hipEvent_t last_h2d = nullptr, last_dev = nullptr;
};

namespace {

G2Scores g_scores = {0, 0, 0, 0, 0, (uint32_t)(N_CODE & 0xF), N_CODE,
#ifdef N_PENALTY
                     1, N_PENALTY
#else
                     0, 0
#endif
};

int g_error_mode = 0;
int g_error_count = 0;
// one-wave-per-pair thresholds: target words, cells (64 per query word x target word)
uint32_t g_wf_min_tw = 48, g_wf_min_cells = 1u << 20, g_wf_max_blocks = 2048;
bool g_wf_user = false;   // set by gasal2_set_wavefront: then it holds for every mode
// WITH_TB default (measured on the uneven batch: the traceback's per-lane
// tiles favour sending more pairs to one wave each)
constexpr uint32_t WF_TB_MIN_TW = 32, WF_TB_MIN_CELLS = 1u << 18;
// Small batches: the lane path runs 64 or 128 pairs per wave, so a batch of a
// few pairs per CU leaves most SIMDs idle and takes one pair's serial time;
// up to this many pairs per CU every pair runs one wave (auto policy only).
// Measured on 150 x 250-base pairs: the one-wave path wins below about 160
// pairs per CU for score-only LOCAL and below about 86 for WITH_TB.
constexpr uint32_t WF_FEW_PER_CU = 128, WF_FEW_PER_CU_TB = 64;
// compute units of the current device (cached per device)
uint32_t g2_cu_count() {
    static std::atomic<uint32_t> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    uint32_t v = cache[dev].load(std::memory_order_relaxed);
    if (!v) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) return 0;
        v = (uint32_t)cu;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}
int g_graph_mode = 1;   // capture the device program per batch shape and replay it
uint64_t g_tb_one_size_max = 64ull << 30;   // lane-path tiles: one size for all groups up to this many bytes
int g_tb_ckpt = 1;   // checkpointed traceback: 1 the fast LOCAL WITH_TB lane path, 2 also one-wave pairs
int g_tb_fused = 0;  // checkpointed lane path: the DP kernel walks each pair's traceback itself
int g_panels = 1;    // transposed sequence panels for the fast LOCAL lane path
// Batch order across storages (default off, the reference's independent
// streams).  The storages on one order chain submit in order: mode 1 makes
// each gasal_aln_async's H2D wait for the previously submitted batch's H2D,
// mode 2 also its device program for that batch's program.  The events belong
// to the storages; a chain remembers the last ones.  Storages not attached to
// a chain of their own (gasal2_storage_set_order) use the process-wide chain
// (gasal2_set_batch_order).
gasal2_order g_chain;
int g_pk16_mode = 1;    // packed 16-bit kernels (and KSW's packed row) when values fit int16; 2: two pairs per lane for every LOCAL mode
int32_t g_ksw_w = -1, g_ksw_zdrop = 0;   // KSW extension (reference: no band, z-drop 0)
int32_t g_ksw_eb = 5;                    // KSW end bonus (reference: PEN_CLIP5)

// Tuned switch thresholds per (algo, start, second best, mean query length
// range); filled by the autotuner (gasal2_tune_set), consulted per batch.
struct TuneEntry {
    int32_t algo, start, second;
    uint32_t lq_lo, lq_hi;   // mean query length in [lq_lo, lq_hi)
    uint32_t min_tw, min_cells, max_blocks;
};
std::vector<TuneEntry> g_tune;
std::mutex g_tune_mu;
std::string g_last_error;
std::mutex g_err_mu;

struct Impl {
    uint32_t *keys = nullptr, *perm = nullptr;
    size_t n_cap = 0;
    uint32_t *hist = nullptr;
    uint32_t *bnd = nullptr;
    size_t bnd_bytes = 0;
    uint32_t *tb = nullptr;
    size_t tb_bytes = 0;
    int32_t *eh = nullptr;
    size_t eh_bytes = 0;
    uint32_t *nops = nullptr;
    size_t nops_cap = 0;
    uint8_t *cigar = nullptr;
    size_t cigar_bytes = 0;
    uint8_t *cigar_q = nullptr;
    size_t cigar_q_bytes = 0;
    int32_t *scratch_i32 = nullptr;  // stand-ins for result arrays the mode does not expose
    size_t scratch_i32_cap = 0;
    uint32_t *aux = nullptr;
    uint32_t *wf = nullptr;          // strip state of the one-wave-per-pair path
    size_t wf_bytes = 0;
    // last batch
    uint32_t last_n = 0, last_max_q = 0, last_max_t = 0, last_cigar_cap = 0;
    bool last_tb = false;
    G2Timing tm;
    bool tm_init = false;
    // batch order across storages (gasal2_set_batch_order): this storage's
    // last H2D and last device program, for the next submitted batch to wait on
    hipEvent_t ev_h2d = nullptr, ev_dev = nullptr;
    gasal2_order *chain = nullptr;   // nullptr: the process-wide chain
    // KSW settings of this storage (gasal2_storage_set_ksw); off: the process-wide ones
    bool ksw_own = false;
    int32_t ksw_w = -1, ksw_zdrop = 0, ksw_eb = 5;
    // captured device program of the last batch shape (see run_async)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    G2Dev g_dev;
    G2Scores g_sc;
    G2Opts g_op;
    bool g_valid = false;
    uint64_t graph_launches = 0, graph_captures = 0;
    // banded runs (gasal2_run_banded) and sub-batches (gasal2_run_subset)
    int32_t *band_lo = nullptr, *band_hi = nullptr;
    uint32_t *band_bw = nullptr;
    uint64_t *band_tbo = nullptr;
    uint8_t *band_flags = nullptr;
    size_t band_cap = 0;
    int32_t *band_wres = nullptr;
    size_t band_wres_bytes = 0;
    uint64_t *tb_gbase = nullptr;    // per-group tile regions of the lane path
    uint32_t *tb_gqw = nullptr;
    size_t tb_gbase_bytes = 0, tb_gqw_bytes = 0;
    std::vector<uint32_t> h_hist, h_gqw;
    std::vector<uint64_t> h_gbase;
    uint32_t *qpan = nullptr, *tpan = nullptr;   // transposed sequence panels of the lane path
    size_t qpan_bytes = 0, tpan_bytes = 0;
    uint32_t *wf_tb = nullptr;       // direction tiles of one-wave-per-pair pairs
    size_t wf_tb_bytes = 0;
    uint64_t *wf_tbo = nullptr;
    size_t wf_tbo_bytes = 0;
    std::vector<uint64_t> h_wf_tbo;
    std::vector<int32_t> h_lo, h_hi;
    std::vector<uint32_t> h_bw;
    std::vector<uint64_t> h_tbo;
    uint32_t *sub_u32 = nullptr;     // qoff, toff, qlen, tlen, idx: 5 x sub_cap
    int32_t *sub_res = nullptr;      // score, qend, tend, qstart, tstart: 5 x sub_cap
    uint32_t *sub_nops = nullptr;
    uint8_t *sub_cigar = nullptr;
    size_t sub_cap = 0, sub_cigar_bytes = 0;
    std::vector<uint32_t> h_sub;     // 6 x chunk: qoff, toff, qlen, tlen, idx (+ spare)
    G2Dev main_dev;                  // device view of the last full or banded run
    bool main_valid = false;
    bool prepped = false;            // resident batch packed/reverse-complemented by gasal2_prep_resident
    int32_t last_paths = 0;          // gasal2_last_paths of the last launched batch
};

Impl *impl_of(gasal_gpu_storage_t *s) {
    if (!s->impl) s->impl = new Impl();
    return (Impl *)s->impl;
}

void report(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    fputs(buf, stderr);
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error = buf;
    g_error_count++;
}

}  // namespace

// Error exit (src/gasal.h:20).  In error mode 1 it returns; callers return too.
extern "C" void gasal2_fail(int code) {
    if (g_error_mode == 0) exit(code);
}

#define G2_FAIL_RET(...)                 \
    do {                                 \
        report(__VA_ARGS__);             \
        gasal2_fail(EXIT_FAILURE);       \
        return;                          \
    } while (0)

// A failed HIP call reports and leaves through gasal2_fail (exit in error
// mode 0).  In error mode 1 it also raises t_failed, which the entry points
// test between their stages so that nothing is launched or copied after a
// failed allocation or copy.
namespace {
thread_local bool t_failed = false;
}

#define G2_CHECK(call)                                                                                  \
    do {                                                                                                \
        hipError_t e__ = (call);                                                                        \
        if (e__ != hipSuccess) {                                                                        \
            report("[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n",              \
                   hipGetErrorString(e__), (int)e__, __LINE__, __FILE__);                               \
            t_failed = true;                                                                            \
            gasal2_fail(EXIT_FAILURE);                                                                  \
        }                                                                                               \
    } while (0)

extern "C" void gasal2_set_error_mode(int32_t mode) { g_error_mode = mode; }
extern "C" int32_t gasal2_error_count(void) { return g_error_count; }
extern "C" const char *gasal2_last_error(void) { return g_last_error.c_str(); }
extern "C" void gasal2_set_n_code(int32_t n_code) { g_scores.n_code = n_code; g_scores.n_value = (uint32_t)n_code & 0xFu; }
extern "C" int32_t gasal2_get_n_code(void) { return g_scores.n_code; }
extern "C" void gasal2_set_n_penalty(int32_t has_penalty, int32_t penalty) {
    g_scores.has_np = has_penalty ? 1 : 0;
    g_scores.np = penalty;
}
extern "C" void gasal2_set_dispatch(int32_t policy) { g2_set_dispatch(policy); }
extern "C" int32_t gasal2_get_dispatch(void) { return g2_get_dispatch(); }
extern "C" void gasal2_tune_set(int32_t algo, int32_t start, int32_t second, uint32_t lq_lo, uint32_t lq_hi,
                                uint32_t min_target_words, uint32_t min_cells, uint32_t max_blocks) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (TuneEntry &e : g_tune)
        if (e.algo == algo && e.start == start && e.second == (second ? 1 : 0) && e.lq_lo == lq_lo && e.lq_hi == lq_hi) {
            e.min_tw = min_target_words ? min_target_words : 1u; e.min_cells = min_cells; e.max_blocks = max_blocks ? max_blocks : 1u;
            return;
        }
    g_tune.push_back({algo, start, second ? 1 : 0, lq_lo, lq_hi, min_target_words ? min_target_words : 1u, min_cells,
                      max_blocks ? max_blocks : 1u});
}
extern "C" void gasal2_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}
extern "C" int32_t gasal2_tune_count(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int32_t)g_tune.size();
}

extern "C" void gasal2_set_tile_budget(uint64_t bytes) { g_tb_one_size_max = bytes; }
extern "C" void gasal2_set_tb_checkpoint(int32_t mode) { g_tb_ckpt = mode < 0 ? 0 : (mode > 2 ? 2 : mode); }
extern "C" void gasal2_set_tb_fused(int32_t on) { g_tb_fused = on ? 1 : 0; }
extern "C" void gasal2_set_panels(int32_t on) { g_panels = on ? 1 : 0; }
extern "C" int32_t gasal2_last_paths(gasal_gpu_storage_t *s) { return s && s->impl ? impl_of(s)->last_paths : 0; }
extern "C" void gasal2_set_batch_order(int32_t on) {
    std::lock_guard<std::mutex> lk(g_chain.mu);
    g_chain.mode = on < 0 ? 0 : (on > 2 ? 2 : on);
    g_chain.last_h2d = g_chain.last_dev = nullptr;
}
extern "C" gasal2_order_t *gasal2_order_new(int32_t mode) {
    gasal2_order *o = new (std::nothrow) gasal2_order();
    if (o) o->mode = mode < 0 ? 0 : (mode > 2 ? 2 : mode);
    return o;
}
extern "C" void gasal2_set_pk16(int32_t mode) { g_pk16_mode = mode < 0 ? 0 : (mode > 2 ? 2 : mode); }
extern "C" void gasal2_set_ksw_extension(int32_t w, int32_t zdrop) {
    g_ksw_w = w;
    g_ksw_zdrop = zdrop > 0 ? zdrop : 0;
}
extern "C" void gasal2_set_ksw_end_bonus(int32_t end_bonus) {
    g_ksw_eb = end_bonus;
}

extern "C" void gasal2_set_graph_mode(int32_t on) { g_graph_mode = on ? 1 : 0; }
extern "C" int32_t gasal2_get_graph_mode(void) { return g_graph_mode; }
extern "C" void gasal2_set_wavefront(uint32_t min_target_words, uint32_t min_cells, uint32_t max_blocks) {
    if (!min_target_words && !min_cells && !max_blocks) {   // back to the per-mode defaults
        g_wf_min_tw = 48;
        g_wf_min_cells = 1u << 20;
        g_wf_max_blocks = 2048;
        g_wf_user = false;
        return;
    }
    g_wf_min_tw = min_target_words ? min_target_words : 1u;
    g_wf_min_cells = min_cells;
    g_wf_max_blocks = max_blocks ? max_blocks : 1u;
    g_wf_user = true;
}

extern "C" int32_t gasal2_device_info(int32_t *n_cu, int32_t *clock_khz, char *name, int32_t name_cap) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (name && name_cap > 0) {
        // marketing name when the runtime reports one, else the ISA name (gfx950...)
        const char *src = prop.name[0] ? prop.name : prop.gcnArchName;
        strncpy(name, src, (size_t)name_cap - 1);
        name[name_cap - 1] = 0;
    }
    return 0;
}

// ------------------------------------------------------------ Parameters ---
// src/args_parser.cpp:8-39 (defaults), :41-44, :46-52, :54-75, :77-95, :97-240, :242-250
Parameters::Parameters(int argc_, char **argv_) {
    sa = 1; sb = 4; gapo = 6; gape = 1;
    start_pos = WITHOUT_START;
    print_out = 0;
    n_threads = 1;
    k_band = 0;
    isPacked = false;
    isReverseComplement = false;
    secondBest = FALSE;
    semiglobal_skipping_head = TARGET;
    semiglobal_skipping_tail = TARGET;
    algo = UNKNOWN;
    query_batch_fasta_filename = "";
    target_batch_fasta_filename = "";
    argc = argc_;
    argv = argv_;
}

Parameters::~Parameters() {
    query_batch_fasta.close();
    target_batch_fasta.close();
}

void Parameters::print() {
    std::cerr << "sa=" << sa << " , sb=" << sb << " , gapo=" << gapo << " , gape=" << gape << std::endl;
    std::cerr << "start_pos=" << start_pos << " , print_out=" << print_out << " , n_threads=" << n_threads << std::endl;
    std::cerr << "semiglobal_skipping_head=" << semiglobal_skipping_head
              << " , semiglobal_skipping_tail=" << semiglobal_skipping_tail << " , algo=" << algo << std::endl;
    std::cerr << std::boolalpha << "isPacked = " << isPacked << " , secondBest = " << secondBest << std::endl;
    std::cerr << "query_batch_fasta_filename=" << query_batch_fasta_filename
              << " , target_batch_fasta_filename=" << target_batch_fasta_filename << std::endl;
}

void Parameters::failure(fail_type f) {
    switch (f) {
        case NOT_ENOUGH_ARGS:
            std::cerr << "Not enough Parameters. Required: -y AL_TYPE file1.fasta file2.fasta. See help (--help, -h) for usage. " << std::endl;
            break;
        case WRONG_ARG:
            std::cerr << "Wrong argument. See help (--help, -h) for usage. " << std::endl;
            break;
        case WRONG_FILES:
            std::cerr << "File error: either a file doesn't exist, or cannot be opened." << std::endl;
            break;
        default:
            break;
    }
    exit(1);
}

void Parameters::help() {
    std::cerr << "Usage: ./test_prog.out [-a] [-b] [-q] [-r] [-s] [-t] [-p] [-n] [-y] <query_batch.fasta> <target_batch.fasta>" << std::endl;
    std::cerr << "Options: -a INT    match score [" << sa << "]" << std::endl;
    std::cerr << "         -b INT    mismatch penalty [" << sb << "]" << std::endl;
    std::cerr << "         -q INT    gap open penalty [" << gapo << "]" << std::endl;
    std::cerr << "         -r INT    gap extension penalty [" << gape << "]" << std::endl;
    std::cerr << "         -s        find the start position" << std::endl;
    std::cerr << "         -t        compute traceback. With this option enabled, \"-s\" has no effect as start position will always be computed with traceback" << std::endl;
    std::cerr << "         -p        print the alignment results" << std::endl;
    std::cerr << "         -n INT    Number of threads [" << n_threads << "]" << std::endl;
    std::cerr << "         -y AL_TYPE       Alignment type . Must be \"local\", \"semi_global\", \"global\", \"ksw\" " << std::endl;
    std::cerr << "         -x HEAD TAIL     specifies, for semi-global alignment, wha should be skipped for heads and tails of the sequences. (NONE, QUERY, TARGET, BOTH)" << std::endl;
    std::cerr << "         -k INT    Band width in case \"banded\" is selected." << std::endl;
    std::cerr << "         --help, -h : displays this message." << std::endl;
    std::cerr << "         --second-best   displays second best score (WITHOUT_START only)." << std::endl;
    std::cerr << "Single-pack multi-Parameters (e.g. -sp) is not supported." << std::endl;
    std::cerr << "		  " << std::endl;
}

static bool parse_source(const std::string &s, data_source &out) {
    if (s == "NONE") out = NONE;
    else if (s == "TARGET") out = TARGET;
    else if (s == "QUERY") out = QUERY;
    else if (s == "BOTH") out = BOTH;
    else return false;
    return true;
}

void Parameters::parse() {
    for (int c = 1; c < argc; c++) {
        std::string a(argv[c]);
        if (a == "--help" || a == "-h") { help(); exit(0); }
    }
    if (argc < 4) failure(NOT_ENOUGH_ARGS);
    int c;
    for (c = 1; c < argc - 2; c++) {
        std::string cur(argv[c]);
        if (cur.size() >= 2 && cur[0] == '-' && cur[1] == '-') {
            if (cur == "--help") { help(); exit(0); }
            if (cur == "--second-best") secondBest = TRUE;
        } else if (!cur.empty() && cur[0] == '-') {
            if (cur.size() > 2) failure(WRONG_ARG);
            char param = cur.size() > 1 ? cur[1] : 0;
            auto next = [&]() -> std::string { c++; return c < argc ? std::string(argv[c]) : std::string(); };
            switch (param) {
                case 'y': {
                    std::string v = next();
                    if (v == "local") algo = LOCAL;
                    else if (v == "semi_global") algo = SEMI_GLOBAL;
                    else if (v == "global") algo = GLOBAL;
                    else if (v == "ksw") algo = KSW;
                    break;
                }
                case 'a': sa = std::stoi(next()); break;
                case 'b': sb = std::stoi(next()); break;
                case 'q': gapo = std::stoi(next()); break;
                case 'r': gape = std::stoi(next()); break;
                case 's': start_pos = WITH_START; break;
                case 't': start_pos = WITH_TB; break;
                case 'p': print_out = 1; break;
                case 'n': n_threads = std::stoi(next()); break;
                case 'k': k_band = std::stoi(next()); break;
                case 'x': {
                    if (!parse_source(next(), semiglobal_skipping_head)) failure(WRONG_ARG);
                    if (!parse_source(next(), semiglobal_skipping_tail)) failure(WRONG_ARG);
                    break;
                }
                default: break;
            }
        } else {
            failure(WRONG_ARG);
        }
    }
    query_batch_fasta_filename = std::string(argv[c]);
    c++;
    target_batch_fasta_filename = std::string(argv[c]);
    fileopen();
}

void Parameters::fileopen() {
    query_batch_fasta.open(query_batch_fasta_filename, std::ifstream::in);
    if (!query_batch_fasta) failure(WRONG_FILES);
    target_batch_fasta.open(target_batch_fasta_filename);
    if (!target_batch_fasta) failure(WRONG_FILES);
}

extern "C" Parameters *gasal2_parameters_new(void) { return new Parameters(0, nullptr); }
extern "C" void gasal2_parameters_free(Parameters *p) { delete p; }
extern "C" void gasal2_parameters_get(const Parameters *p, gasal2_params_t *o) {
    o->sa = p->sa; o->sb = p->sb; o->gapo = p->gapo; o->gape = p->gape;
    o->start_pos = p->start_pos; o->print_out = p->print_out; o->n_threads = p->n_threads;
    o->k_band = p->k_band; o->secondBest = p->secondBest; o->isPacked = p->isPacked ? 1 : 0;
    o->isReverseComplement = p->isReverseComplement ? 1 : 0;
    o->semiglobal_skipping_head = p->semiglobal_skipping_head;
    o->semiglobal_skipping_tail = p->semiglobal_skipping_tail;
    o->algo = p->algo;
}
extern "C" void gasal2_parameters_set(Parameters *p, const gasal2_params_t *i) {
    p->sa = i->sa; p->sb = i->sb; p->gapo = i->gapo; p->gape = i->gape;
    p->start_pos = (comp_start)i->start_pos; p->print_out = i->print_out; p->n_threads = i->n_threads;
    p->k_band = i->k_band; p->secondBest = (Bool)(i->secondBest ? 1 : 0); p->isPacked = i->isPacked != 0;
    p->isReverseComplement = i->isReverseComplement != 0;
    p->semiglobal_skipping_head = (data_source)i->semiglobal_skipping_head;
    p->semiglobal_skipping_tail = (data_source)i->semiglobal_skipping_tail;
    p->algo = (algo_type)i->algo;
}

// ------------------------------------------------------------ host batch ---
// src/host_batch.cpp:11-22
extern "C" host_batch_t *gasal_host_batch_new(uint32_t batch_bytes, uint32_t offset) {
    host_batch_t *res = (host_batch_t *)calloc(1, sizeof(host_batch_t));
    G2_CHECK(hipHostMalloc((void **)&(res->data), (size_t)batch_bytes * sizeof(uint8_t), hipHostMallocDefault));
    res->page_size = batch_bytes;
    res->data_size = 0;
    res->is_locked = 0;
    res->offset = offset;
    res->next = NULL;
    return res;
}

// src/host_batch.cpp:24-41
extern "C" void gasal_host_batch_destroy(host_batch_t *res) {
    if (res == NULL) {
        report("[GASAL ERROR] Trying to free a NULL pointer\n");
        gasal2_fail(1);
        return;
    }
    while (res) {
        host_batch_t *nx = res->next;
        if (res->data != NULL) G2_CHECK(hipHostFree(res->data));
        free(res);
        res = nx;
    }
}

// src/host_batch.cpp:43-47
extern "C" host_batch_t *gasal_host_batch_getlast(host_batch_t *arg) {
    while (arg->next != NULL) arg = arg->next;
    return arg;
}

// src/host_batch.cpp:49-75
extern "C" void gasal_host_batch_reset(gasal_gpu_storage_t *gpu_storage) {
    host_batch_t *lists[2] = {gpu_storage->extensible_host_unpacked_query_batch,
                              gpu_storage->extensible_host_unpacked_target_batch};
    for (int i = 0; i < 2; i++)
        for (host_batch_t *p = lists[i]; p != NULL; p = p->next) {
            p->data_size = 0;
            p->offset = 0;
            p->is_locked = 0;
        }
}

// src/host_batch.cpp:79-153
extern "C" uint32_t gasal_host_batch_fill(gasal_gpu_storage_t *gpu_storage, uint32_t idx, const char *data, uint32_t size,
                                          data_source SRC) {
    host_batch_t *cur_page = NULL;
    uint32_t *p_batch_bytes = NULL;
    if (SRC == QUERY) {
        cur_page = gpu_storage->extensible_host_unpacked_query_batch;
        p_batch_bytes = &(gpu_storage->host_max_query_batch_bytes);
    } else if (SRC == TARGET) {
        cur_page = gpu_storage->extensible_host_unpacked_target_batch;
        p_batch_bytes = &(gpu_storage->host_max_target_batch_bytes);
    } else {
        return idx;
    }
    uint32_t nbr_N = 0;
    while ((size + nbr_N) % 8) nbr_N++;
    while (cur_page->is_locked) cur_page = cur_page->next;
    if (cur_page->next == NULL && cur_page->page_size - cur_page->data_size < size + nbr_N) {
        fprintf(stderr,
                "[GASAL WARNING:] Trying to write %d bytes while only %d remain (%s) (block size %d, filled %d bytes).\n"
                "                 Allocating a new block of size %d, total size available reaches %d. Doing this repeadtedly slows down the execution.\n",
                size + nbr_N, cur_page->page_size - cur_page->data_size, (SRC == QUERY ? "query" : "target"),
                cur_page->page_size, cur_page->data_size, cur_page->page_size * 2, *p_batch_bytes + cur_page->page_size * 2);
        uint32_t newsize = cur_page->page_size * 2;
        while (newsize < size + nbr_N) newsize *= 2;
        host_batch_t *res = gasal_host_batch_new(newsize, cur_page->offset + cur_page->data_size);
        cur_page->next = res;
        cur_page->is_locked = 1;
        *p_batch_bytes = *p_batch_bytes + newsize;
        cur_page = cur_page->next;
    }
    if (cur_page->next != NULL && cur_page->page_size - cur_page->data_size < size + nbr_N) {
        cur_page->next->offset = cur_page->offset + cur_page->data_size;
        cur_page->is_locked = 1;
        cur_page = cur_page->next;
    }
    if (cur_page->page_size - cur_page->data_size >= size + nbr_N) {
        memcpy(&(cur_page->data[idx - cur_page->offset]), data, size);
        memset(&(cur_page->data[idx + size - cur_page->offset]), g_scores.n_code, nbr_N);
        idx = idx + size + nbr_N;
        cur_page->data_size += size + nbr_N;
    }
    return idx;
}

// src/host_batch.cpp:162-223
extern "C" uint32_t gasal_host_batch_add(gasal_gpu_storage_t *gpu_storage, uint32_t idx, const char *data, uint32_t size,
                                         data_source SRC) {
    host_batch_t *cur_page = NULL;
    uint32_t *p_batch_bytes = NULL;
    if (SRC == QUERY) {
        cur_page = gpu_storage->extensible_host_unpacked_query_batch;
        p_batch_bytes = &(gpu_storage->host_max_query_batch_bytes);
    } else if (SRC == TARGET) {
        cur_page = gpu_storage->extensible_host_unpacked_target_batch;
        p_batch_bytes = &(gpu_storage->host_max_target_batch_bytes);
    } else {
        return idx;
    }
    for (;;) {
        if (*p_batch_bytes >= idx + size && (cur_page->next == NULL || cur_page->next->offset >= idx + size)) {
            memcpy(&(cur_page->data[idx - cur_page->offset]), data, size);
            idx = idx + size;
            if (idx - cur_page->offset > cur_page->data_size) cur_page->data_size = idx - cur_page->offset;
            return idx;
        } else if (*p_batch_bytes >= idx + size && cur_page->next != NULL && cur_page->next->offset < idx + size) {
            cur_page = cur_page->next;
        } else {
            fprintf(stderr,
                    "[GASAL WARNING:] Trying to write %d bytes at position %d on host memory (%s) while only  %d bytes are available. Therefore, allocating %d bytes more on CPU. Repeating this many times can provoke a degradation of performance.\n",
                    size, idx, (SRC == QUERY ? "query" : "target"), *p_batch_bytes, *p_batch_bytes * 2);
            *p_batch_bytes += *p_batch_bytes;
            while (*p_batch_bytes < size) *p_batch_bytes += *p_batch_bytes;
            host_batch_t *res = gasal_host_batch_new(*p_batch_bytes, idx);
            // appended after the last page: the reference links it after the
            // page the walk stopped at (src/host_batch.cpp:211-213), which
            // drops the pages after it and the bases written there
            while (cur_page->next != NULL) cur_page = cur_page->next;
            cur_page->next = res;
            cur_page = cur_page->next;
        }
    }
}

// src/host_batch.cpp:156-159
extern "C" uint32_t gasal_host_batch_addbase(gasal_gpu_storage_t *gpu_storage, uint32_t idx, const char base,
                                             data_source SRC) {
    return gasal_host_batch_add(gpu_storage, idx, &base, 1, SRC);
}

// src/host_batch.cpp:228-244
extern "C" void gasal_host_batch_print(host_batch_t *res) {
    fprintf(stderr, "[GASAL PRINT] Page data: offset=%d, next_offset=%d, data size=%d, page size=%d\n", res->offset,
            (res->next != NULL ? (int)res->next->offset : -1), res->data_size, res->page_size);
}
extern "C" void gasal_host_batch_printall(host_batch_t *res) {
    for (; res != NULL; res = res->next) {
        gasal_host_batch_print(res);
        if (res->next != NULL) fprintf(stderr, "+--->");
    }
}

// ------------------------------------------------------------ results ------
// src/res.cpp:8-73
extern "C" gasal_res_t *gasal_res_new_host(uint32_t max_n_alns, Parameters *params) {
    gasal_res_t *res = (gasal_res_t *)calloc(1, sizeof(gasal_res_t));
    G2_CHECK(hipHostMalloc((void **)&(res->aln_score), (size_t)max_n_alns * sizeof(int32_t), hipHostMallocDefault));
    if (params->algo != GLOBAL) {
        if (params->start_pos == WITH_START || params->start_pos == WITH_TB) {
            G2_CHECK(hipHostMalloc((void **)&(res->query_batch_start), (size_t)max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
            G2_CHECK(hipHostMalloc((void **)&(res->target_batch_start), (size_t)max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
        }
        G2_CHECK(hipHostMalloc((void **)&(res->query_batch_end), (size_t)max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
        G2_CHECK(hipHostMalloc((void **)&(res->target_batch_end), (size_t)max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
    }
    if (params->start_pos == WITH_TB)
        G2_CHECK(hipHostMalloc((void **)&(res->n_cigar_ops), (size_t)max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
    return res;
}

// src/res.cpp:76-101
extern "C" gasal_res_t *gasal_res_new_device(gasal_res_t *device_cpy) {
    gasal_res_t *d_c = NULL;
    G2_CHECK(hipMalloc((void **)&d_c, sizeof(gasal_res_t)));
    G2_CHECK(hipMemcpy(d_c, device_cpy, sizeof(gasal_res_t), hipMemcpyHostToDevice));
    return d_c;
}

// src/res.cpp:106-140
extern "C" gasal_res_t *gasal_res_new_device_cpy(uint32_t max_n_alns, Parameters *params) {
    gasal_res_t *res = (gasal_res_t *)calloc(1, sizeof(gasal_res_t));
    G2_CHECK(hipMalloc((void **)&(res->aln_score), (size_t)max_n_alns * sizeof(int32_t)));
    if (params->algo != GLOBAL) {
        if (params->start_pos == WITH_START || params->start_pos == WITH_TB) {
            G2_CHECK(hipMalloc((void **)&(res->query_batch_start), (size_t)max_n_alns * sizeof(uint32_t)));
            G2_CHECK(hipMalloc((void **)&(res->target_batch_start), (size_t)max_n_alns * sizeof(uint32_t)));
        }
        G2_CHECK(hipMalloc((void **)&(res->query_batch_end), (size_t)max_n_alns * sizeof(uint32_t)));
        G2_CHECK(hipMalloc((void **)&(res->target_batch_end), (size_t)max_n_alns * sizeof(uint32_t)));
    }
    return res;
}

// src/res.cpp:143-158
extern "C" void gasal_res_destroy_host(gasal_res_t *res) {
    if (res == NULL) return;
    if (res->aln_score) G2_CHECK(hipHostFree(res->aln_score));
    if (res->query_batch_start) G2_CHECK(hipHostFree(res->query_batch_start));
    if (res->target_batch_start) G2_CHECK(hipHostFree(res->target_batch_start));
    if (res->query_batch_end) G2_CHECK(hipHostFree(res->query_batch_end));
    if (res->target_batch_end) G2_CHECK(hipHostFree(res->target_batch_end));
    if (res->n_cigar_ops) G2_CHECK(hipHostFree(res->n_cigar_ops));
    if (res->cigar) G2_CHECK(hipHostFree(res->cigar));
    free(res);
}

// src/res.cpp:160-177
extern "C" void gasal_res_destroy_device(gasal_res_t *device_res, gasal_res_t *device_cpy) {
    if (device_cpy == NULL || device_res == NULL) return;
    if (device_cpy->aln_score) G2_CHECK(hipFree(device_cpy->aln_score));
    if (device_cpy->query_batch_start) G2_CHECK(hipFree(device_cpy->query_batch_start));
    if (device_cpy->target_batch_start) G2_CHECK(hipFree(device_cpy->target_batch_start));
    if (device_cpy->query_batch_end) G2_CHECK(hipFree(device_cpy->query_batch_end));
    if (device_cpy->target_batch_end) G2_CHECK(hipFree(device_cpy->target_batch_end));
    if (device_cpy->cigar) G2_CHECK(hipFree(device_cpy->cigar));
    G2_CHECK(hipFree(device_res));
    free(device_cpy);
}

// ------------------------------------------------------------ interfaces ---
// src/interfaces.cpp:81-96
extern "C" void gasal_op_fill(gasal_gpu_storage_t *gpu_storage_t, uint8_t *data, uint32_t nbr_seqs_in_stream,
                              data_source SRC) {
    uint8_t *host_op = NULL;
    if (SRC == QUERY) host_op = gpu_storage_t->host_query_op;
    else if (SRC == TARGET) host_op = gpu_storage_t->host_target_op;
    if (host_op) memcpy(host_op, data, nbr_seqs_in_stream);
}

template <typename T>
static T *host_realloc(T *src, uint32_t new_n, uint32_t old_n) {
    if (new_n < old_n) {
        report("[GASAL ERROR] cudoHostRealloc: invalid sizes. New size < old size (%d < %d)", new_n, old_n);
        gasal2_fail(EXIT_FAILURE);
        return src;
    }
    T *dst = NULL;
    G2_CHECK(hipHostMalloc((void **)&dst, (size_t)new_n * sizeof(T), hipHostMallocDefault));
    if (src) {
        memcpy(dst, src, (size_t)old_n * sizeof(T));
        G2_CHECK(hipHostFree(src));
    }
    return dst;
}

// src/interfaces.cpp:26-78
extern "C" void gasal_host_alns_resize(gasal_gpu_storage_t *gpu_storage, int new_max_alns, Parameters *params) {
    fprintf(stderr, "[GASAL WARNING] Resizing gpu_storage from %d sequences to %d sequences... ",
            gpu_storage->host_max_n_alns, new_max_alns);
    uint32_t old = gpu_storage->host_max_n_alns, nn = (uint32_t)new_max_alns;
    gpu_storage->host_query_op = host_realloc(gpu_storage->host_query_op, nn, old);
    gpu_storage->host_target_op = host_realloc(gpu_storage->host_target_op, nn, old);
    if (params->algo == KSW) gpu_storage->host_seed_scores = host_realloc(gpu_storage->host_seed_scores, nn, old);
    gpu_storage->host_query_batch_lens = host_realloc(gpu_storage->host_query_batch_lens, nn, old);
    gpu_storage->host_target_batch_lens = host_realloc(gpu_storage->host_target_batch_lens, nn, old);
    gpu_storage->host_query_batch_offsets = host_realloc(gpu_storage->host_query_batch_offsets, nn, old);
    gpu_storage->host_target_batch_offsets = host_realloc(gpu_storage->host_target_batch_offsets, nn, old);
    uint8_t *cig = gpu_storage->host_res ? gpu_storage->host_res->cigar : NULL;
    if (gpu_storage->host_res) gpu_storage->host_res->cigar = NULL;
    gasal_res_destroy_host(gpu_storage->host_res);
    gpu_storage->host_res = gasal_res_new_host(nn, params);
    gpu_storage->host_res->cigar = cig;
    if (params->secondBest) {
        gasal_res_destroy_host(gpu_storage->host_res_second);
        gpu_storage->host_res_second = gasal_res_new_host(nn, params);
    }
    gpu_storage->host_max_n_alns = nn;
    fprintf(stderr, " done. This can harm performance.\n");
}

// src/interfaces.cpp:98-128
extern "C" void gasal_set_device(int gpu_select, bool isPrintingProp) {
    if (isPrintingProp) {
        int num_devices = 0;
        (void)hipGetDeviceCount(&num_devices);
        fprintf(stderr, "Found %d GPUs\n", num_devices);
        if (gpu_select > num_devices - 1) {
            report("Error: can't select device %d when only %d devices are selected (range from 0 to %d)\n", gpu_select,
                   num_devices, num_devices - 1);
            gasal2_fail(EXIT_FAILURE);
            return;
        }
        if (num_devices > 0) {
            hipDeviceProp_t prop;
            for (int dev = 0; dev < num_devices; dev++) {
                (void)hipGetDeviceProperties(&prop, dev);
                fprintf(stderr, "\tGPU %d: %s\n", dev, prop.name);
            }
            (void)hipGetDeviceProperties(&prop, gpu_select);
            fprintf(stderr, "Selected device %d : %s\n", gpu_select, prop.name);
            (void)hipSetDevice(gpu_select);
        }
    } else {
        (void)hipSetDevice(gpu_select);
    }
}

// ------------------------------------------------------------ ctors --------
// src/ctors.cpp:17-23
extern "C" gasal_gpu_storage_v gasal_init_gpu_storage_v(int n_streams) {
    gasal_gpu_storage_v v;
    v.a = (gasal_gpu_storage_t *)calloc((size_t)n_streams, sizeof(gasal_gpu_storage_t));
    v.n = n_streams;
    return v;
}

static inline uint32_t pad8u(uint32_t n) { return n % 8 ? n + (8 - (n % 8)) : n; }

// src/ctors.cpp:26-128
extern "C" void gasal_init_streams(gasal_gpu_storage_v *gpu_storage_vec, int max_query_len, int max_target_len,
                                   int max_n_alns, Parameters *params) {
    uint32_t mq8 = pad8u((uint32_t)max_query_len), mt8 = pad8u((uint32_t)max_target_len);
    uint32_t host_max_query_batch_bytes = (uint32_t)max_n_alns * mq8;
    uint32_t gpu_max_query_batch_bytes = (uint32_t)max_n_alns * mq8;
    uint32_t host_max_target_batch_bytes = (uint32_t)max_n_alns * mt8;
    uint32_t gpu_max_target_batch_bytes = (uint32_t)max_n_alns * mt8;
    uint32_t host_max_n_alns = (uint32_t)max_n_alns, gpu_max_n_alns = (uint32_t)max_n_alns;
    for (int i = 0; i < gpu_storage_vec->n; i++) {
        gasal_gpu_storage_t *s = &(gpu_storage_vec->a[i]);
        s->extensible_host_unpacked_query_batch = gasal_host_batch_new(host_max_query_batch_bytes, 0);
        s->extensible_host_unpacked_target_batch = gasal_host_batch_new(host_max_target_batch_bytes, 0);
        G2_CHECK(hipMalloc((void **)&(s->unpacked_query_batch), (size_t)gpu_max_query_batch_bytes + 16));
        G2_CHECK(hipMalloc((void **)&(s->unpacked_target_batch), (size_t)gpu_max_target_batch_bytes + 16));
        G2_CHECK(hipHostMalloc((void **)&(s->host_query_op), host_max_n_alns, hipHostMallocDefault));
        G2_CHECK(hipHostMalloc((void **)&(s->host_target_op), host_max_n_alns, hipHostMallocDefault));
        memset(s->host_query_op, 0, host_max_n_alns);
        memset(s->host_target_op, 0, host_max_n_alns);
        G2_CHECK(hipMalloc((void **)&(s->query_op), gpu_max_n_alns));
        G2_CHECK(hipMalloc((void **)&(s->target_op), gpu_max_n_alns));
        if (params->isPacked) {
            s->packed_query_batch = (uint32_t *)s->unpacked_query_batch;
            s->packed_target_batch = (uint32_t *)s->unpacked_target_batch;
        } else {
            G2_CHECK(hipMalloc((void **)&(s->packed_query_batch), (size_t)(gpu_max_query_batch_bytes / 8 + 4) * sizeof(uint32_t)));
            G2_CHECK(hipMalloc((void **)&(s->packed_target_batch), (size_t)(gpu_max_target_batch_bytes / 8 + 4) * sizeof(uint32_t)));
        }
        if (params->algo == KSW) {
            G2_CHECK(hipHostMalloc((void **)&(s->host_seed_scores), (size_t)host_max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
            G2_CHECK(hipMalloc((void **)&(s->seed_scores), (size_t)host_max_n_alns * sizeof(uint32_t)));
        } else {
            s->host_seed_scores = NULL;
            s->seed_scores = NULL;
        }
        G2_CHECK(hipHostMalloc((void **)&(s->host_query_batch_lens), (size_t)host_max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
        G2_CHECK(hipHostMalloc((void **)&(s->host_target_batch_lens), (size_t)host_max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
        G2_CHECK(hipHostMalloc((void **)&(s->host_query_batch_offsets), (size_t)host_max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
        G2_CHECK(hipHostMalloc((void **)&(s->host_target_batch_offsets), (size_t)host_max_n_alns * sizeof(uint32_t), hipHostMallocDefault));
        G2_CHECK(hipMalloc((void **)&(s->query_batch_lens), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
        G2_CHECK(hipMalloc((void **)&(s->target_batch_lens), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
        G2_CHECK(hipMalloc((void **)&(s->query_batch_offsets), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
        G2_CHECK(hipMalloc((void **)&(s->target_batch_offsets), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
        s->host_res = gasal_res_new_host(host_max_n_alns, params);
        if (params->start_pos == WITH_TB)
            G2_CHECK(hipHostMalloc((void **)&(s->host_res->cigar), gpu_max_query_batch_bytes, hipHostMallocDefault));
        s->device_cpy = gasal_res_new_device_cpy(gpu_max_n_alns, params);
        s->device_res = gasal_res_new_device(s->device_cpy);
        if (params->secondBest) {
            s->host_res_second = gasal_res_new_host(host_max_n_alns, params);
            s->device_cpy_second = gasal_res_new_device_cpy(host_max_n_alns, params);
            s->device_res_second = gasal_res_new_device(s->device_cpy_second);
        } else {
            s->host_res_second = NULL;
            s->device_cpy_second = NULL;
            s->device_res_second = NULL;
        }
        s->packed_tb_matrix_size = 0;   // direction words are sized per batch (impl)
        s->packed_tb_matrices = NULL;
        G2_CHECK(hipStreamCreate(&(s->str)));
        s->is_free = 1;
        s->host_max_query_batch_bytes = host_max_query_batch_bytes;
        s->host_max_target_batch_bytes = host_max_target_batch_bytes;
        s->host_max_n_alns = host_max_n_alns;
        s->gpu_max_query_batch_bytes = gpu_max_query_batch_bytes;
        s->gpu_max_target_batch_bytes = gpu_max_target_batch_bytes;
        s->gpu_max_n_alns = gpu_max_n_alns;
        s->current_n_alns = 0;
        s->impl = new Impl();
    }
}

static void order_forget(Impl *m);

static void impl_free(Impl *m) {
    if (!m) return;
    order_forget(m);
    if (m->ev_h2d) (void)hipEventDestroy(m->ev_h2d);
    if (m->ev_dev) (void)hipEventDestroy(m->ev_dev);
    void *ptrs[] = {m->qpan, m->tpan, m->tb_gbase, m->tb_gqw, m->wf_tb, m->wf_tbo, m->band_wres, m->band_lo, m->band_hi, m->band_bw, m->band_tbo, m->band_flags, m->sub_u32, m->sub_res,
                    m->sub_nops, m->sub_cigar,
                    m->keys, m->perm, m->hist, m->bnd, m->tb, m->eh,
                    m->nops, m->cigar, m->cigar_q, m->scratch_i32, m->aux, m->wf};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->tm_init)
        for (int i = 0; i < 6; i++) (void)hipEventDestroy(m->tm.ev[i]);
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    delete m;
}

static void order_forget(Impl *m) {
    gasal2_order *o = m->chain ? m->chain : &g_chain;
    std::lock_guard<std::mutex> lk(o->mu);
    if (o->last_h2d && o->last_h2d == m->ev_h2d) o->last_h2d = nullptr;
    if (o->last_dev && o->last_dev == m->ev_dev) o->last_dev = nullptr;
}

extern "C" void gasal2_storage_set_order(gasal_gpu_storage_t *s, gasal2_order_t *o) {
    if (!s || !s->impl) return;
    Impl *m = impl_of(s);
    if (m->chain == o) return;
    order_forget(m);
    m->chain = o;
}

extern "C" void gasal2_order_free(gasal2_order_t *o) { delete o; }

extern "C" void gasal2_storage_set_ksw(gasal_gpu_storage_t *s, int32_t own, int32_t w, int32_t zdrop, int32_t end_bonus) {
    if (!s || !s->impl) return;
    Impl *m = impl_of(s);
    m->ksw_own = own != 0;
    m->ksw_w = w;
    m->ksw_zdrop = zdrop > 0 ? zdrop : 0;
    m->ksw_eb = end_bonus;
}

// src/ctors.cpp:130-190
extern "C" void gasal_destroy_streams(gasal_gpu_storage_v *gpu_storage_vec, Parameters *params) {
    for (int i = 0; i < gpu_storage_vec->n; i++) {
        gasal_gpu_storage_t *s = &(gpu_storage_vec->a[i]);
        if (s->str) (void)hipStreamSynchronize(s->str);
        gasal_host_batch_destroy(s->extensible_host_unpacked_query_batch);
        gasal_host_batch_destroy(s->extensible_host_unpacked_target_batch);
        gasal_res_destroy_host(s->host_res);
        gasal_res_destroy_device(s->device_res, s->device_cpy);
        if (params->secondBest) {
            gasal_res_destroy_host(s->host_res_second);
            gasal_res_destroy_device(s->device_res_second, s->device_cpy_second);
        }
        if (s->seed_scores) G2_CHECK(hipFree(s->seed_scores));
        if (s->host_seed_scores) G2_CHECK(hipHostFree(s->host_seed_scores));
        if (s->query_op) G2_CHECK(hipFree(s->query_op));
        if (s->target_op) G2_CHECK(hipFree(s->target_op));
        if (s->host_query_op) G2_CHECK(hipHostFree(s->host_query_op));
        if (s->host_target_op) G2_CHECK(hipHostFree(s->host_target_op));
        if (s->host_query_batch_offsets) G2_CHECK(hipHostFree(s->host_query_batch_offsets));
        if (s->host_target_batch_offsets) G2_CHECK(hipHostFree(s->host_target_batch_offsets));
        if (s->host_query_batch_lens) G2_CHECK(hipHostFree(s->host_query_batch_lens));
        if (s->host_target_batch_lens) G2_CHECK(hipHostFree(s->host_target_batch_lens));
        if (s->unpacked_query_batch) G2_CHECK(hipFree(s->unpacked_query_batch));
        if (s->unpacked_target_batch) G2_CHECK(hipFree(s->unpacked_target_batch));
        if (!params->isPacked) {
            if (s->packed_query_batch) G2_CHECK(hipFree(s->packed_query_batch));
            if (s->packed_target_batch) G2_CHECK(hipFree(s->packed_target_batch));
        }
        if (s->query_batch_offsets) G2_CHECK(hipFree(s->query_batch_offsets));
        if (s->target_batch_offsets) G2_CHECK(hipFree(s->target_batch_offsets));
        if (s->query_batch_lens) G2_CHECK(hipFree(s->query_batch_lens));
        if (s->target_batch_lens) G2_CHECK(hipFree(s->target_batch_lens));
        impl_free((Impl *)s->impl);
        s->impl = NULL;
        if (s->str) G2_CHECK(hipStreamDestroy(s->str));
        s->str = NULL;
    }
}

// src/ctors.cpp:193-196
extern "C" void gasal_destroy_gpu_storage_v(gasal_gpu_storage_v *gpu_storage_vec) {
    if (gpu_storage_vec->a != NULL) free(gpu_storage_vec->a);
    gpu_storage_vec->a = NULL;
}

// src/ctors.cpp:202-231 (deprecated in the reference)
extern "C" void gasal_gpu_mem_alloc(gasal_gpu_storage_t *gpu_storage, int gpu_max_query_batch_bytes,
                                    int gpu_max_target_batch_bytes, int gpu_max_n_alns, Parameters *params) {
    G2_CHECK(hipMalloc((void **)&(gpu_storage->unpacked_query_batch), (size_t)gpu_max_query_batch_bytes + 16));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->unpacked_target_batch), (size_t)gpu_max_target_batch_bytes + 16));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->packed_query_batch), (size_t)(gpu_max_query_batch_bytes / 8 + 4) * sizeof(uint32_t)));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->packed_target_batch), (size_t)(gpu_max_target_batch_bytes / 8 + 4) * sizeof(uint32_t)));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->query_batch_lens), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->target_batch_lens), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->query_batch_offsets), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
    G2_CHECK(hipMalloc((void **)&(gpu_storage->target_batch_offsets), (size_t)gpu_max_n_alns * sizeof(uint32_t)));
    if (gpu_storage->device_cpy == NULL) gpu_storage->device_cpy = gasal_res_new_device_cpy((uint32_t)gpu_max_n_alns, params);
    gpu_storage->device_res = gasal_res_new_device(gpu_storage->device_cpy);
    gpu_storage->gpu_max_query_batch_bytes = (uint32_t)gpu_max_query_batch_bytes;
    gpu_storage->gpu_max_target_batch_bytes = (uint32_t)gpu_max_target_batch_bytes;
    gpu_storage->gpu_max_n_alns = (uint32_t)gpu_max_n_alns;
}

// src/ctors.cpp:234-252 (deprecated in the reference)
extern "C" void gasal_gpu_mem_free(gasal_gpu_storage_t *gpu_storage, Parameters *params) {
    if (gpu_storage->unpacked_query_batch) G2_CHECK(hipFree(gpu_storage->unpacked_query_batch));
    if (gpu_storage->unpacked_target_batch) G2_CHECK(hipFree(gpu_storage->unpacked_target_batch));
    if (gpu_storage->packed_query_batch) G2_CHECK(hipFree(gpu_storage->packed_query_batch));
    if (gpu_storage->packed_target_batch) G2_CHECK(hipFree(gpu_storage->packed_target_batch));
    if (gpu_storage->query_batch_offsets) G2_CHECK(hipFree(gpu_storage->query_batch_offsets));
    if (gpu_storage->target_batch_offsets) G2_CHECK(hipFree(gpu_storage->target_batch_offsets));
    if (gpu_storage->query_batch_lens) G2_CHECK(hipFree(gpu_storage->query_batch_lens));
    if (gpu_storage->target_batch_lens) G2_CHECK(hipFree(gpu_storage->target_batch_lens));
    gasal_res_destroy_device(gpu_storage->device_res, gpu_storage->device_cpy);
    if (params->secondBest) gasal_res_destroy_device(gpu_storage->device_res_second, gpu_storage->device_cpy_second);
}

// Packed and table kernels for a batch: the choices depend only on the
// mode, the scores and the largest shape the lane path runs (eq query words,
// et target words); shared by the storage path and gasal2_aln_noalloc.
void g2_choose_packed(G2Dev &d, const G2Opts &op, const G2Scores &sc, uint32_t eq, uint32_t et, bool banded) {
    auto bl = [](uint64_t v) { int b = 0; while (v) { b++; v >>= 1; } return b; };
    // Packed 16-bit LOCAL (score and ends; WITH_START then runs the int32
    // reverse pass on the packed forward's end): every value of the pairs the lane
    // path runs fits int16 -- H <= match * min(8 eq, 8 et); E, F >= -(mismatch,
    // N penalty) - open - 9 extend -- and the policy allows it.
    d.pk16 = 0;
    if (d.fast && op.algo == LOCAL && (op.start == WITHOUT_START || op.start == WITH_START) && !op.second && g_pk16_mode) {
        const int64_t hmax = (int64_t)std::max(sc.match, 0) * std::min<int64_t>(8ll * eq, 8ll * et);
        const int64_t lowest = -(int64_t)std::max(sc.mismatch, sc.has_np ? sc.np : 0) - sc.gapoe - 9ll * sc.gape - 8;
        if (hmax <= 32000 && lowest >= -32000) d.pk16 = 1;
    }
    // Two pairs per lane in packed 16-bit halves (k_local_pk2): LOCAL with
    // second best or the checkpointed traceback (policy 2: every LOCAL mode).
    // Cells are ranked as h * 8 + (7 - row) in 16 bits, so H <= match *
    // min(8 eq, 8 et) must stay <= 4095; column indices i * 8 eq + t must fit
    // 16 bits; E and F stay above -32000 as for pk16.
    d.pk2 = 0;
    if (d.fast && op.algo == LOCAL && !banded && g_pk16_mode) {
        const bool tb_ok = op.start != WITH_TB || (d.tbh && !d.tb_fused);
        const bool want = g_pk16_mode == 2 || op.second || op.start == WITH_TB;
        const int64_t hmax = (int64_t)std::max(sc.match, 0) * std::min<int64_t>(8ll * eq, 8ll * et);
        const int64_t lowest = -(int64_t)std::max(sc.mismatch, sc.has_np ? sc.np : 0) - sc.gapoe - 9ll * sc.gape - 8;
        if (tb_ok && want && hmax <= 4095 && lowest >= -32000 && (uint64_t)et * 8u * eq <= 65535u) {
            d.pk2 = 1;
            d.pk16 = 0;
        }
    }
    // Packed 16-bit GLOBAL score: every value of the lane path fits int16 with
    // room below for the packed "minus infinity" (-30000): |H| <= match and
    // penalty times the shorter side plus open and extend over both sides,
    // E and F at most one open and one penalty lower.
    d.pk16g = 0;
    if (d.fast && op.algo == GLOBAL && op.start != WITH_TB && g_pk16_mode && sc.gape <= 255) {
        const int64_t pen = std::max<int64_t>(std::max(sc.mismatch, sc.has_np ? sc.np : 0), 0);
        const int64_t sh = 8ll * std::min(eq, et), lg = 8ll * ((int64_t)eq + et);
        const int64_t bound = (pen + std::max(sc.match, 0)) * sh + 2ll * sc.gapoe + (int64_t)sc.gape * lg + pen;
        if (bound <= 16000) d.pk16g = 1;
    }
    // Packed 16-bit SEMI_GLOBAL: the same bound (head gaps as GLOBAL's
    // boundaries, classic Gotoh) for tails that end on the target or nowhere.
    d.pk16s = 0;
    if (d.fast && op.algo == SEMI_GLOBAL && op.start != WITH_TB && (op.tail == NONE || op.tail == TARGET) &&
        g_pk16_mode && sc.gape <= 255) {
        const int64_t pen = std::max<int64_t>(std::max(sc.mismatch, sc.has_np ? sc.np : 0), 0);
        const int64_t sh = 8ll * std::min(eq, et), lg = 8ll * ((int64_t)eq + et);
        const int64_t bound = (pen + std::max(sc.match, 0)) * sh + 2ll * sc.gapoe + (int64_t)sc.gape * lg + pen;
        if (bound <= 16000) d.pk16s = 1;
    }
    // LOCAL lane-per-pair score-table kernel: position encoding
    // h << S | ~strip | ~column | row must fit 31 bits for the pairs it runs
    // (all pairs, or those the one-wave path leaves).
    if (d.fast && op.algo == LOCAL) {
        const int cb = std::max(1, bl(8ull * eq - 1)), sbits = std::max(1, bl(et - 1));
        const uint64_t maxh = (uint64_t)std::max(sc.match, 0) * std::min<uint64_t>(8ull * eq, 8ull * et);
        const int hb = bl(maxh) + 1;
        const int S = 3 + cb + sbits;
        if (S + hb <= 31) {
            d.enc_ok = 1;
            d.enc_s = S;
            d.enc_cb = cb;
            d.enc_cm = (1u << cb) - 1u;
            d.enc_sm = (1u << sbits) - 1u;
        }
    }
}

// The process-wide scores (gasal_copy_subst_scores, gasal2_set_n_code /
// _n_penalty) as the kernels take them.
G2Scores g2_scores() { return g_scores; }

// Score-table kernels (4-bit signed substitution scores) for the batch.
void g2_choose_fast(G2Dev &d, const G2Opts &op, const G2Scores &sc, uint32_t qw, uint32_t tw) {
    d.fast = 0;
    d.enc_ok = 0;
    const bool lut_algo = op.algo == LOCAL || op.algo == GLOBAL || op.algo == SEMI_GLOBAL || op.algo == KSW || op.algo == BANDED;
    if (lut_algo && g2_get_dispatch() != 1 && qw > 0 && tw > 0) {
        auto fits4 = [](int32_t v) { return v >= -8 && v <= 7; };
        // the N score of the substitution macros (src/gasal_kernels.h:37-56)
        const int32_t nval = sc.has_np ? -sc.np : 0;
        if (fits4(sc.match) && fits4(-sc.mismatch) && fits4(nval)) d.fast = 1;
    }
}

// ------------------------------------------------------------ launcher -----
extern "C" void gasal_copy_subst_scores(gasal_subst_scores *subst) {                 // src/gasal_align.cu:329-339
    g_scores.gapo = subst->gap_open;
    g_scores.gape = subst->gap_extend;
    g_scores.gapoe = subst->gap_open + subst->gap_extend;
    g_scores.match = subst->match;
    g_scores.mismatch = subst->mismatch;
}

namespace {

// Validation of src/gasal_align.cu:32-67.  Returns false after reporting.
bool validate(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n) {
    if (n <= 0) { report("[GASAL ERROR:] actual_n_alns <= 0\n"); gasal2_fail(EXIT_FAILURE); return false; }
    if (qb <= 0) { report("[GASAL ERROR:] actual_query_batch_bytes <= 0\n"); gasal2_fail(EXIT_FAILURE); return false; }
    if (tb <= 0) { report("[GASAL ERROR:] actual_target_batch_bytes <= 0\n"); gasal2_fail(EXIT_FAILURE); return false; }
    if (qb % 8) { report("[GASAL ERROR:] actual_query_batch_bytes=%d is not a multiple of 8\n", qb); gasal2_fail(EXIT_FAILURE); return false; }
    if (tb % 8) { report("[GASAL ERROR:] actual_target_batch_bytes=%d is not a multiple of 8\n", tb); gasal2_fail(EXIT_FAILURE); return false; }
    if (qb > s->host_max_query_batch_bytes) {
        report("[GASAL ERROR:] actual_query_batch_bytes(%d) > host_max_query_batch_bytes(%d)\n", qb, s->host_max_query_batch_bytes);
        gasal2_fail(EXIT_FAILURE); return false;
    }
    if (tb > s->host_max_target_batch_bytes) {
        report("[GASAL ERROR:] actual_target_batch_bytes(%d) > host_max_target_batch_bytes(%d)\n", tb, s->host_max_target_batch_bytes);
        gasal2_fail(EXIT_FAILURE); return false;
    }
    if (n > s->host_max_n_alns) {
        report("[GASAL ERROR:] actual_n_alns(%d) > host_max_n_alns(%d)\n", n, s->host_max_n_alns);
        gasal2_fail(EXIT_FAILURE); return false;
    }
    return true;
}

// Checks made where a batch goes to the device (gasal_aln_async,
// gasal2_upload), beyond the reference's: the host pages fit the byte counts
// (with isPacked the counts are in bases, 8 per packed word, and the pages
// hold 4 bytes per 8 bases), and every pair's sequence lies inside its batch
// at a multiple of 8 bases, so no kernel can read past the copied batch.
bool validate_batch(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, const Parameters *params) {
    const bool packed = params && params->isPacked;
    host_batch_t *lists[2] = {s->extensible_host_unpacked_query_batch, s->extensible_host_unpacked_target_batch};
    const uint32_t counts[2] = {qb, tb};
    const char *what[2] = {"query", "target"};
    for (int k = 0; k < 2; k++) {
        uint64_t ext = 0;
        for (host_batch_t *c = lists[k]; c != NULL; c = c->next)
            if (c->data_size) ext = std::max<uint64_t>(ext, (uint64_t)c->offset + c->data_size);
        const uint64_t cap = packed ? (uint64_t)counts[k] / 2 : (uint64_t)counts[k];
        if (ext > cap) {
            if (packed)
                report("[GASAL ERROR:] isPacked: the %s host batch holds %llu bytes of packed words (%llu bases) but actual_%s_batch_bytes is %u; with isPacked the batch byte counts count bases\n",
                       what[k], (unsigned long long)ext, (unsigned long long)(2 * ext), what[k], counts[k]);
            else
                report("[GASAL ERROR:] the %s host batch holds %llu bytes, more than actual_%s_batch_bytes(%u)\n",
                       what[k], (unsigned long long)ext, what[k], counts[k]);
            gasal2_fail(EXIT_FAILURE);
            return false;
        }
    }
    const uint32_t *off[2] = {s->host_query_batch_offsets, s->host_target_batch_offsets};
    const uint32_t *len[2] = {s->host_query_batch_lens, s->host_target_batch_lens};
    for (int k = 0; k < 2; k++) {
        if (!off[k] || !len[k]) continue;
        const uint32_t *o = off[k], *l = len[k];
        uint32_t bad_align = 0;
        uint64_t worst = 0;
        for (uint32_t i = 0; i < n; i++) {
            bad_align |= o[i] & 7u;
            worst = std::max<uint64_t>(worst, (uint64_t)o[i] + (((uint64_t)l[i] + 7) & ~7ull));
        }
        if (!bad_align && worst <= counts[k]) continue;
        for (uint32_t i = 0; i < n; i++) {
            const uint64_t end = (uint64_t)o[i] + (((uint64_t)l[i] + 7) & ~7ull);
            if ((o[i] & 7u) || end > counts[k]) {
                report("[GASAL ERROR:] pair %u: %s offset %u, length %u does not lie inside actual_%s_batch_bytes(%u) at a multiple of 8\n",
                       i, what[k], o[i], l[i], what[k], counts[k]);
                gasal2_fail(EXIT_FAILURE);
                return false;
            }
        }
    }
    return true;
}

// Device-buffer growth of src/gasal_align.cu:70-145.  The new buffers are
// allocated before the old ones are released, so a failed allocation (error
// mode 1) leaves the storage with its old buffers and recorded sizes.
void hip_free_quiet(void *p) {
    if (p) (void)hipFree(p);
}

bool grow_device(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params) {
    struct Side {
        uint32_t *max;
        uint8_t **unpacked;
        uint32_t **packed;
        uint32_t need;
        const char *what;
    } sides[2] = {{&s->gpu_max_query_batch_bytes, &s->unpacked_query_batch, &s->packed_query_batch, qb, "query"},
                  {&s->gpu_max_target_batch_bytes, &s->unpacked_target_batch, &s->packed_target_batch, tb, "target"}};
    for (int k = 0; k < 2; k++) {
        Side &sd = sides[k];
        if (*sd.max >= sd.need) continue;
        uint32_t i = 2;
        while (*sd.max * i < sd.need) i++;
        const uint32_t nmax = *sd.max * i;
        fprintf(stderr, "[GASAL WARNING:] actual_%s_batch_bytes(%d) > Allocated GPU memory (gpu_max_%s_batch_bytes=%d). Therefore, allocating %d bytes on GPU (gpu_max_%s_batch_bytes=%d). Performance may be lost if this is repeated many times.\n",
                sd.what, sd.need, sd.what, *sd.max, nmax, sd.what, nmax);
        uint8_t *nu = nullptr;
        uint32_t *np = nullptr;
        uint8_t *ncig = nullptr;
        t_failed = false;
        G2_CHECK(hipMalloc((void **)&nu, (size_t)nmax + 16));
        if (!t_failed && !params->isPacked) G2_CHECK(hipMalloc((void **)&np, (size_t)(nmax / 8 + 4) * sizeof(uint32_t)));
        const bool cig = k == 0 && params->start_pos == WITH_TB;
        if (!t_failed && cig) {
            fprintf(stderr, "[GASAL WARNING:] actual_query_batch_bytes(%d) > Allocated HOST memory for CIGAR (gpu_max_query_batch_bytes=%d). Therefore, allocating %d bytes on the host (gpu_max_query_batch_bytes=%d). Performance may be lost if this is repeated many times.\n",
                    qb, nmax, nmax, nmax);
            G2_CHECK(hipHostMalloc((void **)&ncig, nmax, hipHostMallocDefault));
        }
        if (t_failed) {
            hip_free_quiet(nu);
            hip_free_quiet(np);
            if (ncig) (void)hipHostFree(ncig);
            return false;
        }
        const bool aliased = (void *)*sd.packed == (void *)*sd.unpacked;
        if (*sd.unpacked) G2_CHECK(hipFree(*sd.unpacked));
        if (!aliased && *sd.packed) G2_CHECK(hipFree(*sd.packed));
        *sd.unpacked = nu;
        *sd.packed = params->isPacked ? (uint32_t *)nu : np;
        if (cig) {
            if (s->host_res->cigar != NULL) G2_CHECK(hipHostFree(s->host_res->cigar));
            s->host_res->cigar = ncig;
        }
        *sd.max = nmax;
    }
    if (s->gpu_max_n_alns < n) {
        uint32_t i = 2;
        while (s->gpu_max_n_alns * i < n) i++;
        const uint32_t nmax = s->gpu_max_n_alns * i;
        fprintf(stderr, "[GASAL WARNING:] actual_n_alns(%d) > gpu_max_n_alns(%d). Therefore, allocating memory for %d alignments on  GPU (gpu_max_n_alns=%d). Performance may be lost if this is repeated many times.\n",
                n, s->gpu_max_n_alns, nmax, nmax);
        uint32_t **arrs[] = {&s->query_batch_offsets, &s->target_batch_offsets, &s->query_batch_lens,
                             &s->target_batch_lens, &s->seed_scores};
        uint32_t *na[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        uint8_t *nop[2] = {nullptr, nullptr};
        t_failed = false;
        for (int k = 0; k < 5 && !t_failed; k++) G2_CHECK(hipMalloc((void **)&na[k], (size_t)nmax * sizeof(uint32_t)));
        for (int k = 0; k < 2 && !t_failed; k++) G2_CHECK(hipMalloc((void **)&nop[k], nmax));
        gasal_res_t *ncpy = nullptr, *nres = nullptr, *ncpy2 = nullptr, *nres2 = nullptr;
        if (!t_failed) ncpy = gasal_res_new_device_cpy(nmax, params);
        if (!t_failed) nres = gasal_res_new_device(ncpy);
        if (!t_failed && params->secondBest) ncpy2 = gasal_res_new_device_cpy(nmax, params);
        if (!t_failed && params->secondBest) nres2 = gasal_res_new_device(ncpy2);
        if (t_failed) {
            for (uint32_t *p : na) hip_free_quiet(p);
            for (uint8_t *p : nop) hip_free_quiet(p);
            for (gasal_res_t *c : {ncpy, ncpy2}) {
                if (!c) continue;
                for (void *f : {(void *)c->aln_score, (void *)c->query_batch_start, (void *)c->target_batch_start,
                                (void *)c->query_batch_end, (void *)c->target_batch_end})
                    hip_free_quiet(f);
                free(c);
            }
            hip_free_quiet(nres);
            hip_free_quiet(nres2);
            return false;
        }
        for (int k = 0; k < 5; k++) {
            if (*arrs[k]) G2_CHECK(hipFree(*arrs[k]));
            *arrs[k] = na[k];
        }
        if (s->query_op) G2_CHECK(hipFree(s->query_op));
        if (s->target_op) G2_CHECK(hipFree(s->target_op));
        s->query_op = nop[0];
        s->target_op = nop[1];
        gasal_res_destroy_device(s->device_res, s->device_cpy);
        s->device_cpy = ncpy;
        s->device_res = nres;
        if (params->secondBest) {
            gasal_res_destroy_device(s->device_res_second, s->device_cpy_second);
            s->device_cpy_second = ncpy2;
            s->device_res_second = nres2;
        }
        s->gpu_max_n_alns = nmax;
    }
    return true;
}

template <typename T>
void ensure(T *&p, size_t &cap_bytes, size_t need_bytes) {
    if (cap_bytes >= need_bytes && p) return;
    if (p) G2_CHECK(hipFree(p));
    p = nullptr;
    cap_bytes = 0;
    size_t b = need_bytes + need_bytes / 8 + 256;
    hipError_t e = hipMalloc((void **)&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        G2_CHECK(e);
        return;
    }
    cap_bytes = b;
}

G2Opts opts_of(const Parameters *p, const gasal_gpu_storage_t *s) {
    G2Opts o;
    o.algo = p->algo;
    o.start = p->start_pos;
    o.second = p->secondBest ? 1 : 0;
    o.head = p->semiglobal_skipping_head;
    o.tail = p->semiglobal_skipping_tail;
    o.k_band = p->k_band;
    o.is_packed = p->isPacked ? 1 : 0;
    o.is_rc = p->isReverseComplement ? 1 : 0;
    o.banded = 0;
    o.skip_prep = 0;
    o.prep_only = 0;
    o.ksw_w = g_ksw_w;
    o.ksw_zdrop = g_ksw_zdrop;
    o.ksw_eb = g_ksw_eb;
    const Impl *m = s && s->impl ? static_cast<const Impl *>(s->impl) : nullptr;
    if (m && m->ksw_own) {
        o.ksw_w = m->ksw_w;
        o.ksw_zdrop = m->ksw_zdrop;
        o.ksw_eb = m->ksw_eb;
    }
    return o;
}

// Sizes the workspace for this batch and fills the device view.
struct PrepExtra {
    const uint32_t *hq = nullptr, *ht = nullptr;   // host lengths of the pairs run (default: the storage's)
    bool banded = false;                            // direction tiles sized per band by the caller
    bool sub = false;                               // sub-batch: result, CIGAR and scratch buffers come from the caller
};

void prepare(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, const Parameters *params, G2Dev &d,
             const PrepExtra &px = PrepExtra()) {
    Impl *m = impl_of(s);
    const uint32_t *hq = px.hq ? px.hq : s->host_query_batch_lens;
    const uint32_t *ht = px.ht ? px.ht : s->host_target_batch_lens;
    uint32_t mq = 0, mt = 0;
    for (uint32_t k = 0; k < n; k++) {
        mq = std::max(mq, hq[k]);
        mt = std::max(mt, ht[k]);
    }
    if (!px.sub) { m->last_max_q = mq; m->last_max_t = mt; m->last_n = n; }
    const G2Opts op = opts_of(params, s);
    const uint32_t qw = g2_regs(mq), tw = g2_regs(mt);
    const uint32_t groups = (n + 63) / 64;
    size_t b = 0;
    size_t ncap = (size_t)n * sizeof(uint32_t);
    if (m->n_cap < n || !m->keys) {
        for (uint32_t **pp : {&m->keys, &m->perm}) {
            if (*pp) G2_CHECK(hipFree(*pp));
            G2_CHECK(hipMalloc((void **)pp, ncap + 256));
        }
        m->n_cap = n;
    }
    if (!m->hist) G2_CHECK(hipMalloc((void **)&m->hist, (size_t)G2_SHAPE_BUCKETS * 4));
    const uint32_t bnd_cols = 8 * qw + 8;
    ensure(m->bnd, m->bnd_bytes, (size_t)groups * bnd_cols * 64 * 4);
    const bool tb_on = op.start == WITH_TB;
    const bool tb_dp = tb_on && (op.algo == LOCAL || op.algo == GLOBAL);
    const uint32_t eh_cols = (mq + 2 + 7) & ~7u;   // KSW row: whole blocks of 8 columns
    if (op.algo == KSW) ensure(m->eh, m->eh_bytes, (size_t)groups * eh_cols * 64 * 8);
    const uint32_t cap = g2_cigar_capacity(mq, mt);
    if (tb_on && !px.sub) {
        ensure(m->nops, m->nops_cap, (size_t)n * 4);
        ensure(m->cigar, m->cigar_bytes, (size_t)n * cap);
        ensure(m->cigar_q, m->cigar_q_bytes, (size_t)qb + 16);
    }
    if (!px.sub) ensure(m->scratch_i32, m->scratch_i32_cap, (size_t)n * 4 * 8);
    (void)b;
    if (!px.sub) {
        m->last_cigar_cap = cap;
        m->last_tb = tb_on;
    }
    s->packed_tb_matrices = (uint4 *)m->tb;
    s->packed_tb_matrix_size = m->tb_bytes / 16;

    memset(&d, 0, sizeof(d));
    d.uq = s->unpacked_query_batch; d.ut = s->unpacked_target_batch;
    d.pq = s->packed_query_batch; d.pt = s->packed_target_batch;
    d.qoff = s->query_batch_offsets; d.toff = s->target_batch_offsets;
    d.qlen = s->query_batch_lens; d.tlen = s->target_batch_lens;
    d.qop = s->query_op; d.top = s->target_op;
    d.seeds = s->seed_scores;
    d.q_bytes = qb; d.t_bytes = tb; d.n = n;
    d.max_qw = qw; d.max_tw = tw;
    int32_t *scr = m->scratch_i32;
    auto pick = [&](int32_t *p, int idx) { return p ? p : scr + (size_t)idx * n; };
    d.score = pick(s->device_cpy->aln_score, 0);
    d.qend = pick(s->device_cpy->query_batch_end, 1);
    d.tend = pick(s->device_cpy->target_batch_end, 2);
    d.qstart = pick(s->device_cpy->query_batch_start, 3);
    d.tstart = pick(s->device_cpy->target_batch_start, 4);
    if (params->secondBest && s->device_cpy_second) {
        d.s_score = pick(s->device_cpy_second->aln_score, 5);
        d.s_qend = pick(s->device_cpy_second->query_batch_end, 6);
        d.s_tend = pick(s->device_cpy_second->target_batch_end, 7);
    } else {
        d.s_score = scr + (size_t)5 * n; d.s_qend = scr + (size_t)6 * n; d.s_tend = scr + (size_t)7 * n;
    }
    d.n_ops = m->nops;
    d.cigar = m->cigar;
    d.cigar_cap = cap;
    d.cigar_qlayout = tb_on ? m->cigar_q : nullptr;
    d.keys = m->keys; d.perm = m->perm; d.hist = m->hist;
    d.bnd = m->bnd; d.bnd_cols = bnd_cols;
    d.tb = m->tb; d.tb_tw = tw; d.tb_qw8 = 8 * qw;
    d.eh = m->eh; d.eh_cols = eh_cols;
    d.ksw16 = 0;
    if (op.algo == KSW && s->host_seed_scores && !px.sub) {   // every KSW value lies in [0, h0 + match * min(Lq, Lt)]
        uint64_t hmax = 0;
        for (uint32_t k = 0; k < n; k++)
            hmax = std::max<uint64_t>(hmax, (uint64_t)s->host_seed_scores[k] +
                                                (uint64_t)std::max(g_scores.match, 0) * std::min(hq[k], ht[k]));
        if (hmax <= 65535 && g_pk16_mode) d.ksw16 = 1;
    }
    if (!m->aux) G2_CHECK(hipMalloc((void **)&m->aux, 256 + 4096));
    d.aux = m->aux;
    d.sink = m->aux ? m->aux + 64 : nullptr;

    const G2Scores &sc = g_scores;
    g2_choose_fast(d, op, sc, qw, tw);
    // Long LOCAL pairs (score and ends only) run one wave per pair.  The
    // device applies the same shape test in k_keys and also needs the nibble
    // census to allow the score tables; the host count only sizes the launch.
    uint32_t eq = qw, et = tw;   // largest shape the one-lane-per-pair kernels see
    uint32_t nlong = 0;
    if (d.fast && op.algo == LOCAL && (op.start == WITHOUT_START || (op.start == WITH_TB && !px.banded)) && !op.second &&
        g2_get_dispatch() != 2) {
        uint32_t min_tw = g_wf_min_tw, min_cells = g_wf_min_cells, max_blocks = g_wf_max_blocks;
        bool tuned = false;
        if (!g_wf_user && op.start == WITH_TB) { min_tw = WF_TB_MIN_TW; min_cells = WF_TB_MIN_CELLS; }
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            uint64_t sum_q = 0;
            if (!g_tune.empty())
                for (uint32_t k = 0; k < n; k++) sum_q += hq[k];
            const uint64_t mean_q = n ? sum_q / n : 0;
            for (const TuneEntry &e : g_tune)
                if (e.algo == op.algo && e.start == op.start && e.second == op.second && mean_q >= e.lq_lo && mean_q < e.lq_hi) {
                    min_tw = e.min_tw; min_cells = e.min_cells; max_blocks = e.max_blocks;
                    tuned = true;
                    break;
                }
        }
        // the small-batch rule yields to user and tuned thresholds
        const bool few = g2_get_dispatch() == 0 && !g_wf_user && !tuned &&
                         (uint64_t)n <= (uint64_t)g2_cu_count() * (op.start == WITH_TB ? WF_FEW_PER_CU_TB : WF_FEW_PER_CU);
        const bool all = g2_get_dispatch() == 3 || few;
        d.wf_min_tw = all ? 1u : min_tw;
        d.wf_min_cells = all ? 0u : min_cells;
        uint32_t ltw = 0, lqw = 0, sq = 0, st = 0;
        const bool wtb = op.start == WITH_TB;
        if (wtb) m->h_wf_tbo.resize(n);
        uint64_t wtot = 0;   // 32-byte tiles of the long pairs
        // no pair can pass the shape test when the largest shapes do not
        const bool any_long = tw >= d.wf_min_tw && (uint64_t)qw * tw * 64u >= (uint64_t)d.wf_min_cells;
        for (uint32_t k = 0; any_long && k < n; k++) {
            const uint32_t kq = g2_regs(hq[k]), kt = g2_regs(ht[k]);
            if (kt >= d.wf_min_tw && (uint64_t)kq * kt * 64u >= (uint64_t)d.wf_min_cells) {
                nlong++;
                ltw = std::max(ltw, kt);
                lqw = std::max(lqw, kq);
                if (wtb) { m->h_wf_tbo[k] = wtot; wtot += (uint64_t)kq * kt; }
            } else {
                sq = std::max(sq, kq);
                st = std::max(st, kt);
            }
        }
        if (nlong) {
            eq = std::max(sq, 1u);
            et = std::max(st, 1u);
            d.wf_blocks = std::min<uint32_t>(nlong, max_blocks);
            d.wf_stride = ((ltw + 63) / 64) * 17 * 64;
            if (8 * qw > 1024) {   // more than one column block: strip state is saved between blocks
                ensure(m->wf, m->wf_bytes, (size_t)d.wf_blocks * d.wf_stride * 4);
                d.wf_scratch = m->wf;
            } else {
                d.wf_stride = 0;
                d.wf_scratch = nullptr;
            }
            if (wtb) {   // per-pair direction tiles (or checkpoints) of the long pairs and their offsets
                // checkpoints take twice the bytes of direction words: only while
                // they fit in 80% of the device memory free now (this storage's
                // own buffer counted as free: it is reused)
                size_t free_b = 0, total_b = 0;
                const bool room = g_tb_ckpt == 2 && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                                  wtot * 64 <= ((uint64_t)free_b + m->wf_tb_bytes) / 10 * 8;
                const bool wck = g_tb_ckpt == 2 && room &&
                                 (uint64_t)std::max(g_scores.match, 0) * 8u * std::min(lqw, ltw) <= 32767u &&
                                 (uint64_t)g_scores.gapoe +
                                         (uint64_t)std::max(std::max(g_scores.mismatch, g_scores.has_np ? g_scores.np : 0), 0) <=
                                     32767u;
                ensure(m->wf_tb, m->wf_tb_bytes, (size_t)wtot * 32 * (wck ? 2 : 1) + 256);
                ensure(m->wf_tbo, m->wf_tbo_bytes, (size_t)n * 8);
                G2_CHECK(hipMemcpyAsync(m->wf_tbo, m->h_wf_tbo.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->str));
                d.wf_tb = m->wf_tb;
                d.wf_tbo = m->wf_tbo;
                if (wck) d.wf_tbh = m->wf_tb + wtot * 8;
            }
        }
    }
    // Transposed sequence panels for the fast LOCAL lane path (k_panel): groups
    // after the long pairs, eq query words and et target words each.
    if (g_panels && d.fast && op.algo == LOCAL && !px.banded && n > nlong) {
        const uint64_t lgroups = ((uint64_t)(n - nlong) + 63) / 64;
        ensure(m->qpan, m->qpan_bytes, (size_t)(lgroups * eq * 256) + 256);
        ensure(m->tpan, m->tpan_bytes, (size_t)(lgroups * et * 256) + 256);
        d.qpan = m->qpan;
        d.tpan = m->tpan;
        d.qpan_w = eq;
        d.tpan_w = et;
    }
    // direction tiles of the one-lane-per-pair kernels: groups after the long pairs
    if (tb_dp && !px.banded) {
        const uint64_t lgroups = ((uint64_t)(n - nlong) + 63) / 64;
        const uint64_t one_size = lgroups * et * (8 * (uint64_t)eq) * 64 * 4;
        // Checkpointed traceback: two regions of the tile layout (left column
        // states, strip bottom rows) instead of direction words.  Stored
        // states are 16-bit, so H must fit: H <= match * min(Lq, Lt).
        const bool ck = g_tb_ckpt && op.algo == LOCAL && op.start == WITH_TB && d.fast &&
                        (uint64_t)std::max(sc.match, 0) * 8u * std::min(eq, et) <= 32767u &&
                        (uint64_t)sc.gapoe + (uint64_t)std::max(std::max(sc.mismatch, sc.has_np ? sc.np : 0), 0) <= 32767u;
        const uint64_t nreg = ck ? 2 : 1;
        // One region size for all groups while it is within the budget and
        // both regions take at most half of the device memory free now (this
        // storage's own tile buffer counted as free: it is reused), so that
        // several storages of a large batch fit beside each other; else
        // regions sized per group (about half the bytes on mixed lengths).
        bool one = one_size <= g_tb_one_size_max;
        if (one) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
                one = nreg * one_size <= ((uint64_t)free_b + m->tb_bytes) / 2;
        }
        if (one) {
            ensure(m->tb, m->tb_bytes, (size_t)(nreg * one_size) + 256);
            if (ck) d.tbh = m->tb + one_size / 4;
        } else {
            // Per-group regions.  The device orders the lane-path pairs by
            // descending shape bucket (k_keys), so the buckets falling in each
            // 64-slot group are fixed; each group gets the largest bucket
            // edges among them.
            auto shape8 = [](uint32_t x) -> uint32_t {
                if (x < 128u) return x;
                uint32_t e = 31u - (uint32_t)__builtin_clz(x);
                uint32_t v = 128u + (e - 7u) * 16u + ((x >> (e - 4u)) & 15u);
                return v > 255u ? 255u : v;
            };
            auto edge8 = [](uint32_t v) -> uint32_t {   // largest x with shape8(x) == v
                if (v < 128u) return v;
                if (v == 255u) return 0xFFFFFFFFu;
                uint32_t e = 7u + (v - 128u) / 16u, mant = (v - 128u) % 16u;
                return ((17u + mant) << (e - 4u)) - 1u;
            };
            m->h_hist.assign(1u << 16, 0u);
            const bool wfa = d.wf_blocks > 0;
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t kq = g2_regs(hq[k]), kt = g2_regs(ht[k]);
                if (wfa && kt >= d.wf_min_tw && (uint64_t)kq * kt * 64u >= (uint64_t)d.wf_min_cells) continue;
                m->h_hist[(shape8(kq) << 8) | shape8(kt)]++;
            }
            m->h_gbase.assign(lgroups, 0);
            m->h_gqw.assign(lgroups, 0);
            std::vector<uint32_t> gtw(lgroups, 0);
            uint64_t slot = 0;
            for (int32_t key = (1 << 16) - 1; key >= 0; key--) {
                uint32_t c = m->h_hist[key];
                if (!c) continue;
                const uint32_t qe = std::min(edge8((uint32_t)key >> 8), eq), te = std::min(edge8((uint32_t)key & 255u), et);
                const uint64_t g0 = slot / 64, g1 = (slot + c - 1) / 64;
                for (uint64_t g = g0; g <= g1 && g < lgroups; g++) {
                    m->h_gqw[g] = std::max(m->h_gqw[g], qe);
                    gtw[g] = std::max(gtw[g], te);
                }
                slot += c;
            }
            uint64_t tot = 0;   // dwords
            for (uint64_t g = 0; g < lgroups; g++) {
                m->h_gbase[g] = tot;
                tot += (uint64_t)gtw[g] * m->h_gqw[g] * 64 * 8;
            }
            ensure(m->tb, m->tb_bytes, (size_t)(nreg * tot) * 4 + 256);
            if (ck) d.tbh = m->tb + tot;
            ensure(m->tb_gbase, m->tb_gbase_bytes, (size_t)lgroups * 8 + 256);
            ensure(m->tb_gqw, m->tb_gqw_bytes, (size_t)lgroups * 4 + 256);
            G2_CHECK(hipMemcpyAsync(m->tb_gbase, m->h_gbase.data(), (size_t)lgroups * 8, hipMemcpyHostToDevice, s->str));
            G2_CHECK(hipMemcpyAsync(m->tb_gqw, m->h_gqw.data(), (size_t)lgroups * 4, hipMemcpyHostToDevice, s->str));
            d.tb_gbase = m->tb_gbase;
            d.tb_gqw = m->tb_gqw;
        }
        d.tb = m->tb;
        d.tb_tw = et;
        d.tb_qw8 = 8 * eq;
        d.tb_fused = d.tbh && g_tb_fused ? 1u : 0u;
        s->packed_tb_matrices = (uint4 *)m->tb;
        s->packed_tb_matrix_size = m->tb_bytes / 16;
    }
    g2_choose_packed(d, op, sc, eq, et, px.banded);
}

G2Scores scores_now() { return g_scores; }

void upload_async(gasal_gpu_storage_t *s, uint32_t n, Parameters *params) {
    for (host_batch_t *cur = s->extensible_host_unpacked_query_batch; cur != NULL; cur = cur->next)
        if (cur->data_size) G2_CHECK(hipMemcpyAsync(&(s->unpacked_query_batch[cur->offset]), cur->data, cur->data_size, hipMemcpyHostToDevice, s->str));
    for (host_batch_t *cur = s->extensible_host_unpacked_target_batch; cur != NULL; cur = cur->next)
        if (cur->data_size) G2_CHECK(hipMemcpyAsync(&(s->unpacked_target_batch[cur->offset]), cur->data, cur->data_size, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(s->query_batch_lens, s->host_query_batch_lens, (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(s->target_batch_lens, s->host_target_batch_lens, (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(s->query_batch_offsets, s->host_query_batch_offsets, (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(s->target_batch_offsets, s->host_target_batch_offsets, (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    if (params->algo == KSW) {
        if (s->seed_scores == NULL) fprintf(stderr, "seed_scores == NULL\n");
        if (s->host_seed_scores == NULL) fprintf(stderr, "host_seed_scores == NULL\n");
        if (s->seed_scores == NULL || s->host_seed_scores == NULL) { report("[GASAL ERROR:] KSW seed scores missing\n"); gasal2_fail(EXIT_FAILURE); return; }
        G2_CHECK(hipMemcpyAsync(s->seed_scores, s->host_seed_scores, (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    }
    if (params->isReverseComplement) {
        G2_CHECK(hipMemcpyAsync(s->query_op, s->host_query_op, n, hipMemcpyHostToDevice, s->str));
        G2_CHECK(hipMemcpyAsync(s->target_op, s->host_target_op, n, hipMemcpyHostToDevice, s->str));
    }
}

// Graph mode: the device program (a dozen kernels) is captured once per batch
// shape -- every kernel argument is in (G2Dev, G2Scores, G2Opts) -- and
// replayed with one hipGraphLaunch while the shape repeats.  Workspace growth,
// new n or byte counts, scores or dispatch settings change the arguments and
// trigger a new capture.  The program holds kernels only (no memset or copy
// nodes).  Timed runs (per-stage events) launch eagerly.
hipError_t launch_program(Impl *m, const G2Scores &sc, const G2Opts &op, const G2Dev &d, hipStream_t st) {
    m->last_paths = ((d.pk16 || d.pk2 || d.pk16g || d.pk16s || d.ksw16) ? 1 : 0) | (d.wf_blocks ? 2 : 0) | (d.tbh ? 4 : 0);
    if (!g_graph_mode || m->tm.enabled) return g2_run(sc, op, d, st, &m->tm);
    const bool same = m->g_valid && memcmp(&m->g_dev, &d, sizeof(d)) == 0 && memcmp(&m->g_sc, &sc, sizeof(sc)) == 0 &&
                      memcmp(&m->g_op, &op, sizeof(op)) == 0;
    if (!same) {
        if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
        if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
        m->g_valid = false;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t er = g2_run(sc, op, d, st, nullptr);
        hipGraph_t gr = nullptr;
        e = hipStreamEndCapture(st, &gr);
        if (er != hipSuccess) { if (gr) (void)hipGraphDestroy(gr); return er; }
        if (e != hipSuccess) return e;
        m->graph = gr;
        e = hipGraphInstantiate(&m->gexec, gr, nullptr, nullptr, 0);
        if (e != hipSuccess) return e;
        m->g_dev = d; m->g_sc = sc; m->g_op = op;
        m->g_valid = true;
        m->graph_captures++;
    }
    m->graph_launches++;
    return hipGraphLaunch(m->gexec, st);
}

// Enqueues the device program; false (nothing launched) when a workspace
// allocation failed in error mode 1.
//
// With isPacked the caller's words are the packed buffer itself, which the
// reverse/complement step rewrites in place (the reference's aliasing,
// src/ctors.cpp:64-67); a rerun of the same upload (gasal2_run_resident)
// therefore skips that step.
bool run_async(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params) {
    G2Dev d;
    Impl *m = impl_of(s);
    t_failed = false;
    prepare(s, qb, tb, n, params, d);
    if (t_failed) {
        m->main_valid = false;
        return false;
    }
    m->main_dev = d;
    m->main_valid = true;
    G2Scores sc = scores_now();
    G2Opts op = opts_of(params, s);
    const bool in_place = params->isPacked && params->isReverseComplement;
    if (in_place && m->prepped) op.skip_prep = 1;
    hipError_t e = launch_program(m, sc, op, d, s->str);
    if (e != hipSuccess) {
        report("[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n", hipGetErrorString(e), (int)e, __LINE__, __FILE__);
        gasal2_fail(EXIT_FAILURE);
        return false;
    }
    if (in_place) m->prepped = true;   // the resident words are reverse-complemented now
    return true;
}

void download_async(gasal_gpu_storage_t *s, uint32_t qb, uint32_t n, Parameters *params) {
    gasal_res_t *h = s->host_res, *dc = s->device_cpy;
    if (h->aln_score && dc->aln_score) G2_CHECK(hipMemcpyAsync(h->aln_score, dc->aln_score, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    if (h->query_batch_start && dc->query_batch_start) G2_CHECK(hipMemcpyAsync(h->query_batch_start, dc->query_batch_start, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    if (h->target_batch_start && dc->target_batch_start) G2_CHECK(hipMemcpyAsync(h->target_batch_start, dc->target_batch_start, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    if (h->query_batch_end && dc->query_batch_end) G2_CHECK(hipMemcpyAsync(h->query_batch_end, dc->query_batch_end, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    if (h->target_batch_end && dc->target_batch_end) G2_CHECK(hipMemcpyAsync(h->target_batch_end, dc->target_batch_end, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    if (params->start_pos == WITH_TB) {
        Impl *m = impl_of(s);
        if (h->cigar && m->cigar_q) G2_CHECK(hipMemcpyAsync(h->cigar, m->cigar_q, qb, hipMemcpyDeviceToHost, s->str));
        if (h->n_cigar_ops && m->nops) G2_CHECK(hipMemcpyAsync(h->n_cigar_ops, m->nops, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    }
    if (params->secondBest && s->host_res_second && s->device_cpy_second) {
        gasal_res_t *h2 = s->host_res_second, *d2 = s->device_cpy_second;
        if (h2->aln_score && d2->aln_score) G2_CHECK(hipMemcpyAsync(h2->aln_score, d2->aln_score, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
        if (h2->query_batch_end && d2->query_batch_end) G2_CHECK(hipMemcpyAsync(h2->query_batch_end, d2->query_batch_end, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
        if (h2->target_batch_end && d2->target_batch_end) G2_CHECK(hipMemcpyAsync(h2->target_batch_end, d2->target_batch_end, (size_t)n * 4, hipMemcpyDeviceToHost, s->str));
    }
}

}  // namespace

// src/gasal_align.cu:29-307
extern "C" void gasal_aln_async(gasal_gpu_storage_t *gpu_storage, const uint32_t actual_query_batch_bytes,
                                const uint32_t actual_target_batch_bytes, const uint32_t actual_n_alns,
                                Parameters *params) {
    t_failed = false;
    if (!validate(gpu_storage, actual_query_batch_bytes, actual_target_batch_bytes, actual_n_alns)) return;
    if (!validate_batch(gpu_storage, actual_query_batch_bytes, actual_target_batch_bytes, actual_n_alns, params)) return;
    // error mode 1: a failed allocation or copy ends the call before any later stage
    if (!grow_device(gpu_storage, actual_query_batch_bytes, actual_target_batch_bytes, actual_n_alns, params)) return;
    Impl *m = impl_of(gpu_storage);
    m->prepped = false;   // fresh words
    gasal2_order *oc = m->chain ? m->chain : &g_chain;
    const int order = oc->mode.load();
    std::unique_lock<std::mutex> lk(oc->mu, std::defer_lock);
    if (order) {
        lk.lock();
        if (!m->ev_h2d) G2_CHECK(hipEventCreateWithFlags(&m->ev_h2d, hipEventDisableTiming));
        if (!m->ev_dev) G2_CHECK(hipEventCreateWithFlags(&m->ev_dev, hipEventDisableTiming));
        if (t_failed) return;
        if (oc->last_h2d) G2_CHECK(hipStreamWaitEvent(gpu_storage->str, oc->last_h2d, 0));
    }
    upload_async(gpu_storage, actual_n_alns, params);
    if (t_failed) return;
    if (order) {
        G2_CHECK(hipEventRecord(m->ev_h2d, gpu_storage->str));
        oc->last_h2d = m->ev_h2d;
        if (order == 2 && oc->last_dev) G2_CHECK(hipStreamWaitEvent(gpu_storage->str, oc->last_dev, 0));
    }
    if (!run_async(gpu_storage, actual_query_batch_bytes, actual_target_batch_bytes, actual_n_alns, params)) return;
    if (order) {
        G2_CHECK(hipEventRecord(m->ev_dev, gpu_storage->str));
        oc->last_dev = m->ev_dev;
        lk.unlock();
    }
    download_async(gpu_storage, actual_query_batch_bytes, actual_n_alns, params);
    if (t_failed) return;
    gpu_storage->is_free = 0;
}

// src/gasal_align.cu:310-326
extern "C" int gasal_is_aln_async_done(gasal_gpu_storage_t *gpu_storage) {
    if (gpu_storage->is_free == 1) return -2;
    hipError_t err = hipStreamQuery(gpu_storage->str);
    if (err != hipSuccess) {
        if (err == hipErrorNotReady) return -1;
        report("[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n", hipGetErrorString(err), (int)err, __LINE__, __FILE__);
        gasal2_fail(EXIT_FAILURE);
        return -3;
    }
    gasal_host_batch_reset(gpu_storage);
    gpu_storage->is_free = 1;
    gpu_storage->current_n_alns = 0;
    return 0;
}

extern "C" void gasal2_upload(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params) {
    t_failed = false;
    if (!validate(s, qb, tb, n)) return;
    if (!validate_batch(s, qb, tb, n, params)) return;
    impl_of(s)->prepped = false;
    if (!grow_device(s, qb, tb, n, params)) return;
    upload_async(s, n, params);
    if (t_failed) return;
    G2_CHECK(hipStreamSynchronize(s->str));
}

extern "C" void gasal2_run_resident(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params) {
    t_failed = false;
    if (!validate(s, qb, tb, n)) return;
    run_async(s, qb, tb, n, params);
}

extern "C" void gasal2_download(gasal_gpu_storage_t *s, uint32_t qb, uint32_t n, Parameters *params) {
    download_async(s, qb, n, params);
    G2_CHECK(hipStreamSynchronize(s->str));
}

extern "C" void gasal2_synchronize(gasal_gpu_storage_t *s) { G2_CHECK(hipStreamSynchronize(s->str)); }

extern "C" void gasal2_enable_timing(gasal_gpu_storage_t *s, int32_t on) {
    Impl *m = impl_of(s);
    if (on && !m->tm_init) {
        for (int i = 0; i < 6; i++) G2_CHECK(hipEventCreate(&m->tm.ev[i]));
        m->tm_init = true;
    }
    m->tm.enabled = on && m->tm_init;
}

extern "C" float gasal2_stage_ms(gasal_gpu_storage_t *s, int32_t stage) {
    Impl *m = impl_of(s);
    if (!m->tm_init || !m->tm.enabled) return -1.0f;
    static const int a[5] = {0, 0, 1, 2, 3}, b[5] = {5, 1, 2, 3, 4};
    if (stage < 0 || stage > 4) return -1.0f;
    float ms = -1.0f;
    if (hipEventSynchronize(m->tm.ev[b[stage]]) != hipSuccess) return -1.0f;
    if (hipEventElapsedTime(&ms, m->tm.ev[a[stage]], m->tm.ev[b[stage]]) != hipSuccess) return -1.0f;
    return ms;
}

// ------------------------------------------------ banded runs and fallback ---
// Packs and reverse-complements a resident batch (gasal2_upload) in place, so
// that device-side consumers (gasal2_lb_features) can read the packed words
// (gasal2_device_views) before gasal2_run_banded.
extern "C" int32_t gasal2_prep_resident(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params) {
    if (!validate(s, qb, tb, n)) return -1;
    Impl *m = impl_of(s);
    G2Dev d;
    PrepExtra px;
    px.banded = true;   // no direction-tile workspace for this call
    t_failed = false;
    prepare(s, qb, tb, n, params, d, px);
    if (t_failed) return -1;
    G2Opts op = opts_of(params, s);
    op.prep_only = 1;
    hipError_t e = g2_run(scores_now(), op, d, s->str, nullptr);
    if (e != hipSuccess) {
        report("[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n", hipGetErrorString(e), (int)e, __LINE__, __FILE__);
        gasal2_fail(EXIT_FAILURE);
        return -1;
    }
    m->prepped = true;
    return 0;
}

// Device pointers of the resident batch: packed query/target words, offsets
// and lengths (valid until the storage grows or is destroyed).
extern "C" void gasal2_device_views(gasal_gpu_storage_t *s, void **pq, void **pt, void **qoff, void **toff, void **qlen,
                                    void **tlen) {
    if (pq) *pq = s->packed_query_batch;
    if (pt) *pt = s->packed_target_batch;
    if (qoff) *qoff = s->query_batch_offsets;
    if (toff) *toff = s->target_batch_offsets;
    if (qlen) *qlen = s->query_batch_lens;
    if (tlen) *tlen = s->target_batch_lens;
}

// Device pointers of the last run's results: score, query end, target end,
// n_ops and per-pair CIGAR slots of *cap bytes (reverse order).
extern "C" void gasal2_result_views(gasal_gpu_storage_t *s, void **score, void **qend, void **tend, void **nops,
                                    void **cigar, uint32_t *cap) {
    Impl *m = impl_of(s);
    const G2Dev &d = m->main_dev;
    if (score) *score = m->main_valid ? (void *)d.score : nullptr;
    if (qend) *qend = m->main_valid ? (void *)d.qend : nullptr;
    if (tend) *tend = m->main_valid ? (void *)d.tend : nullptr;
    if (nops) *nops = m->main_valid ? (void *)d.n_ops : nullptr;
    if (cigar) *cigar = m->main_valid ? (void *)d.cigar : nullptr;
    if (cap) *cap = m->main_valid ? d.cigar_cap : 0;
}

// Banded LOCAL over per-pair diagonal bands (learned-band path): the batch
// must be resident and prepared (gasal2_upload, gasal2_prep_resident).  Results stay in device memory like
// gasal2_run_resident; flags (host, n bytes) come back synchronously:
// bit0 best cell on the band edge, bit1 traceback left the band, bit2 the
// check band (widened by check_widen each side) disagrees, bit3 not run.
// Any flagged pair should be re-run with gasal2_run_subset.
extern "C" int32_t gasal2_run_banded(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params,
                                     const int32_t *dlo, const int32_t *dhi, uint8_t *flags, int32_t check_widen) {
    if (!validate(s, qb, tb, n)) return -1;
    if (params->algo != LOCAL || params->secondBest || params->start_pos == WITH_START) {
        report("[GASAL ERROR:] banded runs take LOCAL WITHOUT_START or WITH_TB without second best\n");
        gasal2_fail(EXIT_FAILURE);
        return -1;
    }
    Impl *m = impl_of(s);
    PrepExtra px;
    px.banded = true;
    G2Dev d;
    t_failed = false;
    prepare(s, qb, tb, n, params, d, px);
    if (t_failed) {
        m->main_valid = false;
        return -1;
    }
    m->main_dev = d;
    m->main_valid = true;
    if (!m->prepped) {
        report("[GASAL ERROR:] gasal2_run_banded needs gasal2_prep_resident on the uploaded batch first\n");
        gasal2_fail(EXIT_FAILURE);
        return -1;
    }
    if (!d.fast) {   // substitution scores outside the 4-bit tables: every pair goes to the full DP
        if (flags) memset(flags, 8, n);
        return 1;
    }
    const bool tbo_on = params->start_pos == WITH_TB;
    m->h_lo.assign(dlo, dlo + n);
    m->h_hi.assign(dhi, dhi + n);
    m->h_bw.resize(n);
    m->h_tbo.resize(n);
    uint64_t tot = 0;
    for (uint32_t k = 0; k < n; k++) {
        const int64_t QW = g2_regs(s->host_query_batch_lens[k]), TW = g2_regs(s->host_target_batch_lens[k]);
        const int64_t span = (int64_t)dhi[k] - (int64_t)dlo[k];
        int64_t bw = span < 0 ? 1 : ((span + 14) >> 3) + 1;   // >= words any strip of the band touches
        bw = std::max<int64_t>(1, std::min<int64_t>(bw, QW));
        m->h_bw[k] = (uint32_t)bw;
        m->h_tbo[k] = tot;
        if (tbo_on) tot += (uint64_t)TW * (uint64_t)bw;
    }
    if (m->band_cap < n || !m->band_lo) {
        for (void **pp : {(void **)&m->band_lo, (void **)&m->band_hi, (void **)&m->band_bw, (void **)&m->band_tbo,
                          (void **)&m->band_flags})
            if (*pp) G2_CHECK(hipFree(*pp));
        G2_CHECK(hipMalloc((void **)&m->band_lo, (size_t)n * 4 + 256));
        G2_CHECK(hipMalloc((void **)&m->band_hi, (size_t)n * 4 + 256));
        G2_CHECK(hipMalloc((void **)&m->band_bw, (size_t)n * 4 + 256));
        G2_CHECK(hipMalloc((void **)&m->band_tbo, (size_t)n * 8 + 256));
        G2_CHECK(hipMalloc((void **)&m->band_flags, (size_t)n + 256));
        m->band_cap = n;
    }
    G2_CHECK(hipMemcpyAsync(m->band_lo, m->h_lo.data(), (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(m->band_hi, m->h_hi.data(), (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(m->band_bw, m->h_bw.data(), (size_t)n * 4, hipMemcpyHostToDevice, s->str));
    G2_CHECK(hipMemcpyAsync(m->band_tbo, m->h_tbo.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->str));
    if (tbo_on) {
        ensure(m->tb, m->tb_bytes, (size_t)tot * 32 + 256);
        d.tb = m->tb;
    }
    d.band_lo = m->band_lo; d.band_hi = m->band_hi; d.band_bw = m->band_bw; d.band_tbo = m->band_tbo;
    d.band_flags = m->band_flags;
    if (check_widen > 0) {
        ensure(m->band_wres, m->band_wres_bytes, (size_t)n * 12);
        d.band_wres = m->band_wres;
        d.band_widen = check_widen;
    }
    m->main_dev = d;
    G2Opts op = opts_of(params, s);
    op.banded = 1;
    op.skip_prep = 1;
    hipError_t e = launch_program(m, scores_now(), op, d, s->str);
    if (e != hipSuccess) {
        report("[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n", hipGetErrorString(e), (int)e, __LINE__, __FILE__);
        gasal2_fail(EXIT_FAILURE);
        return -1;
    }
    if (flags) G2_CHECK(hipMemcpyAsync(flags, m->band_flags, n, hipMemcpyDeviceToHost, s->str));
    G2_CHECK(hipStreamSynchronize(s->str));
    return 0;
}

// Full DP for pairs idx[0..cnt) of the last run (banded or full) on this
// storage, results written over theirs in device memory (the batch stays
// packed and reverse-complemented on the device).  Pairs go in chunks whose
// direction-tile workspace stays under max_tb_bytes (0: 32 GiB).
extern "C" int32_t gasal2_run_subset(gasal_gpu_storage_t *s, uint32_t qb, uint32_t tb, uint32_t n, Parameters *params,
                                     const uint32_t *idx, uint32_t cnt, uint64_t max_tb_bytes) {
    Impl *m = impl_of(s);
    if (!m->main_valid || m->main_dev.n != n) {
        report("[GASAL ERROR:] gasal2_run_subset needs a preceding run of the same batch on this storage\n");
        gasal2_fail(EXIT_FAILURE);
        return -1;
    }
    if (cnt == 0) return 0;
    for (uint32_t k = 0; k < cnt; k++)
        if (idx[k] >= n) { report("[GASAL ERROR:] gasal2_run_subset: pair index %u >= %u\n", idx[k], n); gasal2_fail(EXIT_FAILURE); return -1; }
    if (!max_tb_bytes) max_tb_bytes = 32ull << 30;
    const bool tb_on = params->start_pos == WITH_TB;
    const G2Dev main = m->main_dev;
    G2Opts op = opts_of(params, s);
    op.skip_prep = 1;
    uint32_t k0 = 0;
    while (k0 < cnt) {
        // grow the chunk while the full-matrix tile workspace fits
        uint32_t k1 = k0, mq = 0, mt = 0;
        while (k1 < cnt) {
            const uint32_t q = std::max(mq, g2_regs(s->host_query_batch_lens[idx[k1]]));
            const uint32_t t = std::max(mt, g2_regs(s->host_target_batch_lens[idx[k1]]));
            const uint64_t groups = (k1 - k0 + 1 + 63) / 64;
            if (k1 > k0 && tb_on && groups * t * (8ull * q) * 64 * 4 > max_tb_bytes) break;
            mq = q; mt = t; k1++;
        }
        const uint32_t c = k1 - k0;
        m->h_sub.resize((size_t)c * 5);
        uint32_t *hq_off = m->h_sub.data(), *ht_off = hq_off + c, *hq = ht_off + c, *ht = hq + c, *hidx = ht + c;
        for (uint32_t k = 0; k < c; k++) {
            const uint32_t p = idx[k0 + k];
            hq_off[k] = s->host_query_batch_offsets[p];
            ht_off[k] = s->host_target_batch_offsets[p];
            hq[k] = s->host_query_batch_lens[p];
            ht[k] = s->host_target_batch_lens[p];
            hidx[k] = p;
        }
        G2Dev ds;
        PrepExtra px;
        px.hq = hq; px.ht = ht; px.sub = true;
        t_failed = false;
        prepare(s, qb, tb, c, params, ds, px);
        if (t_failed) return -1;
        if (m->sub_cap < c || !m->sub_u32) {
            for (void **pp : {(void **)&m->sub_u32, (void **)&m->sub_res, (void **)&m->sub_nops})
                if (*pp) G2_CHECK(hipFree(*pp));
            G2_CHECK(hipMalloc((void **)&m->sub_u32, (size_t)c * 5 * 4 + 256));
            G2_CHECK(hipMalloc((void **)&m->sub_res, (size_t)c * 5 * 4 + 256));
            G2_CHECK(hipMalloc((void **)&m->sub_nops, (size_t)c * 4 + 256));
            m->sub_cap = c;
        }
        G2_CHECK(hipMemcpyAsync(m->sub_u32, m->h_sub.data(), (size_t)c * 5 * 4, hipMemcpyHostToDevice, s->str));
        ds.qoff = m->sub_u32; ds.toff = m->sub_u32 + c; ds.qlen = m->sub_u32 + 2 * c; ds.tlen = m->sub_u32 + 3 * c;
        ds.sub_idx = m->sub_u32 + 4 * c;
        ds.score = m->sub_res; ds.qend = m->sub_res + c; ds.tend = m->sub_res + 2 * c;
        ds.qstart = m->sub_res + 3 * c; ds.tstart = m->sub_res + 4 * c;
        ds.s_score = ds.s_qend = ds.s_tend = nullptr;
        if (tb_on) {
            ensure(m->sub_cigar, m->sub_cigar_bytes, (size_t)c * ds.cigar_cap);
            ds.cigar = m->sub_cigar;
            ds.n_ops = m->sub_nops;
            ds.cigar_qlayout = main.cigar_qlayout;   // the reference layout sits at the pairs' own query offsets
        }
        hipError_t e = g2_run(scores_now(), op, ds, s->str, nullptr);
        if (e == hipSuccess) e = g2_scatter_subset(ds, main, tb_on, s->str);
        if (e != hipSuccess) {
            report("[GASAL CUDA ERROR:] %s(CUDA error no.=%d). Line no. %d in file %s\n", hipGetErrorString(e), (int)e, __LINE__, __FILE__);
            gasal2_fail(EXIT_FAILURE);
            return -1;
        }
        G2_CHECK(hipStreamSynchronize(s->str));   // host staging arrays are reused by the next chunk
        k0 = k1;
    }
    return 0;
}

extern "C" void gasal2_graph_stats(gasal_gpu_storage_t *s, uint64_t *captures, uint64_t *launches) {
    Impl *m = impl_of(s);
    if (captures) *captures = m->graph_captures;
    if (launches) *launches = m->graph_launches;
}

extern "C" uint32_t gasal2_get_cigar(gasal_gpu_storage_t *s, uint32_t k, uint8_t *out, uint32_t cap) {
    Impl *m = impl_of(s);
    if (!m->last_tb || k >= m->last_n || !m->nops) return 0;
    uint32_t nops = 0;
    G2_CHECK(hipMemcpy(&nops, m->nops + k, 4, hipMemcpyDeviceToHost));
    uint32_t nb = std::min(nops, std::min(cap, m->last_cigar_cap));
    if (nb && out) G2_CHECK(hipMemcpy(out, m->cigar + (size_t)k * m->last_cigar_cap, nb, hipMemcpyDeviceToHost));
    return nops;
}
