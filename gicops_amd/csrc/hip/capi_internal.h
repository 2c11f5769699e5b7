// Internal state shared by the C ABI translation units (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "batched.h"
#include "gicops_amd.h"
#include "mlp.h"
#include "search_common.h"

namespace gx {
struct Surrogate;
}

using namespace gx;

struct gx_ctx {
    int device = 0;
    int n_cu = 0;
    char name[256] = {};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    std::vector<hipEvent_t> ev_chunk;  // start/stop per planning chunk
    int tuned_config = 1;
    int tuned_split = 1;
    int tuned_whole = 128;  // whole-bin threshold (ions)
    uint64_t scratch_gen = 0;  // bumped whenever d_slices / d_cand move (captured graphs hold them)
    double *d_lgfact = nullptr;
    int lgfact_n = 0;
    uint32_t *d_counter = nullptr;
    unsigned long long *d_stats = nullptr;
    // per-block scratch, grown on demand
    uint32_t *d_slices = nullptr;
    size_t slices_words = 0;
    uint32_t *d_cand = nullptr;
    size_t cand_words = 0;
    gx::Surrogate *surrogate = nullptr;  // learned e-value model (gx_ctx_set_surrogate)
    std::mutex mu;
};

struct gx_dindex {
    gx_ctx *ctx = nullptr;
    uint64_t gen = 0;  // unique per index created in this process (never reused, unlike the address)
    int n_buckets = 0;
    int maxz = 1;
    uint32_t nbins = 0;
    uint32_t max_speclen = 0;
    uint64_t max_ions = 0;            // largest bucket's ion count (picks the search's addressing)
    uint32_t smax_m = 0, smax_s = 0;  // exact division by max_speclen
    uint32_t cell_shift = 7;         // pass-2 cell table: 2^cell_shift bins per cell
    std::vector<void *> allocs;
    DevBucket *d_buckets = nullptr;
    std::vector<DevBucket> h_buckets;
};

struct gx_dqueries {
    gx_ctx *ctx = nullptr;
    uint32_t n_spectra = 0;
    uint64_t n_peaks = 0;
    uint32_t max_peaks = 0;
    bool has_dm = false;
    float prep_ms = 0.0f;  // device time of gx_queries_preprocess (0 for uploaded batches)
    uint32_t *idx = nullptr;
    int32_t *moz = nullptr;
    int32_t *inten = nullptr;
    float *prec = nullptr;
    float *dm = nullptr;
    gx_psm *psm = nullptr;
    uint32_t *hist = nullptr;  // per-spectrum survival histograms (batched: always; plain: on request)
    bool searched = false;
    bool batched_state = false;               // the last search was batched: best_key/eq/shp, hist, cpsms hold it
    const gx_dindex *searched_index = nullptr;  // the index of the last search
    // batched engine workspace, sized on first use
    int nb = 0;
    int32_t *lq = nullptr, *li = nullptr;
    uint32_t *n_legal = nullptr;
    int32_t *pair_lo = nullptr, *pair_hi = nullptr;
    uint32_t *pair_tiles = nullptr, *pair_pack = nullptr, *pair_pos_off = nullptr, *pair_nsl = nullptr;
    uint32_t *n_items = nullptr, *item_off = nullptr;
    uint64_t *n_pos = nullptr, *pos_off = nullptr;
    uint32_t *cpsms = nullptr;
    uint32_t *counters = nullptr;  // one work counter per chunk
    size_t counters_n = 0;
    unsigned long long *best_key = nullptr;  // per-spectrum top hit of the search kernel
    uint32_t *best_eq = nullptr, *best_shp = nullptr;
    uint32_t *porder = nullptr;            // spectra of each chunk by precursor mass (plan)
    std::vector<uint32_t> h_porder;
    // work units of the search kernel (spectra, or bucket-range parts of heavy
    // spectra), largest first within each chunk, and the folds of the parts
    uint32_t *unit_s = nullptr, *unit_br = nullptr;
    size_t unit_cap = 0;
    uint32_t *fold_s = nullptr, *fold_p0 = nullptr, *fold_n = nullptr;
    size_t fold_cap = 0;
    unsigned long long *part_key = nullptr;
    uint32_t *part_eq = nullptr, *part_shp = nullptr, *part_cps = nullptr, *part_hist = nullptr;
    size_t part_cap = 0;
    std::vector<uint32_t> h_unit_s, h_unit_br, unit_bounds, h_fold_s, h_fold_p0, h_fold_n, fold_bounds;
    std::vector<uint32_t> h_pair_off;  // per-pair slice-position offsets (heavy-spectrum splitting)
    uint32_t n_parts_total = 0;
    std::vector<float> h_prec;             // precursor masses (host copy)
    uint32_t *pos = nullptr;
    size_t pos_cap = 0;  // words
    std::vector<uint64_t> h_pos_off;
    // chunking of the last eager run, and the graph captured from it
    std::vector<uint32_t> bounds;
    uint32_t n_items_total = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<char> graph_key;
    uint64_t buf_gen = 0;  // bumped whenever a chunk buffer moves
    uint64_t graph_ctx_gen = 0, graph_buf_gen = 0;
    const gx_dindex *graph_index = nullptr;  // the index the graph was captured on
    uint64_t graph_index_gen = 0;             // and its generation: a freed index's address may come back
    // learned e-value stage: training features and targets; the fused
    // stage's raw outputs and served flags (n x 4)
    bf16 *l_feat = nullptr;
    float *l_tgt = nullptr, *l_loge = nullptr, *l_out = nullptr;
    uint8_t *l_valid = nullptr, *l_serv = nullptr;
    bool learned_done = false;  // the last search ran the learned stage (l_out holds its outputs)
};

namespace gx {
// text of the last HIP failure on this thread (gx_last_error)
extern thread_local char g_last_error[512];
void note_error(const char *what, int line, hipError_t e);
// records "invalid argument: <what>" as the last error text and returns code
int param_error(int code, const char *what);
}  // namespace gx

#define GX_HIP(call)                                   \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) {                        \
            note_error(#call, __LINE__, e_);           \
            return GX_ERR_DEVICE;                      \
        }                                              \
    } while (0)

namespace gx {
// a device index from bucket metadata; bA/iA from the host buckets, or device
// arrays the index takes over (capi.hip)
int index_create(gx_ctx *c, const gx_bucket *b, int nb, int maxz, uint32_t *const *dev_bA, uint32_t *const *dev_iA,
                 gx_dindex **out);

// the no-alloc search's static device scratch: its address on `device`
// (the current device), looked up once per device (noalloc.hip)
int noalloc_setup(int device);

// learned e-value stage of a batched search (mlp_capi.hip)
int ensure_learned(gx_ctx *c, gx_dqueries *q);
hipError_t enqueue_learned_evalue(gx_ctx *c, gx_dqueries *q, const EvalArgs &ea, float conf, hipStream_t s);

template <class T>
int dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return GX_ERR_BAD_MEM_ALLOC;
    }
    return GX_SUCCESS;
}
}  // namespace gx
