// C ABI of include/gicops_amd.h: contexts, device-resident index and query
// batches, and the search entry points. No function here exits or throws
// across the ABI; HIP failures map to GX_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "capi_internal.h"

using namespace gx;

namespace gx {
thread_local char g_last_error[512] = "";
void note_error(const char *what, int line, hipError_t e) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s (line %d): %s", what, line, hipGetErrorString(e));
}

int param_error(int code, const char *what) {
    std::snprintf(g_last_error, sizeof(g_last_error), "invalid argument: %s", what);
    return code;
}
}  // namespace gx

namespace {

int ensure_scratch(gx_ctx *c, int n_blocks, uint32_t max_slices) {
    const size_t sw = (size_t)n_blocks * 2 * max_slices;
    if (sw > c->slices_words) {
        c->scratch_gen++;
        if (c->d_slices) (void)hipFree(c->d_slices);
        c->d_slices = nullptr;
        c->slices_words = 0;
        int st = dalloc(&c->d_slices, sw);
        if (st) return st;
        c->slices_words = sw;
    }
    const size_t cw = (size_t)n_blocks * kTileEntries;
    if (cw > c->cand_words) {
        c->scratch_gen++;
        if (c->d_cand) (void)hipFree(c->d_cand);
        c->d_cand = nullptr;
        c->cand_words = 0;
        int st = dalloc(&c->d_cand, cw);
        if (st) return st;
        c->cand_words = cw;
    }
    return GX_SUCCESS;
}

// Division by an invariant integer d (2 <= d < 2^12) for every n < 2^31 as a
// multiply-high and a shift (Granlund & Montgomery, PLDI 1994, Thm 4.2): with
// l = ceil(log2 d), k = 31 + l and m = ceil(2^k / d) (< 2^32),
// floor(n / d) = floor(n * m / 2^k) = umulhi(n, m) >> (k - 32).
// Checked here against exact division at the values most likely to break it.
bool make_divisor(uint32_t d, uint32_t &m, uint32_t &s) {
    if (d < 2 || d >= 4096) return false;
    uint32_t l = 0;
    while ((1u << l) < d) l++;
    const uint32_t k = 31 + l;
    const unsigned __int128 two_k = (unsigned __int128)1 << k;
    const unsigned __int128 mm = (two_k + d - 1) / d;
    if (mm >= ((unsigned __int128)1 << 32)) return false;
    m = (uint32_t)mm;
    s = k - 32;
    auto div = [&](uint64_t n) { return (uint32_t)(((n * (uint64_t)m) >> 32) >> s); };
    const uint64_t top = (1ull << 31) - 1;
    for (uint64_t q = 0; q < 4096; q++) {
        for (uint64_t n : {q * d, q * d + d - 1, top - q * d, (top / d - q) * d, (top / d - q) * d + d - 1}) {
            if (n > top) continue;
            if (div(n) != n / d) return false;
        }
    }
    return true;
}

int check_params(const gx_search_params *p, const gx_dindex *idx) {
    if (p->dF < 0 || p->dF > 64) return param_error(GX_ERR_INVLD_PARAM, "dF outside 0..64");
    if (p->scale < 1 || p->max_mass < 1) return param_error(GX_ERR_INVLD_PARAM, "scale and max_mass must be >= 1");
    if (p->min_shp < 0 || p->min_cpsm < 0) return param_error(GX_ERR_INVLD_PARAM, "min_shp and min_cpsm must be >= 0");
    if ((uint64_t)p->max_mass * (uint64_t)p->scale != idx->nbins)
        return param_error(GX_ERR_INVLD_INDEX, "max_mass * scale differs from the index's bin count");
    if (p->evalue_mode != GX_EVALUE_TAILFIT && p->evalue_mode != GX_EVALUE_LEARNED)
        return param_error(GX_ERR_INVLD_PARAM, "unknown evalue_mode");
    if (p->whole_bin < 0 || p->whole_bin > 4096) return param_error(GX_ERR_INVLD_PARAM, "whole_bin outside 0..4096");
    if (p->wide_stream != 0 && p->wide_stream != 1) return param_error(GX_ERR_INVLD_PARAM, "wide_stream not 0 or 1");
    return GX_SUCCESS;
}

}  // namespace

extern "C" int gx_ctx_create(int device, gx_ctx **out) {
    if (!out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GX_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return GX_ERR_INVLD_PARAM;
    gx_ctx *c = new (std::nothrow) gx_ctx();
    if (!c) return GX_ERR_BAD_MEM_ALLOC;
    c->device = device;
    auto fail = [&](int st) {
        gx_ctx_destroy(c);
        return st;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(GX_ERR_DEVICE);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(GX_ERR_DEVICE);
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return fail(GX_ERR_NO_DEVICE);
    c->n_cu = prop.multiProcessorCount;
    std::snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(GX_ERR_DEVICE);
    if (int st = noalloc_setup(device)) return fail(st);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    // lgfact[i] = sum_{j=2..i} log10(j) (utils.h:115-131), computed on the host
    // with the same libm the CPU search uses, long enough for any count
    c->lgfact_n = 4096;
    std::vector<double> lg(c->lgfact_n);
    lg[0] = std::log10(1.0);
    lg[1] = std::log10(1.0);
    for (int i = 2; i < c->lgfact_n; i++) lg[i] = std::log10((double)i) + lg[i - 1];
    if (dalloc(&c->d_lgfact, lg.size())) return fail(GX_ERR_BAD_MEM_ALLOC);
    if (dalloc(&c->d_counter, 1) || dalloc(&c->d_stats, (size_t)kStatRows * ST_N)) return fail(GX_ERR_BAD_MEM_ALLOC);
    if (hipMemcpy(c->d_lgfact, lg.data(), lg.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    *out = c;
    return GX_SUCCESS;
}

extern "C" void gx_ctx_destroy(gx_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->d_lgfact) (void)hipFree(c->d_lgfact);
    if (c->d_counter) (void)hipFree(c->d_counter);
    if (c->d_stats) (void)hipFree(c->d_stats);
    if (c->d_slices) (void)hipFree(c->d_slices);
    if (c->d_cand) (void)hipFree(c->d_cand);
    hipEvent_t evs[] = {c->ev0, c->ev1, c->ev2, c->ev3};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int gx_ctx_device_info(gx_ctx *c, char *name, size_t name_len, int *n_cu) {
    if (!c) return GX_ERR_INVLD_PTR;
    if (name && name_len) std::snprintf(name, name_len, "%s", c->name);
    if (n_cu) *n_cu = c->n_cu;
    return GX_SUCCESS;
}

extern "C" int gx_index_upload(gx_ctx *c, const gx_bucket *b, int nb, int maxz, gx_dindex **out) {
    if (!c || !out || (nb > 0 && !b)) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    if (nb < 1 || nb > kMaxBuckets || maxz < 1) return GX_ERR_INVLD_PARAM;
    std::lock_guard<std::mutex> lk(c->mu);
    return gx::index_create(c, b, nb, maxz, nullptr, nullptr, out);
}

namespace gx {

// A device index from bucket metadata. bA/iA come from the host buckets, or,
// when dev_bA/dev_iA are given, are device arrays the index takes over.
// generations of device indexes: a captured graph replays only on the index
// it was captured on, recognised by (address, generation)
std::atomic<uint64_t> g_index_gen{0};

int index_create(gx_ctx *c, const gx_bucket *b, int nb, int maxz, uint32_t *const *dev_bA, uint32_t *const *dev_iA,
                 gx_dindex **out) {
    const bool on_dev = dev_bA && dev_iA;
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    gx_dindex *d = new (std::nothrow) gx_dindex();
    if (!d) return GX_ERR_BAD_MEM_ALLOC;
    d->ctx = c;
    d->gen = ++g_index_gen;
    d->n_buckets = nb;
    d->maxz = maxz;
    d->h_buckets.resize(nb);
    if (on_dev)  // the index owns the given arrays from here on, also on failure
        for (int i = 0; i < nb; i++) {
            if (dev_bA[i]) d->allocs.push_back(dev_bA[i]);
            if (dev_iA[i]) d->allocs.push_back(dev_iA[i]);
        }
    auto fail = [&](int st) {
        for (void *p : d->allocs) (void)hipFree(p);
        if (d->d_buckets) (void)hipFree(d->d_buckets);
        delete d;
        return st;
    };
    uint32_t nbins = 0;
    for (int i = 0; i < nb; i++) {
        const gx_bucket &x = b[i];
        if (x.n_entries == 0) continue;
        if (!x.mass || (!on_dev && (!x.bA || !x.iA))) return fail(GX_ERR_INVLD_PTR);
        if (nbins == 0) nbins = x.nbins;
        if (x.nbins != nbins || x.nbins < 2) return fail(GX_ERR_INVLD_INDEX);
        if (x.speclen != (uint32_t)((x.peplen - 1) * maxz * 2) || x.speclen < 2) return fail(GX_ERR_INVLD_INDEX);
        if ((uint64_t)x.n_entries * x.speclen != x.n_ions || x.n_ions >= (1ull << 32)) return fail(GX_ERR_INVLD_SIZE);
        if (!on_dev && x.bA[x.nbins] != x.n_ions) return fail(GX_ERR_INVLD_INDEX);
        if (x.n_entries >= (1u << 26)) return fail(GX_ERR_INVLD_SIZE);  // packed top-hit key holds 26-bit psids
        if (x.n_ions >= (1ull << 31)) return fail(GX_ERR_INVLD_SIZE);     // ion ids use the 31-bit divisor
    }
    d->nbins = nbins;
    for (int i = 0; i < nb; i++) {
        const gx_bucket &x = b[i];
        DevBucket db{};
        db.n_entries = x.n_entries;
        db.speclen = x.speclen;
        if (x.n_entries > 0 && !make_divisor(x.speclen, db.div_m, db.div_s)) return fail(GX_ERR_INVLD_INDEX);
        if (x.n_entries > 0) {
            float *m = nullptr;
            uint32_t *bA = on_dev ? dev_bA[i] : nullptr, *iA = on_dev ? dev_iA[i] : nullptr, *ib = nullptr,
                     *is = nullptr;
            if (on_dev && (!bA || !iA)) return fail(GX_ERR_INVLD_PTR);
            if (dalloc(&m, x.n_entries) || (!on_dev && (dalloc(&bA, (size_t)x.nbins + 1) || dalloc(&iA, (size_t)x.n_ions + kIaPad))) ||
                dalloc(&ib, x.n_ions) || dalloc(&is, (x.n_ions + kSampleStep - 1) >> kSampleShift)) {
                for (void *q : {(void *)m, (void *)ib, (void *)is}) if (q) (void)hipFree(q);
                if (!on_dev) {
                    if (bA) (void)hipFree(bA);
                    if (iA) (void)hipFree(iA);
                }
                return fail(GX_ERR_BAD_MEM_ALLOC);
            }
            d->allocs.push_back(m);
            if (!on_dev) {
                d->allocs.push_back(bA);
                d->allocs.push_back(iA);
            }
            d->allocs.push_back(ib);
            d->allocs.push_back(is);
            if (hipMemcpy(m, x.mass, sizeof(float) * x.n_entries, hipMemcpyHostToDevice) != hipSuccess ||
                (!on_dev &&
                 (hipMemcpy(bA, x.bA, sizeof(uint32_t) * ((size_t)x.nbins + 1), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(iA, x.iA, sizeof(uint32_t) * x.n_ions, hipMemcpyHostToDevice) != hipSuccess)))
                return fail(GX_ERR_DEVICE);
            if (launch_ion_bin(bA, iA, ib, x.nbins, x.n_ions, c->stream) != hipSuccess ||
                launch_sample_ids(iA, is, x.n_ions, c->stream) != hipSuccess)
                return fail(GX_ERR_DEVICE);
            // iA[n_ions .. n_ions + kIaPad) hold no ion id (0xffffffff): the search
            // kernel's lanes past a round's ions read them and drop them with the ids
            // outside their tile
            if (hipMemsetAsync(iA + x.n_ions, 0xff, sizeof(uint32_t) * kIaPad, c->stream) != hipSuccess)
                return fail(GX_ERR_DEVICE);
            db.mass = m;
            db.bA = bA;
            db.iA = iA;
            db.ion_bin = ib;
            db.iS = is;
            d->max_speclen = std::max(d->max_speclen, x.speclen);
            d->max_ions = std::max(d->max_ions, (uint64_t)x.n_ions);
        }
        d->h_buckets[i] = db;
    }
    if (!make_divisor(std::max(d->max_speclen, 2u), d->smax_m, d->smax_s)) return fail(GX_ERR_INVLD_INDEX);
    while ((nbins >> d->cell_shift) + 1 > (uint32_t)kPeakIdxWords) d->cell_shift++;
    if (dalloc(&d->d_buckets, nb)) return fail(GX_ERR_BAD_MEM_ALLOC);
    if (hipMemcpy(d->d_buckets, d->h_buckets.data(), sizeof(DevBucket) * nb, hipMemcpyHostToDevice) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(GX_ERR_DEVICE);
    *out = d;
    return GX_SUCCESS;
}

}  // namespace gx

extern "C" void gx_index_free(gx_dindex *d) {
    if (!d) return;
    (void)hipSetDevice(d->ctx->device);
    (void)hipStreamSynchronize(d->ctx->stream);
    for (void *p : d->allocs) (void)hipFree(p);
    if (d->d_buckets) (void)hipFree(d->d_buckets);
    delete d;
}

extern "C" int gx_queries_upload(gx_ctx *c, const gx_query_batch *q, gx_dqueries **out) {
    if (!c || !q || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    if (q->n_spectra > 0 && (!q->idx || !q->precurse)) return GX_ERR_INVLD_PTR;
    if (q->n_peaks > 0 && (!q->moz || !q->intensity)) return GX_ERR_INVLD_PTR;
    if (q->n_spectra > 0 && q->idx[q->n_spectra] != q->n_peaks) return GX_ERR_INVLD_SIZE;
    if (q->n_peaks >= (1ull << 32)) return GX_ERR_INVLD_SIZE;
    uint32_t maxp = 0;
    for (uint32_t s = 0; s < q->n_spectra; s++) {
        if (q->idx[s + 1] < q->idx[s]) return GX_ERR_INVLD_SIZE;
        maxp = std::max(maxp, q->idx[s + 1] - q->idx[s]);
    }
    if (maxp > (uint32_t)kMaxPeaksLds) return GX_ERR_INVLD_SIZE;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    gx_dqueries *d = new (std::nothrow) gx_dqueries();
    if (!d) return GX_ERR_BAD_MEM_ALLOC;
    d->ctx = c;
    d->n_spectra = q->n_spectra;
    d->n_peaks = q->n_peaks;
    d->max_peaks = maxp;
    d->has_dm = q->dm != nullptr;
    d->h_prec.assign(q->precurse, q->precurse + q->n_spectra);
    auto fail = [&](int st) {
        gx_queries_free(d);
        return st;
    };
    const size_t ns = q->n_spectra, np = q->n_peaks;
    if (dalloc(&d->idx, ns + 1) || dalloc(&d->moz, np) || dalloc(&d->inten, np) || dalloc(&d->prec, ns) ||
        dalloc(&d->psm, ns) || (d->has_dm && dalloc(&d->dm, ns)))
        return fail(GX_ERR_BAD_MEM_ALLOC);
    if (ns) {
        if (hipMemcpy(d->idx, q->idx, (ns + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d->prec, q->precurse, ns * 4, hipMemcpyHostToDevice) != hipSuccess ||
            (d->has_dm && hipMemcpy(d->dm, q->dm, ns * 4, hipMemcpyHostToDevice) != hipSuccess))
            return fail(GX_ERR_DEVICE);
    }
    if (np) {
        if (hipMemcpy(d->moz, q->moz, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d->inten, q->intensity, np * 4, hipMemcpyHostToDevice) != hipSuccess)
            return fail(GX_ERR_DEVICE);
    }
    *out = d;
    return GX_SUCCESS;
}

extern "C" void gx_queries_free(gx_dqueries *d) {
    if (!d) return;
    (void)hipSetDevice(d->ctx->device);
    (void)hipStreamSynchronize(d->ctx->stream);
    if (d->graph_exec) (void)hipGraphExecDestroy(d->graph_exec);
    if (d->graph) (void)hipGraphDestroy(d->graph);
    void *ps[] = {d->idx,       d->moz,        d->inten,         d->prec,         d->dm,       d->psm,
                  d->hist,      d->lq,         d->li,            d->n_legal,      d->pair_lo,  d->pair_hi,
                  d->pair_tiles, d->pair_pack, d->pair_pos_off, d->pair_nsl, d->n_items,     d->item_off, d->n_pos,
                  d->pos_off,   d->cpsms,      d->counters,      d->best_key,     d->best_eq,  d->best_shp,
                  d->pos,       d->unit_s,     d->unit_br,       d->fold_s,       d->fold_p0,  d->fold_n,
                  d->part_key,  d->part_eq,    d->part_shp,      d->part_cps,     d->part_hist, d->porder,           d->l_feat,
                  d->l_tgt,      d->l_loge,     d->l_out,         d->l_valid,      d->l_serv};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete d;
}

namespace {

void read_stats(gx_ctx *c, unsigned long long *h) {
    std::vector<unsigned long long> rows((size_t)kStatRows * ST_N);
    for (int k = 0; k < ST_N; k++) h[k] = 0;
    if (hipMemcpy(rows.data(), c->d_stats, rows.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return;
    for (int r = 0; r < kStatRows; r++)
        for (int k = 0; k < ST_N; k++) h[k] += rows[(size_t)r * ST_N + k];
}

void fill_stats(gx_search_stats *stats, const unsigned long long *h) {
    std::memset(stats, 0, sizeof(*stats));
    stats->ions = h[ST_IONS];
    stats->bins = h[ST_BINS];
    stats->window = h[ST_WINDOW];
    stats->candidates = h[ST_CAND];
    stats->spectra = h[ST_SPECTRA];
    stats->peaks = h[ST_PEAKS];
    stats->fallbacks = h[ST_FALLBACK];
    stats->ions_scanned = h[ST_SCANNED];
    stats->errors = h[ST_ERRORS];
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// one block per spectrum, tail fit fused (search.hip)
int search_plain(gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p,
                 gx_search_stats *stats, bool want_hist) {
    if (want_hist && !q->hist) {
        int st = dalloc(&q->hist, (size_t)q->n_spectra * GX_HIST_BINS);
        if (st) return st;
    }
    const uint32_t nb5 = (uint32_t)(2 * p->dF + 1);
    const uint32_t max_slices = std::max<uint32_t>(1, q->max_peaks * nb5);
    const int n_blocks = std::max(1, c->n_cu * 2);
    int st = ensure_scratch(c, n_blocks, max_slices);
    if (st) return st;
    SearchArgs a{};
    a.buckets = idx->d_buckets;
    a.n_buckets = idx->n_buckets;
    a.qidx = q->idx;
    a.moz = q->moz;
    a.inten = q->inten;
    a.prec = q->prec;
    a.dm = q->has_dm ? q->dm : nullptr;
    a.n_spectra = q->n_spectra;
    a.dF = p->dF;
    a.nbins = (int)idx->nbins;
    a.min_shp = p->min_shp;
    a.min_cpsm = p->min_cpsm;
    a.dM = p->dM;
    a.expect_max = p->expect_max;
    a.lgfact = c->d_lgfact;
    a.lgfact_n = c->lgfact_n;
    a.out = q->psm;
    a.hist_out = want_hist ? q->hist : nullptr;
    a.work_counter = c->d_counter;
    a.slice_scratch = c->d_slices;
    a.cand_scratch = c->d_cand;
    a.max_slices = max_slices;
    a.stats = c->d_stats;
    GX_HIP(launch_fill_u32(reinterpret_cast<uint32_t *>(c->d_stats), (uint64_t)kStatRows * ST_N * 2, 0, c->stream));
    GX_HIP(launch_reset(c->d_counter, c->d_stats, c->stream));
    GX_HIP(hipEventRecord(c->ev0, c->stream));
    if (q->n_spectra > 0) GX_HIP(launch_search(a, n_blocks, c->stream));
    GX_HIP(hipEventRecord(c->ev1, c->stream));
    GX_HIP(hipStreamSynchronize(c->stream));
    q->searched = true;
    q->learned_done = false;
    q->batched_state = false;
    q->searched_index = idx;
    if (stats) {
        unsigned long long h[ST_N];
        read_stats(c, h);
        fill_stats(stats, h);
        stats->kernel_ms = elapsed(c->ev0, c->ev1);
        stats->total_ms = stats->kernel_ms;
        stats->config = -1;
    }
    return GX_SUCCESS;
}

// size the per-spectrum workspace of the batched engine
int ensure_batched(gx_dqueries *q, const gx_dindex *idx) {
    if (q->lq && q->nb == idx->n_buckets) return GX_SUCCESS;
    if (q->lq) return param_error(GX_ERR_INVLD_PARAM, "a batch serves one index shape");
    const size_t ns = q->n_spectra, np = q->n_peaks, npairs = ns * (size_t)idx->n_buckets;
    q->nb = idx->n_buckets;
    if (dalloc(&q->lq, np) || dalloc(&q->li, np) || dalloc(&q->n_legal, ns) || dalloc(&q->pair_lo, npairs) ||
        dalloc(&q->pair_hi, npairs) || dalloc(&q->pair_tiles, npairs) || dalloc(&q->pair_pack, npairs) ||
        dalloc(&q->pair_pos_off, npairs) || dalloc(&q->pair_nsl, npairs) || dalloc(&q->n_items, ns) || dalloc(&q->best_key, ns) || dalloc(&q->porder, ns) ||
        dalloc(&q->best_eq, ns) || dalloc(&q->best_shp, ns) ||
        dalloc(&q->item_off, ns + 1) || dalloc(&q->n_pos, ns) || dalloc(&q->pos_off, ns + 1) ||
        dalloc(&q->cpsms, ns) || (!q->hist && dalloc(&q->hist, ns * GX_HIST_BINS)))
        return GX_ERR_BAD_MEM_ALLOC;
    q->h_pos_off.resize(ns + 1);
    return GX_SUCCESS;
}

// replace a device array by a new one of n elements (contents not kept)
template <class T>
int realloc_dev(T **p, size_t n) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    return dalloc(p, n);
}

template <class T>
int grow(T **p, size_t &cap, size_t n, uint64_t &gen) {
    if (n <= cap && *p) return GX_SUCCESS;
    gen++;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    size_t want = n + n / 8 + 64;
    int st = dalloc(p, want);
    if (!st) cap = want;
    return st;
}

constexpr uint64_t kPosChunkWords = 1ull << 32;  // 16 GB of slice positions per planning chunk

struct Batched {
    int cfg_id;
    int split;
    BatchConfig cfg;
    PlanArgs pa;
    ItemArgs ia;
    EvalArgs ea;
    int n_blocks;
    bool learned;  // e-values from the surrogate, plain fit as fallback
    float conf;
};

Batched setup_batched(gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p) {
    Batched B{};
    B.cfg_id = p->config >= 0 ? std::min(p->config, kNumBatchConfigs - 1) : c->tuned_config;
    B.split = p->split > 0 ? p->split : c->tuned_split;
    if (B.split != 1 && B.split != 2 && B.split != 4) B.split = 1;
    B.cfg = batch_config(B.cfg_id);
    PlanArgs &pa = B.pa;
    pa.buckets = idx->d_buckets;
    pa.n_buckets = idx->n_buckets;
    pa.qidx = q->idx;
    pa.moz = q->moz;
    pa.inten = q->inten;
    pa.prec = q->prec;
    pa.dm = q->has_dm ? q->dm : nullptr;
    pa.n_spectra = q->n_spectra;
    pa.dF = p->dF;
    pa.nbins = (int)idx->nbins;
    pa.dM = p->dM;
    pa.tile_words = B.cfg.tile_words / B.split;
    pa.whole_bin = (uint32_t)(p->whole_bin > 0 ? p->whole_bin : c->tuned_whole);
    {
        // larger blocks plan faster on indexes with large buckets (measured: 8 % on
        // a 244M-entry index) and slower on small ones (27 % on a 5.6M-entry one)
        uint32_t max_ent = 0;
        for (const DevBucket &bk : idx->h_buckets) max_ent = std::max(max_ent, bk.n_entries);
        pa.slice_threads = max_ent > (1u << 22) ? 256 : 128;
    }
    pa.lq = q->lq;
    pa.li = q->li;
    pa.n_legal = q->n_legal;
    pa.pair_lo = q->pair_lo;
    pa.pair_hi = q->pair_hi;
    pa.pair_tiles = q->pair_tiles;
    pa.pair_pack = q->pair_pack;
    pa.pair_pos_off = q->pair_pos_off;
    pa.pair_nsl = q->pair_nsl;
    pa.n_items = q->n_items;
    pa.n_pos = q->n_pos;
    pa.item_off = q->item_off;
    pa.pos_off = q->pos_off;
    pa.plan_order = q->porder;
    pa.stats = c->d_stats;
    B.n_blocks = std::max(1, c->n_cu * B.cfg.blocks_per_cu);
    B.learned = p->evalue_mode == GX_EVALUE_LEARNED;
    B.conf = p->surrogate_conf;
    return B;
}

void fill_item_args(Batched &B, gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p) {
    ItemArgs &ia = B.ia;
    ia.buckets = idx->d_buckets;
    ia.n_buckets = idx->n_buckets;
    ia.qidx = q->idx;
    ia.dF = p->dF;
    ia.min_shp = p->min_shp;
    ia.nbins = (int)idx->nbins;
    ia.lgfact = c->d_lgfact;
    ia.lgfact_n = c->lgfact_n;
    ia.lq = q->lq;
    ia.li = q->li;
    ia.n_legal = q->n_legal;
    ia.pair_lo = q->pair_lo;
    ia.pair_hi = q->pair_hi;
    ia.pair_tiles = q->pair_tiles;
    ia.pair_pack = q->pair_pack;
    ia.pair_pos_off = q->pair_pos_off;
    ia.pair_nsl = q->pair_nsl;
    ia.pos_off = q->pos_off;
    ia.pos = q->pos;
    ia.unit_s = q->unit_s;
    ia.unit_br = q->unit_br;
    ia.part_key = q->part_key;
    ia.part_eq = q->part_eq;
    ia.part_shp = q->part_shp;
    ia.part_cps = q->part_cps;
    ia.part_hist = q->part_hist;
    ia.cand_scratch = c->d_cand;
    ia.tile_entries_max = B.cfg.tile_words * 4;
    ia.split = B.split;
    ia.max_speclen = idx->max_speclen;
    ia.wide = (p->wide_stream || !narrow_stream_fits(idx->max_ions)) ? 1 : 0;
    ia.smax_m = idx->smax_m;
    ia.smax_s = idx->smax_s;
    ia.cell_shift = idx->cell_shift;
    ia.hist = q->hist;
    ia.cpsms = q->cpsms;
    ia.best_key = q->best_key;
    ia.best_eq = q->best_eq;
    ia.best_shp = q->best_shp;
    ia.stats = c->d_stats;
    EvalArgs &ea = B.ea;
    ea.buckets = idx->d_buckets;
    ea.n_spectra = q->n_spectra;
    ea.min_cpsm = p->min_cpsm;
    ea.expect_max = p->expect_max;
    ea.best_key = q->best_key;
    ea.best_eq = q->best_eq;
    ea.best_shp = q->best_shp;
    ea.hist = q->hist;
    ea.cpsms = q->cpsms;
    ea.out = q->psm;
    ea.stats = c->d_stats;
}

// the device work before chunking: clear accumulators, plan spectra, scan
hipError_t enqueue_prefix(gx_ctx *c, gx_dqueries *q, const Batched &B, hipStream_t sm) {
    hipError_t e;
    if ((e = launch_fill_u32(reinterpret_cast<uint32_t *>(c->d_stats), (uint64_t)kStatRows * ST_N * 2, 0, sm)))
        return e;
    if ((e = launch_plan_spectra(B.pa, sm))) return e;
    return launch_scan(B.pa, sm);
}

// the per-chunk plan + search launches and the e-value kernel; per-chunk
// events only on eager runs
hipError_t enqueue_chunks(gx_ctx *c, gx_dqueries *q, Batched &B, hipStream_t sm, bool events) {
    hipError_t e;
    const size_t n_chunks = q->bounds.size() - 1;
    if ((e = launch_fill_u32(q->counters, std::max<size_t>(64, n_chunks), 0, sm))) return e;
    B.pa.pos = q->pos;
    for (size_t k = 0; k < n_chunks; k++) {
        B.pa.chunk_s0 = q->bounds[k];
        B.pa.chunk_s1 = q->bounds[k + 1];
        if (events && (e = hipEventRecord(c->ev_chunk[2 * k], sm))) return e;
        if ((e = launch_plan_slices(B.pa, sm))) return e;
        if (events && (e = hipEventRecord(c->ev_chunk[2 * k + 1], sm))) return e;
        B.ia.chunk_s0 = q->bounds[k];
        B.ia.chunk_s1 = q->bounds[k + 1];
        B.ia.unit0 = q->unit_bounds[k];
        B.ia.n_units = q->unit_bounds[k + 1] - q->unit_bounds[k];
        B.ia.work_counter = q->counters + k;
        // whole spectra are written by the search kernel, the parts of heavy
        // ones by the fold that follows it
        if ((e = launch_search_spectra(B.cfg_id, B.ia, B.n_blocks, sm))) return e;
        FoldArgs fa{};
        fa.f_s = q->fold_s + q->fold_bounds[k];
        fa.f_slot0 = q->fold_p0 + q->fold_bounds[k];
        fa.f_n = q->fold_n + q->fold_bounds[k];
        fa.n_folds = q->fold_bounds[k + 1] - q->fold_bounds[k];
        fa.part_key = q->part_key;
        fa.part_eq = q->part_eq;
        fa.part_shp = q->part_shp;
        fa.part_cps = q->part_cps;
        fa.part_hist = q->part_hist;
        fa.hist = q->hist;
        fa.cpsms = q->cpsms;
        fa.best_eq = q->best_eq;
        fa.best_shp = q->best_shp;
        fa.best_key = q->best_key;
        if ((e = launch_fold(fa, sm))) return e;
    }
    if (events && (e = hipEventRecord(c->ev2, sm))) return e;
    if (B.learned) return enqueue_learned_evalue(c, q, B.ea, B.conf, sm);
    return launch_evalue(B.ea, sm);
}

std::vector<char> graph_key_of(const Batched &B, const gx_search_params *p) {
    std::vector<char> k(sizeof(gx_search_params) + 2 * sizeof(int));
    gx_search_params pp = *p;
    pp.config = B.cfg_id;
    pp.split = B.split;
    pp.whole_bin = (int32_t)B.pa.whole_bin;
    pp.use_graph = 0;
    std::memcpy(k.data(), &pp, sizeof(pp));
    return k;
}

int search_batched(gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p,
                   gx_search_stats *stats) {
    int st = ensure_batched(q, idx);
    if (st) return st;
    if (p->evalue_mode == GX_EVALUE_LEARNED && (st = ensure_learned(c, q))) return st;
    Batched B = setup_batched(c, idx, q, p);
    const uint32_t ns = q->n_spectra;
    hipStream_t sm = c->stream;
    const std::vector<char> key = graph_key_of(B, p);

    // ---- replay a captured graph of the same batch and parameters ----
    if (p->use_graph && q->graph_exec && q->graph_key == key && q->graph_ctx_gen == c->scratch_gen &&
        q->graph_buf_gen == q->buf_gen && q->graph_index == idx && q->graph_index_gen == idx->gen) {
        GX_HIP(hipEventRecord(c->ev0, sm));
        GX_HIP(hipGraphLaunch(q->graph_exec, sm));
        GX_HIP(hipEventRecord(c->ev3, sm));
        GX_HIP(hipStreamSynchronize(sm));
        q->searched = true;
        q->learned_done = B.learned;
        q->batched_state = true;
        q->searched_index = idx;
        if (stats) {
            unsigned long long h[ST_N];
            read_stats(c, h);
            fill_stats(stats, h);
            stats->items = q->n_items_total;
            stats->chunks = (uint32_t)(q->bounds.size() - 1);
            stats->config = B.cfg_id;
            stats->total_ms = elapsed(c->ev0, c->ev3);
        }
        return GX_SUCCESS;
    }

    // ---- eager run: plan, size the chunk buffers, search, model ----
    GX_HIP(hipEventRecord(c->ev0, sm));
    GX_HIP(enqueue_prefix(c, q, B, sm));
    GX_HIP(hipEventRecord(c->ev1, sm));
    // chunking needs the slice-position offsets on the host
    uint32_t n_items_total = 0;
    GX_HIP(hipMemcpyAsync(q->h_pos_off.data(), q->pos_off, (ns + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, sm));
    GX_HIP(hipMemcpyAsync(&n_items_total, q->item_off + ns, sizeof(uint32_t), hipMemcpyDeviceToHost, sm));
    GX_HIP(hipStreamSynchronize(sm));
    q->bounds.assign(1, 0);
    uint64_t max_chunk = 1;
    {
        uint32_t s0 = 0;
        while (s0 < ns) {
            uint32_t s1 = s0 + 1;
            while (s1 < ns && q->h_pos_off[s1 + 1] - q->h_pos_off[s0] <= kPosChunkWords) s1++;
            max_chunk = std::max<uint64_t>(max_chunk, q->h_pos_off[s1] - q->h_pos_off[s0]);
            q->bounds.push_back(s1);
            s0 = s1;
        }
    }
    const size_t n_chunks = q->bounds.size() - 1;
    // Work units. The slice positions a spectrum needs stand in for its work.
    // A spectrum heavier than an eighth of one block's fair share of its chunk
    // is cut into parts of contiguous buckets, each at most that heavy, searched
    // by different blocks into slots of the part scratch and folded after the
    // chunk's search. Units are handed out largest first (a counting sort by
    // the bit length of the work), so that a chunk ends on small units.
    {
        const uint32_t nb = (uint32_t)idx->n_buckets;
        std::vector<uint64_t> cut(n_chunks);
        bool any_heavy = false;
        for (size_t k = 0; k < n_chunks; k++) {
            const uint32_t s0 = q->bounds[k], s1 = q->bounds[k + 1];
            cut[k] = std::max<uint64_t>(1, (q->h_pos_off[s1] - q->h_pos_off[s0]) / (8 * (uint64_t)B.n_blocks));
            for (uint32_t s = s0; s < s1 && !any_heavy; s++)
                any_heavy = q->h_pos_off[s + 1] - q->h_pos_off[s] > cut[k];
        }
        if (any_heavy && nb > 1) {
            q->h_pair_off.resize((size_t)ns * nb);
            GX_HIP(hipMemcpyAsync(q->h_pair_off.data(), q->pair_pos_off, sizeof(uint32_t) * (size_t)ns * nb,
                                  hipMemcpyDeviceToHost, sm));
            GX_HIP(hipStreamSynchronize(sm));
        }
        q->h_unit_s.clear();
        q->h_unit_br.clear();
        q->h_fold_s.clear();
        q->h_fold_p0.clear();
        q->h_fold_n.clear();
        q->unit_bounds.assign(1, 0);
        q->fold_bounds.assign(1, 0);
        uint32_t max_parts = 0, parts_total = 0;
        std::vector<uint32_t> us, ubr;
        std::vector<uint64_t> uw;
        std::vector<uint32_t> pb;  // part boundaries (bucket indices) of one spectrum
        for (size_t k = 0; k < n_chunks; k++) {
            const uint32_t s0 = q->bounds[k], s1 = q->bounds[k + 1];
            us.clear();
            ubr.clear();
            uw.clear();
            uint32_t slot = 0;
            for (uint32_t s = s0; s < s1; s++) {
                const uint64_t w = q->h_pos_off[s + 1] - q->h_pos_off[s];
                pb.clear();
                if (w > cut[k] && nb > 1 && any_heavy) {
                    const uint32_t *po = q->h_pair_off.data() + (size_t)s * nb;
                    uint64_t acc = 0;
                    pb.push_back(0);
                    for (uint32_t b = 0; b < nb; b++) {
                        const uint64_t wb = (b + 1 < nb ? (uint64_t)po[b + 1] : w) - po[b];
                        if (acc > 0 && acc + wb > cut[k]) {
                            pb.push_back(b);
                            acc = 0;
                        }
                        acc += wb;
                    }
                    pb.push_back(nb);
                }
                const uint32_t np = pb.empty() ? 0u : (uint32_t)pb.size() - 1;
                if (np >= 2 && slot + np < kWholeUnit) {
                    const uint32_t *po = q->h_pair_off.data() + (size_t)s * nb;
                    q->h_fold_s.push_back(s);
                    q->h_fold_p0.push_back(slot);
                    q->h_fold_n.push_back(np);
                    for (uint32_t i = 0; i < np; i++) {
                        const uint32_t b0 = pb[i], b1 = pb[i + 1];
                        const uint64_t e = b1 < nb ? (uint64_t)po[b1] : w;
                        us.push_back(s);
                        ubr.push_back(b0 | (b1 << 8) | (slot << 16));
                        uw.push_back(e - po[b0]);
                        slot++;
                    }
                } else {
                    us.push_back(s);
                    ubr.push_back(0u | (nb << 8) | (kWholeUnit << 16));
                    uw.push_back(w);
                }
            }
            const uint32_t nu = (uint32_t)us.size();
            uint32_t cnt[66] = {};
            auto cls = [&](uint32_t i) { return 64 - __builtin_clzll(uw[i] + 1); };
            for (uint32_t i = 0; i < nu; i++) cnt[65 - cls(i)]++;
            uint32_t run = 0;
            for (int c = 0; c < 66; c++) {
                const uint32_t n = cnt[c];
                cnt[c] = run;
                run += n;
            }
            const size_t base = q->h_unit_s.size();
            q->h_unit_s.resize(base + nu);
            q->h_unit_br.resize(base + nu);
            for (uint32_t i = 0; i < nu; i++) {
                const uint32_t r = cnt[65 - cls(i)]++;
                q->h_unit_s[base + r] = us[i];
                q->h_unit_br[base + r] = ubr[i];
            }
            q->unit_bounds.push_back((uint32_t)q->h_unit_s.size());
            q->fold_bounds.push_back((uint32_t)q->h_fold_s.size());
            max_parts = std::max(max_parts, slot);
            parts_total += slot;
        }
        q->n_parts_total = parts_total;
        const size_t nu = q->h_unit_s.size(), nf = std::max<size_t>(1, q->h_fold_s.size());
        const size_t npart = std::max<uint32_t>(1, max_parts);
        if (nu > q->unit_cap || !q->unit_s) {
            q->buf_gen++;
            const size_t want = nu + nu / 8 + 64;
            q->unit_cap = 0;
            if (realloc_dev(&q->unit_s, want) || realloc_dev(&q->unit_br, want)) return GX_ERR_BAD_MEM_ALLOC;
            q->unit_cap = want;
        }
        if (nf > q->fold_cap || !q->fold_s) {
            q->buf_gen++;
            const size_t want = nf + nf / 8 + 64;
            q->fold_cap = 0;
            if (realloc_dev(&q->fold_s, want) || realloc_dev(&q->fold_p0, want) || realloc_dev(&q->fold_n, want))
                return GX_ERR_BAD_MEM_ALLOC;
            q->fold_cap = want;
        }
        if (npart > q->part_cap || !q->part_key) {
            q->buf_gen++;
            const size_t want = npart + npart / 8 + 16;
            q->part_cap = 0;
            if (realloc_dev(&q->part_key, want) || realloc_dev(&q->part_eq, want) ||
                realloc_dev(&q->part_shp, want) || realloc_dev(&q->part_cps, want) ||
                realloc_dev(&q->part_hist, want * GX_HIST_BINS))
                return GX_ERR_BAD_MEM_ALLOC;
            q->part_cap = want;
        }
        if (nu) {
            GX_HIP(hipMemcpyAsync(q->unit_s, q->h_unit_s.data(), 4 * nu, hipMemcpyHostToDevice, sm));
            GX_HIP(hipMemcpyAsync(q->unit_br, q->h_unit_br.data(), 4 * nu, hipMemcpyHostToDevice, sm));
        }
        if (!q->h_fold_s.empty()) {
            const size_t n = q->h_fold_s.size();
            GX_HIP(hipMemcpyAsync(q->fold_s, q->h_fold_s.data(), 4 * n, hipMemcpyHostToDevice, sm));
            GX_HIP(hipMemcpyAsync(q->fold_p0, q->h_fold_p0.data(), 4 * n, hipMemcpyHostToDevice, sm));
            GX_HIP(hipMemcpyAsync(q->fold_n, q->h_fold_n.data(), 4 * n, hipMemcpyHostToDevice, sm));
        }
    }
    // the plan visits each chunk's spectra by precursor mass (1 Da classes):
    // neighbouring blocks then share windows, and their boundary searches
    // share cache lines
    q->h_porder.resize(ns);
    for (size_t k = 0; k < n_chunks; k++) {
        const uint32_t s0 = q->bounds[k], s1 = q->bounds[k + 1];
        constexpr uint32_t kCls = 8192;
        std::vector<uint32_t> cnt(kCls + 1, 0);
        auto cls = [&](uint32_t s) {
            const float m = q->h_prec[s];
            return m > 0.0f ? std::min<uint32_t>((uint32_t)m, kCls - 1) : 0u;
        };
        for (uint32_t s = s0; s < s1; s++) cnt[cls(s) + 1]++;
        for (uint32_t c = 0; c < kCls; c++) cnt[c + 1] += cnt[c];
        for (uint32_t s = s0; s < s1; s++) q->h_porder[s0 + cnt[cls(s)]++] = s;
    }
    if (ns) GX_HIP(hipMemcpyAsync(q->porder, q->h_porder.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice, sm));
    if (grow(&q->pos, q->pos_cap, max_chunk, q->buf_gen)) return GX_ERR_BAD_MEM_ALLOC;
    if (!q->counters || n_chunks > q->counters_n) {
        q->buf_gen++;
        if (q->counters) (void)hipFree(q->counters);
        q->counters = nullptr;
        q->counters_n = 0;
        if (dalloc(&q->counters, std::max<size_t>(64, n_chunks))) return GX_ERR_BAD_MEM_ALLOC;
        q->counters_n = std::max<size_t>(64, n_chunks);
    }
    // candidate overflow scratch, sized once for every configuration so that
    // it never moves under a captured graph
    size_t cand_need = 0;
    for (int k = 0; k < kNumBatchConfigs; k++) {
        const BatchConfig bc = batch_config(k);
        cand_need = std::max(cand_need, (size_t)std::max(1, c->n_cu * bc.blocks_per_cu) * bc.tile_words * 4);
    }
    if (c->cand_words < cand_need) {
        c->scratch_gen++;
        if (c->d_cand) (void)hipFree(c->d_cand);
        c->d_cand = nullptr;
        c->cand_words = 0;
        if (dalloc(&c->d_cand, cand_need)) return GX_ERR_BAD_MEM_ALLOC;
        c->cand_words = cand_need;
    }
    while (c->ev_chunk.size() < 2 * n_chunks) {
        hipEvent_t e;
        GX_HIP(hipEventCreate(&e));
        c->ev_chunk.push_back(e);
    }
    q->n_items_total = n_items_total;
    fill_item_args(B, c, idx, q, p);
    GX_HIP(enqueue_chunks(c, q, B, sm, true));
    GX_HIP(hipEventRecord(c->ev3, sm));
    GX_HIP(hipStreamSynchronize(sm));
    q->searched = true;
    q->learned_done = B.learned;
    q->batched_state = true;
    q->searched_index = idx;
    if (stats) {
        unsigned long long h[ST_N];
        read_stats(c, h);
        fill_stats(stats, h);
        stats->items = n_items_total;
        stats->chunks = (uint32_t)n_chunks;
        stats->config = B.cfg_id;
        double plan = elapsed(c->ev0, c->ev1), search = 0;
        for (size_t k = 0; k < n_chunks; k++) {
            plan += elapsed(c->ev_chunk[2 * k], c->ev_chunk[2 * k + 1]);
            const hipEvent_t stop = (k + 1 < n_chunks) ? c->ev_chunk[2 * k + 2] : c->ev2;
            search += elapsed(c->ev_chunk[2 * k + 1], stop);
        }
        stats->plan_ms = plan;
        stats->kernel_ms = search;
        stats->evalue_ms = elapsed(c->ev2, c->ev3);
        stats->total_ms = elapsed(c->ev0, c->ev3);
    }

    // ---- capture the same launches (no host synchronisation) as a graph ----
    if (p->use_graph) {
        if (q->graph_exec) (void)hipGraphExecDestroy(q->graph_exec);
        if (q->graph) (void)hipGraphDestroy(q->graph);
        q->graph_exec = nullptr;
        q->graph = nullptr;
        q->graph_key.clear();
        GX_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_prefix(c, q, B, sm);
        if (e == hipSuccess) e = enqueue_chunks(c, q, B, sm, false);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(sm, &g);
        if (e != hipSuccess || e2 != hipSuccess || !g) {
            note_error(e != hipSuccess ? "enqueue during capture" : "hipStreamEndCapture", __LINE__,
                       e != hipSuccess ? e : e2);
            if (g) (void)hipGraphDestroy(g);
            return GX_ERR_DEVICE;
        }
        hipGraphExec_t ge = nullptr;
        const hipError_t e3 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            note_error("hipGraphInstantiate", __LINE__, e3);
            (void)hipGraphDestroy(g);
            return GX_ERR_DEVICE;
        }
        q->graph = g;
        q->graph_exec = ge;
        q->graph_key = key;
        q->graph_ctx_gen = c->scratch_gen;
        q->graph_buf_gen = q->buf_gen;
        q->graph_index = idx;
        q->graph_index_gen = idx->gen;
    }
    return GX_SUCCESS;
}

int search_locked(gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p,
                  gx_search_stats *stats, bool want_hist) {
    int st = check_params(p, idx);
    if (st) return st;
    if (idx->ctx != c || q->ctx != c) return param_error(GX_ERR_INVLD_PARAM, "index or batch of another context");
    if (p->engine != GX_ENGINE_BATCHED && p->engine != GX_ENGINE_SPECTRUM)
        return param_error(GX_ERR_INVLD_PARAM, "unknown engine");
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    // every count bc or yc stays below (2dF+1)*speclen/2 + 1
    if ((int)((2 * p->dF + 1) * (idx->max_speclen / 2) + 1) > c->lgfact_n) return GX_ERR_INVLD_SIZE;
    if (p->evalue_mode == GX_EVALUE_LEARNED) {
        // the surrogate reads the batched engine's per-spectrum histograms
        if (p->engine != GX_ENGINE_BATCHED)
            return param_error(GX_ERR_INVLD_PARAM, "the learned e-value needs the batched engine");
        if (!c->surrogate) return param_error(GX_ERR_NOT_READY, "no surrogate attached (gx_ctx_set_surrogate)");
        if (!(p->surrogate_conf >= 0.0f && p->surrogate_conf <= 1.0f))
            return param_error(GX_ERR_INVLD_PARAM, "surrogate_conf outside 0..1");
    }
    if (p->engine == GX_ENGINE_SPECTRUM) return search_plain(c, idx, q, p, stats, want_hist);
    return search_batched(c, idx, q, p, stats);
}

}  // namespace

extern "C" int gx_search_device(gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p,
                                gx_search_stats *stats) {
    if (!c || !idx || !q || !p) return GX_ERR_INVLD_PTR;
    std::lock_guard<std::mutex> lk(c->mu);
    return search_locked(c, idx, q, p, stats, q->hist != nullptr);
}

extern "C" int gx_results_download(gx_ctx *c, gx_dqueries *q, gx_psm *psms, uint32_t *hist) {
    if (!c || !q) return GX_ERR_INVLD_PTR;
    if (!q->searched) return GX_ERR_NOT_READY;
    if (hist && !q->hist) return GX_ERR_NOT_READY;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    if (psms && q->n_spectra)
        GX_HIP(hipMemcpy(psms, q->psm, sizeof(gx_psm) * q->n_spectra, hipMemcpyDeviceToHost));
    if (hist && q->n_spectra)
        GX_HIP(hipMemcpy(hist, q->hist, sizeof(uint32_t) * GX_HIST_BINS * q->n_spectra, hipMemcpyDeviceToHost));
    return GX_SUCCESS;
}

extern "C" int gx_search(gx_ctx *c, const gx_dindex *idx, const gx_query_batch *qb, const gx_search_params *p,
                         gx_psm *psms, uint32_t *hist, gx_search_stats *stats) {
    if (!c || !idx || !qb || !p || !psms) return GX_ERR_INVLD_PTR;
    gx_dqueries *q = nullptr;
    int st = gx_queries_upload(c, qb, &q);
    if (st) return st;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        st = search_locked(c, idx, q, p, stats, hist != nullptr);
    }
    if (!st) st = gx_results_download(c, q, psms, hist);
    gx_queries_free(q);
    return st;
}

extern "C" int gx_autotune(gx_ctx *c, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p,
                           gx_tune_result *out) {
    if (!c || !idx || !q || !p || !out) return GX_ERR_INVLD_PTR;
    std::lock_guard<std::mutex> lk(c->mu);
    std::memset(out, 0, sizeof(*out));
    gx_search_params pp = *p;
    pp.engine = GX_ENGINE_BATCHED;
    pp.use_graph = 0;
    double best = 1e300;
    int bc = c->tuned_config, bs = c->tuned_split, bw = c->tuned_whole;
    auto trial = [&](int cfg, int sp, int wb, double &t) -> int {
        pp.config = cfg;
        pp.split = sp;
        pp.whole_bin = wb;
        t = 1e300;
        for (int rep = 0; rep < 2; rep++) {  // the first run also sizes buffers
            gx_search_stats st;
            const int r = search_locked(c, idx, q, &pp, &st, false);
            if (r) return r;
            t = std::min(t, st.total_ms);
        }
        if (out->n_tried < GX_TUNE_MAX) {
            out->tried_config[out->n_tried] = cfg;
            out->tried_split[out->n_tried] = sp;
            out->tried_whole_bin[out->n_tried] = wb;
            out->tried_ms[out->n_tried] = t;
            out->n_tried++;
        }
        if (t < best) {
            best = t;
            bc = cfg;
            bs = sp;
            bw = wb;
        }
        return GX_SUCCESS;
    };
    // launch configuration x window split at the current threshold, then the
    // threshold for the two best pairs (two launch shapes can be within a few
    // percent of each other at one threshold and apart at another)
    const int splits[] = {1, 2};
    const int w0 = bw;
    std::vector<std::pair<double, std::pair<int, int>>> first;
    for (int cfg = 0; cfg < kNumBatchConfigs; cfg++)
        for (int sp : splits) {
            double t;
            const int r = trial(cfg, sp, w0, t);
            if (r) return r;
            first.push_back({t, {cfg, sp}});
        }
    std::sort(first.begin(), first.end());
    for (size_t k = 0; k < std::min<size_t>(2, first.size()); k++)
        for (int wb : {8, 16, 32, 64, 128}) {
            if (wb == w0) continue;
            double t;
            const int r = trial(first[k].second.first, first[k].second.second, wb, t);
            if (r) return r;
        }
    c->tuned_config = bc;
    c->tuned_split = bs;
    c->tuned_whole = bw;
    out->config = bc;
    out->split = bs;
    out->whole_bin = bw;
    return GX_SUCCESS;
}

extern "C" int gx_last_error(char *buf, size_t len) {
    if (!buf || !len) return GX_ERR_INVLD_PTR;
    std::snprintf(buf, len, "%s", g_last_error);
    return GX_SUCCESS;
}
