// Index-sharded search: export of one shard's partial results, and the merge
// of the shards' partials into PSMs with the tail fit on the merged
// histogram.
//
// The reference's distributed mode (USE_MPI) partitions the peptides over
// nodes (lbe_internal.cpp:75-103); each node searches every spectrum against
// its share and sends per spectrum a partRes {first and last non-empty
// histogram bin, top hyperscore, cpsms, spectrum id} plus up to Xsamples = 128
// histogram bins as u16 (dslim_query.cpp:973-1000, expeRT.cpp:714-806,
// slm_dsts.h:856-934); the receiver sums the histograms, takes the largest
// hyperscore (first node on ties) and fits (dslim_score.cpp:199-387). That
// encoding loses counts above 65535 (and rescales above 65500 candidates),
// and writes past its 128-bin slot when the non-empty span is longer. Here
// the record carries the shard's packed top-hit key, tie count and shared
// ions, and the whole 802-bin u32 histogram travels beside it, so that the
// merged result is exactly the unsharded one:
//
//   export_kernel  one thread per spectrum: the search's top hit with its
//                  psid made global (+ the shard's offset in the bucket);
//   merge_kernel   one wave per spectrum: histogram sums, the best key over
//                  the shards (ties: the smaller bucket and global psid, the
//                  unsharded rule; tie counts add), the candidate total, then
//                  the tail fit and e-value as in the e-value kernel.
#include <vector>

#include "capi_internal.h"
#include "device_util.h"

namespace gx {
namespace {

constexpr int kMaxParts = 1024;

struct ExportArgs {
    const unsigned long long *best_key;
    const uint32_t *best_eq, *best_shp, *cpsms;
    const uint32_t *offset;   // per bucket
    const DevBucket *buckets;
    uint32_t n_buckets;
    uint32_t n_spectra;
    gx_partial *out;
    unsigned int *overflow;   // set when a global psid needs more than 26 bits
};

__global__ void __launch_bounds__(256) export_kernel(ExportArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n_spectra) return;
    gx_partial r;
    r.key = 0;
    r.eq = a.best_eq[s];
    r.cpsms = a.cpsms[s];
    r.sharedions = 0;
    r.totalions = 0;
    r.reserved = 0;
    const unsigned long long key = a.best_key[s];
    if (key) {
        const uint32_t low = ~(uint32_t)(key & 0xffffffffu);
        const uint32_t b = low >> 26;
        if (b >= a.n_buckets) {  // not a key of this index
            atomicOr(a.overflow, 1u);
            a.out[s] = r;
            return;
        }
        const uint32_t psid = (low & ((1u << 26) - 1)) + a.offset[b];
        if (psid >= (1u << 26)) atomicOr(a.overflow, 1u);
        r.key = (key & 0xffffffff00000000ull) | (unsigned long long)(~((b << 26) | (psid & ((1u << 26) - 1))));
        r.sharedions = (uint16_t)a.best_shp[s];
        r.totalions = (uint16_t)a.buckets[b].speclen;
    }
    a.out[s] = r;
}

struct MergeArgs {
    const gx_partial *const *rec;
    const uint32_t *const *hist;
    int n_parts;
    uint32_t n_spectra;
    int min_cpsm;
    double expect_max;
    gx_psm *out;
    uint32_t *out_hist;  // merged histograms, or null
};

__global__ void __launch_bounds__(64) merge_kernel(MergeArgs a) {
    __shared__ uint32_t hist[GX_HIST_BINS];
    __shared__ double sx[GX_HIST_BINS];
    __shared__ FitScratch F;
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x;
    for (int w = lane; w < GX_HIST_BINS; w += 64) {
        uint32_t h = 0;
        for (int k = 0; k < a.n_parts; k++) h += a.hist[k][(size_t)s * GX_HIST_BINS + w];
        hist[w] = h;
        if (a.out_hist) a.out_hist[(size_t)s * GX_HIST_BINS + w] = h;
    }
    // every lane reduces the few records the same way (no divergence)
    unsigned long long key = 0;
    uint32_t eq = 0, shp = 0, cps = 0, tot = 0;
    for (int k = 0; k < a.n_parts; k++) {
        const gx_partial p = a.rec[k][s];
        cps += p.cpsms;
        const uint32_t h1 = (uint32_t)(key >> 32), h2 = (uint32_t)(p.key >> 32);
        if (h1 == h2) {
            eq += p.eq;
            if (p.key > key) {
                key = p.key;
                shp = p.sharedions;
                tot = p.totalions;
            }
        } else if (h2 > h1) {
            key = p.key;
            eq = p.eq;
            shp = p.sharedions;
            tot = p.totalions;
        }
    }
    __syncthreads();
    const uint32_t need = (uint32_t)max(a.min_cpsm, 1);
    gx_psm r;
    r.hyperscore = 0.0f;
    r.psid = -1;
    r.idxoffset = 0;
    r.sharedions = 0;
    r.totalions = 0;
    r.flags = 0;
    r.cpsms = (int32_t)cps;
    r.hyp = 0;
    r.mu = 0;
    r.beta = 0;
    r.evalue = INFINITY;
    if (cps >= 1 && key) {
        const uint32_t low = ~(uint32_t)(key & 0xffffffffu);
        r.flags |= GX_PSM_HAS_HIT;
        r.hyperscore = __uint_as_float((uint32_t)(key >> 32));
        r.psid = (int32_t)(low & ((1u << 26) - 1));
        r.idxoffset = (uint16_t)(low >> 26);
        r.sharedions = (uint16_t)shp;
        r.totalions = (uint16_t)tot;
        r.hyp = hyp_bin(r.hyperscore);
        if (eq > 1) r.flags |= GX_PSM_TIE;
    }
    if (cps >= need && r.psid >= 0) {
        r.flags |= GX_PSM_MODELED;
        double sl, ic;
        int32_t mu, beta;
        if (tail_fit_block(hist, sx, F, r.hyp, r.cpsms, sl, ic, mu, beta)) {
            r.flags |= GX_PSM_FIT_OK;
            r.mu = mu;
            r.beta = beta;
            r.evalue = evalue_of(mu, beta, r.hyp, (uint32_t)r.cpsms);
            if (r.evalue < a.expect_max) r.flags |= GX_PSM_REPORTED;
        }
    }
    if (lane == 0) a.out[s] = r;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gx_partials_export(gx_ctx *c, const gx_dindex *idx, const gx_dqueries *q, const uint32_t *psid_offset,
                                  gx_partial *rec, uint32_t *hist, int on_device) {
    if (!c || !idx || !q || (!rec && !hist)) return GX_ERR_INVLD_PTR;
    if (!q->searched || !q->batched_state || !q->best_key || !q->hist) return GX_ERR_NOT_READY;
    if (q->searched_index != idx) return param_error(GX_ERR_INVLD_INDEX, "the batch was last searched on another index");
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    const uint32_t ns = q->n_spectra;
    if (ns == 0) return GX_SUCCESS;
    std::vector<uint32_t> off(idx->n_buckets, 0);
    if (psid_offset)
        for (int b = 0; b < idx->n_buckets; b++) {
            off[b] = psid_offset[b];
            if ((uint64_t)off[b] + idx->h_buckets[b].n_entries > (1ull << 26))
                return param_error(GX_ERR_INVLD_SIZE, "global psids need more than 26 bits");
        }
    hipStream_t st = c->stream;
    uint32_t *d_off = nullptr;
    unsigned int *d_flag = nullptr;
    gx_partial *d_rec = nullptr;
    auto cleanup = [&]() {
        for (void *p : {(void *)d_off, (void *)d_flag, (void *)(on_device ? nullptr : d_rec)})
            if (p) (void)hipFree(p);
    };
    auto fail = [&](int code) {
        cleanup();
        return code;
    };
    if (dalloc(&d_off, off.size()) || dalloc(&d_flag, 1) || (!on_device && rec && dalloc(&d_rec, ns)))
        return fail(GX_ERR_BAD_MEM_ALLOC);
    if (on_device) d_rec = rec;
    if (hipMemcpyAsync(d_off, off.data(), 4 * off.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(d_flag, 0, 4, st) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    if (rec) {
        ExportArgs ea{q->best_key,     q->best_eq, q->best_shp, q->cpsms, d_off, idx->d_buckets,
                      (uint32_t)idx->n_buckets, ns, d_rec, d_flag};
        hipLaunchKernelGGL(export_kernel, dim3((ns + 255) / 256), dim3(256), 0, st, ea);
        if (hipGetLastError() != hipSuccess) return fail(GX_ERR_DEVICE);
        if (!on_device &&
            hipMemcpyAsync(rec, d_rec, sizeof(gx_partial) * ns, hipMemcpyDeviceToHost, st) != hipSuccess)
            return fail(GX_ERR_DEVICE);
    }
    if (hist && hipMemcpyAsync(hist, q->hist, sizeof(uint32_t) * GX_HIST_BINS * (size_t)ns,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    unsigned int flag = 0;
    if (hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    cleanup();
    return flag ? GX_ERR_INVLD_SIZE : GX_SUCCESS;
}

extern "C" int gx_partials_merge(gx_ctx *c, int n_parts, const gx_partial *const *rec, const uint32_t *const *hist,
                                 uint32_t n_spectra, const gx_search_params *p, int on_device, gx_psm *psms,
                                 uint32_t *merged_hist, float *device_ms) {
    if (!c || !rec || !hist || !p || !psms) return GX_ERR_INVLD_PTR;
    if (n_parts < 1 || n_parts > kMaxParts) return param_error(GX_ERR_INVLD_PARAM, "n_parts (1..1024)");
    for (int k = 0; k < n_parts; k++)
        if (n_spectra && (!rec[k] || !hist[k])) return GX_ERR_INVLD_PTR;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    if (device_ms) *device_ms = 0.0f;
    if (n_spectra == 0) return GX_SUCCESS;
    const size_t ns = n_spectra, hw = ns * GX_HIST_BINS;
    hipStream_t st = c->stream;
    std::vector<void *> owned;
    auto cleanup = [&]() {
        for (void *q : owned) (void)hipFree(q);
    };
    auto fail = [&](int code) {
        cleanup();
        return code;
    };
    auto alloc = [&](size_t bytes) -> void * {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return nullptr;
        owned.push_back(q);
        return q;
    };
    // device copies of host inputs, and the pointer tables
    std::vector<const gx_partial *> d_rec(n_parts);
    std::vector<const uint32_t *> d_hist(n_parts);
    for (int k = 0; k < n_parts; k++) {
        if (on_device) {
            d_rec[k] = rec[k];
            d_hist[k] = hist[k];
            continue;
        }
        void *r = alloc(sizeof(gx_partial) * ns), *h = alloc(sizeof(uint32_t) * hw);
        if (!r || !h) return fail(GX_ERR_BAD_MEM_ALLOC);
        if (hipMemcpyAsync(r, rec[k], sizeof(gx_partial) * ns, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(h, hist[k], sizeof(uint32_t) * hw, hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(GX_ERR_DEVICE);
        d_rec[k] = (const gx_partial *)r;
        d_hist[k] = (const uint32_t *)h;
    }
    void *t_rec = alloc(sizeof(void *) * n_parts), *t_hist = alloc(sizeof(void *) * n_parts);
    gx_psm *d_out = (gx_psm *)alloc(sizeof(gx_psm) * ns);
    uint32_t *d_mh = merged_hist ? (uint32_t *)alloc(sizeof(uint32_t) * hw) : nullptr;
    if (!t_rec || !t_hist || !d_out || (merged_hist && !d_mh)) return fail(GX_ERR_BAD_MEM_ALLOC);
    if (hipMemcpyAsync(t_rec, d_rec.data(), sizeof(void *) * n_parts, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(t_hist, d_hist.data(), sizeof(void *) * n_parts, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    MergeArgs ma{(const gx_partial *const *)t_rec, (const uint32_t *const *)t_hist, n_parts, n_spectra,
                 p->min_cpsm, p->expect_max, d_out, d_mh};
    if (hipEventRecord(c->ev0, st) != hipSuccess) return fail(GX_ERR_DEVICE);
    hipLaunchKernelGGL(merge_kernel, dim3(n_spectra), dim3(64), 0, st, ma);
    if (hipGetLastError() != hipSuccess || hipEventRecord(c->ev1, st) != hipSuccess) return fail(GX_ERR_DEVICE);
    if (hipMemcpyAsync(psms, d_out, sizeof(gx_psm) * ns, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (merged_hist &&
         hipMemcpyAsync(merged_hist, d_mh, sizeof(uint32_t) * hw, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    float ms = 0.0f;
    if (device_ms && hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) *device_ms = ms;
    cleanup();
    return GX_SUCCESS;
}
