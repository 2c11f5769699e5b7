// Query preprocessing on the GPU (gx_queries_preprocess): peak picking of a
// raw batch straight into a device-resident query batch.
//
// The reference's GPU path (s2::ArraySort, superstep2/kernel.cu:241-422)
// sorts all peaks of a batch with two stable sorts and scales the top QALEN
// of each spectrum including the base peak; its CPU path
// (MSQuery::pickpeaks, msquery.cpp:338-423) sorts each spectrum with an
// unstable key/value quicksort (keyval.h:49-103), sets the base peak to
// base_int exactly and stops at the first peak below min_int. The two differ
// on ties and on the base peak; this follows the CPU path:
//
//   pick_kernel    one thread per spectrum: the quicksort's own partition
//                  steps on the spectrum's (intensity, m/z) pairs, then the
//                  normalisation and cut; writes the kept count;
//   exclusive scan of the kept counts (rocPRIM) -> the batch's CSR offsets;
//   compact_kernel one wave per spectrum copies its kept tail.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "capi_internal.h"

namespace gx {
namespace {

constexpr int kMaxQalen = kMaxPeaksLds;  // the search keeps <= 2048 peaks per spectrum

// (int) of a double as x86-64 stores it: truncation, INT32_MIN for NaN and
// out-of-range values (what the reference's `int *= double` leaves there)
__device__ __forceinline__ int32_t to_int32_x86(double x) {
    return (x > -2147483649.0 && x < 2147483648.0) ? (int32_t)x : INT32_MIN;
}

struct PickArgs {
    const uint32_t *idx;  // raw CSR offsets, n_spectra + 1
    int32_t *in, *mz;     // raw peaks, sorted and normalised in place
    uint32_t *kept;       // n_spectra + 1 (the last stays 0 for the scan)
    uint32_t n_spectra;
    int32_t base_int, min_int, qalen;
};

__global__ void __launch_bounds__(64) pick_kernel(PickArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n_spectra) return;
    const uint32_t off = a.idx[s];
    const int32_t n = (int32_t)(a.idx[s + 1] - off);
    int32_t *key = a.in + off, *val = a.mz + off;
    // quicksort with the partition steps of KeyVal_Parallel_Internal: middle
    // pivot, Hoare scan, swap and step both cursors. The smaller subrange is
    // taken first, so at most 32 ranges wait on the stack.
    if (n >= 2) {
        int32_t stk[2 * 34];
        int sp = 0;
        stk[sp++] = 0;
        stk[sp++] = n - 1;
        while (sp > 0) {
            const int32_t right = stk[--sp], left = stk[--sp];
            int32_t i = left, j = right;
            const int32_t pivot = key[(int32_t)(((int64_t)left + right) / 2)];
            while (i <= j) {
                while (key[i] < pivot) i++;
                while (key[j] > pivot) j--;
                if (i <= j) {
                    const int32_t k = key[i], v = val[i];
                    key[i] = key[j];
                    val[i] = val[j];
                    key[j] = k;
                    val[j] = v;
                    i++;
                    j--;
                }
            }
            const bool lo = left < j, hi = i < right;
            if (lo && hi) {
                if (j - left > right - i) {
                    stk[sp++] = left, stk[sp++] = j;
                    stk[sp++] = i, stk[sp++] = right;
                } else {
                    stk[sp++] = i, stk[sp++] = right;
                    stk[sp++] = left, stk[sp++] = j;
                }
            } else if (lo) {
                stk[sp++] = left, stk[sp++] = j;
            } else if (hi) {
                stk[sp++] = i, stk[sp++] = right;
            }
        }
    }
    // normalisation and cut (msquery.cpp:350-376)
    uint32_t kept = 0;
    if (n > 0) {
        const double factor = (double)a.base_int / (double)key[n - 1];
        key[n - 1] = a.base_int;
        int32_t k = 1;
        for (int32_t j = n - 2; j >= n - a.qalen && j >= 0; j--) {
            const int32_t v = to_int32_x86((double)key[j] * factor);
            key[j] = v;
            if (v >= a.min_int)
                k++;
            else
                break;
        }
        kept = (uint32_t)(k >= a.qalen ? a.qalen : k);
    }
    a.kept[s] = kept;
}

struct CompactArgs {
    const uint32_t *idx, *kept, *out_idx;
    const int32_t *in, *mz;
    int32_t *out_in, *out_mz;
    uint32_t n_spectra;
};

__global__ void __launch_bounds__(256) compact_kernel(CompactArgs a) {
    const uint32_t s = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= a.n_spectra) return;
    const uint32_t k = a.kept[s], src = a.idx[s + 1] - k, dst = a.out_idx[s];
    for (uint32_t i = threadIdx.x & 63; i < k; i += 64) {
        a.out_in[dst + i] = a.in[src + i];
        a.out_mz[dst + i] = a.mz[src + i];
    }
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gx_queries_preprocess(gx_ctx *c, const gx_query_batch *q, const gx_pick_params *pp,
                                     gx_dqueries **out) {
    if (!c || !q || !pp || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    if (pp->qalen < 1 || pp->qalen > kMaxQalen || pp->base_int < 1)
        return param_error(GX_ERR_INVLD_PARAM, "pick parameters (qalen 1..2048, base_int >= 1)");
    const uint32_t ns = q->n_spectra;
    if (ns > 0 && (!q->idx || !q->precurse)) return GX_ERR_INVLD_PTR;
    if (q->n_peaks > 0 && (!q->moz || !q->intensity)) return GX_ERR_INVLD_PTR;
    if (ns > 0 && q->idx[ns] != q->n_peaks) return GX_ERR_INVLD_SIZE;
    if (q->n_peaks >= (1ull << 31)) return GX_ERR_INVLD_SIZE;
    uint64_t cap = 0;
    uint32_t maxp = 0;
    for (uint32_t s = 0; s < ns; s++) {
        if (q->idx[s + 1] < q->idx[s]) return GX_ERR_INVLD_SIZE;
        const uint32_t n = q->idx[s + 1] - q->idx[s];
        cap += std::min<uint32_t>(n, (uint32_t)pp->qalen);
        maxp = std::max(maxp, std::min<uint32_t>(n, (uint32_t)pp->qalen));
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    gx_dqueries *d = new (std::nothrow) gx_dqueries();
    if (!d) return GX_ERR_BAD_MEM_ALLOC;
    d->ctx = c;
    d->n_spectra = ns;
    d->max_peaks = maxp;
    d->has_dm = q->dm != nullptr;
    d->h_prec.assign(q->precurse, q->precurse + ns);
    uint32_t *r_idx = nullptr, *kept = nullptr;
    int32_t *r_in = nullptr, *r_mz = nullptr;
    void *tmp = nullptr;
    auto release = [&]() {
        for (void *p : {(void *)r_idx, (void *)kept, (void *)r_in, (void *)r_mz, tmp})
            if (p) (void)hipFree(p);
    };
    auto fail = [&](int st) {
        release();
        gx_queries_free(d);
        return st;
    };
    const size_t np = q->n_peaks;
    if (dalloc(&r_idx, (size_t)ns + 1) || dalloc(&kept, (size_t)ns + 1) || dalloc(&r_in, np) || dalloc(&r_mz, np) ||
        dalloc(&d->idx, (size_t)ns + 1) || dalloc(&d->moz, cap) || dalloc(&d->inten, cap) || dalloc(&d->prec, ns) ||
        dalloc(&d->psm, ns) || (d->has_dm && dalloc(&d->dm, ns)))
        return fail(GX_ERR_BAD_MEM_ALLOC);
    hipStream_t st = c->stream;
    size_t tmp_bytes = 0;
    if (rocprim::exclusive_scan(nullptr, tmp_bytes, kept, d->idx, 0u, (size_t)ns + 1, rocprim::plus<uint32_t>(), st) !=
        hipSuccess)
        return fail(GX_ERR_DEVICE);
    if (hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess) return fail(GX_ERR_BAD_MEM_ALLOC);
    if ((q->idx && hipMemcpyAsync(r_idx, q->idx, 4 * ((size_t)ns + 1), hipMemcpyHostToDevice, st) != hipSuccess) ||
        (ns && hipMemcpyAsync(d->prec, q->precurse, 4 * (size_t)ns, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (d->has_dm && hipMemcpyAsync(d->dm, q->dm, 4 * (size_t)ns, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (np && (hipMemcpyAsync(r_in, q->intensity, 4 * np, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(r_mz, q->moz, 4 * np, hipMemcpyHostToDevice, st) != hipSuccess)))
        return fail(GX_ERR_DEVICE);
    if (hipMemsetAsync(kept, 0, 4 * ((size_t)ns + 1), st) != hipSuccess) return fail(GX_ERR_DEVICE);
    if (hipEventRecord(c->ev0, st) != hipSuccess) return fail(GX_ERR_DEVICE);
    if (ns > 0) {
        PickArgs pa{r_idx, r_in, r_mz, kept, ns, pp->base_int, pp->min_int, pp->qalen};
        hipLaunchKernelGGL(pick_kernel, dim3((ns + 63) / 64), dim3(64), 0, st, pa);
        if (hipGetLastError() != hipSuccess) return fail(GX_ERR_DEVICE);
    }
    if (rocprim::exclusive_scan(tmp, tmp_bytes, kept, d->idx, 0u, (size_t)ns + 1, rocprim::plus<uint32_t>(), st) !=
        hipSuccess)
        return fail(GX_ERR_DEVICE);
    if (ns > 0) {
        CompactArgs ca{r_idx, kept, d->idx, r_in, r_mz, d->inten, d->moz, ns};
        hipLaunchKernelGGL(compact_kernel, dim3((ns + 3) / 4), dim3(256), 0, st, ca);
        if (hipGetLastError() != hipSuccess) return fail(GX_ERR_DEVICE);
    }
    if (hipEventRecord(c->ev1, st) != hipSuccess) return fail(GX_ERR_DEVICE);
    uint32_t total = 0;
    if (hipMemcpyAsync(&total, d->idx + ns, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) d->prep_ms = ms;
    release();
    if (total > cap) return fail(GX_ERR_DEVICE);  // cannot happen: each kept count is <= min(n, qalen)
    d->n_peaks = total;
    *out = d;
    return GX_SUCCESS;
}

extern "C" int gx_queries_shape(const gx_dqueries *d, uint32_t *n_spectra, uint64_t *n_peaks, float *prep_ms) {
    if (!d) return GX_ERR_INVLD_PTR;
    if (n_spectra) *n_spectra = d->n_spectra;
    if (n_peaks) *n_peaks = d->n_peaks;
    if (prep_ms) *prep_ms = d->prep_ms;
    return GX_SUCCESS;
}

extern "C" int gx_queries_download(const gx_dqueries *d, uint32_t *idx, int32_t *moz, int32_t *intensity) {
    if (!d) return GX_ERR_INVLD_PTR;
    std::lock_guard<std::mutex> lk(d->ctx->mu);
    if (hipSetDevice(d->ctx->device) != hipSuccess) return GX_ERR_DEVICE;
    GX_HIP(hipStreamSynchronize(d->ctx->stream));
    if (idx) GX_HIP(hipMemcpy(idx, d->idx, 4 * ((size_t)d->n_spectra + 1), hipMemcpyDeviceToHost));
    if (d->n_peaks) {
        if (moz) GX_HIP(hipMemcpy(moz, d->moz, 4 * d->n_peaks, hipMemcpyDeviceToHost));
        if (intensity) GX_HIP(hipMemcpy(intensity, d->inten, 4 * d->n_peaks, hipMemcpyDeviceToHost));
    }
    return GX_SUCCESS;
}
