// gx_search_spectra_noalloc: the per-spectrum search engine (search_kernel)
// with every buffer beyond the caller's inputs and outputs in one static
// device array, for callers that cannot allocate device memory. No
// hipMalloc/hipFree and no allocating library call happens on this path.
//
// The array (`scratch`) lives in this file. Its name is only a host-side
// stand-in: each device's copy is reached through the address that
// hipGetSymbolAddress returns with that device current, looked up once per
// device (at gx_ctx_create) and kept here. Kernels take that address as an
// argument. Layout (32-bit words):
//
//   [0, 64)                       work counter (word 0) and the batch check (words 1..3)
//   [64, 64 + 2*ST_N)             search counters (u64)
//   [kSliceBase, + blocks*2*max)  per-block slice cursors/ends beyond the LDS
//   [kCandBase, + blocks*tile)    per-block candidates beyond the LDS
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "batched.h"
#include "capi_internal.h"

namespace gx {
namespace {

constexpr int kNoAllocBlocks = 512;                // 2 blocks per CU on 256 CUs
constexpr uint32_t kNoAllocMaxSlices = 4096;      // (legal peaks) x (2 dF + 1) per spectrum and bucket
constexpr uint64_t kSliceBase = 256;
constexpr uint64_t kCandBase = kSliceBase + (uint64_t)kNoAllocBlocks * 2 * kNoAllocMaxSlices;
constexpr uint64_t kScratchWords = kCandBase + (uint64_t)kNoAllocBlocks * kTileEntries;  // 15.7M words, 60 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

uint32_t *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;               // guards g_scratch_ptr
std::mutex g_device_mu[kMaxDevices];   // one no-alloc call at a time per device (they share the array)

// largest peak count of the batch (word 1) and a malformed-CSR flag (word 2):
// idx must be non-decreasing and end at n_peaks
__global__ void check_batch_kernel(const uint32_t *idx, uint32_t n_spectra, uint64_t n_peaks, uint32_t *chk) {
    uint32_t mx = 0, bad = 0;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_spectra; s += gridDim.x * blockDim.x) {
        const uint32_t a = idx[s], b = idx[s + 1];
        if (b < a) bad = 1;
        else mx = max(mx, b - a);
        if (s + 1 == n_spectra && (uint64_t)b != n_peaks) bad = 1;
    }
    if (mx) atomicMax(&chk[1], mx);
    if (bad) atomicOr(&chk[2], 1u);
}

}  // namespace

// Called by gx_ctx_create with `device` current: the device address of this
// device's copy of `scratch`, looked up once and kept.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return GX_ERR_INVLD_PARAM;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return GX_SUCCESS;
    uint32_t *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        note_error("hipGetSymbolAddress(scratch)", __LINE__, e != hipSuccess ? e : hipErrorInvalidSymbol);
        return GX_ERR_DEVICE;
    }
    g_scratch_ptr[device] = p;
    return GX_SUCCESS;
}

}  // namespace gx

extern "C" int gx_search_spectra_noalloc(gx_ctx *c, const gx_dindex *idx, const gx_query_batch *q,
                                         const gx_search_params *p, gx_psm *psms, uint32_t *hist,
                                         gx_search_stats *stats) {
    if (!c || !idx || !q || !p || !psms) return GX_ERR_INVLD_PTR;
    if (q->n_spectra > 0 && (!q->idx || !q->precurse)) return GX_ERR_INVLD_PTR;
    if (q->n_peaks > 0 && (!q->moz || !q->intensity)) return GX_ERR_INVLD_PTR;
    if (q->n_peaks >= (1ull << 32)) return param_error(GX_ERR_INVLD_SIZE, "n_peaks must be below 2^32");
    if (idx->ctx != c) return param_error(GX_ERR_INVLD_PARAM, "index of another context");
    if (p->engine != GX_ENGINE_SPECTRUM || p->evalue_mode != GX_EVALUE_TAILFIT)
        return param_error(GX_ERR_INVLD_PARAM, "the no-alloc search runs the per-spectrum engine with the plain fit");
    if (p->dF < 0 || p->dF > 64) return param_error(GX_ERR_INVLD_PARAM, "dF outside 0..64");
    if (p->min_shp < 0 || p->min_cpsm < 0) return param_error(GX_ERR_INVLD_PARAM, "min_shp and min_cpsm must be >= 0");
    if (p->scale < 1 || p->max_mass < 1 || (uint64_t)p->max_mass * (uint64_t)p->scale != idx->nbins)
        return param_error(GX_ERR_INVLD_INDEX, "max_mass * scale differs from the index's bin count");
    if ((int)((2 * p->dF + 1) * (idx->max_speclen / 2) + 1) > c->lgfact_n) return GX_ERR_INVLD_SIZE;
    if (c->device < 0 || c->device >= kMaxDevices) return GX_ERR_INVLD_PARAM;
    uint32_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[c->device];
    }
    if (!base) return param_error(GX_ERR_NOT_READY, "device scratch not set up (gx_ctx_create)");
    std::lock_guard<std::mutex> dev_lk(g_device_mu[c->device]);
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    hipStream_t sm = c->stream;
    uint32_t *counter = base;
    uint32_t *chk = base;  // words 1..3
    unsigned long long *st_dev = reinterpret_cast<unsigned long long *>(base + 64);
    // the array keeps the previous call's contents: clear the counter, the
    // check words and the counters (kernels, not memset nodes)
    GX_HIP(launch_fill_u32(base, 64 + 2 * (uint64_t)ST_N, 0, sm));
    uint32_t maxp = 0;
    if (q->n_spectra > 0) {
        const uint32_t blocks = std::min<uint32_t>(1024, (q->n_spectra + 255) / 256);
        hipLaunchKernelGGL(check_batch_kernel, dim3(blocks), dim3(256), 0, sm, q->idx, q->n_spectra, q->n_peaks, chk);
        GX_HIP(hipGetLastError());
        uint32_t h[4] = {};
        GX_HIP(hipMemcpyAsync(h, chk, sizeof(h), hipMemcpyDeviceToHost, sm));
        GX_HIP(hipStreamSynchronize(sm));
        if (h[2]) return param_error(GX_ERR_INVLD_SIZE, "query idx is not a CSR offset array ending at n_peaks");
        maxp = h[1];
    }
    if (maxp > (uint32_t)kMaxPeaksLds) return param_error(GX_ERR_INVLD_SIZE, "more than 2048 peaks in a spectrum");
    const uint32_t max_slices = std::max<uint32_t>(1, maxp * (uint32_t)(2 * p->dF + 1));
    if (max_slices > kNoAllocMaxSlices)
        return param_error(GX_ERR_INVLD_SIZE, "peaks x (2 dF + 1) above the no-alloc limit of 4096 slices");
    const int n_blocks = std::max(1, std::min(kNoAllocBlocks, c->n_cu * 2));
    SearchArgs a{};
    a.buckets = idx->d_buckets;
    a.n_buckets = idx->n_buckets;
    a.qidx = q->idx;
    a.moz = q->moz;
    a.inten = q->intensity;
    a.prec = q->precurse;
    a.dm = q->dm;
    a.n_spectra = q->n_spectra;
    a.dF = p->dF;
    a.nbins = (int)idx->nbins;
    a.min_shp = p->min_shp;
    a.min_cpsm = p->min_cpsm;
    a.dM = p->dM;
    a.expect_max = p->expect_max;
    a.lgfact = c->d_lgfact;
    a.lgfact_n = c->lgfact_n;
    a.out = psms;
    a.hist_out = hist;
    a.work_counter = counter;
    // per-block regions at the fixed maximum stride, so that a smaller batch
    // never reaches past what the array holds
    a.slice_scratch = base + kSliceBase;
    a.max_slices = kNoAllocMaxSlices;
    a.cand_scratch = base + kCandBase;
    a.stats = st_dev;
    GX_HIP(hipEventRecord(c->ev0, sm));
    if (q->n_spectra > 0) GX_HIP(launch_search(a, n_blocks, sm));
    GX_HIP(hipEventRecord(c->ev1, sm));
    unsigned long long h[ST_N] = {};
    if (stats) GX_HIP(hipMemcpyAsync(h, st_dev, sizeof(h), hipMemcpyDeviceToHost, sm));
    GX_HIP(hipStreamSynchronize(sm));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ions = h[ST_IONS];
        stats->bins = h[ST_BINS];
        stats->window = h[ST_WINDOW];
        stats->candidates = h[ST_CAND];
        stats->spectra = h[ST_SPECTRA];
        stats->peaks = h[ST_PEAKS];
        stats->ions_scanned = h[ST_SCANNED];
        stats->errors = h[ST_ERRORS];
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) stats->kernel_ms = ms;
        stats->total_ms = stats->kernel_ms;
        stats->config = -1;
    }
    return GX_SUCCESS;
}
