// Fragment-ion index construction on the GPU (gx_index_build): the work of
// s1::ConstructIndexChunk (superstep1/kernel.cu:484-733), or on the CPU of
// DSLIM_ConstructChunk + DSLIM_SLMTransform (dslim.cpp:89-97), for every
// length bucket of mass-sorted entries:
//
//   fragment_kernel  one thread per entry: the fragment bins of its b and y
//                    ions in the CPU build's arithmetic (float prefix sums of
//                    UTILS_GenerateSpectrum, utils.cpp:161-226; the double
//                    per-site additions of UTILS_GenerateModSpectrum,
//                    utils.cpp:347-474; the precursor-mass filter and clamp
//                    of dslim.cpp:318-337), keyed by ion id;
//   radix sort       rocPRIM's device radix sort of (bin, ion id) pairs over
//                    the bin's bits: stable, so ion ids ascend within each
//                    bin (the SLM transform, dslim.cpp:376-402, 474-504);
//   bin_starts       bA[b] = first position of bin b in the sorted keys.
//
// The reference's own GPU build sums in another order (superstep1/kernel.cu
// :635-733), which moves some ions across bin boundaries; this one is
// bit-identical to the CPU index.
#include <rocprim/device/device_radix_sort.hpp>

#include <vector>

#include "../host/slm_constants.h"
#include "capi_internal.h"

namespace gx {
namespace {

struct FragArgs {
    float aa[26];        // residue + static modification masses, per letter
    float res[26];       // residue masses alone (UTILS_CalculateModMass adds them separately)
    float smod[26];      // static modification masses
    float h2o, proton;
    const char *seqs;
    const uint32_t *seqid, *modnum;
    const uint64_t *sites;
    uint32_t n_entries;
    int L, maxz, scale;
    double min_mass, max_mass;
    uint32_t nbins;
    uint32_t mod_mass[GX_MAX_MOD_TYPES];
    int all_mods;
    uint32_t *keys;      // bin of ion id i
    uint32_t *vals;      // i
};

__device__ __forceinline__ int letter(char c) { return c - 'A'; }

__global__ void __launch_bounds__(256) fragment_kernel(FragArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_entries) return;
    const int L = a.L, n = L - 1, maxz = a.maxz;
    const uint32_t speclen = 2u * (uint32_t)maxz * (uint32_t)n;
    const char *seq = a.seqs + (size_t)a.seqid[k] * (size_t)L;
    const uint32_t modnum = a.modnum[k];
    const uint64_t sites = a.sites[k];
    uint32_t *keys = a.keys + (size_t)k * speclen;
    uint32_t *vals = a.vals + (size_t)k * speclen;
    const uint32_t id0 = k * speclen;
    // precursor mass and the fill / range decisions of DSLIM_Construct
    bool keep;
    if (modnum == 0) {
        float m = a.h2o;  // UTILS_CalculatePepMass
        for (int l = 0; l < L; l++) m += a.aa[letter(seq[l])];
        keep = m > 0.0f && (double)m >= a.min_mass && (double)m <= a.max_mass;
    } else {
        float m = a.h2o;  // UTILS_CalculateModMass
        for (int l = 0; l < L; l++) {
            m += a.res[letter(seq[l])];
            m += a.smod[letter(seq[l])];
        }
        for (uint32_t w = modnum; (w & 0xFu) != 0; w >>= 4) {
            m += (float)a.mod_mass[(w & 0xFu) - 1u] / (float)a.scale;
            if (!a.all_mods) break;
        }
        keep = (double)m > a.min_mass && (double)m < a.max_mass;
    }
    if (!keep) {
        for (uint32_t r = 0; r < speclen; r++) {
            keys[r] = 0;
            vals[r] = id0 + r;
        }
        return;
    }
    // modification numbers of the set sites, in order
    int mods[8];
    int seen = 0;
    for (int i = 0; i < 8; i++) {
        mods[i] = (int)((modnum >> (4 * i)) & 0xFu) - 1;
        seen += mods[i] != -1 ? 1 : 0;
    }
    const float fs = (float)a.scale;
    const uint32_t top = a.nbins - 1u;
    for (int z = 0; z < maxz; z++) {
        const int bs = z * n, ys = z * n + maxz * n;
        const float zf = (float)(z + 1);
        // b ions: prefix sums from the N-terminus; y ions: from the C-terminus
        float fb = a.aa[letter(seq[0])] + a.proton * (float)(z + 1);
        float fy = (a.aa[letter(seq[L - 1])] + a.proton * (float)(z + 1)) + a.h2o;
        int cb = 0, cy = 0;
        for (int l = 0; l < n; l++) {
            if (l > 0) {
                fb = fb + (a.aa[letter(seq[l])] + a.proton * 0.0f);
                fy = fy + (a.aa[letter(seq[L - 1 - l])] + a.proton * 0.0f);
            }
            float vb = fb, vy = fy;
            if (modnum != 0) {
                cb = min(cb + (int)((sites >> l) & 1u), seen);
                for (int q = 0; q < cb; q++) vb = (float)((double)vb + (double)a.mod_mass[mods[q]] / a.scale);
                cy = min(cy + (int)((sites >> (L - 1 - l)) & 1u), seen);
                for (int q = 0; q < cy; q++)
                    vy = (float)((double)vy + (double)a.mod_mass[mods[seen - 1 - q]] / a.scale);
            }
            uint32_t bb = (uint32_t)((vb * fs) / zf), by = (uint32_t)((vy * fs) / zf);
            keys[bs + l] = bb > top ? top : bb;
            keys[ys + l] = by > top ? top : by;
            vals[bs + l] = id0 + (uint32_t)(bs + l);
            vals[ys + l] = id0 + (uint32_t)(ys + l);
        }
    }
}

// bA[b] = number of keys < b: a lower bound per bin in the sorted keys
__global__ void __launch_bounds__(256) bin_starts_kernel(const uint32_t *keys, uint32_t n, uint32_t nbins,
                                                         uint32_t *bA) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbins) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    bA[b] = lo;
}

// every index the fragment kernel derives from an entry stays in bounds:
// letters 'A'..'Z', seqid < n_peps, modification nibbles of defined types
// packed from the low end, no more set sites than modifications
bool entries_valid(const gx_bucket &x, int n_mod_types) {
    const size_t nseq = (size_t)x.n_peps * (size_t)x.peplen;
    for (size_t i = 0; i < nseq; i++)
        if (x.seqs[i] < 'A' || x.seqs[i] > 'Z') return false;
    for (uint32_t k = 0; k < x.n_entries; k++) {
        if (x.seqid[k] >= x.n_peps) return false;
        int seen = 0;
        bool ended = false;
        for (int i = 0; i < 8; i++) {
            const uint32_t m = (x.modnum[k] >> (4 * i)) & 0xFu;
            if (m == 0) {
                ended = true;
                continue;
            }
            if (ended || (int)m > n_mod_types) return false;
            seen++;
        }
        if (x.sites[k] >> x.peplen) return false;
        if (__builtin_popcountll(x.sites[k]) > seen) return false;
    }
    return true;
}

int build_bucket_device(gx_ctx *c, const gx_bucket &x, const gx_index_build_params &bp, uint32_t **bA_out,
                        uint32_t **iA_out) {
    *bA_out = *iA_out = nullptr;
    const uint64_t n = x.n_ions;
    const uint32_t nbins = x.nbins;
    hipStream_t s = c->stream;
    char *d_seqs = nullptr;
    uint32_t *d_seqid = nullptr, *d_modnum = nullptr, *keys = nullptr, *vals = nullptr, *keys2 = nullptr,
             *iA = nullptr, *bA = nullptr;
    uint64_t *d_sites = nullptr;
    void *tmp = nullptr;
    auto cleanup = [&]() {
        for (void *p : {(void *)d_seqs, (void *)d_seqid, (void *)d_modnum, (void *)keys, (void *)vals,
                        (void *)keys2, (void *)d_sites, tmp})
            if (p) (void)hipFree(p);
    };
    auto fail = [&](int st) {
        cleanup();
        if (iA) (void)hipFree(iA);
        if (bA) (void)hipFree(bA);
        return st;
    };
    const size_t nseq = (size_t)x.n_peps * (size_t)x.peplen;
    if (dalloc(&d_seqs, nseq) || dalloc(&d_seqid, x.n_entries) || dalloc(&d_modnum, x.n_entries) ||
        dalloc(&d_sites, x.n_entries) || dalloc(&keys, n) || dalloc(&vals, n) || dalloc(&keys2, n) ||
        dalloc(&iA, n + kIaPad) || dalloc(&bA, (size_t)nbins + 1))  // iA's padding: set by index_create
        return fail(GX_ERR_BAD_MEM_ALLOC);
    if (hipMemcpy(d_seqs, x.seqs, nseq, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_seqid, x.seqid, 4 * (size_t)x.n_entries, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_modnum, x.modnum, 4 * (size_t)x.n_entries, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_sites, x.sites, 8 * (size_t)x.n_entries, hipMemcpyHostToDevice) != hipSuccess)
        return fail(GX_ERR_DEVICE);
    FragArgs fa{};
    for (int i = 0; i < 26; i++) {
        fa.aa[i] = kResidueMass[i] + kStaticMod[i];
        fa.res[i] = kResidueMass[i];
        fa.smod[i] = kStaticMod[i];
    }
    fa.h2o = kH2O;
    fa.proton = kProton;
    fa.seqs = d_seqs;
    fa.seqid = d_seqid;
    fa.modnum = d_modnum;
    fa.sites = d_sites;
    fa.n_entries = x.n_entries;
    fa.L = x.peplen;
    fa.maxz = bp.maxz;
    fa.scale = bp.scale;
    fa.min_mass = bp.min_mass;
    fa.max_mass = bp.max_mass;
    fa.nbins = nbins;
    for (int t = 0; t < GX_MAX_MOD_TYPES; t++) fa.mod_mass[t] = bp.mod_mass[t];
    fa.all_mods = bp.modmass_all;
    fa.keys = keys;
    fa.vals = vals;
    hipLaunchKernelGGL(fragment_kernel, dim3((x.n_entries + 255) / 256), dim3(256), 0, s, fa);
    if (hipGetLastError() != hipSuccess) return fail(GX_ERR_DEVICE);
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < nbins) end_bit++;
    size_t tmp_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, vals, iA, (size_t)n, 0u, (unsigned)end_bit, s) !=
        hipSuccess)
        return fail(GX_ERR_DEVICE);
    if (hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess) return fail(GX_ERR_BAD_MEM_ALLOC);
    if (rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, vals, iA, (size_t)n, 0u, (unsigned)end_bit, s) !=
        hipSuccess)
        return fail(GX_ERR_DEVICE);
    hipLaunchKernelGGL(bin_starts_kernel, dim3((nbins + 1 + 255) / 256), dim3(256), 0, s, keys2, (uint32_t)n, nbins,
                       bA);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return fail(GX_ERR_DEVICE);
    cleanup();
    *bA_out = bA;
    *iA_out = iA;
    return GX_SUCCESS;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gx_index_build(gx_ctx *c, const gx_bucket *b, int nb, const gx_index_build_params *bp,
                              gx_dindex **out) {
    if (!c || !b || !bp || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    if (nb < 1 || nb > kMaxBuckets || bp->maxz < 1 || bp->maxz > 6 || bp->scale < 1 || bp->n_mod_types < 0 ||
        bp->n_mod_types > GX_MAX_MOD_TYPES)
        return param_error(GX_ERR_INVLD_PARAM, "build parameters");
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    std::vector<uint32_t *> bA(nb, nullptr), iA(nb, nullptr);
    auto release = [&]() {
        for (int i = 0; i < nb; i++) {
            if (bA[i]) (void)hipFree(bA[i]);
            if (iA[i]) (void)hipFree(iA[i]);
        }
    };
    for (int i = 0; i < nb; i++) {
        const gx_bucket &x = b[i];
        if (x.n_entries == 0) continue;
        if (!x.seqs || !x.seqid || !x.modnum || !x.sites || !x.mass) {
            release();
            return GX_ERR_INVLD_PTR;
        }
        if (x.peplen < 2 || x.peplen > 62 || x.speclen != (uint32_t)((x.peplen - 1) * bp->maxz * 2) ||
            (uint64_t)x.n_entries * x.speclen != x.n_ions || x.n_ions >= (1ull << 31) ||
            x.nbins != (uint32_t)(bp->max_mass * bp->scale) || x.nbins < 2) {
            release();
            return param_error(GX_ERR_INVLD_INDEX, "bucket metadata");
        }
        if (!entries_valid(x, bp->n_mod_types)) {
            release();
            return param_error(GX_ERR_INVLD_INDEX, "bucket entries (sequence letters, seqid, modnum or sites)");
        }
        const int st = build_bucket_device(c, x, *bp, &bA[i], &iA[i]);
        if (st) {
            release();
            return st;
        }
    }
    // empty buckets carry no arrays; index_create skips them
    return index_create(c, b, nb, bp->maxz, bA.data(), iA.data(), out);
}

extern "C" int gx_index_download(gx_dindex *d, int bucket, uint32_t *bA, uint32_t *iA) {
    if (!d) return GX_ERR_INVLD_PTR;
    if (bucket < 0 || bucket >= d->n_buckets) return GX_ERR_INVLD_INDEX;
    std::lock_guard<std::mutex> lk(d->ctx->mu);
    if (hipSetDevice(d->ctx->device) != hipSuccess) return GX_ERR_DEVICE;
    const DevBucket &db = d->h_buckets[bucket];
    if (db.n_entries == 0) return GX_SUCCESS;
    if (bA) GX_HIP(hipMemcpy(bA, db.bA, sizeof(uint32_t) * ((size_t)d->nbins + 1), hipMemcpyDeviceToHost));
    if (iA) GX_HIP(hipMemcpy(iA, db.iA, sizeof(uint32_t) * (size_t)db.n_entries * db.speclen, hipMemcpyDeviceToHost));
    return GX_SUCCESS;
}
