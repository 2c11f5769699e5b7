// Fragment-ion-index search + tail-fit e-value, one persistent kernel for gfx950.
//
// What it computes is DSLIM_QuerySpectrum (dslim_query.cpp:749-1081) followed
// by expeRT::ModelTailFit (expeRT.cpp:396-559) for every spectrum of a batch;
// how it computes it is MI355X-first:
//
//   * persistent blocks (2 per CU) pull spectra from a global atomic counter;
//   * per spectrum the legal peaks are compacted and rank-sorted by m/z in
//     LDS, and the precursor window of every length bucket is found by one
//     lane per bucket;
//   * per (spectrum, bucket) every (peak, bin) slice of iA is located with one
//     lower_bound on the window start; slices are then streamed tile by tile
//     over the psid window: short slices lane-serially, long ones (the
//     y1/b1 "hub" bins that hold a large share of a bucket) by a whole wave
//     reading 64 consecutive ion ids and keeping the ballot prefix below the
//     tile end;
//   * pass 1 counts shared peaks per entry with u32 LDS atomics on packed
//     16-bit counters (2 entries per word, so a 44 KB tile covers 22.5K
//     psids);
//   * entries with count >= min_shp become candidates; pass 2 rescans each
//     candidate's own fragment bins (ion_bin, the inverse of iA) against the
//     sorted peaks to get bc, yc, ibc, iyc exactly as the scatter sums them,
//     then the hyperscore (double log10, stored as f32), the survival
//     histogram (LDS atomics) and the top hit;
//   * the tail fit runs on the LDS histogram at the end of the spectrum.
#include "device_util.h"

namespace gx {
namespace {

constexpr int kWaves = kSearchThreads / 64;
constexpr uint32_t kLight = 16;  // slices with at most this many ions left are walked by one lane

struct Lds {
    uint32_t cnt[kTileWords];          // count tile; also scratch for the peak sort and the tail fit
    int32_t pk_q[kMaxPeaksLds];        // legal peaks sorted by m/z
    int32_t pk_i[kMaxPeaksLds];
    uint32_t sl_cur[kSliceLds];        // slice cursor (next ion index)
    uint32_t sl_end[kSliceLds];        // slice end (bin end, exclusive)
    uint32_t cand[kCandLds];
    uint32_t hist[GX_HIST_BINS];
    int32_t win_lo[kMaxBuckets];
    int32_t win_hi[kMaxBuckets];
    // per-wave top hits for the block reduction
    float w_best_h[kWaves];
    int32_t w_best_b[kWaves];
    int32_t w_best_p[kWaves];
    uint32_t w_best_shp[kWaves];
    uint32_t w_eq[kWaves];
    uint32_t w_cpsms[kWaves];
    uint32_t n_legal;
    uint32_t n_cand;
    uint32_t spec;
    FitScratch fit;
    unsigned long long st[ST_N];
};

__global__ void reset_kernel(uint32_t *counter, unsigned long long *stats) {
    const int t = threadIdx.x;
    if (t == 0) *counter = 0;
    if (t < ST_N) stats[t] = 0;
}

// iS[k] = iA[k * kSampleStep]: a thinner copy of the ion ids, for the plan's
// boundary brackets
__global__ void sample_ids_kernel(const uint32_t *__restrict__ iA, uint32_t *__restrict__ iS, uint64_t n) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        iS[k] = iA[k << kSampleShift];
}

__global__ void ion_bin_kernel(const uint32_t *__restrict__ bA, const uint32_t *__restrict__ iA,
                               uint32_t *__restrict__ ion_bin, uint32_t nbins, uint64_t n_ions) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_ions;
         i += (uint64_t)gridDim.x * blockDim.x) {
        // bin b holds positions [bA[b], bA[b+1]); find the last b with bA[b] <= i
        uint32_t lo = 0, n = nbins + 1;
        while (n > 0) {
            uint32_t half = n >> 1;
            if (!((uint32_t)i < bA[lo + half])) {
                lo += half + 1;
                n -= half + 1;
            } else {
                n = half;
            }
        }
        ion_bin[iA[i]] = lo - 1;
    }
}

__global__ void __launch_bounds__(kSearchThreads, 4) search_kernel(SearchArgs a) {
    __shared__ Lds L;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int dF = a.dF;
    const int nbin5 = 2 * dF + 1;
    const uint32_t min_cnt = (uint32_t)max(a.min_shp, 1);
    uint32_t *g_slices = a.slice_scratch + (size_t)blockIdx.x * 2 * a.max_slices;
    uint32_t *g_cand = a.cand_scratch + (size_t)blockIdx.x * kTileEntries;

    for (int w = tid; w < kTileWords; w += kSearchThreads) L.cnt[w] = 0;
    if (tid < ST_N) L.st[tid] = 0;
    unsigned long long my_ions = 0, my_bins = 0;

    for (;;) {
        if (tid == 0) L.spec = atomicAdd(a.work_counter, 1u);
        __syncthreads();
        const uint32_t s = L.spec;
        if (s >= a.n_spectra) break;
        const uint32_t off = a.qidx[s];
        const int npk = (int)(a.qidx[s + 1] - off);
        const float prec = a.prec[s];
        const double dM = a.dm ? (double)a.dm[s] : a.dM;

        // ---- legal peaks, compacted into the count tile (scratch) ------
        if (tid == 0) L.n_legal = 0;
        for (int w = tid; w < GX_HIST_BINS; w += kSearchThreads) L.hist[w] = 0;
        __syncthreads();
        int32_t *tmp_q = reinterpret_cast<int32_t *>(L.cnt);
        int32_t *tmp_i = tmp_q + kMaxPeaksLds;
        for (int base = 0; base < npk; base += kSearchThreads) {
            const int k = base + tid;
            int32_t q = 0, in = 0;
            bool legal = false;
            if (k < npk) {
                q = a.moz[off + k];
                in = a.inten[off + k];
                legal = q > dF && (int64_t)q < (int64_t)a.nbins - 1 - dF;
            }
            const uint64_t m = __ballot(legal);
            uint32_t wbase = 0;
            if (lane == 0 && m) wbase = atomicAdd(&L.n_legal, (uint32_t)__popcll(m));
            wbase = __shfl(wbase, 0, 64);
            if (legal) {
                const uint32_t pos = wbase + (uint32_t)__popcll(m & ((1ull << lane) - 1));
                tmp_q[pos] = q;
                tmp_i[pos] = in;
            }
        }
        __syncthreads();
        const int P = (int)L.n_legal;
        // rank sort by (m/z, position): unique ranks, stable
        for (int k = tid; k < P; k += kSearchThreads) {
            const int32_t q = tmp_q[k];
            int r = 0;
            for (int j = 0; j < P; j++) {
                const int32_t o = tmp_q[j];
                r += (o < q) || (o == q && j < k);
            }
            L.pk_q[r] = q;
            L.pk_i[r] = tmp_i[k];
        }
        // ---- precursor windows, one lane per bucket (DSLIM_BinarySearch) ----
        for (int b = tid; b < a.n_buckets; b += kSearchThreads) {
            const DevBucket bk = a.buckets[b];
            int32_t lo = 0, hi = -1;
            const float pm1 = (float)((double)prec - dM);
            const float pm2 = (float)((double)prec + dM);
            const uint32_t n = bk.n_entries;
            bool ok = n > 0 && !(dM < 0.0);
            if (ok) {
                const float m0 = bk.mass[0], mN = bk.mass[n - 1];
                if (pm1 < m0)
                    lo = 0;
                else if (pm1 > mN)
                    ok = false;
                else
                    lo = (int32_t)lower_bound_f32(bk.mass, n, pm1);
                if (ok) {
                    if (pm2 > mN)
                        hi = (int32_t)n - 1;
                    else if (pm2 < m0)
                        ok = false;
                    else
                        hi = (int32_t)upper_bound_f32(bk.mass, n, pm2) - 1;
                }
                if (ok) ok = hi >= 0 && lo < (int32_t)n && bk.mass[hi] <= pm2 && bk.mass[lo] >= pm1 && hi >= lo;
            }
            L.win_lo[b] = ok ? lo : 0;
            L.win_hi[b] = ok ? hi : -1;
        }
        __syncthreads();
        // the sort used the count tile as scratch
        for (int w = tid; w < 2 * kMaxPeaksLds; w += kSearchThreads) L.cnt[w] = 0;
        if (tid < kWaves) {
            L.w_best_h[tid] = 0.0f;
            L.w_best_b[tid] = 0x7fffffff;
            L.w_best_p[tid] = 0x7fffffff;
            L.w_best_shp[tid] = 0;
            L.w_eq[tid] = 0;
            L.w_cpsms[tid] = 0;
        }
        __syncthreads();

        // per-wave top hit, kept in lane 0 registers
        float bh = 0.0f;
        int bb = 0x7fffffff, bp = 0x7fffffff;
        uint32_t bshp = 0, beq = 0, bcps = 0;
        uint64_t my_window = 0;
        const int nsl = P * nbin5;
        uint32_t *sl_cur = (nsl <= kSliceLds) ? L.sl_cur : g_slices;
        uint32_t *sl_end = (nsl <= kSliceLds) ? L.sl_end : g_slices + a.max_slices;

        for (int b = 0; b < a.n_buckets; b++) {
            const int32_t lo = L.win_lo[b], hi = L.win_hi[b];
            if (hi < lo) continue;
            const DevBucket bk = a.buckets[b];
            const uint32_t speclen = bk.speclen;
            const uint32_t lo_id = (uint32_t)lo * speclen;
            if (tid == 0) my_window += (uint64_t)(hi - lo + 1);
            // ---- slices: [lower_bound(window start), bin end) ----
            for (int j = tid; j < nsl; j += kSearchThreads) {
                const int k = j / nbin5;
                const uint32_t bin = (uint32_t)(L.pk_q[k] - dF + (j - k * nbin5));
                const uint32_t st = gptr(bk.bA)[bin], en = gptr(bk.bA)[bin + 1];
                uint32_t c = en;
                if (en > st) c = st + lower_bound_u32(gptr(bk.iA) + st, en - st, lo_id);
                sl_cur[j] = c;
                sl_end[j] = en;
                my_bins++;
            }
            __syncthreads();
            for (int32_t tlo = lo; tlo <= hi; tlo += kTileEntries) {
                const int32_t thi = min(hi, tlo + kTileEntries - 1);
                const uint32_t id_end = (uint32_t)(thi + 1) * speclen;
                const uint32_t tbase = (uint32_t)tlo;
                // ---- pass 1a: short slices, one lane each ----
                for (int j = tid; j < nsl; j += kSearchThreads) {
                    uint32_t c = sl_cur[j];
                    const uint32_t e = sl_end[j];
                    if (e - c > kLight || c >= e) continue;
                    while (c < e) {
                        const uint32_t raw = gptr(bk.iA)[c];
                        if (raw >= id_end) break;
                        const uint32_t el = raw / speclen - tbase;
                        atomicAdd(&L.cnt[el >> 1], 1u << ((el & 1u) << 4));
                        c++;
                        my_ions++;
                    }
                    sl_cur[j] = c;
                }
                __syncthreads();
                // ---- pass 1b: long slices, one wave each, 64 ids per step ----
                for (int j = wave; j < nsl; j += kWaves) {
                    uint32_t c = sl_cur[j];
                    const uint32_t e = sl_end[j];
                    if (e - c <= kLight) continue;
                    for (;;) {
                        const uint32_t idx = c + (uint32_t)lane;
                        const bool inb = idx < e;
                        const uint32_t raw = inb ? gptr(bk.iA)[idx] : 0xffffffffu;
                        const bool valid = inb && raw < id_end;
                        if (valid) {
                            const uint32_t el = raw / speclen - tbase;
                            atomicAdd(&L.cnt[el >> 1], 1u << ((el & 1u) << 4));
                            my_ions++;
                        }
                        const uint32_t nv = (uint32_t)__popcll(__ballot(valid));
                        c += nv;
                        if (nv < 64) break;
                    }
                    if (lane == 0) sl_cur[j] = c;
                }
                __syncthreads();
                // ---- scan the tile: candidates, clear ----
                if (tid == 0) L.n_cand = 0;
                __syncthreads();
                const uint32_t nent = (uint32_t)(thi - tlo + 1);
                const uint32_t nw = (nent + 1) >> 1;
                for (uint32_t w = tid; w < nw; w += kSearchThreads) {
                    const uint32_t v = L.cnt[w];
                    if (v == 0) continue;
                    L.cnt[w] = 0;
                    const uint32_t c0 = v & 0xffffu, c1 = v >> 16;
                    if (c0 >= min_cnt) {
                        const uint32_t pos = atomicAdd(&L.n_cand, 1u);
                        const uint32_t psid = tbase + 2 * w;
                        if (pos < kCandLds) L.cand[pos] = psid; else g_cand[pos - kCandLds] = psid;
                    }
                    if (c1 >= min_cnt && 2 * w + 1 < nent) {
                        const uint32_t pos = atomicAdd(&L.n_cand, 1u);
                        const uint32_t psid = tbase + 2 * w + 1;
                        if (pos < kCandLds) L.cand[pos] = psid; else g_cand[pos - kCandLds] = psid;
                    }
                }
                __syncthreads();
                // ---- pass 2: rescore candidates exactly, one wave each ----
                const uint32_t ncand = L.n_cand;
                const uint32_t half = speclen >> 1;
                for (uint32_t ci = wave; ci < ncand; ci += kWaves) {
                    const uint32_t psid = ci < kCandLds ? L.cand[ci] : g_cand[ci - kCandLds];
                    const auto bins = gptr(bk.ion_bin) + (size_t)psid * speclen;
                    uint32_t bc = 0, yc = 0, ibc = 0, iyc = 0;
                    for (uint32_t r = lane; r < speclen; r += 64) {
                        const int32_t x = (int32_t)bins[r];
                        int k = lower_bound_lds(L.pk_q, P, x - dF);
                        uint32_t n = 0, sum = 0;
                        for (; k < P && L.pk_q[k] <= x + dF; k++) {
                            n++;
                            sum += (uint32_t)L.pk_i[k];
                        }
                        if (r < half) {
                            bc += n;
                            ibc += sum;
                        } else {
                            yc += n;
                            iyc += sum;
                        }
                    }
                    bc = wave_sum_u32(bc);
                    yc = wave_sum_u32(yc);
                    ibc = wave_sum_u32(ibc);
                    iyc = wave_sum_u32(iyc);
                    if (lane == 0) {
                        const uint32_t shpk = bc + yc;
                        if (shpk >= (uint32_t)a.min_shp && (int)bc < a.lgfact_n && (int)yc < a.lgfact_n) {
                            const double h1 = a.lgfact[bc] + a.lgfact[yc];
                            float hs = (float)(h1 + log10((double)(1u + ibc)) + log10((double)(1u + iyc)) - 4);
                            if (hs > 0) {
                                if (hs >= 80.0f) hs = 80.0f - 1;
                                bcps++;
                                atomicAdd(&L.hist[(int)((double)(hs * 10.0f) + 0.5)], 1u);
                                if (better(hs, b, (int)psid, bh, bb, bp)) {
                                    if (hs != bh) beq = 0;
                                    bh = hs;
                                    bb = b;
                                    bp = (int)psid;
                                    bshp = shpk;
                                }
                                if (hs == bh) beq++;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        // ---- per-spectrum reduction of the top hit ----
        if (lane == 0) {
            L.w_best_h[wave] = bh;
            L.w_best_b[wave] = bb;
            L.w_best_p[wave] = bp;
            L.w_best_shp[wave] = bshp;
            L.w_eq[wave] = beq;
            L.w_cpsms[wave] = bcps;
        }
        if (tid == 0) atomicAdd(&L.st[ST_WINDOW], (unsigned long long)my_window);
        __syncthreads();
        float th = 0.0f;
        int tb = 0x7fffffff, tp = 0x7fffffff;
        uint32_t tshp = 0, teq = 0, cps = 0;
        for (int w = 0; w < kWaves; w++) {
            cps += L.w_cpsms[w];
            if (L.w_cpsms[w] == 0) continue;
            if (better(L.w_best_h[w], L.w_best_b[w], L.w_best_p[w], th, tb, tp)) {
                th = L.w_best_h[w];
                tb = L.w_best_b[w];
                tp = L.w_best_p[w];
                tshp = L.w_best_shp[w];
            }
        }
        for (int w = 0; w < kWaves; w++)
            if (L.w_cpsms[w] && L.w_best_h[w] == th) teq += L.w_eq[w];
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
        if (cps >= 1) {
            r.flags |= GX_PSM_HAS_HIT;
            r.hyperscore = th;
            r.psid = tp;
            r.idxoffset = (uint16_t)tb;
            r.sharedions = (uint16_t)tshp;
            r.totalions = (uint16_t)a.buckets[tb].speclen;
            r.hyp = (int32_t)((double)(th * 10.0f) + 0.5);
            if (teq > 1) r.flags |= GX_PSM_TIE;
        }
        const uint32_t need = (uint32_t)max(a.min_cpsm, 1);
        if (cps >= need) {
            double sl, ic;
            int32_t mu, beta;
            r.flags |= GX_PSM_MODELED;
            const bool fit_ok = tail_fit_block(L.hist, reinterpret_cast<double *>(L.cnt), L.fit, r.hyp, (int)cps, sl,
                                               ic, mu, beta);
            // the fit used the count tile as scratch: clear what it touched
            for (int w = tid; w < 2 * GX_HIST_BINS; w += kSearchThreads) L.cnt[w] = 0;
            if (fit_ok) {
                r.flags |= GX_PSM_FIT_OK;
                r.mu = mu;
                r.beta = beta;
                const double e = evalue_of(mu, beta, r.hyp, cps);
                r.evalue = e;
                if (e < a.expect_max) r.flags |= GX_PSM_REPORTED;
            }
        }
        if (tid == 0) {
            a.out[s] = r;
            atomicAdd(&L.st[ST_CAND], (unsigned long long)cps);
            atomicAdd(&L.st[ST_PEAKS], (unsigned long long)P);
            atomicAdd(&L.st[ST_SPECTRA], 1ull);
        }
        if (a.hist_out)
            for (int w = tid; w < GX_HIST_BINS; w += kSearchThreads)
                a.hist_out[(size_t)s * GX_HIST_BINS + w] = L.hist[w];
        __syncthreads();
    }
    atomicAdd(&L.st[ST_IONS], my_ions);
    atomicAdd(&L.st[ST_SCANNED], my_ions);
    atomicAdd(&L.st[ST_BINS], my_bins);
    __syncthreads();
    if (tid < ST_N && L.st[tid]) atomicAdd(&a.stats[tid], L.st[tid]);
}

}  // namespace

hipError_t launch_reset(uint32_t *counter, unsigned long long *stats, hipStream_t s) {
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, counter, stats);
    return hipGetLastError();
}

hipError_t launch_search(const SearchArgs &a, int n_blocks, hipStream_t s) {
    hipLaunchKernelGGL(search_kernel, dim3(n_blocks), dim3(kSearchThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sample_ids(const uint32_t *iA, uint32_t *iS, uint64_t n_ions, hipStream_t s) {
    const uint64_t n = (n_ions + kSampleStep - 1) >> kSampleShift;
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(sample_ids_kernel, dim3((unsigned)blocks), dim3(256), 0, s, iA, iS, n);
    return hipGetLastError();
}

hipError_t launch_ion_bin(const uint32_t *bA, const uint32_t *iA, uint32_t *ion_bin, uint32_t nbins,
                          uint64_t n_ions, hipStream_t s) {
    if (n_ions == 0) return hipSuccess;
    uint64_t blocks = (n_ions + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(ion_bin_kernel, dim3((unsigned)blocks), dim3(256), 0, s, bA, iA, ion_bin, nbins, n_ions);
    return hipGetLastError();
}

}  // namespace gx
