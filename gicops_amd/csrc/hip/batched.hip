// Batched fragment-ion search for gfx950: plan kernels + one persistent,
// work-stealing search launch per planning chunk + the e-value kernel. Same
// arithmetic as DSLIM_QuerySpectrum (dslim_query.cpp:749-1081) and
// ModelTailFit (expeRT.cpp:396-559); the order of work is what changes:
//
//   plan_spectra   one block per spectrum: legal peaks compacted and sorted by
//                  m/z, precursor windows of every bucket, the per-spectrum
//                  bound on shared-peak counts (which picks 8-, 16- or 32-bit
//                  LDS counters) and the number of psid tiles per window;
//   scan           item and slice-position offsets per spectrum;
//   plan_slices    one block per (bucket, spectrum): every (peak, bin) slice
//                  of iA is cut to the window and to its tiles (every such
//                  boundary bracketed within 16 ids from the 16x ion-id
//                  samples; small bins are flagged to be streamed whole), and
//                  only slices with ions left get slice positions;
//   search_spectra persistent blocks pull spectra from a counter; each
//                  (bucket, tile) item of the spectrum streams its slices'
//                  ion ids in 64-id pieces (several loads in flight per wave)
//                  and counts shared peaks with LDS atomics; entries with
//                  enough shared peaks become candidates, rescored once per
//                  spectrum for exact bc/yc/ibc/iyc; the spectrum's
//                  histogram, candidate count and top hit are written once;
//   evalue         one wave per spectrum: top hit, tail fit, e-value.
#include "batched.h"
#include "device_util.h"

namespace gx {
namespace {

constexpr int kPlanThreads = 256;
// plan_slices block: 128 threads keep more blocks in flight; 256 when the
// index's buckets are large (longer windows, more tile boundaries per slice)
constexpr int kSliceThreads = 128, kSliceThreadsLarge = 256;
constexpr int kPeakCap = kMaxPeaksLds;  // 2048 legal peaks per spectrum
constexpr int kBitonicMin = 512;        // above this many legal peaks plan_spectra sorts them bitonically
static_assert(kPeakCap <= 2048, "the bitonic sort keys hold an 11-bit peak index");
constexpr int kStatSpread = kStatRows;
constexpr int kUnroll = 6;              // stream chunks in flight per wave: 2 * kUnroll
// The ion stream addresses iA as a 32-bit byte offset from a scalar base plus
// the instruction's immediate offset 256 u (chunk u of a batch). The offsets
// are kept kStreamBias bytes high, so that offset - 256 u never goes below
// zero. An index with a bucket whose iA (with its padding) is too large for
// that is searched by the kernel's WIDE instance: ion positions, 64-bit
// addresses.
constexpr uint32_t kStreamBias = 256u * 2u * kUnroll;
constexpr uint64_t kNarrowIons = ((1ull << 32) - kStreamBias) / 4u - kIaPad;
// lanes past a round read at most 64 (2 kUnroll) - 1 ids into the padding
static_assert(64u * 2u * kUnroll <= kIaPad, "iA's padding covers a batch of dummy chunks");
constexpr int kCandLdsSpec = 640;       // candidates kept in LDS per spectrum (the rest in global scratch)
constexpr uint32_t kPsidMask = (1u << 26) - 1u;
constexpr uint32_t kWholeFlag = 0x80000000u;  // marks a whole-bin slice start (ion positions < 2^31)

__device__ __forceinline__ void stat_add(unsigned long long *stats, int which, unsigned long long v) {
    if (v) atomicAdd(&stats[(blockIdx.x % kStatSpread) * ST_N + which], v);
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kPlanThreads) plan_spectra_kernel(PlanArgs a) {
    __shared__ int32_t tq[kPeakCap], ti[kPeakCap], sq[kPeakCap];
    __shared__ uint32_t s_nlegal;
    __shared__ int s_maxwin;
    __shared__ uint32_t s_tiles[kMaxBuckets];
    __shared__ unsigned long long s_win;
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t s = blockIdx.x;
    const int dF = a.dF;
    const uint32_t off = a.qidx[s];
    const int npk = (int)(a.qidx[s + 1] - off);
    if (tid == 0) {
        s_nlegal = 0;
        s_maxwin = 0;
        s_win = 0;
    }
    __syncthreads();
    for (int base = 0; base < npk; base += kPlanThreads) {
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
        if (lane == 0 && m) wbase = atomicAdd(&s_nlegal, (uint32_t)__popcll(m));
        wbase = (uint32_t)__builtin_amdgcn_readlane((int)wbase, 0);
        if (legal) {
            const uint32_t pos = wbase + (uint32_t)__popcll(m & ((1ull << lane) - 1));
            tq[pos] = q;
            ti[pos] = in;
        }
    }
    __syncthreads();
    const int P = (int)s_nlegal;
    if (P > kBitonicMin && a.nbins <= (1 << 21)) {
        // many peaks: a bitonic sort in LDS of (bin << 11 | compacted index),
        // the index breaking ties as the rank sort below does
        int np2 = kBitonicMin;
        while (np2 < P) np2 <<= 1;
        uint32_t *sk = reinterpret_cast<uint32_t *>(sq);
        for (int k = tid; k < np2; k += kPlanThreads) sk[k] = k < P ? ((uint32_t)tq[k] << 11) | (uint32_t)k : 0xffffffffu;
        __syncthreads();
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < (np2 >> 1); i += kPlanThreads) {
                    const int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1)), hi = lo + stride;
                    const uint32_t x = sk[lo], y = sk[hi];
                    if ((x > y) == ((lo & size) == 0)) {
                        sk[lo] = y;
                        sk[hi] = x;
                    }
                }
                __syncthreads();
            }
        }
        for (int r = tid; r < P; r += kPlanThreads) {
            const uint32_t key = sk[r];
            const int32_t q = (int32_t)(key >> 11);
            sq[r] = q;
            a.lq[off + r] = q;
            a.li[off + r] = ti[key & 2047u];
        }
    } else {
        for (int k = tid; k < P; k += kPlanThreads) {
            const int32_t q = tq[k];
            int r = 0;
            for (int j = 0; j < P; j++) {
                const int32_t o = tq[j];
                r += (o < q) || (o == q && j < k);
            }
            sq[r] = q;
            a.lq[off + r] = q;
            a.li[off + r] = ti[k];
        }
    }
    __syncthreads();
    // most peaks any 2dF+1 consecutive bins hold: each fragment ion is counted
    // at most this many times per entry
    int mw = 0;
    for (int k = tid; k < P; k += kPlanThreads) mw = max(mw, upper_bound_lds(sq, P, sq[k] + 2 * dF) - k);
    mw = wave_max_i32(mw);
    if (lane == 0) atomicMax(&s_maxwin, mw);
    __syncthreads();
    const int maxwin = s_maxwin;
    const float prec = a.prec[s];
    const double dM = a.dm ? (double)a.dm[s] : a.dM;
    for (int b = tid; b < a.n_buckets; b += kPlanThreads) {
        const DevBucket bk = a.buckets[b];
        int32_t lo, hi;
        const bool ok = precursor_window(bk, prec, dM, lo, hi);
        const uint64_t bound = (uint64_t)bk.speclen * (uint64_t)maxwin;
        const uint32_t pack = bound <= 255 ? 4u : (bound <= 65535 ? 2u : 1u);
        const uint64_t T = (uint64_t)pack * (uint64_t)a.tile_words;
        uint32_t tiles = 0;
        if (ok) {
            atomicAdd(&s_win, (unsigned long long)(hi - lo + 1));
            if (P > 0) tiles = (uint32_t)(((uint64_t)(hi - lo + 1) + T - 1) / T);
        }
        const uint32_t pi = s * (uint32_t)a.n_buckets + b;
        a.pair_lo[pi] = ok ? lo : 0;
        a.pair_hi[pi] = ok ? hi : -1;
        a.pair_tiles[pi] = tiles;
        a.pair_pack[pi] = pack;
        s_tiles[b] = tiles;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t it = 0;
        uint64_t pw = 0;
        const uint64_t nsl = (uint64_t)P * (uint64_t)(2 * dF + 1);
        for (int b = 0; b < a.n_buckets; b++) {
            const uint32_t pi = s * (uint32_t)a.n_buckets + b;
            a.pair_pos_off[pi] = (uint32_t)pw;
            it += s_tiles[b];
            if (s_tiles[b]) pw += nsl * (uint64_t)(s_tiles[b] + 1);
        }
        a.n_legal[s] = (uint32_t)P;
        a.n_items[s] = it;
        a.n_pos[s] = pw;
        stat_add(a.stats, ST_WINDOW, s_win);
        stat_add(a.stats, ST_PEAKS, (unsigned long long)P);
        stat_add(a.stats, ST_SPECTRA, 1ull);
    }
}

// exclusive scans of the per-spectrum item and slice-position counts
__global__ void __launch_bounds__(1024) scan_kernel(PlanArgs a) {
    __shared__ uint32_t wsum_i[16];
    __shared__ unsigned long long wsum_p[16];
    __shared__ uint32_t carry_i;
    __shared__ unsigned long long carry_p;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        carry_i = 0;
        carry_p = 0;
    }
    __syncthreads();
    for (uint32_t base = 0; base < a.n_spectra; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t vi = i < a.n_spectra ? a.n_items[i] : 0u;
        const unsigned long long vp = i < a.n_spectra ? a.n_pos[i] : 0ull;
        // inclusive wave scans
        uint32_t si = vi;
        unsigned long long sp = vp;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t ti = __shfl_up(si, o, 64);
            const unsigned long long tp = __shfl_up(sp, o, 64);
            if (lane >= o) {
                si += ti;
                sp += tp;
            }
        }
        if (lane == 63) {
            wsum_i[wave] = si;
            wsum_p[wave] = sp;
        }
        __syncthreads();
        uint32_t wi = 0;
        unsigned long long wp = 0;
        for (int w = 0; w < wave; w++) {
            wi += wsum_i[w];
            wp += wsum_p[w];
        }
        if (i < a.n_spectra) {
            a.item_off[i] = carry_i + wi + si - vi;
            a.pos_off[i] = carry_p + wp + sp - vp;
        }
        __syncthreads();
        if (tid == 1023) {
            carry_i += wi + si;
            carry_p += wp + sp;
        }
        __syncthreads();
    }
    if (tid == 0) {
        a.item_off[a.n_spectra] = carry_i;
        a.pos_off[a.n_spectra] = carry_p;
    }
}

// brackets of the first positions in [st, en) with iA[p] >= x0 and >= x1 from
// the bin's samples iS (every kSampleStep-th id) alone: p0 in [lo0, hi0], p1 in
// [lo1, hi1], each with hi - lo < kSampleStep. The two searches' loads are
// independent, so each step waits for one latency instead of two. A search
// turned off (do0 / do1 false) reads nothing; its bracket is then unused.
__device__ __forceinline__ void bracket_sampled2(const DevBucket &bk, uint32_t st, uint32_t en, uint32_t x0, uint32_t x1,
                                                 uint32_t &lo0, uint32_t &hi0, uint32_t &lo1, uint32_t &hi1,
                                                 bool do0 = true, bool do1 = true) {
    lo0 = st;
    hi0 = en;
    lo1 = st;
    hi1 = en;
    const uint32_t k0 = (st + kSampleStep - 1u) >> kSampleShift;
    if (en > 0 && k0 <= ((en - 1u) >> kSampleShift)) {
        const uint32_t k1 = (en - 1u) >> kSampleShift;
        const auto is = gptr(bk.iS);
        uint32_t b0 = k0, b1 = k0, n0 = do0 ? k1 - k0 + 1u : 0u, n1 = do1 ? k1 - k0 + 1u : 0u;
        while (n0 > 0 || n1 > 0) {
            const uint32_t h0 = n0 >> 1, h1 = n1 >> 1;
            const uint32_t v0 = n0 > 0 ? is[b0 + h0] : 0u, v1 = n1 > 0 ? is[b1 + h1] : 0u;
            if (n0 > 0) {
                if (v0 < x0) {
                    b0 += h0 + 1u;
                    n0 -= h0 + 1u;
                } else {
                    n0 = h0;
                }
            }
            if (n1 > 0) {
                if (v1 < x1) {
                    b1 += h1 + 1u;
                    n1 -= h1 + 1u;
                } else {
                    n1 = h1;
                }
            }
        }
        if (b0 <= k1) hi0 = min(hi0, b0 << kSampleShift);
        if (b0 > k0) lo0 = max(lo0, ((b0 - 1u) << kSampleShift) + 1u);
        if (b1 <= k1) hi1 = min(hi1, b1 << kSampleShift);
        if (b1 > k0) lo1 = max(lo1, ((b1 - 1u) << kSampleShift) + 1u);
    }
}

// one block per (bucket, spectrum) of the chunk, bucket-major and by precursor
// mass within a bucket: the blocks in flight read one bucket's bA and iA with
// similar windows, which stay in L2
template <int NT>
__global__ void __launch_bounds__(NT) plan_slices_kernel(PlanArgs a) {
    __shared__ unsigned long long s_st[2];
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t nb = (uint32_t)a.n_buckets;
    const uint32_t nspec = a.chunk_s1 - a.chunk_s0;
    const uint32_t s = a.plan_order[a.chunk_s0 + blockIdx.x % nspec];
    const uint32_t b = blockIdx.x / nspec;
    const uint32_t pi = s * nb + b;
    const uint32_t tiles = a.pair_tiles[pi];
    if (tiles == 0) {
        if (tid == 0) a.pair_nsl[pi] = 0;
        return;
    }
    if (tid < 2) s_st[tid] = 0;
    __syncthreads();
    const int dF = a.dF, nb5 = 2 * dF + 1;
    const int P = (int)a.n_legal[s];
    const int nsl = P * nb5;
    const DevBucket bk = a.buckets[b];
    const uint32_t speclen = bk.speclen;
    const uint32_t lo = (uint32_t)a.pair_lo[pi];
    const uint32_t end_id = (uint32_t)(a.pair_hi[pi] + 1) * speclen;
    const bool from_start = lo == 0, to_end = (uint32_t)(a.pair_hi[pi] + 1) == bk.n_entries;
    const uint32_t T = a.pair_pack[pi] * (uint32_t)a.tile_words;  // psids per tile
    const uint64_t base = (a.pos_off[s] - a.pos_off[a.chunk_s0]) + a.pair_pos_off[pi];
    const int32_t *lq = a.lq + a.qidx[s];
    unsigned long long bins = 0;
    // tile-major rows: row t holds boundary t of every slice, so that the
    // search kernel's item for tile t reads rows t and t+1 coalesced. Row 0 and
    // row `tiles` bracket the window's ends (the first ion at or below it, the
    // end at or above it); an interior row t is the lower end of the bracket
    // of the first ion with id >= (lo + t*T)*speclen, which lies less than
    // kSampleStep ions past it. Only non-empty slices get a column (in any
    // order: the search does not depend on it); their number goes to pair_nsl.
    uint32_t *out = a.pos + base;
    const unsigned long long below = lane ? ((1ull << (uint32_t)lane) - 1ull) : 0ull;
    for (int j0 = 0; j0 < nsl; j0 += NT) {
        const int j = j0 + tid;
        const bool act = j < nsl;
        uint32_t st = 0, en = 0, c0 = 0, cz = 0;
        bool whole = false;
        if (act) {
            const int k = j / nb5;
            const uint32_t bin = (uint32_t)(lq[k] - dF + (j - k * nb5));
            {
                // the bin's start and end in one 8-byte load at a 4-byte aligned
                // address (gfx9 global loads run in unaligned-access mode)
                typedef unsigned long long __attribute__((aligned(4))) u64_a4;
                const unsigned long long se = *(const gmem<u64_a4> *)(gptr(bk.bA) + bin);
                st = (uint32_t)se;
                en = (uint32_t)(se >> 32);
            }
            bins++;
            // bins of at most whole_bin ions are streamed whole by the search
            // kernel (every tile filters by id). Larger ones get rows bracketed
            // from the samples: the search kernel reads fewer than kSampleStep
            // ids past each end of a tile's part of the slice and drops the ids
            // outside the tile. A window from the bucket's first entry starts at
            // the bin's start, one to its last entry ends at the bin's end.
            whole = en - st <= a.whole_bin;
            if (!whole) {
                uint32_t l0, h0, l1, h1;
                bracket_sampled2(bk, st, en, lo * speclen, end_id, l0, h0, l1, h1, !from_start, !to_end);
                c0 = from_start ? st : l0;
                cz = to_end ? en : h1;
            }
        }
        const bool keep = act && (whole ? en > st : cz > c0);
        // the column of each kept slice: a wave-aggregated LDS counter
        const unsigned long long km = __ballot(keep);
        uint32_t wb0 = 0;
        if (lane == 0 && km) wb0 = (uint32_t)atomicAdd(&s_st[0], (unsigned long long)__popcll(km));
        wb0 = (uint32_t)__builtin_amdgcn_readlane((int)wb0, 0);
        if (!keep) continue;
        const uint32_t jc = wb0 + (uint32_t)__popcll(km & below);
        if (whole) {
            for (uint32_t t = 0; t < tiles; t++) out[(uint64_t)t * nsl + jc] = st | kWholeFlag;
            out[(uint64_t)tiles * nsl + jc] = en;
            continue;
        }
        out[jc] = c0;
        out[(uint64_t)tiles * nsl + jc] = cz;
        // interior boundaries two at a time, each pair bracketed from the
        // previous one (boundaries ascend with t, and so do their brackets)
        uint32_t cur = c0;
        for (uint32_t t = 1; t < tiles; t += 2) {
            const uint32_t id0 = (lo + t * T) * speclen;
            if (t + 1 < tiles) {
                const uint32_t id1 = (lo + (t + 1) * T) * speclen;
                uint32_t b0 = cur, b1 = cur;
                if (cz > cur) {
                    uint32_t h0, h1;
                    bracket_sampled2(bk, cur, cz, id0, id1, b0, h0, b1, h1);
                }
                out[(uint64_t)t * nsl + jc] = b0;
                out[(uint64_t)(t + 1) * nsl + jc] = b1;
                cur = b1;
            } else {
                if (cz > cur) {
                    uint32_t l0, h0, l1, h1;
                    bracket_sampled2(bk, cur, cz, id0, id0, l0, h0, l1, h1);
                    cur = l0;
                }
                out[(uint64_t)t * nsl + jc] = cur;
            }
        }
    }
    bins = wave_sum_u64(bins);
    if (lane == 0) atomicAdd(&s_st[1], bins);
    __syncthreads();
    if (tid == 0) {
        stat_add(a.stats, ST_BINS, s_st[1]);
        a.pair_nsl[pi] = (uint32_t)s_st[0];
    }
}

// a slice starting at flat index f marks its 64-id chunk's mask: lane k >= 1 of
// the chunk sets bit k - 1 and lane 0 sets bit 63, so that a lane's count of
// slice starts at or before it is mbcnt over the mask plus bit 63, and the
// mask's popcount is still the chunk's number of slice starts
__device__ __forceinline__ unsigned long long chunk_bit(uint32_t f) {
    const uint32_t k = f & 63u;
    return k ? 1ull << (k - 1u) : 1ull << 63;
}

// ---------------------------------------------------------------------------
template <int THREADS, int TILE_WORDS, int SL>
struct SpecLds {
    // non-empty slices of the tile: first ion minus the slice's flat offset, as
    // a biased byte offset (an ion position in the WIDE instance); entry ne is
    // the virtual slice past the last one and ne + 1 the round's dummy slot,
    // both of which lead into iA's padding. First in LDS, so that its byte
    // addresses (chunk base + 4 * lane's slice) need no separate base register.
    uint32_t s_a[SL + 2];
    alignas(16) uint32_t cnt[TILE_WORDS];  // count tile; in pass 2 its accumulators, candidate words, peaks and cell table
    uint32_t s_flat[SL + 1];   // exclusive prefix of their lengths
    uint32_t dummy[64];        // per-lane word that out-of-tile lanes add 0 to
    alignas(16) uint32_t wsum[THREADS / 64];   // per-wave partial sums (read back as 16-byte vectors)
    alignas(16) uint32_t wsum2[THREADS / 64];
    uint16_t cj0[THREADS];     // per chunk of the round: 4 x (index of the slice at its lane 0)
    unsigned long long cmask[THREADS];  // lanes of each chunk where a slice starts
    uint32_t ne;
    uint32_t cand[kCandLdsSpec];  // the spectrum's candidates, b << 26 | psid (the rest in global scratch)
    uint32_t hist[GX_HIST_BINS];
    unsigned long long w_key[THREADS / 64];
    uint32_t w_eq[THREADS / 64];
    uint32_t w_shp[THREADS / 64];
    uint32_t w_cps[THREADS / 64];
    uint32_t w_scan[THREADS / 64];  // per wave: the unit's consumed ion ids and self-check errors
    uint32_t w_err[THREADS / 64];
    uint32_t b_speclen[kMaxBuckets];
    // the candidates gathered since the last pass 2 come in runs, one per
    // bucket (items go bucket by bucket): run k starts at seg_start[k]
    uint32_t seg_start[kMaxBuckets];
    uint8_t seg_b[kMaxBuckets];
    uint32_t n_seg;
    uint32_t next;             // prefetched next ticket
    uint32_t n_cand;
};

// merge two top-hit summaries (key, count of candidates equal to its
// hyperscore, shared ions of the key's hit)
__device__ __forceinline__ void merge_best(unsigned long long &k, uint32_t &e, uint32_t &shp, unsigned long long k2,
                                           uint32_t e2, uint32_t shp2) {
    const uint32_t h1 = (uint32_t)(k >> 32), h2 = (uint32_t)(k2 >> 32);
    if (h1 == h2) {
        e += e2;
        if (k2 > k) {
            k = k2;
            shp = shp2;
        }
    } else if (h2 > h1) {
        k = k2;
        e = e2;
        shp = shp2;
    }
}

// Pass 2 over the spectrum's candidates gathered so far: exact bc/yc/ibc/iyc
// of each candidate from its own fragment bins (ion_bin) against the sorted
// peaks, flattened over (candidate, fragment) run by run (a run's candidates
// share a bucket, so its stride is that bucket's speclen) and summed with LDS
// atomics; then hyperscore, histogram and top hit. Peaks near a bin are found
// through a cell table (first peak of each 2^cell_shift-bin cell and its bin)
// kept in the tail of the count tile: most bins have no peak within dF and are
// ruled out by one table read; the others walk the sorted peaks, the lane's U
// tasks one peak per step together.
template <int THREADS, int TILE_WORDS, int SL>
__device__ void rescore_candidates(const ItemArgs &a, SpecLds<THREADS, TILE_WORDS, SL> &L, int P, uint32_t qoff,
                                   uint32_t *g_cand, uint32_t &n_err) {
    // the tile during pass 2 (it is clear between items): per candidate of a
    // round three accumulators and its word; the spectrum's peaks (copied here
    // per call) and a sentinel; the cell table
    constexpr int kAccWords = TILE_WORDS - kPeakIdxWords - 2 * (kPeakCap + 1);
    static_assert(kAccWords > 0, "count tile too small for pass 2");
    constexpr int CB = kAccWords / 4;  // candidates per round (3 accumulators and the word each)
    const int tid = threadIdx.x;
    // the peaks as (m/z bin, intensity) pairs, P first, a sentinel (bin past every
    // fragment) at P
    int2 *pk = reinterpret_cast<int2 *>(L.cnt + kAccWords);
    for (int k = tid; k <= P; k += THREADS)
        pk[k] = k < P ? make_int2(a.lq[qoff + k], a.li[qoff + k]) : make_int2(0x7fffffff, 0);
    lds_barrier();
    // this call's top hit and candidate count; merged into the wave's slot at
    // the end, so that nothing of it stays live in registers between calls
    unsigned long long bkey = 0;
    uint32_t beq = 0, bshp = 0, bcps = 0;
    const int dF = a.dF;
    const uint32_t ncand = L.n_cand;
    // per 2^cs-bin cell: the first peak at or after the cell's start, as its index
    // << 16 | its bin's offset from the start (0xffff: 65535 bins or more, or none)
    uint32_t *cell_tab = L.cnt + (TILE_WORDS - kPeakIdxWords);
    const uint32_t cs = a.cell_shift;
    const uint32_t ncell = ((uint32_t)a.nbins >> cs) + 1;
    for (uint32_t c = tid; c < ncell; c += THREADS) {
        int lo = 0, n = P;
        const int32_t v = (int32_t)(c << cs);
        while (n > 0) {
            const int h = n >> 1;
            if (pk[lo + h].x < v) {
                lo += h + 1;
                n -= h + 1;
            } else {
                n = h;
            }
        }
        const uint32_t off = lo < P ? (uint32_t)min(pk[lo].x - v, 0xffff) : 0xffffu;
        cell_tab[c] = ((uint32_t)lo << 16) | off;
    }
    uint32_t *acc_n = L.cnt, *acc_ib = L.cnt + CB, *acc_iy = L.cnt + 2 * CB;
    const uint32_t nseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.n_seg);
    lds_barrier();
    uint32_t *cw = L.cnt + 3 * CB;
    for (uint32_t c0 = 0; c0 < ncand; c0 += CB) {
        const uint32_t nc = min((uint32_t)CB, ncand - c0);
        // the round's candidate words, from LDS or global scratch
        for (uint32_t c = tid; c < nc; c += THREADS) {
            const uint32_t ci = c0 + c;
            cw[c] = ci < (uint32_t)kCandLdsSpec ? L.cand[ci] : gptr(g_cand)[ci - kCandLdsSpec];
        }
        lds_barrier();
        for (uint32_t k = 0; k < nseg; k++) {
            // the run's candidates in this round, (candidate, fragment) tasks
            // with the run's bucket's speclen as stride
            const uint32_t sb = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.seg_start[k]);
            const uint32_t se =
                k + 1 < nseg ? (uint32_t)__builtin_amdgcn_readfirstlane((int)L.seg_start[k + 1]) : ncand;
            const uint32_t s0 = max(sb, c0), s1 = min(se, c0 + nc);
            if (s1 <= s0) continue;
            const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.seg_b[k]);
            const DevBucket &bk = a.buckets[b];
            const uint32_t spl = bk.speclen, dm = bk.div_m, dsh = bk.div_s;
            const auto ib = gptr(bk.ion_bin);
            const uint32_t nrun = s1 - s0, tasks = nrun * spl;  // < 2^31: the division below is exact
            constexpr int U = 4;
            for (uint32_t tb = tid; tb < tasks; tb += U * THREADS) {
                int32_t x[U];
                uint32_t cl[U], rr[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint32_t tk = tb + (uint32_t)u * THREADS;
                    const uint32_t c = __umulhi(tk, dm) >> dsh;
                    const uint32_t r = tk - c * spl;
                    const uint32_t ci = s0 + min(c, nrun - 1);
                    const uint32_t cd = cw[ci - c0];
                    const bool ok = tk < tasks;
                    // every candidate of the run is of its bucket (a self-check)
                    n_err += ok && (cd >> 26) != b ? 1u : 0u;
                    cl[u] = ok ? ci - c0 : 0xffffffffu;
                    rr[u] = r;
                    // every task loads (a task past the run reads bin 0 of the bucket,
                    // unused: cl marks it), so no select waits for the load
                    x[u] = (int32_t)ib[ok ? (size_t)(cd & kPsidMask) * spl + r : 0];
                }
                // peaks within +-dF of each fragment bin: the cell table rules out
                // most bins at once; the rest walk the peaks from the cell's first,
                // all U tasks one peak per step (a finished task reads the sentinel)
                uint32_t kk[U], n[U], sm[U];
                bool any = false;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int32_t lo = x[u] - dF;
                    const uint32_t c = (uint32_t)max(lo, 0) >> cs;
                    const uint32_t e = cell_tab[c];
                    const bool walk = cl[u] != 0xffffffffu && (int32_t)((c << cs) + (e & 0xffffu)) <= x[u] + dF;
                    kk[u] = walk ? e >> 16 : (uint32_t)P;
                    n[u] = 0;
                    sm[u] = 0;
                    any |= walk;
                }
                while (__ballot(any)) {
                    any = false;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int2 q = pk[kk[u]];
                        const bool le = q.x <= x[u] + dF;
                        const bool in = le && q.x >= x[u] - dF;
                        n[u] += in ? 1u : 0u;
                        sm[u] += in ? (uint32_t)q.y : 0u;
                        kk[u] = le ? kk[u] + 1u : (uint32_t)P;
                        any |= le;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (n[u]) {
                        const bool isy = rr[u] >= (spl >> 1);
                        atomicAdd(&acc_n[cl[u]], isy ? (n[u] << 16) : n[u]);
                        atomicAdd(isy ? &acc_iy[cl[u]] : &acc_ib[cl[u]], sm[u]);
                    }
                }
            }
        }
        lds_barrier();
        for (uint32_t c = tid; c < nc; c += THREADS) {
            const uint32_t v = acc_n[c];
            const uint32_t bc = v & 0xffffu, yc = v >> 16;
            const uint32_t ibc = acc_ib[c], iyc = acc_iy[c];
            acc_n[c] = 0;
            acc_ib[c] = 0;
            acc_iy[c] = 0;
            const uint32_t cd = cw[c];
            cw[c] = 0;
            const uint32_t shpk = bc + yc;
            // pass 1 counted at least min_shp matches for this entry; the exact
            // recount must agree
            n_err += shpk < (uint32_t)max(a.min_shp, 1) ? 1u : 0u;
            if (shpk >= (uint32_t)a.min_shp && (int)bc < a.lgfact_n && (int)yc < a.lgfact_n) {
                float hs = hyperscore(bc, yc, ibc, iyc, a.lgfact);
                if (hs > 0) {
                    if (hs >= 80.0f) hs = 80.0f - 1;
                    bcps++;
                    atomicAdd(&L.hist[hyp_bin(hs)], 1u);
                    merge_best(bkey, beq, bshp, hit_key(hs, cd >> 26, cd & kPsidMask), 1u, shpk);
                }
            }
        }
        lds_barrier();
    }
    for (uint32_t c = tid; c < ncell; c += THREADS) cell_tab[c] = 0;
    for (int k = tid; k <= P; k += THREADS) pk[k] = make_int2(0, 0);
    if (tid == 0) {
        L.n_cand = 0;
        L.n_seg = 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long k2 = __shfl_xor(bkey, o, 64);
        const uint32_t e2 = __shfl_xor(beq, o, 64), s2 = __shfl_xor(bshp, o, 64);
        merge_best(bkey, beq, bshp, k2, e2, s2);
    }
    bcps = wave_sum_u32(bcps);
    if ((tid & 63) == 0) {
        const int wave = tid >> 6;
        merge_best(L.w_key[wave], L.w_eq[wave], L.w_shp[wave], bkey, beq, bshp);
        L.w_cps[wave] += bcps;
    }
    lds_barrier();
}

// Persistent search: blocks pull whole spectra of the chunk from a counter.
// For each of the spectrum's (bucket, tile) items the block streams the
// tile's slices of ion ids (pass 1: shared-peak counts in the packed LDS
// tile), then collects the tile's entries with count >= min_shp as
// candidates. Pass 2 rescans all candidates of the spectrum at once; the
// histogram, candidate count and top hit are written once per spectrum.
template <int THREADS, int TILE_WORDS, int SL, int WAVES_PER_SIMD, bool WIDE>
__global__ void __launch_bounds__(THREADS, WAVES_PER_SIMD) search_spectra_kernel(ItemArgs a) {
    constexpr int NW = THREADS / 64;
    __shared__ SpecLds<THREADS, TILE_WORDS, SL> L;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int dF = a.dF, nb5 = 2 * dF + 1;
    const uint32_t min_cnt = (uint32_t)max(a.min_shp, 1);
    const uint32_t nbk = (uint32_t)a.n_buckets;
    uint32_t *g_cand = a.cand_scratch + (size_t)blockIdx.x * a.tile_entries_max;
    const uint32_t cand_cap = (uint32_t)kCandLdsSpec + (uint32_t)a.tile_entries_max;
    for (int w = tid; w < TILE_WORDS; w += THREADS) L.cnt[w] = 0;
    for (int w = tid; w < GX_HIST_BINS; w += THREADS) L.hist[w] = 0;
    L.cmask[tid] = 0;
    // never changed (out-of-tile lanes add 0). Such a lane's check reads the
    // field its (meaningless) entry index selects, at bit 0, 8, 16 or 24 with
    // width 8, 16 or 31; every byte of the word is the low byte of
    // (min_cnt - 1) ^ 1, so the low byte of any such field differs from that of
    // min_cnt - 1 and no such field equals it.
    if (tid < 64) L.dummy[tid] = (((min_cnt - 1u) ^ 1u) & 0xffu) * 0x01010101u;
    for (int b = tid; b < a.n_buckets; b += THREADS) L.b_speclen[b] = a.buckets[b].speclen;
    const uint32_t nunits = a.n_units;
    if (tid == 0) {
        L.next = atomicAdd(a.work_counter, 1u);
        L.n_cand = 0;
    }

    for (;;) {
        lds_barrier();
        const uint32_t ticket = L.next;
        if (ticket >= nunits) break;
        const uint32_t s = a.unit_s[a.unit0 + ticket];
        const uint32_t br = a.unit_br[a.unit0 + ticket];
        const uint32_t ub0 = br & 0xffu, ub1 = (br >> 8) & 0xffu, slot = br >> 16;
        lds_barrier();
        // fetch the next ticket now; its latency overlaps this spectrum
        if (tid == 0) L.next = atomicAdd(a.work_counter, 1u);
        const int P = (int)a.n_legal[s];
        const uint32_t qoff = a.qidx[s];  // the legal peaks, read by pass 2
        const int nsl = P * nb5;
        // the unit's top hit and candidate count, per wave (merged by every
        // pass-2 call; ordered before the first one by the barriers between)
        if (lane == 0) {
            L.w_key[wave] = 0;
            L.w_eq[wave] = 0;
            L.w_shp[wave] = 0;
            L.w_cps[wave] = 0;
        }
        if (tid == 0) L.n_seg = 0;  // (the unit's candidate list is empty)
        // quads of the count tile left to clear by the last item (cleared in the
        // next item's setup, or before pass 2, which uses the tile's words)
        uint32_t dirty_q = 0;
        auto clear_tile = [&]() {
            uint4 *cq = reinterpret_cast<uint4 *>(L.cnt);
            for (uint32_t q = tid; q < dirty_q; q += THREADS) cq[q] = make_uint4(0, 0, 0, 0);
            dirty_q = 0;
        };
        uint32_t n_scanned = 0, n_errors = 0;
        const uint64_t pos_s = a.pos_off[s] - a.pos_off[a.chunk_s0];
        // the spectrum's per-bucket plan, bucket b in lane b of every wave (one
        // round trip per spectrum instead of a chain per item)
        uint32_t w_tiles = 0, w_pack = 0, w_poff = 0, w_nsl = 0;
        int32_t w_lo = 0, w_hi = -1;
        if ((uint32_t)lane < nbk) {
            const uint32_t pi = s * nbk + (uint32_t)lane;
            w_tiles = a.pair_tiles[pi];
            w_pack = a.pair_pack[pi];
            w_poff = a.pair_pos_off[pi];
            w_nsl = a.pair_nsl[pi];
            w_lo = a.pair_lo[pi];
            w_hi = a.pair_hi[pi];
        }
        // the slice words of the next (bucket, tile, slice chunk) step are
        // loaded one step ahead, so that their latency overlaps this step. A
        // pair's rows have nsl columns, of which the first w_nsl (its non-empty
        // slices) are filled.
        uint32_t nx_pt = 0, nx_pt1 = 0, nx_pe = 0;
        auto fetch = [&](uint32_t fb, uint32_t ft, int fc0) {
            const uint32_t ftiles = (uint32_t)__builtin_amdgcn_readlane((int)w_tiles, (int)fb);
            const uint32_t fpoff = (uint32_t)__builtin_amdgcn_readlane((int)w_poff, (int)fb);
            const int fnsl = __builtin_amdgcn_readlane((int)w_nsl, (int)fb);
            const uint32_t *fpos = a.pos + pos_s + fpoff;
            if (tid < min(SL, fnsl - fc0)) {
                const uint32_t j = (uint32_t)(fc0 + tid);
                nx_pt = fpos[(uint64_t)ft * nsl + j];
                nx_pt1 = fpos[(uint64_t)(ft + 1) * nsl + j];
                nx_pe = fpos[(uint64_t)ftiles * nsl + j];
            }
        };
        const unsigned long long todo0 = __ballot(w_tiles > 0 && (uint32_t)lane >= ub0 && (uint32_t)lane < ub1);
        if (todo0) fetch((uint32_t)__builtin_ctzll(todo0), 0, 0);
        for (unsigned long long todo = todo0; todo; todo &= todo - 1) {
            const uint32_t b = (uint32_t)__builtin_ctzll(todo);
            const int nslb = __builtin_amdgcn_readlane((int)w_nsl, (int)b);  // non-empty slices
            if (nslb == 0) {  // nothing of the window in this spectrum's bins
                const unsigned long long rest = todo & (todo - 1);
                if (rest) fetch((uint32_t)__builtin_ctzll(rest), 0, 0);
                continue;
            }
            const uint32_t tiles = (uint32_t)__builtin_amdgcn_readlane((int)w_tiles, (int)b);
            const int32_t lo = __builtin_amdgcn_readlane(w_lo, (int)b), hi = __builtin_amdgcn_readlane(w_hi, (int)b);
            const uint32_t pack = (uint32_t)__builtin_amdgcn_readlane((int)w_pack, (int)b);
            const uint32_t lp = (uint32_t)__builtin_ctz(pack);  // log2 entries per word (pack is 1, 2 or 4)
            const uint32_t fbits = 32u >> lp;                             // counter width
            const uint32_t fbw = fbits == 32 ? 31u : fbits;               // field width for ubfe (counts < 2^31)
            const uint32_t fsh = 5u - lp;                                 // log2 of the counter width
            const uint32_t T = ((uint32_t)TILE_WORDS / (uint32_t)a.split) << lp;
            const DevBucket bk = a.buckets[b];
            const uint32_t speclen = bk.speclen;
            const auto ia = gptr(bk.iA);
            const uint32_t sent = bk.n_entries * speclen;  // iA's padding starts here
            // slice bases as biased byte offsets (loads from a scalar base, a 32-bit
            // offset and the immediate), or as ion positions in the WIDE instance
            const uint32_t so_shift = WIDE ? 0u : 2u, so_bias = WIDE ? 0u : kStreamBias;
            const auto iab = (const __attribute__((address_space(1))) char *)ia - kStreamBias;
            const uint32_t dm = bk.div_m, dsh = bk.div_s;
            const uint32_t btag = b << 26;
            for (uint32_t t = 0; t < tiles; t++) {
                const int32_t tlo = lo + (int32_t)(t * T);
                const int32_t thi = min(hi, tlo + (int32_t)T - 1);
                const uint32_t nent = (uint32_t)(thi - tlo + 1);
                const uint32_t tlo_id = (uint32_t)tlo * speclen;
                const uint32_t span_id = nent * speclen;
                // the last item's stream is done (every wave's increments and
                // candidate appends); room for every entry of this tile as a
                // candidate? An LDS-only barrier, so that the next step's slice
                // words loaded by the last step stay in flight; a full one before
                // pass 2, which reads candidates that pass 1 wrote to global scratch
                lds_barrier();
                if (L.n_cand + nent > cand_cap) {
                    clear_tile();
                    __syncthreads();
                    rescore_candidates<THREADS, TILE_WORDS, SL>(a, L, P, qoff, g_cand, n_errors);
                }
                // a new run of candidates where the bucket changes (ordered before
                // pass 1's appends by its barriers, and before pass 2 by the
                // barrier above its call)
                if (tid == 0 && (L.n_seg == 0 || L.seg_b[L.n_seg - 1] != (uint8_t)b)) {
                    L.seg_start[L.n_seg] = L.n_cand;
                    L.seg_b[L.n_seg] = (uint8_t)b;
                    L.n_seg++;
                }
                // ---- pass 1: shared-peak counts of every entry in the tile ----
                for (int c0 = 0; c0 < nslb; c0 += SL) {
                    const int ns = min(SL, nslb - c0);
                    // slice ranges of this tile; non-empty ones compacted in order
                    static_assert(SL == THREADS, "one slice per thread");
                    uint32_t at = 0, len = 0;
                    // this step's words (rows t, t+1 and the last of its slices),
                    // then the next step's load
                    const uint32_t pt = nx_pt, pt1 = nx_pt1, pe = nx_pe;
                    if (c0 + SL < nslb) {
                        fetch(b, t, c0 + SL);
                    } else if (t + 1 < tiles) {
                        fetch(b, t + 1, 0);
                    } else {
                        const unsigned long long rest = todo & (todo - 1);
                        if (rest) fetch((uint32_t)__builtin_ctzll(rest), 0, 0);
                    }
                    if (tid < ns) {
                        // a small bin is [start | flag, end in the last row), a larger
                        // one has the lower ends of its tile boundaries' brackets in rows t
                        // and t+1: the tile's part ends before row t+1 + kSampleStep (the
                        // window's end in the last row)
                        if (pt & kWholeFlag) {  // small bin: the whole bin, filtered by id below
                            at = pt & ~kWholeFlag;
                            len = pe - at;
                        } else {
                            at = pt;
                            len = min(pt1 + kSampleStep, pe) - at;
                        }
                    }
                    const uint32_t fl = len ? 1u : 0u;
                    const uint32_t il = wave_incl_scan(len, lane), ifl = wave_incl_scan(fl, lane);
                    if (lane == 63) {
                        L.wsum[wave] = il;
                        L.wsum2[wave] = ifl;
                    }
                    lds_barrier();
                    if (dirty_q) clear_tile();  // (ordered before the stream by the barrier below)
                    uint32_t bl = 0, bf = 0;
                    static_assert(NW % 4 == 0, "per-wave sums read four at a time");
#pragma unroll
                    for (int g = 0; g < NW / 4; g++) {
                        const uint4 x = reinterpret_cast<const uint4 *>(L.wsum)[g];
                        const uint4 y = reinterpret_cast<const uint4 *>(L.wsum2)[g];
                        bl += (4 * g + 0 < wave ? x.x : 0u) + (4 * g + 1 < wave ? x.y : 0u) +
                              (4 * g + 2 < wave ? x.z : 0u) + (4 * g + 3 < wave ? x.w : 0u);
                        bf += (4 * g + 0 < wave ? y.x : 0u) + (4 * g + 1 < wave ? y.y : 0u) +
                              (4 * g + 2 < wave ? y.z : 0u) + (4 * g + 3 < wave ? y.w : 0u);
                    }
                    // the tile's ions, flattened over the non-empty slices, in rounds
                    // of THREADS 64-id chunks. Each slice sets the bit of its first
                    // lane in its chunk's mask (for the first round, right here); a
                    // scan of the masks' popcounts gives the number of slices
                    // starting before each chunk. The masks are clear on entry.
                    if (fl) {
                        const uint32_t f = bl + il - len;
                        // ion position = s_a + flat index (mod 2^32), as bytes + bias or not
                        L.s_a[bf + ifl - 1] = ((at - f) << so_shift) + so_bias;
                        L.s_flat[bf + ifl - 1] = f;
                        if ((f >> 6) < (uint32_t)THREADS) atomicOr(&L.cmask[f >> 6], chunk_bit(f));
                    }
                    if (tid == THREADS - 1) {
                        // past the last slice: a virtual slice starting at N whose
                        // lanes read iA's padding (its chunk bit is set when N is not
                        // a multiple of 64, i.e. when a real chunk holds it)
                        const uint32_t n_fl = bl + il;
                        L.s_flat[bf + ifl] = n_fl;
                        L.ne = bf + ifl;
                        L.s_a[bf + ifl] = ((sent - n_fl) << so_shift) + so_bias;
                        if ((n_fl & 63u) && (n_fl >> 6) < (uint32_t)THREADS) atomicOr(&L.cmask[n_fl >> 6], chunk_bit(n_fl));
                    }
                    lds_barrier();
                    const uint32_t ne = L.ne;
                    const uint32_t N = L.s_flat[ne];
                    const uint32_t nch = (N + 63) >> 6;
                    uint32_t base = 0;  // slices starting before the round
                    for (uint32_t r0 = 0; r0 < nch; r0 += THREADS) {
                        const uint32_t nr = min((uint32_t)THREADS, nch - r0);
                        if (r0 > 0) {  // later rounds: the previous round's readers cleared the masks
                            lds_barrier();
                            if ((uint32_t)tid < ne) {
                                const uint32_t f = L.s_flat[tid];
                                const uint32_t c = (f >> 6) - r0;
                                if (f >= (r0 << 6) && c < nr) atomicOr(&L.cmask[c], chunk_bit(f));
                            }
                            if (tid == THREADS - 1 && (N & 63u)) {  // the virtual slice
                                const uint32_t c = (N >> 6) - r0;
                                if (N >= (r0 << 6) && c < nr) atomicOr(&L.cmask[c], chunk_bit(N));
                            }
                            lds_barrier();
                        }
                        {
                            const unsigned long long cm = (uint32_t)tid < nr ? L.cmask[tid] : 0ull;
                            const uint32_t v = (uint32_t)__popcll(cm);
                            const uint32_t inc = wave_incl_scan(v, lane);
                            if (lane == 63) L.wsum[wave] = inc;
                            lds_barrier();
                            uint32_t wb = 0, tot = 0;
#pragma unroll
                            for (int g = 0; g < NW / 4; g++) {
                                const uint4 x = reinterpret_cast<const uint4 *>(L.wsum)[g];
                                wb += (4 * g + 0 < wave ? x.x : 0u) + (4 * g + 1 < wave ? x.y : 0u) +
                                      (4 * g + 2 < wave ? x.z : 0u) + (4 * g + 3 < wave ? x.w : 0u);
                                tot += (x.x + x.y) + (x.z + x.w);
                            }
                            // a lane's slice in chunk c is jb + (slice starts at lanes 1..L of
                            // the chunk), with jb = (slices starting before the chunk) - 1, + 1
                            // when one starts at lane 0 (bit 63); kept as the byte offset 4 jb
                            if ((uint32_t)tid < nr)
                                L.cj0[tid] = (uint16_t)((base + wb + inc - v + (uint32_t)(cm >> 63) - 1u) << 2);
                            // the dummy slot: chunks past the round read iA's padding
                            if (tid == 0) L.s_a[ne + 1] = ((sent - ((r0 + nr) << 6)) << so_shift) + so_bias;
                            base += tot;
                        }
                        lds_barrier();
                        // U chunks per wave in three batched LDS/HBM stages: chunk
                        // metadata, slice bases, ion ids; lane L's slice is the chunk's
                        // first plus the slice starts at lanes <= L
                        constexpr int U = 2 * kUnroll;
                        for (uint32_t c = (uint32_t)wave * U; c < nr; c += NW * U) {
                            uint32_t j0[U], sa[U], raw[U];
                            unsigned long long mk[U];
                            {
                                // lane u reads chunk c + u's metadata (uniform per chunk) and,
                                // as the chunk's only reader, clears its mask for the next
                                // round or step (their mask writes come after at least one
                                // barrier); the wave takes the metadata from that lane. A chunk
                                // past the round is a dummy whose lanes all read the padding.
                                uint32_t my_j0 = (ne + 1) << 2;
                                unsigned long long my_mk = 1ull << 63;
                                if ((lane < U) & (c + (uint32_t)lane < nr)) {
                                    const uint32_t cc = c + (uint32_t)lane;
                                    my_j0 = L.cj0[cc];
                                    my_mk = L.cmask[cc];
                                    L.cmask[cc] = 0;
                                }
#pragma unroll
                                for (int u = 0; u < U; u++) {
                                    j0[u] = (uint32_t)__builtin_amdgcn_readlane((int)my_j0, u);
                                    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)my_mk, u);
                                    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(my_mk >> 32), u);
                                    mk[u] = ((unsigned long long)hi << 32) | lo;
                                }
                            }
#pragma unroll
                            for (int u = 0; u < U; u++) {
                                // slice starts at lanes 1..L of the chunk (bit k - 1 for lane k;
                                // bit 63, lane 0, is never below a lane)
                                const uint32_t j = __builtin_amdgcn_mbcnt_hi(
                                    (uint32_t)(mk[u] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk[u], 0u));
                                sa[u] = *(const uint32_t *)((const char *)L.s_a + (j0[u] + (j << 2)));
                            }
                            // every lane loads, with no select: a lane past the round's
                            // ions reads the padding (no ion id), dropped with the ids
                            // outside the tile below
                            if constexpr (!WIDE) {
                                const uint32_t f4 = (((r0 + c) << 6) + (uint32_t)lane) << 2;
#pragma unroll
                                for (int u = 0; u < U; u++)
                                    raw[u] = *(gmem<uint32_t> *)(iab + (size_t)(uint32_t)(sa[u] + f4) + 256u * (uint32_t)u);
                            } else {
                                const uint32_t f0 = ((r0 + c) << 6) + (uint32_t)lane;
#pragma unroll
                                for (int u = 0; u < U; u++) raw[u] = ia[sa[u] + f0 + 64u * (uint32_t)u];
                            }
                            // the batch's counter increments are issued together, without
                            // a branch: a lane outside the tile adds 0 to its dummy word.
                            // One wait then covers all their returned old values.
                            uint32_t old[U], xv[U];
#pragma unroll
                            for (int u = 0; u < U; u++) {
                                // ids of the tile's entries are [tlo_id, tlo_id + span_id);
                                // with the exact division (checked on the host for every
                                // id < 2^31) that gives 0 <= el < nent
                                const uint32_t dr = raw[u] - tlo_id;
                                const bool in_tile = dr < span_id;
                                // the wave's count of consumed ids (uniform: kept in a scalar register)
                                n_scanned += (uint32_t)__popcll(__ballot(in_tile));
                                // floor(raw / speclen) - tlo, as tlo_id = tlo * speclen
                                const uint32_t el = __umulhi(dr, dm) >> dsh;
                                // the counter's bit offset in its word is x mod 32 (shifts and
                                // bit-field extracts read the low 5 bits of the offset); x >> fsh
                                // gives el back
                                const uint32_t x = el << fsh;
                                uint32_t *w = in_tile ? &L.cnt[el >> lp] : &L.dummy[lane];
                                const uint32_t one = 1u << (x & 31u);
                                old[u] = atomicAdd(w, in_tile ? one : 0u);
                                xv[u] = x;
                            }
                            // the increment that reaches min_shp makes the entry a candidate,
                            // exactly once. One uniform test for the batch (a field equal to
                            // min_shp - 1 makes the xor, and so the minimum, zero); the rare
                            // batch that makes candidates tests its chunks again (the empty
                            // asm keeps the compiler from holding 12 lane masks for that)
                            uint32_t dmin = 0xffffffffu;
#pragma unroll
                            for (int u = 0; u < U; u++) dmin = min(dmin, __builtin_amdgcn_ubfe(old[u], xv[u], fbw) ^ (min_cnt - 1u));
                            if (__ballot(dmin == 0u))
#pragma unroll
                            for (int u = 0; u < U; u++) {
                                asm volatile("" : "+v"(old[u]));
                                if (__builtin_amdgcn_ubfe(old[u], xv[u], fbw) == min_cnt - 1u) {
                                    const uint32_t el = xv[u] >> fsh;
                                    const uint32_t p = atomicAdd(&L.n_cand, 1u);
                                    const uint32_t cd = btag | ((uint32_t)tlo + el);
                                    if (p < (uint32_t)kCandLdsSpec)
                                        L.cand[p] = cd;
                                    else
                                        g_cand[p - kCandLdsSpec] = cd;
                                }
                            }
                        }
                    }
                }
                // the candidates were collected in pass 1; the tile is cleared later
                dirty_q = (((nent + pack - 1) >> lp) + 3) >> 2;
            }
        }
        lds_barrier();
        clear_tile();
        __syncthreads();  // candidates in global scratch, see above
        if (L.n_cand) rescore_candidates<THREADS, TILE_WORDS, SL>(a, L, P, qoff, g_cand, n_errors);
        {
            const uint32_t werr = wave_sum_u32(n_errors);
            if (lane == 0) {
                L.w_scan[wave] = n_scanned;  // already the wave's total
                L.w_err[wave] = werr;
            }
        }
        // ---- block reduction of the top hit; the unit's results ----
        lds_barrier();
        const bool whole = slot == kWholeUnit;
        uint32_t *hout = whole ? a.hist + (size_t)s * GX_HIST_BINS : a.part_hist + (size_t)slot * GX_HIST_BINS;
        for (int w = tid; w < GX_HIST_BINS; w += THREADS) {
            hout[w] = L.hist[w];
            L.hist[w] = 0;
        }
        if (tid == 0) {
            unsigned long long key = 0, scanned = 0, errors = 0;
            uint32_t eq = 0, shp = 0, cps = 0;
            for (int w = 0; w < NW; w++) {
                cps += L.w_cps[w];
                scanned += L.w_scan[w];
                errors += L.w_err[w];
                merge_best(key, eq, shp, L.w_key[w], L.w_eq[w], L.w_shp[w]);
            }
            if (whole) {
                a.best_key[s] = key;
                a.best_eq[s] = eq;
                a.best_shp[s] = shp;
                a.cpsms[s] = cps;
            } else {
                a.part_key[slot] = key;
                a.part_eq[slot] = eq;
                a.part_shp[slot] = shp;
                a.part_cps[slot] = cps;
            }
            // the unit's statistics (H, the window-restricted ion ids, is
            // counted where they are consumed); per unit, so that no counter
            // stays live in registers across the launch
            stat_add(a.stats, ST_CAND, cps);
            stat_add(a.stats, ST_IONS, scanned);
            stat_add(a.stats, ST_SCANNED, scanned);
            stat_add(a.stats, ST_ERRORS, errors);
        }
    }
}

// one wave per heavy spectrum of the chunk: its parts' histograms and
// candidate counts add up, and their top hits merge with the same rule as
// within a block (the parts are disjoint bucket ranges of one spectrum)
__global__ void __launch_bounds__(64) fold_kernel(FoldArgs a) {
    const int lane = threadIdx.x;
    const uint32_t i = blockIdx.x;
    const uint32_t s = a.f_s[i], p0 = a.f_slot0[i], n = a.f_n[i];
    for (int w = lane; w < GX_HIST_BINS; w += 64) {
        uint32_t h = 0;
        for (uint32_t k = 0; k < n; k++) h += a.part_hist[(size_t)(p0 + k) * GX_HIST_BINS + w];
        a.hist[(size_t)s * GX_HIST_BINS + w] = h;
    }
    if (lane == 0) {
        unsigned long long key = 0;
        uint32_t eq = 0, shp = 0, cps = 0;
        for (uint32_t k = 0; k < n; k++) {
            cps += a.part_cps[p0 + k];
            merge_best(key, eq, shp, a.part_key[p0 + k], a.part_eq[p0 + k], a.part_shp[p0 + k]);
        }
        a.best_key[s] = key;
        a.best_eq[s] = eq;
        a.best_shp[s] = shp;
        a.cpsms[s] = cps;
    }
}

// ---------------------------------------------------------------------------
// one 64-thread block per spectrum: the top hit, the plain tail fit and the
// e-value (dslim_query.cpp:1017-1061, expeRT.cpp:396-559)
__global__ void __launch_bounds__(64) evalue_kernel(EvalArgs a) {
    __shared__ uint32_t hist[GX_HIST_BINS];
    __shared__ double sx[GX_HIST_BINS];
    __shared__ FitScratch F;
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x;
    const uint32_t need = (uint32_t)max(a.min_cpsm, 1);
    gx_psm r = top_hit_psm(a.best_key[s], a.best_eq[s], a.best_shp[s], a.cpsms[s], a.buckets, need);
    if (!(r.flags & GX_PSM_MODELED)) {
        if (lane == 0) a.out[s] = r;
        return;
    }
    for (int w = lane; w < GX_HIST_BINS; w += 64) hist[w] = a.hist[(size_t)s * GX_HIST_BINS + w];
    __syncthreads();
    double sl, ic;
    int32_t mu, beta;
    if (tail_fit_block(hist, sx, F, r.hyp, r.cpsms, sl, ic, mu, beta)) {
        r.flags |= GX_PSM_FIT_OK;
        r.mu = mu;
        r.beta = beta;
        r.evalue = evalue_of(mu, beta, r.hyp, (uint32_t)r.cpsms);
        if (r.evalue < a.expect_max) r.flags |= GX_PSM_REPORTED;
    }
    if (lane == 0) a.out[s] = r;
}

// WAVES_PER_SIMD = waves of the resident blocks per SIMD (4 SIMDs per CU):
// it caps the VGPRs so that blocks_per_cu blocks fit
template <int THREADS, int TILE_WORDS, int SL, int WAVES_PER_SIMD>
hipError_t launch_spectra(const ItemArgs &a, int n_blocks, hipStream_t s) {
    if (a.wide)
        hipLaunchKernelGGL((search_spectra_kernel<THREADS, TILE_WORDS, SL, WAVES_PER_SIMD, true>), dim3(n_blocks),
                           dim3(THREADS), 0, s, a);
    else
        hipLaunchKernelGGL((search_spectra_kernel<THREADS, TILE_WORDS, SL, WAVES_PER_SIMD, false>), dim3(n_blocks),
                           dim3(THREADS), 0, s, a);
    return hipGetLastError();
}

}  // namespace

// LDS per block: 4*TILE_WORDS + 13*SL + 2.5 KB candidates + 3.2 KB histogram
// (the spectrum's peaks live in the count tile during pass 2)
BatchConfig batch_config(int id) {
    switch (id) {
        case 0: return {1024, 32768, 1024, 1};  // 128 KB tile, one block per CU
        case 1: return {512, 16384, 512, 2};    // 64 KB tile, two blocks per CU
        case 2: return {256, 8192, 256, 3};     // 32 KB tile, three blocks per CU
        default: return {768, 28672, 768, 1};   // 112 KB tile, one block of 12 waves per CU
    }
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_plan_spectra(const PlanArgs &a, hipStream_t s) {
    if (a.n_spectra == 0) return hipSuccess;
    hipLaunchKernelGGL(plan_spectra_kernel, dim3(a.n_spectra), dim3(kPlanThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scan(const PlanArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_plan_slices(const PlanArgs &a, hipStream_t s) {
    const uint64_t nblk = (uint64_t)(a.chunk_s1 - a.chunk_s0) * (uint64_t)a.n_buckets;
    if (nblk == 0) return hipSuccess;
    if (a.slice_threads == kSliceThreadsLarge)
        hipLaunchKernelGGL(plan_slices_kernel<kSliceThreadsLarge>, dim3((unsigned)nblk), dim3(kSliceThreadsLarge), 0, s, a);
    else
        hipLaunchKernelGGL(plan_slices_kernel<kSliceThreads>, dim3((unsigned)nblk), dim3(kSliceThreads), 0, s, a);
    return hipGetLastError();
}

bool narrow_stream_fits(uint64_t n_ions) { return n_ions < kNarrowIons; }

hipError_t launch_search_spectra(int config, const ItemArgs &a, int n_blocks, hipStream_t s) {
    switch (config) {
        case 0: return launch_spectra<1024, 32768, 1024, 4>(a, n_blocks, s);
        case 1: return launch_spectra<512, 16384, 512, 4>(a, n_blocks, s);
        case 2: return launch_spectra<256, 8192, 256, 3>(a, n_blocks, s);
        default: return launch_spectra<768, 28672, 768, 3>(a, n_blocks, s);
    }
}

hipError_t launch_fold(const FoldArgs &a, hipStream_t s) {
    if (a.n_folds == 0) return hipSuccess;
    hipLaunchKernelGGL(fold_kernel, dim3(a.n_folds), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_evalue(const EvalArgs &a, hipStream_t s) {
    if (a.n_spectra == 0) return hipSuccess;
    hipLaunchKernelGGL(evalue_kernel, dim3(a.n_spectra), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace gx
