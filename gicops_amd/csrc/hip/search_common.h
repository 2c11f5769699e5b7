// Shared device-side declarations of the search kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gicops_amd.h"

namespace gx {

// One length bucket as the kernels see it. mass/bA/iA are the reference
// layout; ion_bin is derived on the device at upload time (ion id -> bin,
// the inverse of iA, i.e. the reference's SpecArr, dslim.cpp:327-333).
struct DevBucket {
    const float *mass;
    const uint32_t *bA;
    const uint32_t *iA;     // n_entries * speclen ion ids, then kIaPad sentinels (0xffffffff)
    const uint32_t *ion_bin;
    const uint32_t *iS;     // every kSampleStep-th ion id: iS[k] = iA[k kSampleStep] (boundary brackets)
    uint32_t n_entries;
    uint32_t speclen;
    // exact division by speclen for n < 2^31: n / speclen == umulhi(n, div_m) >> div_s
    uint32_t div_m;
    uint32_t div_s;
};

// Work counters accumulated by the search kernel (u64 each).
enum { ST_IONS = 0, ST_BINS, ST_WINDOW, ST_CAND, ST_PEAKS, ST_SPECTRA, ST_FALLBACK, ST_SCANNED, ST_ERRORS, ST_N };

struct SearchArgs {
    const DevBucket *buckets;
    int n_buckets;
    // query batch (CSR)
    const uint32_t *qidx;
    const int32_t *moz;
    const int32_t *inten;
    const float *prec;
    const float *dm;  // may be null
    uint32_t n_spectra;
    // parameters
    int dF;
    int nbins;        // max_mass*scale
    int min_shp;
    int min_cpsm;
    double dM;
    double expect_max;
    const double *lgfact;  // lgfact[i], i < lgfact_n
    int lgfact_n;
    // outputs
    gx_psm *out;
    uint32_t *hist_out;  // may be null: n_spectra*GX_HIST_BINS
    // scheduling + scratch
    uint32_t *work_counter;
    uint32_t *slice_scratch;   // per block: 2*max_slices words
    uint32_t *cand_scratch;    // per block: tile_entries words
    uint32_t max_slices;
    unsigned long long *stats; // ST_N counters
};

// Limits of the LDS-resident per-spectrum state.
constexpr int kMaxPeaksLds = 2048;   // legal peaks per spectrum
// the plan brackets tile boundaries from every kSampleStep-th ion id
constexpr uint32_t kSampleShift = 4;
constexpr uint32_t kSampleStep = 1u << kSampleShift;
// iA is followed by kIaPad sentinel slots (0xffffffff, no ion id): the batched
// search's lanes past a round's ions read them instead of branching
constexpr uint32_t kIaPad = 1024;
constexpr int kMaxBuckets = 64;      // length buckets
constexpr int kSearchThreads = 512;  // 8 waves of 64
constexpr int kTileWords = 11264;    // count tile: 2 entries (u16) per word -> 22528 entries
constexpr int kTileEntries = 2 * kTileWords;
constexpr int kSliceLds = 1024;      // slices kept in LDS, more go to scratch
constexpr int kCandLds = 1024;       // candidates kept in LDS, more go to scratch
constexpr int kPeakIdxWords = 2048;  // cell table of the batched pass 2 (one word per cell, 2048 cells)
constexpr int kStatRows = 64;        // stat counters are spread over rows to avoid one hot address

hipError_t launch_reset(uint32_t *counter, unsigned long long *stats, hipStream_t s);
hipError_t launch_search(const SearchArgs &a, int n_blocks, hipStream_t s);
hipError_t launch_sample_ids(const uint32_t *iA, uint32_t *iS, uint64_t n_ions, hipStream_t s);
hipError_t launch_ion_bin(const uint32_t *bA, const uint32_t *iA, uint32_t *ion_bin, uint32_t nbins,
                          uint64_t n_ions, hipStream_t s);

}  // namespace gx
