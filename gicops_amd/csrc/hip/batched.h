// Batched engine: plan kernels, the persistent work-stealing search over
// (spectrum, bucket, tile) items, and the e-value kernel.
#pragma once

#include "search_common.h"

namespace gx {

// One search configuration (the autotuner's choice): block size and the width
// of the LDS count tile. Counts are packed 4 (u8) or 2 (u16) per 32-bit word.
struct BatchConfig {
    int threads;     // 256, 512, 768 or 1024
    int tile_words;  // count-tile words per block
    int slice_chunk; // slices staged in LDS at a time
    int blocks_per_cu;
};

constexpr int kNumBatchConfigs = 4;
BatchConfig batch_config(int id);

struct PlanArgs {
    const DevBucket *buckets;
    int n_buckets;
    const uint32_t *qidx;
    const int32_t *moz;
    const int32_t *inten;
    const float *prec;
    const float *dm;
    uint32_t n_spectra;
    int dF;
    int nbins;
    double dM;
    int tile_words;          // count-tile words per item (the configuration's tile / split)
    uint32_t whole_bin;      // bins of at most this many ions are streamed whole
    int slice_threads;       // plan_slices block size: 256 for an index with a bucket of more than 2^22 entries, else 128
    // outputs of plan_spectra
    int32_t *lq;        // sorted legal peaks (CSR at qidx offsets)
    int32_t *li;
    uint32_t *n_legal;  // per spectrum
    int32_t *pair_lo;   // [s*nb + b]
    int32_t *pair_hi;
    uint32_t *pair_tiles;    // tiles of the pair (0 = no window)
    uint32_t *pair_pack;     // entries per count word (4 or 2)
    uint32_t *pair_pos_off;  // spectrum-local slice-position offset (words)
    uint32_t *pair_nsl;      // non-empty slices of the pair (written by plan_slices)
    uint32_t *n_items;       // per spectrum, scanned in place to offsets by scan_kernel
    uint64_t *n_pos;         // per spectrum, scanned in place to offsets
    // scanned offsets (n_spectra + 1 entries)
    uint32_t *item_off;
    uint64_t *pos_off;
    // outputs of plan_slices
    uint32_t *pos;           // chunk-local slice positions: [a, z) per slice
    const uint32_t *plan_order;  // the order in which plan_slices visits each chunk's spectra
    uint32_t chunk_s0;       // first spectrum of the current chunk
    uint32_t chunk_s1;
    unsigned long long *stats;
};

struct ItemArgs {
    const DevBucket *buckets;
    int n_buckets;
    const uint32_t *qidx;
    int dF;
    int min_shp;
    int nbins;
    const double *lgfact;
    int lgfact_n;
    const int32_t *lq;
    const int32_t *li;
    const uint32_t *n_legal;
    const int32_t *pair_lo;
    const int32_t *pair_hi;
    const uint32_t *pair_tiles;
    const uint32_t *pair_pack;
    const uint32_t *pair_pos_off;
    const uint32_t *pair_nsl;
    const uint64_t *pos_off;
    const uint32_t *pos;
    uint32_t chunk_s0, chunk_s1;  // spectra of the chunk
    // the chunk's work units in the order they are handed out: a spectrum, or
    // a part of a heavy spectrum (a bucket range) whose results go to a slot
    // of the part scratch and are folded afterwards
    const uint32_t *unit_s;  // spectrum of each unit
    const uint32_t *unit_br; // b0 | b1 << 8 | slot << 16; slot 0xffff = the whole spectrum
    uint32_t unit0, n_units; // this chunk's units
    uint32_t *work_counter;  // this chunk's counter
    uint32_t *cand_scratch;  // per block overflow candidates
    int tile_entries_max;
    int split;               // window split: psids per item = tile / split
    int wide;                // 1: ion stream by 64-bit addresses (the WIDE kernel instance)
    uint32_t max_speclen;    // pass-2 task stride, with its exact divisor
    uint32_t smax_m, smax_s;
    uint32_t cell_shift;     // bins per pass-2 cell-table cell = 2^cell_shift
    // per-spectrum results
    uint32_t *hist;          // n_spectra * GX_HIST_BINS
    uint32_t *cpsms;
    unsigned long long *best_key;  // top hit: h bits << 32 | ~(b << 26 | psid)
    uint32_t *best_eq;       // candidates with the top hyperscore
    uint32_t *best_shp;      // shared ions of the top hit
    // results of the parts of heavy spectra, by slot
    unsigned long long *part_key;
    uint32_t *part_eq, *part_shp, *part_cps, *part_hist;
    unsigned long long *stats;
};

// folding of the parts of the chunk's heavy spectra into their results
struct FoldArgs {
    const uint32_t *f_s, *f_slot0, *f_n;  // spectrum, first slot and number of parts
    uint32_t n_folds;
    const unsigned long long *part_key;
    const uint32_t *part_eq, *part_shp, *part_cps, *part_hist;
    uint32_t *hist, *cpsms, *best_eq, *best_shp;
    unsigned long long *best_key;
};

constexpr uint32_t kWholeUnit = 0xffffu;  // unit slot of a spectrum searched whole

struct EvalArgs {
    const DevBucket *buckets;
    uint32_t n_spectra;
    int min_cpsm;
    double expect_max;
    const unsigned long long *best_key;
    const uint32_t *best_eq;
    const uint32_t *best_shp;
    const uint32_t *hist;
    const uint32_t *cpsms;
    gx_psm *out;
    unsigned long long *stats;
};

hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_plan_spectra(const PlanArgs &a, hipStream_t s);
hipError_t launch_scan(const PlanArgs &a, hipStream_t s);
hipError_t launch_plan_slices(const PlanArgs &a, hipStream_t s);
hipError_t launch_search_spectra(int config, const ItemArgs &a, int n_blocks, hipStream_t s);
// whether the search kernel's 32-bit ion-stream offsets reach every ion of a
// bucket of n_ions ions (and its padding)
bool narrow_stream_fits(uint64_t n_ions);
hipError_t launch_evalue(const EvalArgs &a, hipStream_t s);
hipError_t launch_fold(const FoldArgs &a, hipStream_t s);

}  // namespace gx
