/*
 * gicops_amd.h - C ABI of the MI355X (gfx950) fragment-ion-index search.
 *
 * This is the drop-in boundary for the reference's search + e-value stage:
 *
 *   reference (C++/CUDA)                                this ABI
 *   -------------------------------------------------   -----------------------------
 *   hcp::gpu::cuda::s3::initialize()                    gx_ctx_create()
 *     (include/cuda/superstep3/kernel.hpp:96,
 *      superstep3/kernel.cu:251-261)
 *   hcp::gpu::cuda::s3::deinitialize()                  gx_ctx_destroy()
 *     (kernel.hpp:100)
 *   hcp::gpu::cuda::s1::getdIndex(Index*)               gx_index_upload()
 *     (superstep1/kernel.cu:310-358)
 *   s3::search(Queries<int>*, Index*, uint, int, hCell*) gx_search()  (H2D + search + D2H)
 *     (kernel.hpp:98) + s4::processResults(Index*,        gx_search_device() (device-resident)
 *     Queries<int>*, int) (superstep4/kernel.hpp:49)
 *   DSLIM_QuerySpectrum(Queries<int>*, Index*, uint, int) gx_search() with the same
 *     (include/dslim.h:196, dslim_query.cpp:749)         per-spectrum semantics
 *   expeRT::ModelTailFit(Results*) (expeRT.cpp:396)      in-kernel (GX_EVALUE_TAILFIT)
 *
 * Status codes are the reference's (source/core/include/slmerr.h:24-46) plus
 * GX_ERR_DEVICE / GX_ERR_NOT_READY / GX_ERR_NO_DEVICE. No function exits the
 * process or throws across the ABI: device errors come back as GX_ERR_DEVICE
 * (the reference calls exit(), cuda/driver.hpp:41-48).
 *
 * Host buffers are caller-owned; device buffers are library-owned. A context
 * serialises its own calls (internal mutex); use one context per device.
 */
#ifndef GICOPS_AMD_H
#define GICOPS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (slmerr.h:24-46) ---------------------------------- */
#define GX_SUCCESS 0
#define GX_ERR_BASE (-3000)
#define GX_ERR_INVLD_MEMORY (GX_ERR_BASE - 1)
#define GX_ERR_INVLD_PARAM (GX_ERR_BASE - 2)
#define GX_ERR_INVLD_INDEX (GX_ERR_BASE - 3)
#define GX_ERR_INVLD_SIZE (GX_ERR_BASE - 4)
#define GX_ERR_ALREADY_INIT (GX_ERR_BASE - 5)
#define GX_ERR_FILE_NOT_FOUND (GX_ERR_BASE - 6)
#define GX_ERR_BAD_MEM_ALLOC (GX_ERR_BASE - 7)
#define GX_ERR_INVLD_PTR (GX_ERR_BASE - 8)
#define GX_ERR_INVLD_SORT (GX_ERR_BASE - 9)
#define GX_ERR_INVLD_MOD (GX_ERR_BASE - 10)
#define GX_ERR_NOT_ENOUGH_DATA (GX_ERR_BASE - 12)
#define GX_ERR_FILE_ERROR (GX_ERR_BASE - 13)
/* extensions of this library */
#define GX_ERR_DEVICE (GX_ERR_BASE - 50)    /* HIP runtime / launch failure */
#define GX_ERR_NOT_READY (GX_ERR_BASE - 51) /* object used before it was built */
#define GX_ERR_NO_DEVICE (GX_ERR_BASE - 52) /* no gfx950 device visible */

/* ---- index: one bucket per peptide length (slm_dsts.h:298-346) ------- */
/* Read-only view of one length bucket of the SLM index, in the reference's
 * layout: entries sorted ascending by float precursor mass; bA[b] = number of
 * ions in bins < b (nbins+1 words); iA = ion ids ascending within each bin;
 * ion id = psid*speclen + r with speclen = (L-1)*maxz*2, r < speclen/2 a
 * b-ion. The fields after `iA` are metadata the search does not read. */
typedef struct gx_bucket {
    int32_t peplen;        /* L */
    uint32_t n_peps;       /* unmodified sequences (lclpepCnt) */
    uint32_t n_entries;    /* peptides + variants (lcltotCnt) */
    uint32_t speclen;      /* (L-1)*maxz*2 */
    uint64_t n_ions;       /* n_entries*speclen = bA[nbins] */
    uint32_t nbins;        /* max_mass*scale */
    const char *seqs;      /* n_peps*L residues */
    const float *mass;     /* n_entries, ascending */
    const uint32_t *seqid; /* n_entries */
    const uint32_t *modnum;/* n_entries */
    const uint64_t *sites; /* n_entries */
    const uint32_t *bA;    /* nbins+1 */
    const uint32_t *iA;    /* n_ions */
} gx_bucket;

/* ---- query batch (Queries<int>, slm_dsts.h:520-658) ------------------ */
typedef struct gx_query_batch {
    uint32_t n_spectra;
    uint64_t n_peaks;        /* idx[n_spectra] */
    const uint32_t *idx;     /* n_spectra+1 CSR row offsets */
    const int32_t *moz;      /* n_peaks, (int)(m/z*scale) */
    const int32_t *intensity;/* n_peaks */
    const float *precurse;   /* n_spectra, precursor mass */
    const float *dm;         /* optional n_spectra per-spectrum tolerance (Da);
                                NULL = params.dM for all (an extension: the
                                reference has one global dM) */
} gx_query_batch;

/* ---- parameters (gParams subset, slm_dsts.h:349-431, argp.hpp:272-370) */
enum { GX_EVALUE_TAILFIT = 0, GX_EVALUE_LEARNED = 1 };
/* search engines: the batched plan + persistent item kernel (default), or one
 * block per spectrum (the plain kernel) */
enum { GX_ENGINE_BATCHED = 0, GX_ENGINE_SPECTRUM = 1 };

typedef struct gx_search_params {
    int32_t dF;            /* fragment tolerance in bins ((uint)(0.02*scale) = 2) */
    int32_t scale;         /* bins per Da (100) */
    int32_t max_mass;      /* Da (5000); legal peaks dF < q < max_mass*scale-1-dF */
    int32_t min_shp;       /* min shared peaks (4) */
    int32_t min_cpsm;      /* min candidate PSMs for an e-value (4) */
    int32_t evalue_mode;   /* GX_EVALUE_* */
    double dM;             /* precursor tolerance, Da (global) */
    double expect_max;     /* report gate: e < expect_max (20) */
    float surrogate_conf;  /* GX_EVALUE_LEARNED: confidence threshold for the MLP */
    int32_t engine;        /* GX_ENGINE_* */
    int32_t config;        /* batched launch configuration, -1 = the context's tuned choice */
    int32_t use_graph;     /* batched: 1 = capture the batch as a hipGraph once, then replay */
    int32_t split;         /* batched: window split, psids per item = tile / split (1, 2 or 4; 0 = tuned) */
    int32_t whole_bin;     /* batched: bins of at most this many ions are streamed whole and filtered by
                              id instead of searched for tile boundaries (1..4096; 0 = tuned) */
    int32_t wide_stream;   /* batched: 1 = address iA with 64-bit ion positions even when 32-bit byte
                              offsets reach every bucket (what an index with a bucket of about 2^30
                              ions or more always uses); 0 = the faster 32-bit form where it fits */
} gx_search_params;

/* ---- per-spectrum result (hCell + Results + e-value) ----------------- */
enum {
    GX_PSM_HAS_HIT = 1,    /* cpsms >= 1: a top hit exists */
    GX_PSM_MODELED = 2,    /* cpsms >= min_cpsm: the tail fit ran */
    GX_PSM_FIT_OK = 4,     /* the fit is defined (see the flat-slice rule) */
    GX_PSM_REPORTED = 8,   /* e < expect_max: DFile_PrintScore would print it */
    GX_PSM_SURROGATE = 16, /* e-value came from the learned surrogate */
    GX_PSM_TIE = 32        /* another candidate had the same top hyperscore */
};

typedef struct gx_psm {
    float hyperscore;    /* top hit hyperscore (f32, clamped to 79), 0 if none */
    int32_t psid;        /* index into the bucket's mass-sorted entries, -1 if none */
    uint16_t idxoffset;  /* bucket (length) index of the top hit */
    uint16_t sharedions; /* bc+yc of the top hit */
    uint16_t totalions;  /* speclen of the top hit's bucket */
    uint16_t flags;      /* GX_PSM_* */
    int32_t cpsms;       /* candidate PSMs (h > 0, shared >= min_shp) */
    int32_t hyp;         /* (int)(hyperscore*10 + 0.5): histogram bin of the top hit */
    int32_t mu;          /* tail-fit slope * 1e6, truncated (Results::mu) */
    int32_t beta;        /* tail-fit intercept * 1e6, truncated (Results::beta) */
    double evalue;       /* cpsms * 10^(mu/1e6*hyp + beta/1e6); +inf if not modeled */
} gx_psm;

#define GX_HIST_BINS 802 /* 2 + MAX_HYPERSCORE*10 (kernel.hpp:32) */

/* Work counters of one search call, summed over spectra (window-restricted
 * ion reads, bA pairs, window entries, candidates, tail-fit fallbacks). */
typedef struct gx_search_stats {
    uint64_t ions;        /* H: iA entries inside the precursor windows */
    uint64_t bins;        /* B: (bucket, peak, bin) triples whose bA pair was read */
    uint64_t window;      /* W: index entries inside the precursor windows */
    uint64_t candidates;  /* sum of cpsms */
    uint64_t spectra;     /* spectra searched */
    uint64_t peaks;       /* query peaks read */
    uint64_t fallbacks;   /* learned mode: spectra routed to the plain tail fit */
    uint64_t ions_scanned;/* ion ids the search kernel streamed (self-check: equals `ions`) */
    uint64_t errors;      /* consistency failures: a candidate whose exact recount (pass 2) is below min_shp although the scatter (pass 1) counted min_shp; always 0 */
    uint64_t items;       /* batched: (spectrum, bucket, tile) work items */
    uint32_t chunks;      /* batched: planning chunks */
    int32_t config;       /* batched: launch configuration used */
    double kernel_ms;     /* device time of the search kernel(s) (HIP events on the context stream) */
    double evalue_ms;     /* device time of the e-value stage (0 when fused into the search kernel) */
    double plan_ms;       /* device time of the plan kernels (batched engine) */
    double total_ms;      /* device time of the whole call, first to last kernel */
} gx_search_stats;

typedef struct gx_ctx gx_ctx;
typedef struct gx_dindex gx_dindex;
typedef struct gx_dqueries gx_dqueries;

/* Create a context on HIP device `device` (lgfact table, stream, events).
 * Replaces s3::initialize(). */
int gx_ctx_create(int device, gx_ctx **out);
void gx_ctx_destroy(gx_ctx *ctx);

/* Upload an index (all length buckets) to HBM; replaces s1::getdIndex().
 * `maxz` is the fragment charge count the index was built with. Buckets with
 * n_entries == 0 are kept but never searched. */
int gx_index_upload(gx_ctx *ctx, const gx_bucket *buckets, int n_buckets, int maxz, gx_dindex **out);
void gx_index_free(gx_dindex *idx);

/* Build the fragment-ion index of every bucket on the GPU, replacing
 * s1::ConstructIndexChunk (superstep1/kernel.cu:484-733) and the CPU's
 * DSLIM_ConstructChunk + DSLIM_SLMTransform (dslim.cpp:89-97). `buckets`
 * are the mass-sorted entries (mass, seqid, modnum, sites, seqs); their bA
 * and iA are ignored (may be NULL). Fragment bins follow the CPU build's
 * float arithmetic (utils.cpp:161-474, dslim.cpp:318-337), and iA holds ion
 * ids ascending within each bin, so the result is bit-identical to the CPU
 * index. The device index is ready for search, like gx_index_upload's. */
#define GX_MAX_MOD_TYPES 15
typedef struct gx_index_build_params {
    int32_t maxz;
    int32_t scale;
    float min_mass;
    float max_mass;
    int32_t n_mod_types;
    uint32_t mod_mass[GX_MAX_MOD_TYPES]; /* (uint)(mass * scale) per type (modMass, mods.cpp) */
    int32_t modmass_all;                 /* 0: the reference's first-modification precursor quirk */
} gx_index_build_params;
int gx_index_build(gx_ctx *ctx, const gx_bucket *buckets, int n_buckets, const gx_index_build_params *bp,
                   gx_dindex **out);
/* Copy bucket `bucket`'s bA (nbins+1) and iA (n_ions) of a device index to
 * the host (either pointer may be NULL). */
int gx_index_download(gx_dindex *idx, int bucket, uint32_t *bA, uint32_t *iA);

/* Upload a query batch to HBM (dQueries::H2D, superstep3/kernel.cu:192-206). */
int gx_queries_upload(gx_ctx *ctx, const gx_query_batch *q, gx_dqueries **out);
void gx_queries_free(gx_dqueries *q);

/* Peak-picking parameters: gParams base_int (base_int x YAXISMULTIPLIER),
 * min_int (base_int * cutoff + 0.5) (argp.hpp:304-307) and QALEN. */
typedef struct gx_pick_params {
    int32_t base_int; /* intensity of the most intense peak after normalisation (100000) */
    int32_t min_int;  /* smallest kept normalised intensity (1000) */
    int32_t qalen;    /* most intense peaks kept, 1..2048 (QALEN, 100) */
} gx_pick_params;

/* Preprocess a batch of raw spectra on the GPU into a device-resident query
 * batch ready for gx_search_device, replacing s2::ArraySort +
 * s2::computenewlens + s2::pickpeaks (superstep2/kernel.cu:100-422) with the
 * CPU's MSQuery::pickpeaks semantics (msquery.cpp:338-423): per spectrum, the
 * peaks sorted by intensity with the reference's key/value quicksort (so
 * equal intensities end in its order, keyval.h:49-103), the most intense set
 * to base_int, the next qalen-1 scaled by base_int/max and kept while
 * >= min_int, stored in ascending-intensity order. `raw` holds integer m/z
 * (m/z*scale) and intensity (intensity*100) as the MS2 reader stores them
 * (msquery.cpp:286-287); its dm array, if any, is kept. */
int gx_queries_preprocess(gx_ctx *ctx, const gx_query_batch *raw, const gx_pick_params *pp, gx_dqueries **out);

/* Number of spectra and peaks of a device batch, the device time of its
 * gx_queries_preprocess (sort, scan and compaction on the context stream,
 * HIP events; 0 for an uploaded batch), and a copy of its CSR arrays to the
 * host (idx: n_spectra+1, moz/intensity: n_peaks). Any pointer may be NULL. */
int gx_queries_shape(const gx_dqueries *q, uint32_t *n_spectra, uint64_t *n_peaks, float *prep_ms);
int gx_queries_download(const gx_dqueries *q, uint32_t *idx, int32_t *moz, int32_t *intensity);

/* Search a device-resident batch and model its e-values; results stay on the
 * device until gx_results_download(). Replaces s3::search +
 * s4::processResults for one batch. */
int gx_search_device(gx_ctx *ctx, const gx_dindex *idx, gx_dqueries *q, const gx_search_params *p,
                     gx_search_stats *stats);

/* Copy the results of the last gx_search_device() on `q` to the host.
 * `hist` (optional) receives n_spectra*GX_HIST_BINS u32 survival counts. */
int gx_results_download(gx_ctx *ctx, gx_dqueries *q, gx_psm *psms, uint32_t *hist);

/* One call: upload `q`, search, model, download (s3::search semantics). */
int gx_search(gx_ctx *ctx, const gx_dindex *idx, const gx_query_batch *q, const gx_search_params *p,
              gx_psm *psms, uint32_t *hist, gx_search_stats *stats);

/* The per-spectrum engine's search (GX_ENGINE_SPECTRUM, plain tail fit)
 * without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture
 * elsewhere in the process, tight serving loops, harnesses that forbid
 * hipMalloc). Every pointer of `q` (idx, moz, intensity, precurse, dm) and
 * `psms` (n_spectra records) / `hist` (NULL or n_spectra*GX_HIST_BINS u32)
 * are device memory of the context's GPU, owned by the caller. The work
 * buffers live in one static device array of 60 MiB per GPU, held until the
 * process ends: the per-block slice and candidate overflow of up to 512
 * blocks. Limits: at most 2048 peaks per spectrum and peaks x (2 dF + 1) <=
 * 4096 (GX_ERR_INVLD_SIZE beyond; a malformed idx array is GX_ERR_INVLD_SIZE
 * too). The array keeps the previous call's contents; each call clears the
 * parts that must start at zero. All calls on one GPU share it, so the
 * library serialises them per device within the process (calls from other
 * threads wait); the call returns when the search is complete. Results equal
 * gx_search_device's with engine GX_ENGINE_SPECTRUM. */
int gx_search_spectra_noalloc(gx_ctx *ctx, const gx_dindex *idx, const gx_query_batch *q,
                              const gx_search_params *p, gx_psm *psms, uint32_t *hist, gx_search_stats *stats);

/* Autotuner result: the launch configuration and window split chosen, and
 * the device time of every candidate tried on the sample batch. */
#define GX_TUNE_MAX 16
typedef struct gx_tune_result {
    int32_t config;
    int32_t split;
    int32_t whole_bin;
    int32_t n_tried;
    int32_t tried_config[GX_TUNE_MAX];
    int32_t tried_split[GX_TUNE_MAX];
    int32_t tried_whole_bin[GX_TUNE_MAX];
    double tried_ms[GX_TUNE_MAX];
} gx_tune_result;

/* Time every batched launch configuration x window split on a sample batch
 * (device-resident, as uploaded by gx_queries_upload), then the whole-bin
 * threshold (8, 16, 32, 64, 128 ions) for the two fastest pairs (16 trials), and make the winners the context's
 * defaults (params.config = -1, split = 0, whole_bin = 0). No reference
 * counterpart: the reference's launch shape is fixed (kernel.cu:588-598). */
int gx_autotune(gx_ctx *ctx, const gx_dindex *idx, gx_dqueries *sample, const gx_search_params *p,
                gx_tune_result *out);

/* ---- index-sharded search: partial results and their merge ---------- */
/* One index shard's result for one spectrum, the record an index-sharded
 * search exchanges before the e-value. It replaces the reference's partRes
 * + u16 histogram buffer (slm_dsts.h:856-934, 1005-1040; filled by
 * dslim_query.cpp:973-1000 and expeRT::StoreIResults, expeRT.cpp:714-806),
 * which is lossy above 65535 counts and limited to 128 histogram bins; the
 * whole 802-bin u32 histogram of the spectrum travels beside it. */
typedef struct gx_partial {
    uint64_t key;        /* top hit: f32 hyperscore bits << 32 | ~(bucket << 26 | psid); 0 = none.
                            A larger key is the better hit (ties: smaller bucket, then psid) */
    uint32_t eq;         /* entries of this shard at the top hyperscore */
    uint32_t cpsms;      /* this shard's candidates */
    uint16_t sharedions; /* bc+yc of this shard's top hit */
    uint16_t totalions;  /* speclen of its bucket */
    uint32_t reserved;
} gx_partial;

/* Export the partial results of the last (batched-engine) search of `q` on
 * index shard `idx`: one gx_partial per spectrum into `rec` and the
 * n_spectra x GX_HIST_BINS survival histograms into `hist` (either may be
 * NULL). psid_offset[bucket] (NULL = 0) is added to the shard's psids, so
 * that a shard holding entries [off, off+n) of a bucket's mass-sorted list
 * reports global psids. on_device: the destinations are device memory of
 * the context's GPU (e.g. a collective's send buffer), else host memory.
 * No reference counterpart for the export itself; the reference writes its
 * partials to files (dslim_query.cpp:1005-1012, dslim_comm.cpp). */
int gx_partials_export(gx_ctx *ctx, const gx_dindex *idx, const gx_dqueries *q, const uint32_t *psid_offset,
                       gx_partial *rec, uint32_t *hist, int on_device);

/* Merge n_parts shards' partials of the same n_spectra spectra into PSMs,
 * replacing DSLIM_Score::CombineResults (dslim_score.cpp:199-387): the
 * histograms and candidate counts add, the top hit is the largest key (the
 * unsharded tie rule, so the merge of contiguous psid ranges equals the
 * unsharded search exactly; the reference prefers the first node), and the
 * tail fit and e-value run on the merged histogram. rec[k]/hist[k] are device
 * memory (on_device = 1) or host memory. Writes psms (host) and, if not
 * NULL, the merged histograms (host); device_ms (may be NULL) receives the
 * merge kernel's time. */
int gx_partials_merge(gx_ctx *ctx, int n_parts, const gx_partial *const *rec, const uint32_t *const *hist,
                      uint32_t n_spectra, const gx_search_params *p, int on_device, gx_psm *psms,
                      uint32_t *merged_hist, float *device_ms);

/* Measured read bandwidth of the context's device in GB/s (HIP events): a
 * sequential read of a 4 GiB buffer (segment_lines = 0), or reads of random
 * segments of 1, 2, 4 or 8 consecutive 128-B lines of it, the access shape
 * of the search's ion-id stream. A measured denominator for the search's
 * roofline beside the published HBM peak; no reference counterpart. */
int gx_measure_read_bandwidth(gx_ctx *ctx, int segment_lines, double *gbps);

/* Text of the last failure on this thread: for GX_ERR_DEVICE the call, its
 * source line and the HIP error string (the reference prints the CUDA error
 * and exits, cuda/driver.hpp:41-48); for a rejected search argument, which
 * one and why. */
int gx_last_error(char *buf, size_t len);

/* Name of the device the context runs on and its CU count (diagnostics). */
int gx_ctx_device_info(gx_ctx *ctx, char *name, size_t name_len, int *n_cu);

/* ---- learned e-value surrogate ----------------------------------------
 * No reference counterpart: the reference always runs expeRT::ModelTailFit
 * (expeRT.cpp:396-559), a least-squares line through a spectrum's log10
 * survival curve between two marks, expeRT.cpp:477-515 (one before the first
 * point at or below 22 % of the curve's drop, and the last point at or above
 * 87 % of it), evaluated at the top hit's bin.
 *
 * The surrogate is a 128-256-256-3 GELU MLP (BF16 operands on MFMA, FP32
 * accumulation and master weights) that predicts where those two marks lie,
 * as fractions of the curve, and a confidence logit. With evalue_mode =
 * GX_EVALUE_LEARNED the batched engine runs one fused kernel per 16 spectra:
 * features, the network, then for each spectrum whose curve is servable (at
 * least 8 points, no survival value <= 0) and whose confidence is at least
 * surrogate_conf, each mark is looked for among 32 points around its
 * predicted index; when both marks are found there, the line is fitted on
 * that window with wave-parallel sums and the e-value formed from the
 * truncated mu, beta as the plain path forms it (GX_PSM_SURROGATE set).
 * Every other modeled spectrum gets the plain fit, bit-identical to
 * GX_EVALUE_TAILFIT (stats.fallbacks counts them). Top hits, cpsms and
 * histograms are the same in both modes. A served e-value uses the plain
 * fit's own window; its mu, beta can differ from the plain fit's only through
 * the summation order of the line fit.
 *
 * Features (per spectrum, bf16): the log10 survival 1 - C/cpsms, where C is
 * the prefix count of the histogram from the first non-empty bin below the
 * top hit (stt1) to the last (end1), linearly interpolated at 120 points
 * t = j*(n-1)/119 and divided by 4; then stt1/800, end1/800, hyp/800,
 * (hyp-end1)/400, n/400, log10(cpsms)/4, the last curve value/4 and the first
 * curve value (single precision, log10 of max(x, 1e-7)). Training targets
 * per row: {mark before the mark2 rule, mark2, n-1, servable}. */
typedef struct gx_mlp gx_mlp;

#define GX_TRAIN_MAX_EPOCHS 64
typedef struct gx_train_params {
    int32_t batch;       /* minibatch rows (>= 16) */
    int32_t epochs;      /* passes over the training set (1..GX_TRAIN_MAX_EPOCHS) */
    float lr;            /* Adam step size (b1 0.9, b2 0.999, eps 1e-8) */
    float lr_decay;      /* per-epoch multiplier of lr (0 or 1 = constant) */
    float conf_weight;   /* weight of the confidence cross-entropy */
    int32_t pad;
    uint64_t seed;       /* minibatch shuffling */
} gx_train_params;

typedef struct gx_train_report {
    uint64_t samples;    /* training rows */
    int32_t epochs;      /* epochs run */
    int32_t pad;
    /* per epoch, means over the rows seen: squared errors of the two
     * normalised mark positions, confidence cross-entropy, and the fraction
     * of rows whose two marks lay in the check windows around the predicted
     * indexes (the confidence label) */
    float mse_mark[GX_TRAIN_MAX_EPOCHS];
    float mse_mark2[GX_TRAIN_MAX_EPOCHS];
    float bce[GX_TRAIN_MAX_EPOCHS];
    float in_window[GX_TRAIN_MAX_EPOCHS];
} gx_train_report;

/* Number of fp32 parameters: W1[128][256], b1[256], W2[256][256], b2[256],
 * W3[256][16], b3[16] (row-major, in that order; output columns 3..15 are
 * padding and stay zero). */
int gx_mlp_param_count(void);
int gx_mlp_create(gx_ctx *ctx, gx_mlp **out);
void gx_mlp_destroy(gx_mlp *m);
/* Glorot-uniform weights, zero biases, Adam state cleared. */
int gx_mlp_init(gx_mlp *m, uint64_t seed);
/* Set / read the parameters; `norm` (optional, 4 floats) is the output
 * normalisation {mark shift, mark scale, mark2 shift, mark2 scale}: mark
 * position (fraction of n-1) = output * scale + shift. Setting parameters
 * clears the Adam state. */
int gx_mlp_set_params(gx_mlp *m, const float *params, size_t n, const float *norm);
int gx_mlp_get_params(gx_mlp *m, float *params, size_t n, float *norm);
/* Append the spectra of a batch searched in GX_EVALUE_TAILFIT mode whose
 * curve is servable and whose plain fit is defined to the training set
 * (features and targets stay in HBM). */
int gx_mlp_add_training(gx_mlp *m, gx_dqueries *q, uint64_t *n_added);
/* Train on the GPU (forward, loss, backward and Adam kernels). The output
 * normalisation is fixed from the training targets on the first call unless
 * set by gx_mlp_set_params. */
int gx_mlp_train(gx_mlp *m, const gx_train_params *tp, gx_train_report *rep);
/* One forward + backward pass on host rows (feat n*128, target n*4) with the
 * given normalisation; writes the gradient of the mean loss
 * (gx_mlp_param_count floats), and optionally the raw outputs (n*16) and the
 * loss sums {sq. err mark, sq. err mark2, cross-entropy, rows labelled 1}. */
int gx_mlp_grad(gx_mlp *m, const float *feat, const float *target, uint32_t n, const float *norm,
                float conf_weight, float *grad, float *out, float *loss);
/* Forward pass on host rows: out = n*16 raw outputs (columns 0..2 used). */
int gx_mlp_predict(gx_mlp *m, const float *feat, uint32_t n, float *out);
/* Attach the surrogate used by GX_EVALUE_LEARNED searches (NULL detaches). */
int gx_ctx_set_surrogate(gx_ctx *ctx, gx_mlp *m);
/* Features (n*128), targets (n*4), log10 of the plain e-value (n), and the
 * valid (plain fit defined) / servable flags of a searched batch (any output
 * pointer may be NULL). Targets and log10 e are those of the batch's last
 * search, meaningful after a GX_EVALUE_TAILFIT search. */
int gx_features_download(gx_ctx *ctx, gx_dqueries *q, float *feat, float *target, float *loge, uint8_t *valid,
                         uint8_t *servable);
/* After a GX_EVALUE_LEARNED search of `q`: per spectrum the network's raw
 * outputs {mark, mark2, confidence logit} and 1.0 where the surrogate served
 * the e-value, else 0.0 (n*4 floats). */
int gx_surrogate_outputs_download(gx_ctx *ctx, gx_dqueries *q, float *out);

#ifdef __cplusplus
}
#endif

#endif /* GICOPS_AMD_H */
