/*
 * oracle.h - CPU restatement of the reference's fragment-ion search and tail
 * fit. TEST INFRASTRUCTURE ONLY: the checker for the HIP path (tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg). Nothing in the
 * product links or calls it.
 *
 * Pinned by (tests/test_oracle_*.py):
 *   orc_lower_bound/orc_upper_bound  testapp known-answer vectors
 *                                    (source/apps/testapp/testapp.cpp:45,60-81)
 *                                    and brute force;
 *   orc_window                       brute force over float masses;
 *   orc_search (counts, candidates)  brute-force b/y matching per (spectrum,
 *                                    entry) from the inverted index;
 *   hyperscore                       closed form in (bc, yc, ibc, iyc);
 *   lgfact                           high-precision log10 sums;
 *   orc_tailfit                      textbook least squares on hand-built
 *                                    survival histograms, every branch
 *                                    (n = 2, 3, > 3, stt1 == end1, the
 *                                    replacement paths, the flat slice);
 *   orc_pickpeaks                    hand-traced partition orders of tied
 *                                    intensities, numpy's sort for distinct
 *                                    ones, the normalisation in closed form.
 */
#ifndef GICOPS_ORACLE_H
#define GICOPS_ORACLE_H

#include <stdint.h>

#include "gicops_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-spectrum work counters (the terms of the algorithmic-bytes formula). */
typedef struct orc_counters {
    uint64_t ions;      /* H: sum of window-restricted iA slice lengths */
    uint64_t bin_ions;  /* sum of whole-bin lengths bA[b+1]-bA[b] of the same triples */
    uint32_t bins;      /* B: (bucket, legal peak, bin) triples whose bA pair is read */
    uint32_t window;    /* W: entries inside the precursor windows, all buckets */
    uint32_t legal;     /* legal peaks */
    uint32_t buckets;   /* buckets with a non-empty window */
} orc_counters;

/* first i in [0,n) with a[i] >= v (std::lower_bound) */
int64_t orc_lower_bound(const uint32_t *a, int64_t n, uint32_t v);
/* first i in [0,n) with a[i] > v (std::upper_bound) */
int64_t orc_upper_bound(const uint32_t *a, int64_t n, uint32_t v);

/* DSLIM_BinarySearch (dslim_query.cpp:1102-1242) with dM applied as
 * (float)(p -/+ dM). Returns 1 and the inclusive window when the spectrum is
 * searched in this bucket (the caller still skips maxlimit < minlimit). */
int orc_window(const float *mass, uint32_t n, float precmass, double dM, int32_t *minlimit, int32_t *maxlimit);

/* lgfact[i] = sum_{j=2..i} log10(j) (utils.h:115-131), for i < n. */
void orc_lgfact(double *out, int n);

/* ModelTailFit (expeRT.cpp:396-559) on an 802-bin histogram of counts.
 * Returns 1 when the fit is defined, 0 for the flat-slice / empty-slice
 * cases the reference leaves undefined. mu/beta are the x1e6 truncated ints. */
int orc_tailfit(const uint32_t *hist, int hyp, int cpsms, double *slope, double *intercept, int32_t *mu,
                int32_t *beta);

/* DSLIM_QuerySpectrum (dslim_query.cpp:749-1081) over a whole batch with
 * OpenMP (schedule(dynamic,4)); `threads` <= 0 uses all. hist (optional)
 * receives n_spectra*GX_HIST_BINS counts; cnt (optional) per-spectrum
 * counters. Returns GX_SUCCESS or a GX_ERR_* code. */
int orc_search(const gx_bucket *buckets, int n_buckets, const gx_query_batch *q, const gx_search_params *p,
               gx_psm *out, uint32_t *hist, orc_counters *cnt, int threads);

/* Per-entry b/y counts and intensity sums of one spectrum against one bucket
 * window, straight from the scatter (for fine-grained tests): writes
 * bc/yc/ibc/iyc for psid in [minlimit, maxlimit] into arrays of that size. */
int orc_scorecard(const gx_bucket *b, const int32_t *moz, const int32_t *inten, uint32_t npk,
                  const gx_search_params *p, int32_t minlimit, int32_t maxlimit, uint16_t *bc, uint16_t *yc,
                  uint32_t *ibc, uint32_t *iyc);

/* MSQuery::pickpeaks (msquery.cpp:338-423) on one raw spectrum, in place:
 * KeyVal_Parallel's quicksort by intensity (keyval.h:49-103), the base peak
 * set to base_int, the next qalen-1 scaled by base_int/max (int *= double,
 * x86 conversion) while >= min_int. Returns the kept count; the kept peaks
 * are the last ones of the arrays. */
int64_t orc_pickpeaks(int32_t *intn, int32_t *mz, int64_t n, int32_t base_int, int32_t min_int, int32_t qalen);

#ifdef __cplusplus
}
#endif

#endif /* GICOPS_ORACLE_H */
