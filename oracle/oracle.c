/*
 * oracle.c - CPU restatement of DSLIM_QuerySpectrum + expeRT::ModelTailFit.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the parity checker and the CPU
 * baseline, never the product path.
 *
 * Rules where the reference is undefined or non-deterministic (each one is
 * also what the HIP path implements):
 *   - top hit: the maximum f32 hyperscore; ties go to the smallest
 *     (bucket, psid), i.e. the first one met in bucket-then-psid order
 *     (the reference's top-K heap with topmatches=1 has capacity 0 and reads
 *     out of bounds, dslim.cpp:438-444, minheap.h:105-107);
 *   - tail fit: when the regression start mark would be negative (flat or
 *     empty slice, where the reference reads sx[-1] through lwvector::clip,
 *     expeRT.cpp:484,525-528) or the fit is not finite / does not fit the
 *     reference's int x1e6 storage, the fit is "undefined": no e-value;
 *   - lgfact: extended past MAX_SHDPEAKS (utils.h:110) by the same recurrence,
 *     so counts above 80 do not read out of bounds;
 *   - peaks are legal iff dF < q < max_mass*scale-1-dF as signed integers
 *     (dslim_query.cpp:860; the reference's unsigned comparison lets negative
 *     m/z through to an out-of-bounds bA read);
 *   - a spectrum needs cpsms >= max(1, min_cpsm) for an e-value (with
 *     min_cpsm = 0 the reference would read an empty heap).
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_LGFACT_N 4096

int64_t orc_lower_bound(const uint32_t *a, int64_t n, uint32_t v) {
    int64_t lo = 0, len = n;
    while (len > 0) {
        int64_t half = len >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

int64_t orc_upper_bound(const uint32_t *a, int64_t n, uint32_t v) {
    int64_t lo = 0, len = n;
    while (len > 0) {
        int64_t half = len >> 1;
        if (!(v < a[lo + half])) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

/* first i with mass[i] >= v */
static int64_t lb_f(const float *a, int64_t n, float v) {
    int64_t lo = 0, len = n;
    while (len > 0) {
        int64_t half = len >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

/* first i with mass[i] > v */
static int64_t ub_f(const float *a, int64_t n, float v) {
    int64_t lo = 0, len = n;
    while (len > 0) {
        int64_t half = len >> 1;
        if (!(v < a[lo + half])) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

/* DSLIM_BinarySearch, dslim_query.cpp:1102-1164. BinFindMin returns the
 * first entry with Mass >= pmass1 and BinFindMax the last with Mass <= pmass2
 * (dslim_query.cpp:1167-1242), i.e. lower_bound and upper_bound-1. */
int orc_window(const float *mass, uint32_t n, float precmass, double dM, int32_t *minlimit, int32_t *maxlimit) {
    const float pmass1 = (float)((double)precmass - dM);
    const float pmass2 = (float)((double)precmass + dM);
    if (n == 0) {
        *minlimit = 0;
        *maxlimit = -1;
        return 0;
    }
    const int32_t mn = 0, mx = (int32_t)n - 1;
    if (dM < 0.0) {
        *minlimit = mn;
        *maxlimit = mx;
        return 0;
    }
    if (pmass1 < mass[mn]) {
        *minlimit = mn;
    } else if (pmass1 > mass[mx]) {
        *minlimit = mx;
        *maxlimit = mx;
        return 0;
    } else {
        *minlimit = (int32_t)lb_f(mass, n, pmass1);
    }
    if (pmass2 > mass[mx]) {
        *maxlimit = mx;
    } else if (pmass2 < mass[mn]) {
        *minlimit = mn;
        *maxlimit = mn;
        return 0;
    } else {
        *maxlimit = (int32_t)ub_f(mass, n, pmass2) - 1;
    }
    /* dslim_query.cpp:1160-1161 */
    if (*maxlimit >= 0 && *minlimit < (int32_t)n && mass[*maxlimit] <= pmass2 && mass[*minlimit] >= pmass1) return 1;
    return 0;
}

void orc_lgfact(double *out, int n) {
    if (n > 0) out[0] = log10(1.0);
    if (n > 1) out[1] = log10(1.0);
    for (int i = 2; i < n; i++) out[i] = log10((double)i) + out[i - 1];
}

/* expeRT::rargmax / argmax / largmax, expeRT.cpp:1148-1201 */
static int rargmax(const double *d, int i1, int i2, double v) {
    for (int p = i2; p >= i1; p--)
        if (d[p] >= v) return p;
    return i2;
}
static int argmax_(const double *d, int i1, int i2, double v) {
    for (int p = i1; p <= i2; p++)
        if (d[p] >= v) return p;
    return i1;
}
static int largmax(const double *d, int i1, int i2, double v) {
    for (int p = i1; p <= i2; p++)
        if (d[p] <= v) return p;
    return i1;
}

int orc_tailfit(const uint32_t *hist, int hyp, int cpsms, double *slope, double *intercept, int32_t *mu,
                int32_t *beta) {
    double yy[GX_HIST_BINS];
    double sx[GX_HIST_BINS];
    for (int i = 0; i < GX_HIST_BINS; i++) yy[i] = (double)hist[i];
    *slope = 0;
    *intercept = 0;
    *mu = 0;
    *beta = 0;
    if (hyp < 0 || hyp >= GX_HIST_BINS) return 0;
    /* curve region, expeRT.cpp:414-422 */
    int end1 = rargmax(yy, 0, hyp - 1, 1.0);
    int stt1 = argmax_(yy, 0, end1, 1.0);
    if (stt1 == end1) end1 += 1;
    const int n = end1 - stt1 + 1;
    if (n < 1) return 0; /* hyp == 0: empty slice */
    const double vaa = (double)cpsms;
    if (cpsms < 1) { /* expeRT.cpp:434-441 */
        *slope = 0;
        *intercept = 100;
        *mu = 0;
        *beta = 100000000;
        return 1;
    }
    /* survival s(x) = 1 - cumsum/vaa, expeRT.cpp:453-461 */
    double acc = 0;
    for (int i = 0; i < n; i++) {
        acc += yy[stt1 + i];
        sx[i] = acc;
    }
    for (int i = 0; i < n; i++) {
        sx[i] /= vaa;
        sx[i] += -1.0;
        sx[i] = -sx[i];
    }
    for (int i = 0; i < n; i++)
        if (sx[i] > 1.0) sx[i] = 0.999;
    /* replace s(x) <= 0 by the last value >= 1e-4 (expeRT.cpp:464-468) */
    const int repl = rargmax(sx, 0, n - 1, 1e-4);
    const double rv = sx[repl];
    for (int i = 0; i < n; i++)
        if (sx[i] <= 0.0) sx[i] = rv;
    for (int i = 0; i < n; i++) sx[i] = log10(sx[i]);
    /* marks at 22% / 87% of the height, expeRT.cpp:477-515 */
    int mark = 0, mark2 = 0;
    const double hgt = sx[n - 1] - sx[0];
    if (n > 3) {
        mark = largmax(sx, 0, n - 1, sx[0] + hgt * 0.22) - 1;
        mark2 = rargmax(sx, 0, n - 1, sx[0] + hgt * 0.87);
        if (mark2 == n) mark2 -= 1;
        if (mark >= mark2) mark = mark2 - 1;
    } else if (n == 3) {
        mark = largmax(sx, 0, n - 1, sx[0] + hgt * 0.22) - 1;
        mark2 = n - 1;
        if (mark >= mark2) mark = mark2 - 1;
    } else {
        mark = 0;
        mark2 = n - 1;
    }
    if (mark < 0 || mark2 < mark) return 0; /* flat slice: reference reads sx[-1] */
    /* LinearFit on X = stt1+mark..stt1+mark2, Y = sx[mark..mark2] (expeRT.cpp:1241-1289) */
    const int m = mark2 - mark + 1;
    double a, b;
    if (m == 1) {
        a = 0.0;
        b = sx[mark];
    } else {
        double xbar = 0.0, ybar = 0.0;
        for (int i = 0; i < m; i++) {
            xbar = xbar + (double)(stt1 + mark + i);
            ybar = ybar + sx[mark + i];
        }
        xbar = xbar / (double)m;
        ybar = ybar / (double)m;
        double top = 0.0, bot = 0.0;
        for (int i = 0; i < m; i++) {
            const double xi = (double)(stt1 + mark + i);
            top = top + (xi - xbar) * (sx[mark + i] - ybar);
            bot = bot + (xi - xbar) * (xi - xbar);
        }
        a = top / bot;
        b = ybar - a * xbar;
    }
    *slope = a;
    *intercept = b;
    const double am = a * 1e6, bm = b * 1e6;
    if (!(fabs(am) < 2147483647.0) || !(fabs(bm) < 2147483647.0)) return 0; /* also rejects NaN */
    *mu = (int32_t)am;
    *beta = (int32_t)bm;
    return 1;
}

typedef struct byc {
    uint16_t bc, yc;
    uint32_t ibc, iyc;
} byc_t;

static inline int legal_peak(int32_t q, int32_t dF, int64_t nbins) {
    return q > dF && (int64_t)q < nbins - 1 - dF;
}

/* scatter of dslim_query.cpp:853-917 for one spectrum against one bucket
 * window; counters optional */
static void scatter(const gx_bucket *b, const int32_t *moz, const int32_t *inten, uint32_t npk, int32_t dF,
                    int64_t nbins, int32_t minlimit, int32_t maxlimit, byc_t *sc, orc_counters *c) {
    const uint32_t speclen = b->speclen, half = speclen / 2;
    const uint32_t lo = (uint32_t)minlimit * speclen;
    const uint32_t hi = (uint32_t)(maxlimit + 1) * speclen - 1;
    for (uint32_t k = 0; k < npk; k++) {
        const int32_t q = moz[k];
        const uint32_t intn = (uint32_t)inten[k];
        if (!legal_peak(q, dF, nbins)) continue;
        for (int32_t bin = q - dF; bin <= q + dF; bin++) {
            const uint32_t start = b->bA[bin], end = b->bA[bin + 1];
            if (c) c->bins++;
            if (end - start < 1) continue;
            if (c) c->bin_ions += end - start;
            const int64_t stt = start + orc_lower_bound(b->iA + start, end - start, lo);
            const int64_t ends = stt + orc_upper_bound(b->iA + stt, end - stt, hi) - 1;
            if (c && ends >= stt) c->ions += (uint64_t)(ends - stt + 1);
            for (int64_t ion = stt; ion <= ends; ion++) {
                const uint32_t raw = b->iA[ion];
                const uint32_t ppid = raw / speclen;
                const uint32_t isY = (raw % speclen) / half;
                const uint32_t isB = 1 - isY;
                byc_t *e = sc + (ppid - (uint32_t)minlimit);
                e->bc += isB;
                e->ibc += intn * isB;
                e->yc += isY;
                e->iyc += intn * isY;
            }
        }
    }
}

int orc_scorecard(const gx_bucket *b, const int32_t *moz, const int32_t *inten, uint32_t npk,
                  const gx_search_params *p, int32_t minlimit, int32_t maxlimit, uint16_t *bc, uint16_t *yc,
                  uint32_t *ibc, uint32_t *iyc) {
    if (!b || !p || maxlimit < minlimit || minlimit < 0 || (uint32_t)maxlimit >= b->n_entries) return GX_ERR_INVLD_PARAM;
    const int64_t w = (int64_t)maxlimit - minlimit + 1;
    byc_t *sc = (byc_t *)calloc((size_t)w, sizeof(byc_t));
    if (!sc) return GX_ERR_BAD_MEM_ALLOC;
    scatter(b, moz, inten, npk, p->dF, (int64_t)p->max_mass * p->scale, minlimit, maxlimit, sc, NULL);
    for (int64_t i = 0; i < w; i++) {
        bc[i] = sc[i].bc;
        yc[i] = sc[i].yc;
        ibc[i] = sc[i].ibc;
        iyc[i] = sc[i].iyc;
    }
    free(sc);
    return GX_SUCCESS;
}

int orc_search(const gx_bucket *buckets, int n_buckets, const gx_query_batch *q, const gx_search_params *p,
               gx_psm *out, uint32_t *hist, orc_counters *cnt, int threads) {
    if (!buckets || !q || !p || !out) return GX_ERR_INVLD_PTR;
    if (p->dF < 0 || p->scale < 1 || p->max_mass < 1 || n_buckets < 0 || n_buckets > 65535) return GX_ERR_INVLD_PARAM;
    const int64_t nbins = (int64_t)p->max_mass * p->scale;
    uint32_t maxent = 0;
    int maxlen_cnt = 1;
    for (int i = 0; i < n_buckets; i++) {
        if (buckets[i].n_entries > maxent) maxent = buckets[i].n_entries;
        if (buckets[i].n_entries && buckets[i].nbins != (uint32_t)nbins) return GX_ERR_INVLD_INDEX;
        /* bc or yc is at most min(peaks, (2dF+1)*speclen/2) */
        const int cap = (2 * p->dF + 1) * (int)(buckets[i].speclen / 2) + 1;
        if (cap > maxlen_cnt) maxlen_cnt = cap;
    }
    if (maxlen_cnt > ORC_LGFACT_N) return GX_ERR_INVLD_SIZE;
    double lgf[ORC_LGFACT_N];
    orc_lgfact(lgf, ORC_LGFACT_N);
    const int nthr = threads > 0 ? threads : omp_get_max_threads();
    const uint32_t min_cpsm = p->min_cpsm > 1 ? (uint32_t)p->min_cpsm : 1u;
    int status = GX_SUCCESS;

#pragma omp parallel num_threads(nthr)
    {
        byc_t *sc = (byc_t *)calloc(maxent ? maxent : 1, sizeof(byc_t));
        uint32_t *h = (uint32_t *)calloc(GX_HIST_BINS, sizeof(uint32_t));
        if (!sc || !h) {
#pragma omp atomic write
            status = GX_ERR_BAD_MEM_ALLOC;
        }
#pragma omp for schedule(dynamic, 4)
        for (int64_t s = 0; s < (int64_t)q->n_spectra; s++) {
            if (!sc || !h) continue;
            const uint32_t off = q->idx[s];
            const uint32_t npk = q->idx[s + 1] - off;
            const int32_t *moz = q->moz + off;
            const int32_t *inten = q->intensity + off;
            const float pmass = q->precurse[s];
            const double dM = q->dm ? (double)q->dm[s] : p->dM;
            orc_counters c;
            memset(&c, 0, sizeof(c));
            for (uint32_t k = 0; k < npk; k++) c.legal += legal_peak(moz[k], p->dF, nbins);
            memset(h, 0, sizeof(uint32_t) * GX_HIST_BINS);
            uint32_t cpsms = 0, ties = 0;
            float best = 0.0f;
            int32_t best_psid = -1;
            int best_b = -1;
            uint16_t best_shp = 0, best_tot = 0;
            for (int ixx = 0; ixx < n_buckets; ixx++) {
                const gx_bucket *b = buckets + ixx;
                if (b->n_entries == 0) continue;
                int32_t minlimit, maxlimit;
                const int val = orc_window(b->mass, b->n_entries, pmass, dM, &minlimit, &maxlimit);
                if (!val || maxlimit < minlimit) continue;
                c.buckets++;
                c.window += (uint32_t)(maxlimit - minlimit + 1);
                scatter(b, moz, inten, npk, p->dF, nbins, minlimit, maxlimit, sc, &c);
                /* candidate scan, dslim_query.cpp:919-966 */
                for (int32_t it = minlimit; it <= maxlimit; it++) {
                    byc_t *e = sc + (it - minlimit);
                    const uint32_t bcc = e->bc, ycc = e->yc;
                    const uint32_t shpk = bcc + ycc;
                    if (shpk >= (uint32_t)p->min_shp) {
                        const double h1 = lgf[bcc] + lgf[ycc];
                        float hs = (float)(h1 + log10((double)(1u + e->ibc)) + log10((double)(1u + e->iyc)) - 4);
                        if (hs > 0) {
                            if (hs >= 80.0f) hs = 80.0f - 1;
                            cpsms++;
                            h[(int)((double)(hs * 10.0f) + 0.5)] += 1;
                            if (hs > best) {
                                best = hs;
                                best_psid = it;
                                best_b = ixx;
                                best_shp = (uint16_t)shpk;
                                best_tot = (uint16_t)b->speclen;
                                ties = 0;
                            } else if (hs == best) {
                                ties++;
                            }
                        }
                    }
                }
                memset(sc, 0, sizeof(byc_t) * (size_t)(maxlimit - minlimit + 1));
            }
            gx_psm r;
            memset(&r, 0, sizeof(r));
            r.psid = -1;
            r.evalue = INFINITY;
            r.cpsms = (int32_t)cpsms;
            if (cpsms >= 1) {
                r.flags |= GX_PSM_HAS_HIT;
                r.hyperscore = best;
                r.psid = best_psid;
                r.idxoffset = (uint16_t)best_b;
                r.sharedions = best_shp;
                r.totalions = best_tot;
                r.hyp = (int32_t)((double)(best * 10.0f) + 0.5);
                if (ties) r.flags |= GX_PSM_TIE;
            }
            if (cpsms >= min_cpsm) {
                /* dslim_query.cpp:1017-1061 */
                double sl, ic;
                int32_t mu, beta;
                r.flags |= GX_PSM_MODELED;
                if (orc_tailfit(h, r.hyp, (int)cpsms, &sl, &ic, &mu, &beta)) {
                    r.flags |= GX_PSM_FIT_OK;
                    r.mu = mu;
                    r.beta = beta;
                    double w = mu, bb = beta;
                    w /= 1e6;
                    bb /= 1e6;
                    const double lgs = (w * r.hyp) + bb;
                    double e = pow(10, lgs);
                    e *= (double)cpsms;
                    r.evalue = e;
                    if (e < p->expect_max) r.flags |= GX_PSM_REPORTED;
                }
            }
            out[s] = r;
            if (hist) memcpy(hist + (size_t)s * GX_HIST_BINS, h, sizeof(uint32_t) * GX_HIST_BINS);
            if (cnt) cnt[s] = c;
        }
        free(sc);
        free(h);
    }
    return status;
}

/* ---- peak picking (MSQuery::pickpeaks, msquery.cpp:338-423) ----------- */

/* KeyVal_Parallel_Internal (keyval.h:49-103) without the OpenMP tasks: the
 * recursion visits the same subranges with the same partition steps. */
static void keyval_qsort(int32_t *key, int32_t *val, int64_t left, int64_t right) {
    int64_t i = left, j = right;
    const int32_t pivotkey = key[(left + right) / 2];
    while (i <= j) {
        while (key[i] < pivotkey) i++;
        while (key[j] > pivotkey) j--;
        if (i <= j) {
            const int32_t tk = key[i], tv = val[i];
            key[i] = key[j];
            val[i] = val[j];
            key[j] = tk;
            val[j] = tv;
            i++;
            j--;
        }
    }
    if (left < j) keyval_qsort(key, val, left, j);
    if (i < right) keyval_qsort(key, val, i, right);
}

/* x86-64 cvttsd2si: truncation, INT32_MIN when NaN or out of range */
static int32_t cvt_i32(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

int64_t orc_pickpeaks(int32_t *intn, int32_t *mz, int64_t n, int32_t base_int, int32_t min_int, int32_t qalen) {
    if (n <= 0) return 0;
    keyval_qsort(intn, mz, 0, n - 1);
    const double factor = (double)base_int / intn[n - 1];
    intn[n - 1] = base_int;
    int64_t newlen = 1;
    for (int64_t j = n - 2; j >= n - qalen && j >= 0; j--) {
        intn[j] = cvt_i32(intn[j] * factor);
        if (intn[j] >= min_int)
            newlen++;
        else
            break;
    }
    if (newlen >= qalen) newlen = qalen;
    return newlen;
}

