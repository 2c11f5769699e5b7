// Device helpers shared by the search engines: bound searches, wave
// reductions, the precursor window, the hyperscore and the tail fit.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "gicops_amd.h"
#include "search_common.h"

namespace gx {

// Index arrays reached through a pointer loaded from memory (DevBucket) are
// generic pointers to the compiler, whose loads become FLAT loads; those also
// count in the LDS counter, so every LDS wait then waits for them too. These
// views make them global loads (vector memory counter only).
template <class T>
using gmem = const __attribute__((address_space(1))) T;

template <class T>
__device__ __forceinline__ gmem<T> *gptr(const T *p) {
    return (gmem<T> *)p;
}

// Block barrier that orders LDS only. __syncthreads() also waits for every
// outstanding global load (its release fence covers global memory), which
// would stall loads issued ahead of time; this one leaves them in flight.
// Use it only where no global data written by one thread is read by another.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class P>
__device__ __forceinline__ uint32_t lower_bound_u32(P a, uint32_t n, uint32_t v) {
    uint32_t lo = 0;
    while (n > 0) {
        uint32_t half = n >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

template <class P>
__device__ __forceinline__ uint32_t lower_bound_f32(P a, uint32_t n, float v) {
    uint32_t lo = 0;
    while (n > 0) {
        uint32_t half = n >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

template <class P>
__device__ __forceinline__ uint32_t upper_bound_f32(P a, uint32_t n, float v) {
    uint32_t lo = 0;
    while (n > 0) {
        uint32_t half = n >> 1;
        if (!(v < a[lo + half])) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

template <class T>
__device__ __forceinline__ int lower_bound_lds(const T *a, int n, T v) {
    int lo = 0;
    while (n > 0) {
        int half = n >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

template <class T>
__device__ __forceinline__ int upper_bound_lds(const T *a, int n, T v) {
    int lo = 0;
    while (n > 0) {
        int half = n >> 1;
        if (!(v < a[lo + half])) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    return v;
}

// inclusive prefix sum across a wave
// lane l's double in every lane (two v_readlane, no LDS)
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// inclusive wave scan on DPP moves (no LDS traffic): shifts within each row of
// 16 lanes, then row 15's and row 31's last lanes broadcast into the rows
// above them. Call with every lane of the wave active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
    const int rl = lane & 15;
    uint32_t t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    if (rl >= 1) v += t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    if (rl >= 2) v += t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    if (rl >= 4) v += t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    if (rl >= 8) v += t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xf, 0xf, false);  // row_bcast:15
    if ((lane & 31) >= 16) v += t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xf, 0xf, false);  // row_bcast:31
    if (lane >= 32) v += t;
    return v;
}

// top-hit order: higher f32 hyperscore, then smaller (bucket, psid)
__device__ __forceinline__ bool better(float h, int b, int p, float bh, int bb, int bp) {
    if (h != bh) return h > bh;
    if (b != bb) return b < bb;
    return p < bp;
}

// Packed top-hit key: larger key = better hit. The hyperscore is positive,
// so its bit pattern orders like the value; the low word is the complement
// of (bucket << 26 | psid) so that ties prefer the smaller bucket and psid.
__device__ __forceinline__ unsigned long long hit_key(float h, uint32_t b, uint32_t psid) {
    return ((unsigned long long)__float_as_uint(h) << 32) | (unsigned long long)(~((b << 26) | psid));
}

// DSLIM_BinarySearch (dslim_query.cpp:1102-1242): the inclusive window
// [lo, hi] of entries with (float)(p - dM) <= mass <= (float)(p + dM).
__device__ __forceinline__ bool precursor_window(const DevBucket &bk, float prec, double dM, int32_t &lo,
                                                 int32_t &hi) {
    lo = 0;
    hi = -1;
    const uint32_t n = bk.n_entries;
    if (n == 0 || dM < 0.0) return false;
    const float pm1 = (float)((double)prec - dM);
    const float pm2 = (float)((double)prec + dM);
    const auto mass = gptr(bk.mass);
    const float m0 = mass[0], mN = mass[n - 1];
    if (!(pm1 < m0) && pm1 > mN) return false;
    if (!(pm2 > mN) && pm2 < m0) return false;
    // lower_bound of pm1 and upper_bound of pm2 in one loop, their loads
    // independent (one latency per step instead of two)
    const bool s0 = !(pm1 < m0), s1 = !(pm2 > mN);
    uint32_t b0 = 0, n0 = s0 ? n : 0u, b1 = 0, n1 = s1 ? n : 0u;
    while (n0 > 0 || n1 > 0) {
        const uint32_t h0 = n0 >> 1, h1 = n1 >> 1;
        const float v0 = n0 > 0 ? mass[b0 + h0] : 0.0f, v1 = n1 > 0 ? mass[b1 + h1] : 0.0f;
        if (n0 > 0) {
            if (v0 < pm1) {
                b0 += h0 + 1;
                n0 -= h0 + 1;
            } else {
                n0 = h0;
            }
        }
        if (n1 > 0) {
            if (!(pm2 < v1)) {
                b1 += h1 + 1;
                n1 -= h1 + 1;
            } else {
                n1 = h1;
            }
        }
    }
    lo = s0 ? (int32_t)b0 : 0;
    hi = s1 ? (int32_t)b1 - 1 : (int32_t)n - 1;
    const bool ok = hi >= 0 && lo < (int32_t)n && mass[hi] <= pm2 && mass[lo] >= pm1 && hi >= lo;
    if (!ok) hi = lo - 1;
    return ok;
}

// hyperscore = lgfact[bc] + lgfact[yc] + log10(1+ibc) + log10(1+iyc) - 4 in
// double, stored as f32 (dslim_query.cpp:935-938)
__device__ __forceinline__ float hyperscore(uint32_t bc, uint32_t yc, uint32_t ibc, uint32_t iyc,
                                            const double *lgfact) {
    const double h1 = lgfact[bc] + lgfact[yc];
    return (float)(h1 + log10((double)(1u + ibc)) + log10((double)(1u + iyc)) - 4);
}

// histogram bin / maxhypscore: (int)(h*10 + 0.5) with h*10 in float
__device__ __forceinline__ int hyp_bin(float h) { return (int)((double)(h * 10.0f) + 0.5); }

// e = cpsms * 10^((mu/1e6)*hyp + beta/1e6) (dslim_query.cpp:1023-1041)
__device__ __forceinline__ double evalue_of(int32_t mu, int32_t beta, int32_t hyp, uint32_t cpsms) {
    double w = mu, b = beta;
    w /= 1e6;
    b /= 1e6;
    const double lgs = (w * hyp) + b;
    double e = pow(10.0, lgs);
    e *= (double)cpsms;
    return e;
}

// The per-spectrum result before its e-value: the top hit unpacked from its
// packed key, its tie flag, candidates, and whether the tail fit runs
// (cpsms >= need, dslim_query.cpp:1017-1021).
__device__ inline gx_psm top_hit_psm(unsigned long long key, uint32_t eq, uint32_t shp, uint32_t cps,
                                     const DevBucket *buckets, uint32_t need) {
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
    if (cps >= 1 && key) {
        const uint32_t low = ~(uint32_t)(key & 0xffffffffu);
        const uint32_t b = low >> 26, psid = low & ((1u << 26) - 1);
        r.flags |= GX_PSM_HAS_HIT;
        r.hyperscore = __uint_as_float((uint32_t)(key >> 32));
        r.psid = (int32_t)psid;
        r.idxoffset = (uint16_t)b;
        r.sharedions = (uint16_t)shp;
        r.totalions = (uint16_t)buckets[b].speclen;
        r.hyp = hyp_bin(r.hyperscore);
        if (eq > 1) r.flags |= GX_PSM_TIE;
    }
    if (cps >= need && r.psid >= 0) r.flags |= GX_PSM_MODELED;
    return r;
}

struct FitScratch {
    int32_t i[8];
    double d[4];
};

// Orders one wave's LDS writes before its later LDS reads by other lanes of
// the same wave (the LDS serves a wave's operations in order; the fences keep
// the compiler from moving memory operations across this point).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// expeRT::ModelTailFit (expeRT.cpp:396-559) by one whole wave, on a histogram
// in LDS or global memory. The cumulative sum is over integer counts, so its
// parallel order is exact; the regression sums run in the reference's order
// (every lane runs the same sequential sums). `sx` is LDS scratch for
// GX_HIST_BINS doubles. Returns fit_ok; results are the same in every lane.
__device__ bool tail_fit_wave(const uint32_t *hist, double *sx, int hyp, int cpsms, double &slope, double &icept,
                              int32_t &mu, int32_t &beta) {
    const int lane = threadIdx.x & 63;
    slope = icept = 0;
    mu = beta = 0;
    if (hyp < 0 || hyp >= GX_HIST_BINS || cpsms < 1) return false;
    int e = -1;
    for (int p = lane; p <= hyp - 1; p += 64)
        if (hist[p] >= 1) e = max(e, p);
    e = wave_max_i32(e);
    const int end0 = (e >= 0) ? e : hyp - 1;
    int s = 0x7fffffff;
    for (int p = lane; p <= end0; p += 64)
        if (hist[p] >= 1) s = min(s, p);
    s = wave_min_i32(s);
    const int stt1 = (s != 0x7fffffff) ? s : 0;
    const int end1 = (stt1 == end0) ? end0 + 1 : end0;
    const int n = end1 - stt1 + 1;
    if (n < 1) return false;
    const double vaa = (double)cpsms;
    uint32_t carry = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        uint32_t v = (i < n && stt1 + i < GX_HIST_BINS) ? hist[stt1 + i] : 0u;
        uint32_t inc = wave_incl_scan(v, lane) + carry;
        if (i < n) {
            double x = (double)inc;
            x /= vaa;
            x += -1.0;
            x = -x;
            if (x > 1.0) x = 0.999;
            sx[i] = x;
        }
        carry = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    }
    int r = -1;
    for (int i = lane; i < n; i += 64)
        if (sx[i] >= 1e-4) r = max(r, i);
    r = wave_max_i32(r);
    const int repl = (r >= 0) ? r : n - 1;
    wave_lds_sync();
    const double rv = sx[repl];
    for (int i = lane; i < n; i += 64) {
        double x = sx[i];
        if (x <= 0.0) x = rv;
        sx[i] = log10(x);
    }
    wave_lds_sync();
    const double s0 = sx[0];
    const double hgt = sx[n - 1] - s0;
    int mark = 0, mark2 = 0;
    if (n >= 3) {
        const double t1 = s0 + hgt * 0.22;
        int f = 0x7fffffff;
        for (int i = lane; i < n; i += 64)
            if (sx[i] <= t1) f = min(f, i);
        f = wave_min_i32(f);
        mark = ((f != 0x7fffffff) ? f : 0) - 1;
        if (n > 3) {
            const double t2 = s0 + hgt * 0.87;
            int g = -1;
            for (int i = lane; i < n; i += 64)
                if (sx[i] >= t2) g = max(g, i);
            g = wave_max_i32(g);
            mark2 = (g >= 0) ? g : n - 1;
            if (mark2 == n) mark2 -= 1;
        } else {
            mark2 = n - 1;
        }
        if (mark >= mark2) mark = mark2 - 1;
    } else {
        mark = 0;
        mark2 = n - 1;
    }
    // Burkardt's llsq (expeRT.cpp:1241-1289) in the reference's sequential
    // order. Every lane runs the same sums; the curve values are read 8 at a
    // time (each read is an LDS broadcast: all lanes, one address) ahead of
    // their 8 dependent additions, so the sums wait on the additions, not on
    // a read per step.
    if (!(mark >= 0 && mark2 >= mark)) return false;
    // the bounds are the same in every lane; as scalars the loops below run
    // on scalar control flow
    mark = __builtin_amdgcn_readfirstlane(mark);
    mark2 = __builtin_amdgcn_readfirstlane(mark2);
    const int x0 = __builtin_amdgcn_readfirstlane(stt1) + mark;
    const int m = mark2 - mark + 1;
    double a = 0, b = 0;
    if (m == 1) {
        a = 0.0;
        b = sx[mark];
    } else {
        double xbar = 0.0, ybar = 0.0;
        int i = 0;
        for (; i + 8 <= m; i += 8) {
            double y[8];
#pragma unroll
            for (int u = 0; u < 8; u++) y[u] = sx[mark + i + u];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                xbar = xbar + (double)(x0 + i + u);
                ybar = ybar + y[u];
            }
        }
        for (; i < m; i++) {
            xbar = xbar + (double)(x0 + i);
            ybar = ybar + sx[mark + i];
        }
        xbar = xbar / (double)m;
        ybar = ybar / (double)m;
        double top = 0.0, bot = 0.0;
        i = 0;
        for (; i + 8 <= m; i += 8) {
            double y[8];
#pragma unroll
            for (int u = 0; u < 8; u++) y[u] = sx[mark + i + u];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double xi = (double)(x0 + i + u);
                top = top + (xi - xbar) * (y[u] - ybar);
                bot = bot + (xi - xbar) * (xi - xbar);
            }
        }
        for (; i < m; i++) {
            const double xi = (double)(x0 + i);
            top = top + (xi - xbar) * (sx[mark + i] - ybar);
            bot = bot + (xi - xbar) * (xi - xbar);
        }
        a = top / bot;
        b = ybar - a * xbar;
    }
    slope = a;
    icept = b;
    const double am = a * 1e6, bm = b * 1e6;
    if (!(fabs(am) < 2147483647.0 && fabs(bm) < 2147483647.0)) return false;
    mu = (int32_t)am;
    beta = (int32_t)bm;
    return true;
}

// The same for every thread of a block: wave 0 fits, the others wait and
// read the result.
__device__ bool tail_fit_block(const uint32_t *hist, double *sx, FitScratch &F, int hyp, int cpsms, double &slope,
                               double &icept, int32_t &mu, int32_t &beta) {
    if (threadIdx.x < 64) {
        const bool ok = tail_fit_wave(hist, sx, hyp, cpsms, slope, icept, mu, beta);
        if (threadIdx.x == 0) {
            F.i[4] = ok ? 1 : 0;
            F.i[5] = mu;
            F.i[6] = beta;
            F.d[0] = slope;
            F.d[1] = icept;
        }
    }
    __syncthreads();
    const bool ok = F.i[4] != 0;
    slope = F.d[0];
    icept = F.d[1];
    mu = F.i[5];
    beta = F.i[6];
    __syncthreads();
    return ok;
}

}  // namespace gx
