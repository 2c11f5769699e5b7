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
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
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
    if (pm1 < m0)
        lo = 0;
    else if (pm1 > mN)
        return false;
    else
        lo = (int32_t)lower_bound_f32(mass, n, pm1);
    if (pm2 > mN)
        hi = (int32_t)n - 1;
    else if (pm2 < m0)
        return false;
    else
        hi = (int32_t)upper_bound_f32(mass, n, pm2) - 1;
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

struct FitScratch {
    int32_t i[8];
    double d[4];
};

// expeRT::ModelTailFit (expeRT.cpp:396-559) on a histogram in LDS. Called by
// every thread of the block; wave 0 does the work, separated by block
// barriers. The cumulative sum is over integer counts, so its parallel order
// is exact; the regression sums run in lane 0 in the reference's order.
// `sx` is LDS scratch for GX_HIST_BINS doubles. Returns fit_ok.
__device__ bool tail_fit_block(const uint32_t *hist, double *sx, FitScratch &F, int hyp, int cpsms, double &slope,
                               double &icept, int32_t &mu, int32_t &beta) {
    const int tid = threadIdx.x, lane = tid & 63;
    const bool w0 = tid < 64;
    slope = icept = 0;
    mu = beta = 0;
    if (hyp < 0 || hyp >= GX_HIST_BINS || cpsms < 1) return false;
    if (w0) {
        int e = -1;
        for (int p = lane; p <= hyp - 1; p += 64)
            if (hist[p] >= 1) e = max(e, p);
        e = wave_max_i32(e);
        const int end1 = (e >= 0) ? e : hyp - 1;
        int s = 0x7fffffff;
        for (int p = lane; p <= end1; p += 64)
            if (hist[p] >= 1) s = min(s, p);
        s = wave_min_i32(s);
        const int stt1 = (s != 0x7fffffff) ? s : 0;
        if (lane == 0) {
            F.i[0] = stt1;
            F.i[1] = (stt1 == end1) ? end1 + 1 : end1;
        }
    }
    __syncthreads();
    const int stt1 = F.i[0], end1 = F.i[1];
    const int n = end1 - stt1 + 1;
    if (n < 1) return false;
    const double vaa = (double)cpsms;
    if (w0) {
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
            carry = __shfl(inc, 63, 64);
        }
    }
    __syncthreads();
    if (w0) {
        int r = -1;
        for (int i = lane; i < n; i += 64)
            if (sx[i] >= 1e-4) r = max(r, i);
        r = wave_max_i32(r);
        const int repl = (r >= 0) ? r : n - 1;
        if (lane == 0) F.d[2] = sx[repl];
    }
    __syncthreads();
    if (w0) {
        const double rv = F.d[2];
        for (int i = lane; i < n; i += 64) {
            double x = sx[i];
            if (x <= 0.0) x = rv;
            sx[i] = log10(x);
        }
    }
    __syncthreads();
    if (w0) {
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
        if (lane == 0) {
            int ok = 0;
            double a = 0, b = 0;
            if (mark >= 0 && mark2 >= mark) {
                const int m = mark2 - mark + 1;
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
                const double am = a * 1e6, bm = b * 1e6;
                ok = (fabs(am) < 2147483647.0 && fabs(bm) < 2147483647.0) ? 1 : 0;
                F.i[5] = ok ? (int32_t)am : 0;
                F.i[6] = ok ? (int32_t)bm : 0;
            }
            F.i[4] = ok;
            F.d[0] = a;
            F.d[1] = b;
        }
    }
    __syncthreads();
    const bool ok = F.i[4] != 0;
    slope = F.d[0];
    icept = F.d[1];
    mu = ok ? F.i[5] : 0;
    beta = ok ? F.i[6] : 0;
    __syncthreads();
    return ok;
}

}  // namespace gx
