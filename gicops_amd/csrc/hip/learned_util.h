// Device functions of the learned e-value stage shared by the training-row
// kernel and the fused inference kernel, so that both see the same curve,
// the same features and the same window rules. All are called by one whole
// wave (64 lanes) with wave-uniform arguments.
#pragma once

#include "device_util.h"
#include "mlp.h"

namespace gx {

// GELU in its tanh form, 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
// (Hendrycks & Gimpel, 2016), written as z * sigmoid(2u): one hardware exp and
// one reciprocal instead of an erf series.
constexpr float kGeluK = 1.5957691216057308f;  // 2 * sqrt(2 / pi)
constexpr float kGeluC = 0.044715f;

__device__ __forceinline__ float gelu_sigmoid(float z) {
    const float t = kGeluK * (z + kGeluC * z * z * z);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-t));
}

__device__ __forceinline__ float mlp_gelu(float z) { return z * gelu_sigmoid(z); }

__device__ __forceinline__ float mlp_gelu_grad(float z) {
    const float s = gelu_sigmoid(z);
    return s + z * s * (1.0f - s) * kGeluK * (1.0f + 3.0f * kGeluC * z * z);
}

// The tail fit's curve bounds (expeRT.cpp:414-452): stt1 = first non-empty
// bin below the top hit's bin hyp, end1 = last one (hyp - 1 and 0 when there
// is none); end1 += 1 when the two are equal. n = end1 - stt1 + 1.
struct Curve {
    int stt1, end1, n;
    uint32_t nonempty;  // non-empty bins below hyp
};

__device__ inline Curve curve_bounds(const uint32_t *h, int hyp, int lane) {
    int e = -1, st = 0x7fffffff;
    uint32_t ne = 0;
    for (int p = lane; p < hyp; p += 64) {
        if (h[p] >= 1) {
            e = max(e, p);
            st = min(st, p);
            ne++;
        }
    }
    e = wave_max_i32(e);
    st = wave_min_i32(st);
    Curve c;
    c.nonempty = wave_sum_u32(ne);
    c.end1 = e >= 0 ? e : hyp - 1;
    c.stt1 = st != 0x7fffffff ? st : 0;
    if (c.stt1 == c.end1) c.end1 += 1;
    c.n = c.end1 - c.stt1 + 1;
    return c;
}

constexpr int kHistChunks = (GX_HIST_BINS + 63) / 64;  // 64-bin chunks of a histogram

// cpre[i] = sum of h[stt1 .. stt1 + i] for i < n (exact integer prefix). All
// of the curve's loads are issued before the scan uses any of them.
__device__ inline void curve_prefix(const uint32_t *h, const Curve &c, int lane, uint32_t *cpre) {
    uint32_t v[kHistChunks];
#pragma unroll
    for (int k = 0; k < kHistChunks; k++) {
        const int i = k * 64 + lane;
        v[k] = (i < c.n && c.stt1 + i < GX_HIST_BINS) ? h[c.stt1 + i] : 0u;
    }
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < kHistChunks; k++) {
        if (k * 64 >= c.n) break;
        const int i = k * 64 + lane;
        const uint32_t inc = wave_incl_scan(v[k], lane) + carry;
        if (i < c.n) cpre[i] = inc;
        carry = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    }
}

// The curve in one pass over bins 0..hyp: pre[b] = sum of h[0..b] for b <=
// hyp (b < GX_HIST_BINS), and the bounds. Bins below stt1 are empty, so the
// curve's prefix counts are pre + stt1 (the same integers curve_prefix forms).
// All loads are issued before the scan uses any of them.
__device__ inline Curve curve_load(const uint32_t *h, int hyp, int lane, uint32_t *pre) {
    const int top = min(hyp + 1, GX_HIST_BINS);
    uint32_t v[kHistChunks];
#pragma unroll
    for (int k = 0; k < kHistChunks; k++) {
        const int b = k * 64 + lane;
        v[k] = b < top ? h[b] : 0u;
    }
    int e = -1, st = 0x7fffffff;
    uint32_t ne = 0, carry = 0;
#pragma unroll
    for (int k = 0; k < kHistChunks; k++) {
        if (k * 64 >= top) break;
        const int b = k * 64 + lane;
        if (v[k] >= 1 && b < hyp) {
            e = max(e, b);
            st = min(st, b);
            ne++;
        }
        const uint32_t inc = wave_incl_scan(v[k], lane) + carry;
        if (b < top) pre[b] = inc;
        carry = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    }
    e = wave_max_i32(e);
    st = wave_min_i32(st);
    Curve c;
    c.nonempty = wave_sum_u32(ne);
    c.end1 = e >= 0 ? e : hyp - 1;
    c.stt1 = st != 0x7fffffff ? st : 0;
    if (c.stt1 == c.end1) c.end1 += 1;
    c.n = c.end1 - c.stt1 + 1;
    return c;
}

// log10 of the survival at prefix count `cnt` exactly as the plain fit forms
// it (expeRT.cpp:453-475: x = -(cnt/vaa - 1), then log10), for a curve with
// no replaced values
__device__ __forceinline__ double log_survival(uint32_t cnt, double vaa) {
    double x = (double)cnt;
    x /= vaa;
    x += -1.0;
    x = -x;
    return log10(x);
}

// single-precision survival feature of a (fractional) prefix count
__device__ __forceinline__ float feat_log_survival(float cnt, float cps) {
    const float x = 1.0f - cnt / cps;
    return log10f(fmaxf(x, 1e-7f));
}

// Features of one spectrum into f[0..kFeat): the log10 survival resampled to
// kCurvePts points by linear interpolation of the prefix counts (point j at
// t = j*(n-1)/(kCurvePts-1)), divided by 4; then stt1/800, end1/800, hyp/800,
// (hyp-end1)/400, n/400, log10(cpsms)/4, the last curve value/4 and the first
// curve value. Single precision; the rounding to bf16 is the network's input.
__device__ inline void curve_features(const uint32_t *cpre, const Curve &c, int hyp, uint32_t cps, int lane,
                                      bf16 *f) {
    const float fc = (float)cps;
    for (int j = lane; j < kCurvePts; j += 64) {
        const float t = (float)(j * (c.n - 1)) / (float)(kCurvePts - 1);
        const int i0 = min((int)t, c.n - 1);
        const int i1 = min(i0 + 1, c.n - 1);
        const float fr = t - (float)i0;
        const float cnt = (float)cpre[i0] * (1.0f - fr) + (float)cpre[i1] * fr;
        f[j] = (bf16)(feat_log_survival(cnt, fc) / 4.0f);
    }
    if (lane < kFeat - kCurvePts) {
        float v = 0.0f;
        switch (lane) {
            case 0: v = (float)c.stt1 / 800.0f; break;
            case 1: v = (float)c.end1 / 800.0f; break;
            case 2: v = (float)hyp / 800.0f; break;
            case 3: v = (float)(hyp - c.end1) / 400.0f; break;
            case 4: v = (float)c.n / 400.0f; break;
            case 5: v = log10f(fc) / 4.0f; break;
            case 6: v = feat_log_survival((float)cpre[c.n - 1], fc) / 4.0f; break;
            default: v = feat_log_survival((float)cpre[0], fc); break;
        }
        f[kCurvePts + lane] = (bf16)v;
    }
}

// The surrogate can serve a spectrum only if its curve has at least
// kMinCurve points and no survival value <= 0 (so none is replaced,
// expeRT.cpp:461-468, and the log curve is non-increasing): the last prefix
// count is below cpsms.
__device__ __forceinline__ bool curve_servable(const Curve &c, const uint32_t *cpre, uint32_t cps) {
    return c.n >= kMinCurve && cpre[c.n - 1] < cps;
}

// First point of a kSnapLanes-point check window around a predicted index q
// (clamped to the curve). Shared with the training label.
__device__ __forceinline__ int snap_base(int q, int n, int back) {
    return min(max(q - back, 0), max(n - kSnapLanes, 0));
}

// predicted index of a mark from the network's output (a fraction of n - 1)
__device__ __forceinline__ int predicted_index(float o, float scale, float shift, int n) {
    const float p = (o * scale + shift) * (float)(n - 1);
    if (!(p > 0.0f)) return 0;  // also NaN
    if (p >= (float)(n - 1)) return n - 1;
    return (int)rintf(p);
}

// Training label of a row: both marks would be found in their check windows
// around the predicted indexes q1, q2 (the conditions the inference kernel
// checks on the curve itself). `mark` is the 22 % mark before the rule that
// moves it below mark2 (expeRT.cpp:509-510), i.e. one before the first point
// at or below the 22 % threshold.
__device__ __forceinline__ bool snap_label(int q1, int q2, int mark, int mark2, int n) {
    const int f = mark + 1;  // first point at or below the 22 % threshold
    const int b1 = snap_base(q1, n, kSnapLanes / 2 - 1);
    const bool ok1 = f < b1 + kSnapLanes && (f > b1 || b1 == 0);
    const int b2 = snap_base(q2, n, kSnapLanes / 2);
    const int e2 = min(b2 + kSnapLanes, n);
    const bool ok2 = mark2 >= b2 && mark2 < e2 && (mark2 < e2 - 1 || e2 == n);
    return ok1 && ok2;
}

}  // namespace gx
