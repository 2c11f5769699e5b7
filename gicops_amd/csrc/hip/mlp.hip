// Learned e-value surrogate kernels for gfx950: survival-histogram features,
// a strided BF16 GEMM on v_mfma_f32_16x16x32_bf16 with fused epilogues
// (bias + GELU forward, GELU' backward), loss, bias gradients, Adam and the
// confidence routing. The network predicts the reference tail fit's slope and
// the fitted log-survival at the top hit's bin (so e = cpsms * 10^u), plus a
// confidence logit; low confidence or a sparse histogram falls back to the
// plain fit (expeRT::ModelTailFit, expeRT.cpp:396-559).
#include <math.h>

#include "device_util.h"
#include "mlp.h"

namespace gx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// One wave per 16 x (16*NT) output strip; K in steps of 32. Lane l supplies
// A[row l%16][k0 + 8*(l/16) .. +7] and B[k0 + 8*(l/16) .. +7][col l%16] and
// receives C[4*(l/16) + j][l%16], j < 4 (the 16x16x32 bf16 MFMA layout; the
// sum over k does not depend on which lane holds which k, as long as A and B
// agree). The A fragment is loaded once per k step for all NT column tiles.
// Out-of-range rows, columns and k are zero-filled, so any M, N, K work.
template <int NT>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles_n = (g.N + 15) >> 4;
    const int groups_n = (tiles_n + NT - 1) / NT;
    const long long strip = (long long)blockIdx.x * 4 + wave;
    const int tm = (int)(strip / groups_n), tg = (int)(strip % groups_n);
    if ((long long)tm * 16 >= g.M) return;
    const int r = lane & 15, kg = lane >> 4;
    const int m = tm * 16 + r;
    // 16-byte loads of A when its k run is contiguous and aligned
    const bool a_vec = g.sak == 1 && (g.sam & 7) == 0 && (g.K & 31) == 0 && ((uintptr_t)g.A & 15) == 0;
    floatx4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        bf16x8 va;
        if (a_vec) {
            if (m < g.M) {
                va = *reinterpret_cast<const bf16x8 *>(g.A + (long long)m * g.sam + k0 + 8 * kg);
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) va[j] = (bf16)0.0f;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = k0 + 8 * kg + j;
                va[j] = (k < g.K && m < g.M) ? g.A[(long long)m * g.sam + (long long)k * g.sak] : (bf16)0.0f;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int n = (tg * NT + t) * 16 + r;
            bf16x8 vb;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = k0 + 8 * kg + j;
                vb[j] = (k < g.K && n < g.N) ? g.B[(long long)k * g.sbk + (long long)n * g.sbn] : (bf16)0.0f;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int col = (tg * NT + t) * 16 + r;
        if (col >= g.N) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = tm * 16 + 4 * kg + j;
            if (row >= g.M) continue;
            const float v = acc[t][j];
            if (g.mode == 0) {
                g.C[(long long)row * g.ldc + col] = v + (g.bias ? g.bias[col] : 0.0f);
            } else if (g.mode == 1) {
                const float z = v + g.bias[col];
                g.Z[(long long)row * g.ldz + col] = z;
                g.H[(long long)row * g.ldh + col] = (bf16)gelu(z);
            } else if (g.mode == 2) {
                g.H[(long long)row * g.ldh + col] = (bf16)(v * gelu_grad(g.Z[(long long)row * g.ldz + col]));
            } else {
                g.H[(long long)row * g.ldh + col] = (bf16)gelu(v + g.bias[col]);
            }
        }
    }
}

// One wave per spectrum. The curve is the plain fit's own (expeRT.cpp:414-475):
// bins from the first non-empty one (stt1) to the last non-empty one below
// the top hit (end1), survival s = 1 - cumsum/cpsms from stt1, s > 1 -> 0.999,
// s <= 0 replaced by the last value >= 1e-4, then log10. It is resampled to
// kCurvePts points by linear interpolation; 8 scalars follow.
__global__ void __launch_bounds__(64) features_kernel(FeatArgs a) {
    __shared__ double sx[GX_HIST_BINS];
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x;
    if (s >= a.n) return;
    const gx_psm r = a.psm[s];
    const uint32_t *h = a.hist + (size_t)s * GX_HIST_BINS;
    const int hyp = r.hyp;
    const uint32_t cps = (uint32_t)max(r.cpsms, 0);
    bf16 *f = a.feat + (size_t)s * kFeat;
    const bool modeled = (r.flags & GX_PSM_MODELED) && cps > 0 && hyp > 0 && hyp < GX_HIST_BINS;
    int e = -1, st = 0x7fffffff;
    uint32_t nonempty = 0;
    if (modeled) {
        for (int p = lane; p < hyp; p += 64) {
            if (h[p] >= 1) {
                e = max(e, p);
                st = min(st, p);
                nonempty++;
            }
        }
    }
    e = wave_max_i32(e);
    st = wave_min_i32(st);
    nonempty = wave_sum_u32(nonempty);
    int end1 = e >= 0 ? e : hyp - 1;
    const int stt1 = st != 0x7fffffff ? st : 0;
    if (stt1 == end1) end1 += 1;
    const int n = modeled ? end1 - stt1 + 1 : 0;
    // survival from stt1
    uint32_t carry = 0;
    for (int b0 = 0; b0 < n; b0 += 64) {
        const int i = b0 + lane;
        const uint32_t v = (i < n && stt1 + i < GX_HIST_BINS) ? h[stt1 + i] : 0u;
        const uint32_t inc = wave_incl_scan(v, lane) + carry;
        if (i < n) {
            double x = -((double)inc / (double)cps - 1.0);
            sx[i] = x > 1.0 ? 0.999 : x;
        }
        carry = __shfl(inc, 63, 64);
    }
    __syncthreads();
    int rp = -1;
    for (int i = lane; i < n; i += 64)
        if (sx[i] >= 1e-4) rp = max(rp, i);
    rp = wave_max_i32(rp);
    const double rv = n > 0 ? sx[rp >= 0 ? rp : n - 1] : 1.0;
    __syncthreads();
    int bad = 0;
    for (int i = lane; i < n; i += 64) {
        const double x = sx[i] <= 0.0 ? rv : sx[i];
        const double l = log10(x);
        bad |= isfinite(l) ? 0 : 1;
        sx[i] = isfinite(l) ? l : -8.0;
    }
    bad = wave_max_i32(bad);
    __syncthreads();
    for (int j = lane; j < kCurvePts; j += 64) {
        float v = 0.0f;
        if (n > 0) {
            const double t = n > 1 ? (double)j * (double)(n - 1) / (double)(kCurvePts - 1) : 0.0;
            const int i0 = min((int)t, n - 1);
            const int i1 = min(i0 + 1, n - 1);
            const double fr = t - (double)i0;
            v = (float)((sx[i0] * (1.0 - fr) + sx[i1] * fr) / 4.0);
        }
        f[j] = (bf16)v;
    }
    if (lane == 0) {
        const double anchor = n > 0 ? sx[n - 1] : 0.0;
        const double lc = cps > 0 ? log10((double)cps) : 0.0;
        f[kCurvePts + 0] = (bf16)(n > 0 ? (float)stt1 / 800.0f : 0.0f);
        f[kCurvePts + 1] = (bf16)(n > 0 ? (float)end1 / 800.0f : 0.0f);
        f[kCurvePts + 2] = (bf16)(n > 0 ? (float)hyp / 800.0f : 0.0f);
        f[kCurvePts + 3] = (bf16)(n > 0 ? (float)(hyp - end1) / 400.0f : 0.0f);
        f[kCurvePts + 4] = (bf16)(n > 0 ? (float)n / 400.0f : 0.0f);
        f[kCurvePts + 5] = (bf16)(n > 0 ? (float)(lc / 4.0) : 0.0f);
        f[kCurvePts + 6] = (bf16)(n > 0 ? (float)(anchor / 4.0) : 0.0f);
        f[kCurvePts + 7] = (bf16)(n > 0 ? (float)sx[0] : 0.0f);
        const bool fit = (r.flags & GX_PSM_FIT_OK) != 0;
        const double slope = (double)r.mu / 1e6;
        const double u = slope * hyp + (double)r.beta / 1e6;
        a.anchor[s] = (float)anchor;
        a.target[2 * (size_t)s] = (float)slope;
        a.target[2 * (size_t)s + 1] = (float)(u - anchor);
        a.loge[s] = fit ? (float)(lc + u) : 0.0f;
        a.valid[s] = (modeled && fit) ? 1 : 0;
        a.sparse[s] = (!modeled || bad || nonempty < (uint32_t)kMinTailBins) ? 1 : 0;
    }
}

__global__ void loss_kernel(LossArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float acc[4];
    if (threadIdx.x < 4) acc[threadIdx.x] = 0.0f;
    __syncthreads();
    if (i < a.n) {
        const float *o = a.out + (size_t)i * kOutPad;
        const float ta = (a.target[2 * (size_t)i] - a.norm.a_shift) / a.norm.a_scale;
        const float tu = (a.target[2 * (size_t)i + 1] - a.norm.u_shift) / a.norm.u_scale;
        const float ea = o[0] - ta, eu = o[1] - tu;
        const float err = fabsf(eu) * a.norm.u_scale;  // |u_pred - u| = error of log10 e
        const float tol = a.tol * fmaxf(fabsf(a.loge[i]), 1e-3f);
        const float y = err <= tol ? 1.0f : 0.0f;
        const float p = 1.0f / (1.0f + expf(-o[2]));
        const float bce = -(y * logf(fmaxf(p, 1e-7f)) + (1.0f - y) * logf(fmaxf(1.0f - p, 1e-7f)));
        const float inv = 1.0f / (float)a.n;
        bf16 *d = a.dout + (size_t)i * kOutPad;
        d[0] = (bf16)(2.0f * ea * inv);
        d[1] = (bf16)(2.0f * eu * inv);
        d[2] = (bf16)(a.conf_weight * (p - y) * inv);
        for (int j = 3; j < kOutPad; j++) d[j] = (bf16)0.0f;
        atomicAdd(&acc[0], ea * ea);
        atomicAdd(&acc[1], eu * eu);
        atomicAdd(&acc[2], bce);
        atomicAdd(&acc[3], y);
    }
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&a.loss[threadIdx.x], acc[threadIdx.x]);
}

__global__ void colsum_kernel(const bf16 *x, uint32_t rows, uint32_t cols, float *out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (uint32_t r = 0; r < rows; r++) s += (float)x[(size_t)r * cols + c];
    out[c] = s;
}

__global__ void adam_kernel(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                            float eps, float b1t, float b2t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - b1t), vh = vi / (1.0f - b2t);
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void to_bf16_kernel(const float *x, bf16 *y, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        y[i] = (bf16)x[i];
}

__global__ void gather_rows_kernel(const bf16 *src, const float *tsrc, const float *lsrc, const uint32_t *idx,
                                   uint32_t n, bf16 *dst, float *tdst, float *ldst) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    for (int k = threadIdx.x; k < kFeat; k += blockDim.x) dst[(size_t)i * kFeat + k] = src[(size_t)j * kFeat + k];
    if (threadIdx.x == 0) {
        tdst[2 * (size_t)i] = tsrc[2 * (size_t)j];
        tdst[2 * (size_t)i + 1] = tsrc[2 * (size_t)j + 1];
        ldst[i] = lsrc[j];
    }
}

__global__ void route_kernel(RouteArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long fb = 0;
    if (s < a.n) {
        uint8_t accept = 0;
        const bool modeled = (a.psm[s].flags & GX_PSM_MODELED) != 0;
        if (modeled && !a.sparse[s]) {
            const float *o = a.out + (size_t)s * kOutPad;
            const float p = 1.0f / (1.0f + expf(-o[2]));
            // outputs beyond kMaxZ standard deviations of the training targets
            // are out of distribution, whatever the confidence says
            const bool in_range = fabsf(o[0]) <= kMaxZ && fabsf(o[1]) <= kMaxZ;
            if (p >= a.conf_threshold && in_range) {
                gx_psm r = a.psm[s];
                const double slope = (double)o[0] * a.norm.a_scale + a.norm.a_shift;
                const double u = (double)a.anchor[s] + (double)o[1] * a.norm.u_scale + a.norm.u_shift;
                const double e = (double)r.cpsms * pow(10.0, u);
                const double am = slope * 1e6, bm = (u - slope * r.hyp) * 1e6;
                // a survival probability: u <= 0
                if (u <= 0.0 && isfinite(e) && fabs(am) < 2147483647.0 && fabs(bm) < 2147483647.0) {
                    r.mu = (int32_t)am;
                    r.beta = (int32_t)bm;
                    r.evalue = e;
                    r.flags |= GX_PSM_FIT_OK | GX_PSM_SURROGATE;
                    if (e < a.expect_max) r.flags |= GX_PSM_REPORTED;
                    a.psm[s] = r;
                    accept = 1;
                }
            }
        }
        a.route[s] = accept;
        fb = (modeled && !accept) ? 1ull : 0ull;
    }
    fb = wave_sum_u64(fb);
    if ((threadIdx.x & 63) == 0 && fb) atomicAdd(&a.stats[ST_FALLBACK], fb);
}

}  // namespace

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    if (g.K < 0 || g.mode < 0 || g.mode > 3) return hipErrorInvalidValue;
    if ((g.mode == 0 && !g.C) || ((g.mode == 1 || g.mode == 3) && (!g.bias || !g.H)) ||
        (g.mode == 1 && !g.Z) || (g.mode == 2 && (!g.Z || !g.H)))
        return hipErrorInvalidValue;
    const int tiles_n = (g.N + 15) / 16;
    const long long tiles_m = (g.M + 15) / 16;
    if (tiles_n >= 16) {
        const long long strips = tiles_m * ((tiles_n + 15) / 16);
        hipLaunchKernelGGL(gemm_kernel<16>, dim3((unsigned)((strips + 3) / 4)), dim3(256), 0, s, g);
    } else {
        const long long strips = tiles_m * tiles_n;
        hipLaunchKernelGGL(gemm_kernel<1>, dim3((unsigned)((strips + 3) / 4)), dim3(256), 0, s, g);
    }
    return hipGetLastError();
}

hipError_t launch_features(const FeatArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3(a.n), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_loss(const LossArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(loss_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_colsum(const bf16 *x, uint32_t rows, uint32_t cols, float *out, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 63) / 64), dim3(64), 0, s, x, rows, cols, out);
    return hipGetLastError();
}

hipError_t launch_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                       float eps, float b1t, float b2t, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, b1t, b2t);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float *x, bf16 *y, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const bf16 *src, const float *tsrc, const float *lsrc, const uint32_t *idx, uint32_t n,
                              bf16 *dst, float *tdst, float *ldst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(n), dim3(64), 0, s, src, tsrc, lsrc, idx, n, dst, tdst, ldst);
    return hipGetLastError();
}

hipError_t launch_route(const RouteArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(route_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gx
