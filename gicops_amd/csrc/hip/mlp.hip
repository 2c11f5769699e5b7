// Learned e-value surrogate kernels for gfx950: a strided BF16 GEMM on
// v_mfma_f32_16x16x32_bf16 with fused epilogues (bias + GELU forward, GELU'
// backward) for training, the training rows of searched spectra (features
// and the plain fit's exact marks, expeRT.cpp:477-515), the loss with the
// window-check label, bias gradients, Adam, and the weights rearranged as
// MFMA B fragments for the fused inference kernel (learned.hip).
#include <math.h>

#include "learned_util.h"

namespace gx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) { return mlp_gelu(z); }

__device__ __forceinline__ float gelu_grad(float z) { return mlp_gelu_grad(z); }

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

// One wave per searched spectrum: the features (the inference kernel's own
// function), the plain fit's exact marks over the whole curve in the plain
// fit's arithmetic (expeRT.cpp:477-515), and the targets {mark before the
// mark2 rule, mark2, n - 1, servable}.
__global__ void __launch_bounds__(64) features_kernel(FeatArgs a) {
    __shared__ uint32_t cpre[GX_HIST_BINS];
    __shared__ double lsx[GX_HIST_BINS];
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x;
    if (s >= a.n) return;
    const gx_psm r = a.psm[s];
    const uint32_t *h = a.hist + (size_t)s * GX_HIST_BINS;
    const int hyp = r.hyp;
    const uint32_t cps = (uint32_t)max(r.cpsms, 0);
    bf16 *f = a.feat + (size_t)s * kFeat;
    const bool modeled = (r.flags & GX_PSM_MODELED) && cps > 0 && hyp > 0 && hyp < GX_HIST_BINS;
    Curve c{0, 0, 0, 0};
    int serv = 0, mark = 0, mark2 = 0;
    if (modeled) {
        c = curve_bounds(h, hyp, lane);
        curve_prefix(h, c, lane, cpre);
        __syncthreads();
        curve_features(cpre, c, hyp, cps, lane, f);
        if (curve_servable(c, cpre, cps)) {
            const double vaa = (double)cps;
            for (int i = lane; i < c.n; i += 64) lsx[i] = log_survival(cpre[i], vaa);
            __syncthreads();
            const double s0 = lsx[0], hgt = lsx[c.n - 1] - s0;
            const double t1 = s0 + hgt * 0.22, t2 = s0 + hgt * 0.87;
            int fm = 0x7fffffff, g = -1;
            for (int i = lane; i < c.n; i += 64) {
                if (lsx[i] <= t1) fm = min(fm, i);
                if (lsx[i] >= t2) g = max(g, i);
            }
            fm = wave_min_i32(fm);
            g = wave_max_i32(g);
            mark = (fm != 0x7fffffff ? fm : 0) - 1;
            mark2 = g >= 0 ? g : c.n - 1;
            const int ruled = mark >= mark2 ? mark2 - 1 : mark;
            serv = (fm != 0x7fffffff && fm >= 1 && ruled >= 0) ? 1 : 0;
        }
    } else {
        for (int j = lane; j < kFeat; j += 64) f[j] = (bf16)0.0f;
    }
    if (lane == 0) {
        const bool fit = (r.flags & GX_PSM_FIT_OK) != 0;
        const double u = (double)r.mu / 1e6 * hyp + (double)r.beta / 1e6;
        float *t = a.target + (size_t)s * kTarget;
        t[0] = (float)mark;
        t[1] = (float)mark2;
        t[2] = (float)max(c.n - 1, 0);
        t[3] = (float)serv;
        a.loge[s] = (fit && cps > 0) ? (float)(log10((double)cps) + u) : 0.0f;
        a.valid[s] = (modeled && fit) ? 1 : 0;
        a.servable[s] = (uint8_t)serv;
    }
}

// Mean squared error of the two normalised positions (servable rows) plus
// conf_weight times the cross-entropy of the confidence against the label
// "both marks lie in the check windows around the predicted indexes".
__global__ void loss_kernel(LossArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float acc[4];
    if (threadIdx.x < 4) acc[threadIdx.x] = 0.0f;
    __syncthreads();
    if (i < a.n) {
        const float *o = a.out + (size_t)i * kOutPad;
        const float *t = a.target + (size_t)i * kTarget;
        const bool serv = t[3] > 0.5f;
        float e0 = 0.0f, e1 = 0.0f, y = 0.0f;
        if (serv) {
            const float nm1 = t[2];
            const int n = (int)nm1 + 1;
            e0 = o[0] - (t[0] / nm1 - a.norm.a_shift) / a.norm.a_scale;
            e1 = o[1] - (t[1] / nm1 - a.norm.u_shift) / a.norm.u_scale;
            const int q1 = predicted_index(o[0], a.norm.a_scale, a.norm.a_shift, n);
            const int q2 = predicted_index(o[1], a.norm.u_scale, a.norm.u_shift, n);
            y = snap_label(q1, q2, (int)t[0], (int)t[1], n) ? 1.0f : 0.0f;
        }
        const float p = 1.0f / (1.0f + expf(-o[2]));
        const float bce = -(y * logf(fmaxf(p, 1e-7f)) + (1.0f - y) * logf(fmaxf(1.0f - p, 1e-7f)));
        const float inv = 1.0f / (float)a.n;
        bf16 *d = a.dout + (size_t)i * kOutPad;
        d[0] = (bf16)(2.0f * e0 * inv);
        d[1] = (bf16)(2.0f * e1 * inv);
        d[2] = (bf16)(a.conf_weight * (p - y) * inv);
        for (int j = 3; j < kOutPad; j++) d[j] = (bf16)0.0f;
        atomicAdd(&acc[0], e0 * e0);
        atomicAdd(&acc[1], e1 * e1);
        atomicAdd(&acc[2], bce);
        atomicAdd(&acc[3], y);
    }
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&a.loss[threadIdx.x], acc[threadIdx.x]);
}

// column sums of a bf16 [rows][cols] matrix (bias gradients): one block per
// 64 columns, its 16 waves over interleaved rows (a wave reads 128 B of one
// row), then a reduction over the waves in LDS
__global__ void __launch_bounds__(1024) colsum_kernel(const bf16 *x, uint32_t rows, uint32_t cols, float *out) {
    __shared__ float part[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * 64 + lane;
    float s = 0.0f;
    if (c < cols)
        for (uint32_t r = w; r < rows; r += 16) s += (float)x[(size_t)r * cols + c];
    part[w][lane] = s;
    __syncthreads();
    if (w == 0 && c < cols) {
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; k++) t += part[k][lane];
        out[c] = t;
    }
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
    if (threadIdx.x < kTarget) tdst[(size_t)i * kTarget + threadIdx.x] = tsrc[(size_t)j * kTarget + threadIdx.x];
    if (threadIdx.x == 0) ldst[i] = lsrc[j];
}

// the weights as MFMA B fragments (FragLayout), rounded to bf16 as the
// training GEMMs round them
__global__ void frag_weights_kernel(const float *w, bf16 *frag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= FragLayout::N) return;
    int base, ksteps, ncols, src;
    if (i < FragLayout::L2) {
        base = FragLayout::L1, ksteps = kFeat / 32, ncols = kHidden, src = MlpLayout::W1;
    } else if (i < FragLayout::L3) {
        base = FragLayout::L2, ksteps = kHidden / 32, ncols = kHidden, src = MlpLayout::W2;
    } else {
        base = FragLayout::L3, ksteps = kHidden / 32, ncols = kOutPad, src = MlpLayout::W3;
    }
    const int local = i - base, fid = local / 512, within = local % 512;
    const int ln = within / 8, j = within % 8;
    const int nt = fid / ksteps, ks = fid % ksteps;
    const int k = ks * 32 + 8 * (ln / 16) + j, col = nt * 16 + ln % 16;
    frag[i] = (bf16)w[src + k * ncols + col];
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
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 63) / 64), dim3(1024), 0, s, x, rows, cols, out);
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

hipError_t launch_frag_weights(const float *w, bf16 *frag, hipStream_t s) {
    hipLaunchKernelGGL(frag_weights_kernel, dim3((FragLayout::N + 255) / 256), dim3(256), 0, s, w, frag);
    return hipGetLastError();
}

}  // namespace gx
