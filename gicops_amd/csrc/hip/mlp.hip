// Learned e-value surrogate kernels for gfx950: a strided BF16 GEMM on
// v_mfma_f32_16x16x32_bf16 with fused epilogues (bias + GELU forward, GELU'
// backward; split K and the bias gradients as an extra row of ones for the
// weight gradients) for training, the training rows of searched spectra
// (features and the plain fit's exact marks, expeRT.cpp:477-515), the loss
// with the window-check label, Adam, and the weights rearranged as MFMA B
// fragments for the fused inference kernel (learned.hip).
#include <math.h>

#include <algorithm>

#include "learned_util.h"

namespace gx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) { return mlp_gelu(z); }

__device__ __forceinline__ float gelu_grad(float z) { return mlp_gelu_grad(z); }

// C = A B on MFMA with LDS tiles: one block of 4 waves per 64 x 64 output
// tile, K in steps of 32. Each step stages the 64 x 32 A tile and the 32 x 64
// B tile in LDS with 16-byte loads along whichever dimension of each operand
// is contiguous (row-major, or transposed for the backward pass), stored as
// [row][k] and [col][k] so that every MFMA fragment is one 16-byte LDS read:
// lane l supplies A[row l%16][k0 + 8*(l/16) .. +7] and B[k0 + 8*(l/16) ..
// +7][col l%16] and receives C[4*(l/16) + j][l%16], j < 4. Each wave computes
// a 32 x 32 quarter (2 x 2 MFMA tiles). Elements outside M, N, K are zero, so
// any shape works; the epilogue modes are those of GemmArgs. With ksplit > 1
// (mode 0), block b computes tile b % tiles over the k range of chunk
// b / tiles into its own slice of the workspace, and splitk_reduce_kernel
// adds the slices in chunk order.
constexpr int kGT = 64;        // block tile
constexpr int kGK = 32;        // k step
constexpr int kGPad = kGK + 8; // LDS row stride (bf16): 80 B spreads a fragment read over the banks

__device__ __forceinline__ void gemm_stage(const bf16 *P, long long s_row, long long s_k, int rows, int K, int r0,
                                           int k0, bf16 (*T)[kGPad], int tid, int ones_row) {
    // T[r][k] = P(r0 + r, k0 + k), 64 x 32; row ones_row (>= rows, or -1 for
    // none) is all ones
    if (s_k == 1) {
        // 8 consecutive k of one row per thread
        const int r = tid >> 2, kc = (tid & 3) * 8;
        const int gr = r0 + r, gk = k0 + kc;
        const bf16 *src = P + (long long)gr * s_row + gk;
        bf16x8 v;
        if (gr < rows && gk + 8 <= K && (((uintptr_t)src) & 15) == 0) {
            v = *reinterpret_cast<const bf16x8 *>(src);
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = (gr < rows && gk + j < K) ? src[j] : (bf16)0.0f;
            if (gr == ones_row) {
#pragma unroll
                for (int j = 0; j < 8; j++) v[j] = gk + j < K ? (bf16)1.0f : (bf16)0.0f;
            }
        }
        *reinterpret_cast<bf16x8 *>(&T[r][kc]) = v;
    } else {
        // 8 consecutive rows of one k per thread (s_row == 1 or general strides)
        const int k = tid >> 3, rc = (tid & 7) * 8;
        const int gk = k0 + k, gr = r0 + rc;
        const bf16 *src = P + (long long)gr * s_row + (long long)gk * s_k;
        bf16x8 v;
        if (s_row == 1 && gk < K && gr + 8 <= rows && (((uintptr_t)src) & 15) == 0) {
            v = *reinterpret_cast<const bf16x8 *>(src);
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = (gk < K && gr + j < rows) ? src[(long long)j * s_row] : (bf16)0.0f;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (gr + j == ones_row) v[j] = gk < K ? (bf16)1.0f : (bf16)0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) T[rc + j][k] = v[j];
    }
}

__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __align__(16) bf16 As[kGT][kGPad];
    __shared__ __align__(16) bf16 Bs[kGT][kGPad];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rows_c = g.M + (g.colsum_out ? 1 : 0);  // rows of the product, with the ones row
    const int tiles_n = (g.N + kGT - 1) / kGT;
    const int tiles = ((rows_c + kGT - 1) / kGT) * tiles_n;
    const int split = (int)blockIdx.x / tiles, tile = (int)blockIdx.x % tiles;
    const int m0 = (tile / tiles_n) * kGT, n0 = (tile % tiles_n) * kGT;
    // k range of this block (the whole of K without a split; chunks are whole k steps)
    const int kb = g.ksplit > 1 ? split * g.kchunk : 0;
    const int ke = g.ksplit > 1 ? min(g.K, kb + g.kchunk) : g.K;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // the wave's 32 x 32 quarter
    const int r = lane & 15, kg = lane >> 4;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kb; k0 < ke; k0 += kGK) {
        __syncthreads();
        gemm_stage(g.A, g.sam, g.sak, g.M, g.K, m0, k0, As, tid, g.colsum_out ? g.M : -1);
        // B(k, n): its "rows" are the n columns, its k stride is sbk
        gemm_stage(g.B, g.sbn, g.sbk, g.N, g.K, n0, k0, Bs, tid, -1);
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * i + r][8 * kg]);
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * j + r][8 * kg]);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int col = n0 + wn + 16 * j + r;
            if (col >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = m0 + wm + 16 * i + 4 * kg + q;
                if (row >= rows_c) continue;
                const float v = acc[i][j][q];
                if (g.ksplit > 1) {
                    g.ws[((long long)split * rows_c + row) * g.N + col] = v;
                } else if (row == g.M) {
                    g.colsum_out[col] = v;
                } else if (g.mode == 0) {
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

// C = the sum of the ksplit partial products (+ bias), in a fixed order: four
// running sums over chunks s = 4 q + u, then (u0 + u1) + (u2 + u3), so that
// four loads are in flight per step
__global__ void __launch_bounds__(256) splitk_reduce_kernel(GemmArgs g) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const int rows_c = g.M + (g.colsum_out ? 1 : 0);
    const long long mn = (long long)rows_c * g.N;
    if (i >= mn) return;
    const int row = (int)(i / g.N), col = (int)(i % g.N);
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int s = 0;
    for (; s + 4 <= g.ksplit; s += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] += g.ws[(long long)(s + u) * mn + i];
    }
    for (int u = 0; s < g.ksplit; s++, u++) v[u] += g.ws[(long long)s * mn + i];
    const float t = (v[0] + v[1]) + (v[2] + v[3]);
    if (row == g.M)
        g.colsum_out[col] = t;
    else
        g.C[(long long)row * g.ldc + col] = t + (g.bias ? g.bias[col] : 0.0f);
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
__global__ void __launch_bounds__(256) loss_kernel(LossArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float acc[4];
    if (threadIdx.x < 4) acc[threadIdx.x] = 0.0f;
    __syncthreads();
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
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
        // the row's 16 output gradients (3 used) as two 16-byte stores
        static_assert(kOutPad == 16, "two bf16x8 stores per row");
        bf16x8 d0, d1;
#pragma unroll
        for (int j = 0; j < 8; j++) d0[j] = d1[j] = (bf16)0.0f;
        d0[0] = (bf16)(2.0f * e0 * inv);
        d0[1] = (bf16)(2.0f * e1 * inv);
        d0[2] = (bf16)(a.conf_weight * (p - y) * inv);
        bf16x8 *d = reinterpret_cast<bf16x8 *>(a.dout + (size_t)i * kOutPad);
        d[0] = d0;
        d[1] = d1;
        part[0] = e0 * e0;
        part[1] = e1 * e1;
        part[2] = bce;
        part[3] = y;
    }
    // the block's sums: wave sums, then one LDS add per wave
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part[k] += __shfl_xor(part[k], o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) atomicAdd(&acc[k], part[k]);
    }
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&a.loss[threadIdx.x], acc[threadIdx.x]);
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
    if (g.K < 0 || g.mode < 0 || g.mode > 3 || (g.colsum_out && g.mode != 0)) return hipErrorInvalidValue;
    if ((g.mode == 0 && !g.C) || ((g.mode == 1 || g.mode == 3) && (!g.bias || !g.H)) ||
        (g.mode == 1 && !g.Z) || (g.mode == 2 && (!g.Z || !g.H)))
        return hipErrorInvalidValue;
    const long long rows_c = g.M + (g.colsum_out ? 1 : 0);
    const long long tiles = (long long)((rows_c + kGT - 1) / kGT) * ((g.N + kGT - 1) / kGT);
    GemmArgs a = g;
    a.ksplit = 1;
    a.kchunk = g.K;
    if (g.mode == 0 && g.ws && tiles < 256) {
        // about 512 blocks, at least 4 k steps per chunk, within the workspace
        const long long want = (512 + tiles - 1) / tiles;
        const long long by_k = g.K / (4 * kGK);
        const long long by_ws = g.ws_floats / (rows_c * g.N);
        const long long sp = std::min(want, std::min(by_k, by_ws));
        if (sp > 1) {
            a.kchunk = (int)(((g.K + sp - 1) / sp + kGK - 1) / kGK * kGK);
            a.ksplit = (g.K + a.kchunk - 1) / a.kchunk;
        }
    }
    hipLaunchKernelGGL(gemm_kernel, dim3((unsigned)(tiles * a.ksplit)), dim3(256), 0, s, a);
    if (a.ksplit > 1) {
        const long long mn = rows_c * g.N;
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, a);
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
