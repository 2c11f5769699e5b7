// The learned e-value stage of the batched search for gfx950, two kernels:
//
//   learned_net_kernel, one block of 4 waves per 16 spectra:
//   1. per spectrum (each wave takes 4 of the 16): the top hit (as the plain
//      e-value kernel forms it), the tail fit's curve bounds and exact prefix
//      counts, the features (bf16) into row r of the block's 16 x 256
//      activation tile in LDS;
//   2. the 128-256-256-3 GELU network on v_mfma_f32_16x16x32_bf16 for the 16
//      rows at once, each wave computing 4 of the 16 output tiles of a layer:
//      A fragments from the LDS tile, weights as pre-arranged B fragments (one
//      coalesced 16-byte load per lane per MFMA), FP32 accumulation, bias +
//      GELU epilogues back into the tile; the three outputs per spectrum go
//      to global memory (16 B).
//
//   learned_fit_kernel, one wave per spectrum (the plain e-value kernel's
//   shape and occupancy):
//   3. with enough confidence on a servable curve, the two marks of
//      expeRT::ModelTailFit (expeRT.cpp:477-515) are looked for in 32-point
//      windows around the predicted indexes (one log10 per lane, two
//      ballots); when both are found there, the least-squares line of
//      expeRT.cpp:516-545 (Burkardt's llsq) is fitted on the window by
//      wave-parallel sums and the e-value formed from the truncated mu, beta
//      exactly as the plain path does (dslim_query.cpp:1023-1041). Otherwise
//      the plain fit runs (tail_fit_wave, the same code as the plain stage).
//
// (One fused kernel doing all three measured slower than the plain stage: its
// LDS and registers held it to 8 waves per CU, each fitting 4 spectra in
// turn.)
//
// The log survival is non-increasing on a servable curve (integer prefix
// counts, no replaced values), so "first point at or below t1" and "last
// point at or above t2" are decided inside a window whenever the window holds
// the transition; the checks below accept a window only then. A served
// e-value therefore uses the plain fit's own window; only the summation order
// of the line fit differs from the plain path's sequential one.
#include "batched.h"
#include "learned_util.h"

namespace gx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 16;
constexpr int kWaves = 4;
constexpr int kAct = kHidden + 8;  // bf16 row stride: 528 B puts the 16 rows of a fragment load on distinct banks

struct NetLds {
    bf16 act[kRows][kAct];                // features, then h1, then h2
    uint32_t cpre[kWaves][GX_HIST_BINS];  // prefix counts of each wave's current spectrum
};

struct FitLds {
    uint32_t cpre[GX_HIST_BINS];  // prefix counts, or the histogram for the plain fit
    double sx[GX_HIST_BINS];  // the plain fit's curve (fallback) or the window's log survival
};

// The marks found in the check windows and the line through the window;
// false when a mark is not inside its window or the fit is undefined.
__device__ bool snap_fit(const uint32_t *cpre, int n, int stt1, uint32_t cps, int q1, int q2, int lane, int32_t &mu,
                         int32_t &beta) {
    const double vaa = (double)cps;
    // the thresholds from the curve's ends (expeRT.cpp:477-485)
    double e = 0.0;
    if (lane < 2) e = log_survival(cpre[lane == 0 ? 0 : n - 1], vaa);
    const double s0 = readlane_f64(e, 0), sN = readlane_f64(e, 1);
    const double hgt = sN - s0;
    const double t1 = s0 + hgt * 0.22, t2 = s0 + hgt * 0.87;
    const bool lo_half = lane < kSnapLanes;
    const int b1 = snap_base(q1, n, kSnapLanes / 2 - 1), b2 = snap_base(q2, n, kSnapLanes / 2);
    const int p = (lo_half ? b1 : b2) + (lane & (kSnapLanes - 1));
    const bool in = p < n;
    const double v = in ? log_survival(cpre[p], vaa) : 0.0;
    const uint64_t m1 = __ballot(in && lo_half && v <= t1);
    const uint64_t m2 = __ballot(in && !lo_half && v >= t2) >> kSnapLanes;
    if (!m1 || !m2) return false;
    // mark + 1: the first point at or below t1; the window must not start on it
    // (the crossing could lie before), unless the window starts the curve
    const int f = b1 + __builtin_ctzll(m1);
    if ((f == b1 && b1 > 0) || f == 0) return false;
    // mark2: the last point at or above t2; the window must not end on it,
    // unless the window ends the curve
    const int mark2 = b2 + 63 - __builtin_clzll(m2);
    const int e2 = min(b2 + kSnapLanes, n);
    if (mark2 == e2 - 1 && e2 < n) return false;
    int mark = f - 1;
    if (mark >= mark2) mark = mark2 - 1;
    if (mark < 0) return false;
    // least squares over x = stt1 + mark .. stt1 + mark2 (expeRT.cpp:1241-1289)
    // in one wave-parallel pass: xbar in closed form (an exact integer sum over
    // m), then sum y, sum (x - xbar) y, sum (x - xbar)^2 and sum (x - xbar);
    // top = sum (x - xbar)(y - ybar) = sum (x - xbar) y - ybar sum (x - xbar)
    const int m = mark2 - mark + 1;
    const double x0 = (double)(stt1 + mark);
    const double xbar = ((double)m * x0 + (double)m * (double)(m - 1) * 0.5) / (double)m;
    double sy = 0.0, sxy = 0.0, sxx = 0.0, sd = 0.0;
    for (int i = lane; i < m; i += 64) {
        const double y = log_survival(cpre[mark + i], vaa);
        const double dx = (x0 + (double)i) - xbar;
        sy += y;
        sxy += dx * y;
        sxx += dx * dx;
        sd += dx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sy += __shfl_xor(sy, o, 64);
        sxy += __shfl_xor(sxy, o, 64);
        sxx += __shfl_xor(sxx, o, 64);
        sd += __shfl_xor(sd, o, 64);
    }
    const double ybar = sy / (double)m;
    const double top = sxy - ybar * sd, bot = sxx;
    const double a = top / bot, b = ybar - a * xbar;
    const double am = a * 1e6, bm = b * 1e6;
    if (!(fabs(am) < 2147483647.0 && fabs(bm) < 2147483647.0)) return false;
    mu = (int32_t)am;
    beta = (int32_t)bm;
    return true;
}

__global__ void __launch_bounds__(64 * kWaves) learned_net_kernel(LearnArgs a) {
    __shared__ NetLds L;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t *cpre = L.cpre[wave];
    const uint32_t s0 = blockIdx.x * kRows;
    const uint32_t need = (uint32_t)max(a.min_cpsm, 1);

    // ---- 1. features (wave w: rows w, w + 4, w + 8, w + 12) --------------------
    for (int r = wave; r < kRows; r += kWaves) {
        const uint32_t s = s0 + r;
        bf16 *f = L.act[r];
        int hyp = 0;
        uint32_t cps = 0;
        bool modeled = false;
        if (s < a.n) {
            const gx_psm p = top_hit_psm(a.best_key[s], a.best_eq[s], a.best_shp[s], a.cpsms[s], a.buckets, need);
            modeled = (p.flags & GX_PSM_MODELED) && p.hyp > 0 && p.hyp < GX_HIST_BINS;
            hyp = p.hyp;
            cps = (uint32_t)p.cpsms;
        }
        uint32_t meta = 0;
        if (modeled) {
            const uint32_t *h = a.hist + (size_t)s * GX_HIST_BINS;
            const Curve c = curve_load(h, hyp, lane, cpre);
            wave_lds_sync();
            curve_features(cpre + c.stt1, c, hyp, cps, lane, f);
            // the fit kernel's starting point: stt1, n and whether the curve is servable
            meta = (uint32_t)c.stt1 | ((uint32_t)c.n << 10) | (curve_servable(c, cpre + c.stt1, cps) ? 1u << 20 : 0u);
            wave_lds_sync();  // the next spectrum overwrites cpre
        } else {
            for (int j = lane; j < kFeat; j += 64) f[j] = (bf16)0.0f;
        }
        if (lane == 0 && s < a.n) a.raw[(size_t)s * 4 + 3] = __uint_as_float(meta);
    }
    __syncthreads();

    // ---- 2. the network on MFMA (wave w: output tiles 4w .. 4w+3) ----------------
    const int r16 = lane & 15, kg = lane >> 4;
    const bf16x8 *fr = reinterpret_cast<const bf16x8 *>(a.frag);
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
        constexpr int kMaxK = kHidden / 32;
        const int ksteps = layer == 0 ? kFeat / 32 : kHidden / 32;
        const bf16x8 *wf = fr + (layer == 0 ? FragLayout::L1 : FragLayout::L2) / 8;
        const float *bias = a.params + (layer == 0 ? MlpLayout::B1 : MlpLayout::B2);
        // the layer's whole input row fragment first: the tile is overwritten below
        bf16x8 A[kMaxK];
#pragma unroll
        for (int ks = 0; ks < kMaxK; ks++)
            if (ks < ksteps) A[ks] = *reinterpret_cast<const bf16x8 *>(&L.act[r16][ks * 32 + 8 * kg]);
        floatx4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int nt = wave * 4 + t;
            acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < kMaxK; ks++)
                if (ks < ksteps)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks], wf[(nt * ksteps + ks) * 64 + lane], acc[t],
                                                                     0, 0, 0);
        }
        __syncthreads();  // every wave has read the layer's input
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int col = (wave * 4 + t) * 16 + r16;
            const float b = bias[col];
#pragma unroll
            for (int j = 0; j < 4; j++) L.act[4 * kg + j][col] = (bf16)mlp_gelu(acc[t][j] + b);
        }
        __syncthreads();
    }
    if (wave == 0) {
        floatx4 acc{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < kHidden / 32; ks++) {
            const bf16x8 A = *reinterpret_cast<const bf16x8 *>(&L.act[r16][ks * 32 + 8 * kg]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, fr[FragLayout::L3 / 8 + ks * 64 + lane], acc, 0, 0, 0);
        }
        // lane (r16 = output column, kg) holds rows 4 kg .. 4 kg + 3
        if (r16 < kOut) {
            const float b = a.params[MlpLayout::B3 + r16];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t s = s0 + 4 * kg + j;
                if (s < a.n) a.raw[(size_t)s * 4 + r16] = acc[j] + b;
            }
        }
    }
}

// one wave per spectrum: the window check and fit from the network's outputs,
// or the plain fit
__global__ void __launch_bounds__(64) learned_fit_kernel(LearnArgs a) {
    __shared__ FitLds L;
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x;
    const uint32_t need = (uint32_t)max(a.min_cpsm, 1);
    gx_psm p = top_hit_psm(a.best_key[s], a.best_eq[s], a.best_shp[s], a.cpsms[s], a.buckets, need);
    bool served = false;
    if (p.flags & GX_PSM_MODELED) {
        const uint32_t *h = a.hist + (size_t)s * GX_HIST_BINS;
        const float4 o = reinterpret_cast<const float4 *>(a.raw)[s];
        // the network kernel left stt1 | n << 10 | servable << 20 in the fourth word
        const uint32_t meta = __float_as_uint(o.w);
        if ((meta >> 20) & 1u && o.z >= a.conf_logit) {
            const uint32_t cps = (uint32_t)p.cpsms;
            const int stt1 = (int)(meta & 1023u), n = (int)((meta >> 10) & 1023u);
            const Curve c{stt1, stt1 + n - 1, n, 0};
            curve_prefix(h, c, lane, L.cpre);
            wave_lds_sync();
            const int q1 = predicted_index(o.x, a.norm.a_scale, a.norm.a_shift, n);
            const int q2 = predicted_index(o.y, a.norm.u_scale, a.norm.u_shift, n);
            int32_t mu = 0, beta = 0;
            served = snap_fit(L.cpre, n, stt1, cps, q1, q2, lane, mu, beta);
            if (served) {
                p.flags |= GX_PSM_FIT_OK | GX_PSM_SURROGATE;
                p.mu = mu;
                p.beta = beta;
                p.evalue = evalue_of(mu, beta, p.hyp, cps);
                if (p.evalue < a.expect_max) p.flags |= GX_PSM_REPORTED;
            }
            wave_lds_sync();
        }
        if (!served) {
            // the plain fit on an LDS copy of the histogram, as the plain stage runs it
            for (int w = lane; w < GX_HIST_BINS; w += 64) L.cpre[w] = h[w];
            wave_lds_sync();
            double sl, ic;
            int32_t mu, beta;
            if (tail_fit_wave(L.cpre, L.sx, p.hyp, p.cpsms, sl, ic, mu, beta)) {
                p.flags |= GX_PSM_FIT_OK;
                p.mu = mu;
                p.beta = beta;
                p.evalue = evalue_of(mu, beta, p.hyp, (uint32_t)p.cpsms);
                if (p.evalue < a.expect_max) p.flags |= GX_PSM_REPORTED;
            }
            if (lane == 0) atomicAdd(&a.stats[ST_FALLBACK], 1ull);
        }
    }
    if (lane == 0) {
        a.out[s] = p;
        a.raw[(size_t)s * 4 + 3] = served ? 1.0f : 0.0f;
    }
}

}  // namespace

hipError_t launch_learned_evalue(const LearnArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(learned_net_kernel, dim3((a.n + kRows - 1) / kRows), dim3(64 * kWaves), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(learned_fit_kernel, dim3(a.n), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace gx
