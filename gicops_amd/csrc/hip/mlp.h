// Learned e-value surrogate: a 128-256-256-3 GELU MLP on BF16 MFMA, trained
// on the GPU from (survival-histogram features -> plain tail-fit) pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gicops_amd.h"

namespace gx {

constexpr int kFeat = 128;     // input features
constexpr int kHidden = 256;   // hidden width
constexpr int kOutPad = 16;    // 3 outputs padded to one MFMA tile
constexpr int kOut = 3;        // slope, fitted log-survival at the top bin minus the anchor, confidence logit
constexpr int kCurvePts = 120; // resampled points of the fit's log-survival curve

typedef __bf16 bf16;

// Flat fp32 parameter layout shared with the host (and the CPU oracle).
struct MlpLayout {
    static constexpr int W1 = 0;                              // [kFeat][kHidden]
    static constexpr int B1 = W1 + kFeat * kHidden;           // [kHidden]
    static constexpr int W2 = B1 + kHidden;                   // [kHidden][kHidden]
    static constexpr int B2 = W2 + kHidden * kHidden;         // [kHidden]
    static constexpr int W3 = B2 + kHidden;                   // [kHidden][kOutPad], columns >= kOut are 0
    static constexpr int B3 = W3 + kHidden * kOutPad;         // [kOutPad]
    static constexpr int N = B3 + kOutPad;
};

// Output normalisation: target = out * scale + shift (slope, u).
struct MlpNorm {
    float a_shift, a_scale, u_shift, u_scale;
};

struct GemmArgs {
    int M, N, K;
    const bf16 *A;
    long long sam, sak;  // A(m, k) = A[m*sam + k*sak]
    const bf16 *B;
    long long sbk, sbn;  // B(k, n) = B[k*sbk + n*sbn]
    int mode;            // 0: C = acc (+ bias); 1: Z = acc + bias, H = gelu(Z) (bf16);
                         // 2: H = acc * gelu'(Z) (bf16); 3: H = gelu(acc + bias) (bf16), Z not kept
    float *C;            // mode 0 output [M][N] (row stride ldc)
    long long ldc;
    const float *bias;   // modes 1, 3 (optional in mode 0)
    float *Z;            // mode 1 output / mode 2 input [M][N] (row stride ldz)
    long long ldz;
    bf16 *H;             // modes 1, 2 output [M][N] (row stride ldh)
    long long ldh;
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);

// features of searched spectra: the plain fit's log-survival curve resampled
// to kCurvePts points, then stt1, end1, hyp, hyp - end1, curve length,
// log10 cpsms, the anchor and the first curve value; targets from the plain fit
struct FeatArgs {
    uint32_t n;
    const uint32_t *hist;   // n * GX_HIST_BINS
    const gx_psm *psm;      // top hits (and, for training, plain-fit mu, beta, flags)
    bf16 *feat;             // n * kFeat
    float *anchor;          // n: log10 survival at the curve's last bin (end1)
    float *target;          // n * 2: slope, u - anchor with u = slope*hyp + intercept (truncated mu, beta)
    float *loge;            // n: log10 of the plain e-value
    uint8_t *valid;         // n: a defined plain fit exists
    uint8_t *sparse;        // n: not modeled, or fewer than kMinTailBins non-empty bins below the top bin
};
constexpr int kMinTailBins = 3;

hipError_t launch_features(const FeatArgs &a, hipStream_t s);

// per-sample loss and output gradient (rows of kOutPad; only 3 used)
struct LossArgs {
    uint32_t n;
    const float *out;       // n * kOutPad
    const float *target;    // n * 2 (raw)
    const float *loge;      // n
    MlpNorm norm;
    float conf_weight;
    float tol;              // relative tolerance on log10 e for the confidence label
    bf16 *dout;             // n * kOutPad
    float *loss;            // [4]: sum of squared errors (a, u), bce, count within tolerance
};
hipError_t launch_loss(const LossArgs &a, hipStream_t s);

hipError_t launch_colsum(const bf16 *x, uint32_t rows, uint32_t cols, float *out, hipStream_t s);
hipError_t launch_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                       float eps, float b1t, float b2t, hipStream_t s);
hipError_t launch_to_bf16(const float *x, bf16 *y, uint64_t n, hipStream_t s);
hipError_t launch_gather_rows(const bf16 *src, const float *tsrc, const float *lsrc, const uint32_t *idx, uint32_t n,
                              bf16 *dst, float *tdst, float *ldst, hipStream_t s);

// route: accept the surrogate where confident, not sparse, in distribution
// (normalised outputs within kMaxZ) and with a fitted log-survival u <= 0
constexpr float kMaxZ = 4.0f;
struct RouteArgs {
    uint32_t n;
    const float *out;       // n * kOutPad
    const uint8_t *sparse;  // also set for spectra that are not modeled
    const float *anchor;    // u = anchor + predicted residual
    MlpNorm norm;
    float conf_threshold;
    double expect_max;
    gx_psm *psm;            // updated in place for accepted spectra
    uint8_t *route;         // 1 = surrogate accepted, 0 = plain fit
    unsigned long long *stats;  // [ST_FALLBACK] counter row 0
};
hipError_t launch_route(const RouteArgs &a, hipStream_t s);

}  // namespace gx
