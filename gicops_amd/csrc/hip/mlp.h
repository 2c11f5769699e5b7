// Learned e-value surrogate: a 128-256-256-3 GELU MLP on BF16 MFMA, trained
// on the GPU from (survival-curve features -> plain tail-fit window) pairs.
//
// The plain e-value model (expeRT::ModelTailFit, expeRT.cpp:396-559) is a
// least-squares line through the log10 survival curve between two marks: one
// before the first point at or below 22 % of the curve's drop (`mark`) and
// the last point at or above 87 % of it (`mark2`), expeRT.cpp:477-515. The
// network predicts where those
// two marks lie (as fractions of the curve) and how sure it is. The learned
// e-value stage checks each predicted mark against its defining inequality
// in one wave-wide pass over the points around it, fits the line on the
// window in parallel, and falls back to the plain fit on low confidence, a
// short or degenerate curve, or a failed check.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gicops_amd.h"

namespace gx {

struct DevBucket;

constexpr int kFeat = 128;     // input features
constexpr int kHidden = 256;   // hidden width
constexpr int kOutPad = 16;    // 3 outputs padded to one MFMA tile
constexpr int kOut = 3;        // mark position, mark2 position (fractions of the curve), confidence logit
constexpr int kCurvePts = 120; // resampled points of the log-survival curve
constexpr int kTarget = 4;     // per training row: mark, mark2, n - 1, servable
constexpr int kMinCurve = 8;   // shortest curve the surrogate serves
constexpr int kSnapLanes = 32; // points checked around each predicted mark

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

// Weights as MFMA B fragments (bf16): for layer L, output tile nt and k step
// ks, lane l holds W[ks*32 + 8*(l/16) + j][nt*16 + l%16], j < 8, in 16
// contiguous bytes, so a wave loads one fragment with one coalesced load.
struct FragLayout {
    static constexpr int L1 = 0;                                          // 16 nt x 4 ks
    static constexpr int L2 = L1 + (kHidden / 16) * (kFeat / 32) * 512;   // 16 nt x 8 ks
    static constexpr int L3 = L2 + (kHidden / 16) * (kHidden / 32) * 512; // 1 nt x 8 ks
    static constexpr int N = L3 + (kHidden / 32) * 512;
};

// Output normalisation: position = out * scale + shift (mark, mark2).
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
    // mode 0 only (optional): workspace for splitting K over blocks when the
    // output has few tiles (weight gradients, K = rows); the partial tiles are
    // summed in a fixed order, so results do not depend on scheduling
    float *ws;
    long long ws_floats;
    // mode 0 only (optional): A gets an extra row M of ones, whose output
    // row, the column sums of B (a bias gradient), goes here [N]
    float *colsum_out;
    int ksplit;          // set by launch_gemm
    int kchunk;          // set by launch_gemm
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);

// Training rows of searched spectra: the features, and as targets the plain
// fit's exact marks, the curve length and whether the surrogate could serve
// the spectrum at all (a curve of at least kMinCurve points with no
// replaced values and marks that define a window).
struct FeatArgs {
    uint32_t n;
    const uint32_t *hist;   // n * GX_HIST_BINS
    const gx_psm *psm;      // top hits and the plain fit's mu, beta, flags
    bf16 *feat;             // n * kFeat
    float *target;          // n * kTarget: mark, mark2, n - 1, servable
    float *loge;            // n: log10 of the plain e-value (0 if undefined)
    uint8_t *valid;         // n: a defined plain fit exists
    uint8_t *servable;      // n: target[3]
};

hipError_t launch_features(const FeatArgs &a, hipStream_t s);

// per-sample loss and output gradient (rows of kOutPad; only 3 used)
struct LossArgs {
    uint32_t n;
    const float *out;       // n * kOutPad
    const float *target;    // n * kTarget
    MlpNorm norm;
    float conf_weight;
    bf16 *dout;             // n * kOutPad
    float *loss;            // [4]: squared errors (mark, mark2) over servable rows, bce, rows labelled 1
};
hipError_t launch_loss(const LossArgs &a, hipStream_t s);

hipError_t launch_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                       float eps, float b1t, float b2t, hipStream_t s);
hipError_t launch_to_bf16(const float *x, bf16 *y, uint64_t n, hipStream_t s);
hipError_t launch_frag_weights(const float *w, bf16 *frag, hipStream_t s);
hipError_t launch_gather_rows(const bf16 *src, const float *tsrc, const float *lsrc, const uint32_t *idx, uint32_t n,
                              bf16 *dst, float *tdst, float *ldst, hipStream_t s);

// The learned e-value stage of a batched search, two kernels: per block of
// 16 spectra the features and the three layers on MFMA (outputs to `raw`),
// then per spectrum the window check and parallel fit, or the plain fit.
struct LearnArgs {
    uint32_t n;
    const DevBucket *buckets;
    int min_cpsm;
    double expect_max;
    const unsigned long long *best_key;
    const uint32_t *best_eq;
    const uint32_t *best_shp;
    const uint32_t *cpsms;
    const uint32_t *hist;
    const bf16 *frag;       // FragLayout
    const float *params;    // fp32 MlpLayout (biases)
    MlpNorm norm;
    float conf_logit;       // serve when the confidence logit is >= this
    gx_psm *out;
    float *raw;             // n * 4: outputs 0..2 and 1.0 where the surrogate served
    unsigned long long *stats;  // ST_FALLBACK counter row 0
};
hipError_t launch_learned_evalue(const LearnArgs &a, hipStream_t s);

}  // namespace gx
