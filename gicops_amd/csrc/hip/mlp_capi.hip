// C ABI of the learned e-value surrogate (include/gicops_amd.h, gx_mlp_*):
// parameters, training data taken from searched batches, training on the GPU
// with the project's MFMA forward/backward kernels, one-shot gradients for
// the CPU check, and the fused inference stage inside the batched search.
#include <algorithm>
#include <cmath>
#include <new>
#include <numeric>
#include <random>
#include <vector>

#include "capi_internal.h"
#include "mlp.h"

namespace gx {

struct Surrogate {
    gx_ctx *ctx = nullptr;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wb = nullptr;    // bf16 copy in MlpLayout (training GEMMs)
    bf16 *frag = nullptr;  // bf16 MFMA B fragments (FragLayout, inference)
    float *ws = nullptr;   // split-K partial products of the weight-gradient GEMMs
    static constexpr long long kWsFloats = 1ll << 22;
    MlpNorm norm{0.0f, 1.0f, 0.0f, 1.0f};
    bool norm_set = false;
    long long step = 0;
    // training set
    bf16 *ds_feat = nullptr;
    float *ds_tgt = nullptr, *ds_loge = nullptr;
    uint32_t ds_n = 0, ds_cap = 0;
    // minibatch work buffers
    uint32_t wcap = 0;
    bf16 *x = nullptr, *h1 = nullptr, *h2 = nullptr, *dz1 = nullptr, *dz2 = nullptr, *dout = nullptr;
    float *tg = nullptr, *lg = nullptr, *z1 = nullptr, *z2 = nullptr, *out = nullptr, *loss = nullptr;
    uint32_t *idx = nullptr;
};

}  // namespace gx

struct gx_mlp : gx::Surrogate {};

namespace {

void free_all(std::vector<void *> ps) {
    for (void *p : ps)
        if (p) (void)hipFree(p);
}

int ensure_work(gx_mlp *M, uint32_t n) {
    if (n <= M->wcap && M->x) return GX_SUCCESS;
    free_all({M->x, M->h1, M->h2, M->dz1, M->dz2, M->dout, M->tg, M->lg, M->z1, M->z2, M->out, M->idx});
    M->x = M->h1 = M->h2 = M->dz1 = M->dz2 = M->dout = nullptr;
    M->tg = M->lg = M->z1 = M->z2 = M->out = nullptr;
    M->idx = nullptr;
    M->wcap = 0;
    if (dalloc(&M->x, (size_t)n * kFeat) || dalloc(&M->h1, (size_t)n * kHidden) || dalloc(&M->h2, (size_t)n * kHidden) ||
        dalloc(&M->dz1, (size_t)n * kHidden) || dalloc(&M->dz2, (size_t)n * kHidden) ||
        dalloc(&M->dout, (size_t)n * kOutPad) || dalloc(&M->tg, (size_t)n * kTarget) || dalloc(&M->lg, n) ||
        dalloc(&M->z1, (size_t)n * kHidden) || dalloc(&M->z2, (size_t)n * kHidden) ||
        dalloc(&M->out, (size_t)n * kOutPad) || dalloc(&M->idx, n))
        return GX_ERR_BAD_MEM_ALLOC;
    M->wcap = n;
    return GX_SUCCESS;
}

// forward on rows x[0..n): Z1/H1, Z2/H2 (kept for the backward pass), out
hipError_t forward(gx_mlp *M, const bf16 *x, uint32_t n, hipStream_t s, bool keep_z, bf16 *h1, bf16 *h2, float *z1,
                   float *z2, float *out) {
    hipError_t e;
    GemmArgs g{};
    g.M = (int)n;
    g.N = kHidden;
    g.K = kFeat;
    g.A = x;
    g.sam = kFeat;
    g.sak = 1;
    g.B = M->wb + MlpLayout::W1;
    g.sbk = kHidden;
    g.sbn = 1;
    g.mode = keep_z ? 1 : 3;
    g.bias = M->w + MlpLayout::B1;
    g.Z = z1;
    g.ldz = kHidden;
    g.H = h1;
    g.ldh = kHidden;
    if ((e = launch_gemm(g, s))) return e;
    g.K = kHidden;
    g.A = h1;
    g.sam = kHidden;
    g.B = M->wb + MlpLayout::W2;
    g.bias = M->w + MlpLayout::B2;
    g.Z = z2;
    g.H = h2;
    if ((e = launch_gemm(g, s))) return e;
    g.N = kOutPad;
    g.A = h2;
    g.B = M->wb + MlpLayout::W3;
    g.sbk = kOutPad;
    g.mode = 0;
    g.C = out;
    g.ldc = kOutPad;
    g.bias = M->w + MlpLayout::B3;
    return launch_gemm(g, s);
}

// gradients of the loss whose output gradient is M->dout, into M->g
hipError_t backward(gx_mlp *M, const bf16 *x, uint32_t n, hipStream_t s) {
    hipError_t e;
    GemmArgs g{};
    g.ws = M->ws;
    g.ws_floats = M->ws ? Surrogate::kWsFloats : 0;
    // dW3 = H2^T dO
    g.M = kHidden;
    g.N = kOutPad;
    g.K = (int)n;
    g.A = M->h2;
    g.sam = 1;
    g.sak = kHidden;
    g.B = M->dout;
    g.sbk = kOutPad;
    g.sbn = 1;
    g.mode = 0;
    g.C = M->g + MlpLayout::W3;
    g.ldc = kOutPad;
    g.colsum_out = M->g + MlpLayout::B3;  // db3 = column sums of dO
    if ((e = launch_gemm(g, s))) return e;
    // dZ2 = (dO W3^T) * gelu'(Z2)
    GemmArgs d{};
    d.M = (int)n;
    d.N = kHidden;
    d.K = kOutPad;
    d.A = M->dout;
    d.sam = kOutPad;
    d.sak = 1;
    d.B = M->wb + MlpLayout::W3;
    d.sbk = 1;
    d.sbn = kOutPad;
    d.mode = 2;
    d.Z = M->z2;
    d.ldz = kHidden;
    d.H = M->dz2;
    d.ldh = kHidden;
    if ((e = launch_gemm(d, s))) return e;
    // dW2 = H1^T dZ2
    g.M = kHidden;
    g.N = kHidden;
    g.K = (int)n;
    g.A = M->h1;
    g.sam = 1;
    g.sak = kHidden;
    g.B = M->dz2;
    g.sbk = kHidden;
    g.sbn = 1;
    g.C = M->g + MlpLayout::W2;
    g.ldc = kHidden;
    g.colsum_out = M->g + MlpLayout::B2;  // db2 = column sums of dZ2
    if ((e = launch_gemm(g, s))) return e;
    // dZ1 = (dZ2 W2^T) * gelu'(Z1)
    d.K = kHidden;
    d.A = M->dz2;
    d.sam = kHidden;
    d.B = M->wb + MlpLayout::W2;
    d.sbk = 1;
    d.sbn = kHidden;
    d.Z = M->z1;
    d.H = M->dz1;
    if ((e = launch_gemm(d, s))) return e;
    // dW1 = X^T dZ1
    g.M = kFeat;
    g.N = kHidden;
    g.K = (int)n;
    g.A = x;
    g.sam = 1;
    g.sak = kFeat;
    g.B = M->dz1;
    g.sbk = kHidden;
    g.sbn = 1;
    g.C = M->g + MlpLayout::W1;
    g.ldc = kHidden;
    g.colsum_out = M->g + MlpLayout::B1;  // db1 = column sums of dZ1
    return launch_gemm(g, s);
}

}  // namespace

extern "C" int gx_mlp_param_count(void) { return MlpLayout::N; }

extern "C" int gx_mlp_create(gx_ctx *c, gx_mlp **out) {
    if (!c || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    gx_mlp *M = new (std::nothrow) gx_mlp();
    if (!M) return GX_ERR_BAD_MEM_ALLOC;
    M->ctx = c;
    if (dalloc(&M->w, MlpLayout::N) || dalloc(&M->g, MlpLayout::N) || dalloc(&M->m, MlpLayout::N) ||
        dalloc(&M->v, MlpLayout::N) || dalloc(&M->wb, MlpLayout::N) || dalloc(&M->frag, FragLayout::N) ||
        dalloc(&M->loss, 4) || dalloc(&M->ws, (size_t)Surrogate::kWsFloats)) {
        free_all({M->w, M->g, M->m, M->v, M->wb, M->frag, M->loss, M->ws});
        delete M;
        return GX_ERR_BAD_MEM_ALLOC;
    }
    *out = M;
    return GX_SUCCESS;
}

extern "C" void gx_mlp_destroy(gx_mlp *M) {
    if (!M) return;
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
    if (M->ctx->surrogate == M) {
        M->ctx->surrogate = nullptr;
        M->ctx->scratch_gen++;
    }
    free_all({M->w, M->g, M->m, M->v, M->wb, M->frag, M->loss, M->ws, M->ds_feat, M->ds_tgt, M->ds_loge, M->x, M->h1, M->h2, M->dz1,
              M->dz2, M->dout, M->tg, M->lg, M->z1, M->z2, M->out, M->idx});
    delete M;
}

static int upload_params(gx_mlp *M, const std::vector<float> &h) {
    hipStream_t s = M->ctx->stream;
    GX_HIP(hipStreamSynchronize(s));  // the stream does not wait for the null stream's copies
    GX_HIP(hipMemcpy(M->w, h.data(), sizeof(float) * MlpLayout::N, hipMemcpyHostToDevice));
    std::vector<float> z(MlpLayout::N, 0.0f);
    GX_HIP(hipMemcpy(M->m, z.data(), sizeof(float) * MlpLayout::N, hipMemcpyHostToDevice));
    GX_HIP(hipMemcpy(M->v, z.data(), sizeof(float) * MlpLayout::N, hipMemcpyHostToDevice));
    GX_HIP(launch_to_bf16(M->w, M->wb, MlpLayout::N, s));
    GX_HIP(launch_frag_weights(M->w, M->frag, s));
    GX_HIP(hipStreamSynchronize(s));
    M->step = 0;
    M->ctx->scratch_gen++;
    return GX_SUCCESS;
}

extern "C" int gx_mlp_init(gx_mlp *M, uint64_t seed) {
    if (!M) return GX_ERR_INVLD_PTR;
    std::lock_guard<std::mutex> lk(M->ctx->mu);
    std::mt19937_64 rng(seed);
    auto uni = [&](double lim) { return (float)((((double)(rng() >> 11) * (1.0 / 9007199254740992.0)) * 2 - 1) * lim); };
    std::vector<float> h(MlpLayout::N, 0.0f);
    const double l1 = std::sqrt(6.0 / (kFeat + kHidden)), l2 = std::sqrt(6.0 / (2 * kHidden)),
                 l3 = std::sqrt(6.0 / (kHidden + kOut));
    for (int i = 0; i < kFeat * kHidden; i++) h[MlpLayout::W1 + i] = uni(l1);
    for (int i = 0; i < kHidden * kHidden; i++) h[MlpLayout::W2 + i] = uni(l2);
    for (int i = 0; i < kHidden; i++)
        for (int o = 0; o < kOut; o++) h[MlpLayout::W3 + i * kOutPad + o] = uni(l3);
    return upload_params(M, h);
}

extern "C" int gx_mlp_set_params(gx_mlp *M, const float *params, size_t n, const float *norm) {
    if (!M || !params) return GX_ERR_INVLD_PTR;
    if (n != (size_t)MlpLayout::N) return GX_ERR_INVLD_SIZE;
    std::lock_guard<std::mutex> lk(M->ctx->mu);
    std::vector<float> h(params, params + n);
    for (int i = 0; i < kHidden; i++)  // padding columns stay zero
        for (int o = kOut; o < kOutPad; o++) h[MlpLayout::W3 + i * kOutPad + o] = 0.0f;
    for (int o = kOut; o < kOutPad; o++) h[MlpLayout::B3 + o] = 0.0f;
    if (norm) {
        M->norm = MlpNorm{norm[0], norm[1], norm[2], norm[3]};
        M->norm_set = true;
    }
    return upload_params(M, h);
}

extern "C" int gx_mlp_get_params(gx_mlp *M, float *params, size_t n, float *norm) {
    if (!M || !params) return GX_ERR_INVLD_PTR;
    if (n != (size_t)MlpLayout::N) return GX_ERR_INVLD_SIZE;
    std::lock_guard<std::mutex> lk(M->ctx->mu);
    GX_HIP(hipStreamSynchronize(M->ctx->stream));
    GX_HIP(hipMemcpy(params, M->w, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (norm) {
        norm[0] = M->norm.a_shift;
        norm[1] = M->norm.a_scale;
        norm[2] = M->norm.u_shift;
        norm[3] = M->norm.u_scale;
    }
    return GX_SUCCESS;
}

namespace {

int alloc_features(gx_dqueries *q) {
    if (q->l_feat) return GX_SUCCESS;
    const size_t ns = q->n_spectra;
    if (dalloc(&q->l_feat, ns * kFeat) || dalloc(&q->l_tgt, ns * kTarget) || dalloc(&q->l_loge, ns) ||
        dalloc(&q->l_valid, ns) || dalloc(&q->l_serv, ns))
        return GX_ERR_BAD_MEM_ALLOC;
    q->buf_gen++;
    return GX_SUCCESS;
}

FeatArgs feat_args(gx_dqueries *q) {
    FeatArgs fa{};
    fa.n = q->n_spectra;
    fa.hist = q->hist;
    fa.psm = q->psm;
    fa.feat = q->l_feat;
    fa.target = q->l_tgt;
    fa.loge = q->l_loge;
    fa.valid = q->l_valid;
    fa.servable = q->l_serv;
    return fa;
}

// features and plain-fit targets of a searched batch
int batch_features(gx_ctx *c, gx_dqueries *q) {
    int st = alloc_features(q);
    if (st) return st;
    GX_HIP(launch_features(feat_args(q), c->stream));
    return GX_SUCCESS;
}

}  // namespace

extern "C" int gx_mlp_add_training(gx_mlp *M, gx_dqueries *q, uint64_t *n_added) {
    if (!M || !q) return GX_ERR_INVLD_PTR;
    if (!q->searched || !q->hist) return GX_ERR_NOT_READY;
    gx_ctx *c = M->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    int st = batch_features(c, q);
    if (st) return st;
    // the spectra the surrogate could serve: a servable curve and a defined plain fit
    std::vector<uint8_t> valid(q->n_spectra), serv(q->n_spectra);
    GX_HIP(hipMemcpyAsync(valid.data(), q->l_valid, q->n_spectra, hipMemcpyDeviceToHost, c->stream));
    GX_HIP(hipMemcpyAsync(serv.data(), q->l_serv, q->n_spectra, hipMemcpyDeviceToHost, c->stream));
    GX_HIP(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> sel;
    for (uint32_t i = 0; i < q->n_spectra; i++)
        if (valid[i] && serv[i]) sel.push_back(i);
    const uint32_t add = (uint32_t)sel.size();
    if (M->ds_n + add > M->ds_cap) {
        const uint32_t cap = std::max<uint32_t>(M->ds_n + add, M->ds_cap * 2 + 4096);
        bf16 *f = nullptr;
        float *t = nullptr, *l = nullptr;
        if (dalloc(&f, (size_t)cap * kFeat) || dalloc(&t, (size_t)cap * kTarget) || dalloc(&l, cap)) {
            free_all({f, t, l});
            return GX_ERR_BAD_MEM_ALLOC;
        }
        if (M->ds_n) {
            GX_HIP(hipMemcpyAsync(f, M->ds_feat, sizeof(bf16) * M->ds_n * kFeat, hipMemcpyDeviceToDevice, c->stream));
            GX_HIP(hipMemcpyAsync(t, M->ds_tgt, sizeof(float) * M->ds_n * kTarget, hipMemcpyDeviceToDevice,
                                  c->stream));
            GX_HIP(hipMemcpyAsync(l, M->ds_loge, sizeof(float) * M->ds_n, hipMemcpyDeviceToDevice, c->stream));
            GX_HIP(hipStreamSynchronize(c->stream));
        }
        free_all({M->ds_feat, M->ds_tgt, M->ds_loge});
        M->ds_feat = f;
        M->ds_tgt = t;
        M->ds_loge = l;
        M->ds_cap = cap;
    }
    if (add) {
        uint32_t *d_sel = nullptr;
        if (dalloc(&d_sel, add)) return GX_ERR_BAD_MEM_ALLOC;
        GX_HIP(hipMemcpy(d_sel, sel.data(), sizeof(uint32_t) * add, hipMemcpyHostToDevice));
        const hipError_t e = launch_gather_rows(q->l_feat, q->l_tgt, q->l_loge, d_sel, add,
                                                M->ds_feat + (size_t)M->ds_n * kFeat,
                                                M->ds_tgt + (size_t)M->ds_n * kTarget, M->ds_loge + M->ds_n,
                                                c->stream);
        const hipError_t e2 = hipStreamSynchronize(c->stream);
        (void)hipFree(d_sel);
        if (e != hipSuccess || e2 != hipSuccess) {
            note_error("gather training rows", __LINE__, e != hipSuccess ? e : e2);
            return GX_ERR_DEVICE;
        }
    }
    M->ds_n += add;
    if (n_added) *n_added = add;
    return GX_SUCCESS;
}

extern "C" int gx_mlp_train(gx_mlp *M, const gx_train_params *tp, gx_train_report *rep) {
    if (!M || !tp) return GX_ERR_INVLD_PTR;
    if (tp->batch < 16 || tp->epochs < 1 || tp->epochs > GX_TRAIN_MAX_EPOCHS) return GX_ERR_INVLD_PARAM;
    gx_ctx *c = M->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    if (M->ds_n < (uint32_t)tp->batch) return GX_ERR_NOT_ENOUGH_DATA;
    hipStream_t s = c->stream;
    const uint32_t N = M->ds_n, B = (uint32_t)tp->batch;
    int st = ensure_work(M, B);
    if (st) return st;
    if (!M->norm_set) {  // standardise the two positions (fractions of the curve) over the training rows
        std::vector<float> t((size_t)N * kTarget);
        GX_HIP(hipMemcpy(t.data(), M->ds_tgt, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
        double s1 = 0, s1q = 0, s2 = 0, s2q = 0, cnt = 0;
        for (uint32_t i = 0; i < N; i++) {
            const float *r = &t[(size_t)i * kTarget];
            if (r[3] < 0.5f || r[2] <= 0.0f) continue;
            const double p1 = r[0] / r[2], p2 = r[1] / r[2];
            s1 += p1;
            s1q += p1 * p1;
            s2 += p2;
            s2q += p2 * p2;
            cnt += 1;
        }
        cnt = std::max(cnt, 1.0);
        const double m1 = s1 / cnt, m2 = s2 / cnt;
        M->norm = MlpNorm{(float)m1, (float)std::sqrt(std::max(s1q / cnt - m1 * m1, 1e-8)), (float)m2,
                          (float)std::sqrt(std::max(s2q / cnt - m2 * m2, 1e-8))};
        M->norm_set = true;
    }
    if (rep) std::memset(rep, 0, sizeof(*rep));
    std::mt19937_64 rng(tp->seed);
    std::vector<uint32_t> perm(N);
    uint32_t *d_perm = nullptr;
    if (dalloc(&d_perm, N)) return GX_ERR_BAD_MEM_ALLOC;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int rc = GX_SUCCESS;
    for (int ep = 0; ep < tp->epochs && rc == GX_SUCCESS; ep++) {
        std::iota(perm.begin(), perm.end(), 0u);
        std::shuffle(perm.begin(), perm.end(), rng);
        // the context stream does not synchronise with the null stream: copy
        // on it, and wait until the previous epoch no longer reads d_perm
        if (hipStreamSynchronize(s) != hipSuccess ||
            hipMemcpyAsync(d_perm, perm.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = GX_ERR_DEVICE;
            break;
        }
        if (launch_fill_u32(reinterpret_cast<uint32_t *>(M->loss), 4, 0, s) != hipSuccess) {
            rc = GX_ERR_DEVICE;
            break;
        }
        const float lr = tp->lr * (tp->lr_decay > 0 ? std::pow(tp->lr_decay, (float)ep) : 1.0f);
        for (uint32_t b0 = 0; b0 + B <= N; b0 += B) {
            hipError_t e = launch_gather_rows(M->ds_feat, M->ds_tgt, M->ds_loge, d_perm + b0, B, M->x, M->tg, M->lg, s);
            if (!e) e = forward(M, M->x, B, s, true, M->h1, M->h2, M->z1, M->z2, M->out);
            LossArgs la{B, M->out, M->tg, M->norm, tp->conf_weight, M->dout, M->loss};
            if (!e) e = launch_loss(la, s);
            if (!e) e = backward(M, M->x, B, s);
            M->step++;
            const float b1t = std::pow(b1, (float)M->step), b2t = std::pow(b2, (float)M->step);
            if (!e) e = launch_adam(M->w, M->g, M->m, M->v, MlpLayout::N, lr, b1, b2, eps, b1t, b2t, s);
            if (!e) e = launch_to_bf16(M->w, M->wb, MlpLayout::N, s);
            if (e) {
                note_error("training step", __LINE__, e);
                rc = GX_ERR_DEVICE;
                break;
            }
        }
        float l[4];
        if (rc == GX_SUCCESS && (hipMemcpyAsync(l, M->loss, sizeof(l), hipMemcpyDeviceToHost, s) != hipSuccess ||
                                 hipStreamSynchronize(s) != hipSuccess))
            rc = GX_ERR_DEVICE;
        if (rc == GX_SUCCESS && rep) {
            const float nb = (float)((N / B) * B);
            rep->mse_mark[ep] = l[0] / nb;
            rep->mse_mark2[ep] = l[1] / nb;
            rep->bce[ep] = l[2] / nb;
            rep->in_window[ep] = l[3] / nb;
            rep->epochs = ep + 1;
        }
    }
    if (rc == GX_SUCCESS && launch_frag_weights(M->w, M->frag, s) != hipSuccess) rc = GX_ERR_DEVICE;
    (void)hipStreamSynchronize(s);
    (void)hipFree(d_perm);
    if (rep) rep->samples = N;
    c->scratch_gen++;
    return rc;
}

extern "C" int gx_mlp_grad(gx_mlp *M, const float *feat, const float *target, uint32_t n, const float *norm,
                           float conf_weight, float *grad, float *out, float *loss) {
    if (!M || !feat || !target || !norm || !grad) return GX_ERR_INVLD_PTR;
    if (n == 0) return GX_ERR_INVLD_SIZE;
    gx_ctx *c = M->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    hipStream_t s = c->stream;
    int st = ensure_work(M, n);
    if (st) return st;
    float *fx = nullptr;
    if (dalloc(&fx, (size_t)n * kFeat)) return GX_ERR_BAD_MEM_ALLOC;
    const MlpNorm nm{norm[0], norm[1], norm[2], norm[3]};
    hipError_t e = hipStreamSynchronize(s);
    if (!e) e = hipMemcpy(fx, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(M->tg, target, sizeof(float) * n * kTarget, hipMemcpyHostToDevice);
    if (!e) e = launch_to_bf16(fx, M->x, (uint64_t)n * kFeat, s);
    if (!e) e = launch_fill_u32(reinterpret_cast<uint32_t *>(M->loss), 4, 0, s);
    if (!e) e = forward(M, M->x, n, s, true, M->h1, M->h2, M->z1, M->z2, M->out);
    LossArgs la{n, M->out, M->tg, nm, conf_weight, M->dout, M->loss};
    if (!e) e = launch_loss(la, s);
    if (!e) e = backward(M, M->x, n, s);
    if (!e) e = hipStreamSynchronize(s);
    if (!e) e = hipMemcpy(grad, M->g, sizeof(float) * MlpLayout::N, hipMemcpyDeviceToHost);
    if (!e && out) e = hipMemcpy(out, M->out, sizeof(float) * n * kOutPad, hipMemcpyDeviceToHost);
    if (!e && loss) e = hipMemcpy(loss, M->loss, sizeof(float) * 4, hipMemcpyDeviceToHost);
    (void)hipFree(fx);
    if (e) {
        note_error("gx_mlp_grad", __LINE__, e);
        return GX_ERR_DEVICE;
    }
    return GX_SUCCESS;
}

extern "C" int gx_mlp_predict(gx_mlp *M, const float *feat, uint32_t n, float *out) {
    if (!M || !feat || !out) return GX_ERR_INVLD_PTR;
    if (n == 0) return GX_ERR_INVLD_SIZE;
    gx_ctx *c = M->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    hipStream_t s = c->stream;
    int st = ensure_work(M, n);
    if (st) return st;
    float *fx = nullptr;
    if (dalloc(&fx, (size_t)n * kFeat)) return GX_ERR_BAD_MEM_ALLOC;
    hipError_t e = hipStreamSynchronize(s);
    if (!e) e = hipMemcpy(fx, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice);
    if (!e) e = launch_to_bf16(fx, M->x, (uint64_t)n * kFeat, s);
    if (!e) e = forward(M, M->x, n, s, false, M->h1, M->h2, nullptr, nullptr, M->out);
    if (!e) e = hipStreamSynchronize(s);
    if (!e) e = hipMemcpy(out, M->out, sizeof(float) * n * kOutPad, hipMemcpyDeviceToHost);
    (void)hipFree(fx);
    if (e) {
        note_error("gx_mlp_predict", __LINE__, e);
        return GX_ERR_DEVICE;
    }
    return GX_SUCCESS;
}

extern "C" int gx_ctx_set_surrogate(gx_ctx *c, gx_mlp *M) {
    if (!c) return GX_ERR_INVLD_PTR;
    if (M && M->ctx != c) return GX_ERR_INVLD_PARAM;
    std::lock_guard<std::mutex> lk(c->mu);
    c->surrogate = M;
    c->scratch_gen++;  // graphs captured with the previous model are stale
    return GX_SUCCESS;
}

extern "C" int gx_features_download(gx_ctx *c, gx_dqueries *q, float *feat, float *target, float *loge,
                                    uint8_t *valid, uint8_t *servable) {
    if (!c || !q) return GX_ERR_INVLD_PTR;
    if (!q->searched || !q->hist) return GX_ERR_NOT_READY;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    int st = batch_features(c, q);
    if (st) return st;
    GX_HIP(hipStreamSynchronize(c->stream));
    const size_t ns = q->n_spectra;
    if (feat) {
        std::vector<bf16> f(ns * kFeat);
        GX_HIP(hipMemcpy(f.data(), q->l_feat, sizeof(bf16) * f.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); i++) feat[i] = (float)f[i];
    }
    if (target) GX_HIP(hipMemcpy(target, q->l_tgt, sizeof(float) * ns * kTarget, hipMemcpyDeviceToHost));
    if (loge) GX_HIP(hipMemcpy(loge, q->l_loge, sizeof(float) * ns, hipMemcpyDeviceToHost));
    if (valid) GX_HIP(hipMemcpy(valid, q->l_valid, ns, hipMemcpyDeviceToHost));
    if (servable) GX_HIP(hipMemcpy(servable, q->l_serv, ns, hipMemcpyDeviceToHost));
    return GX_SUCCESS;
}

extern "C" int gx_surrogate_outputs_download(gx_ctx *c, gx_dqueries *q, float *out) {
    if (!c || !q || !out) return GX_ERR_INVLD_PTR;
    if (!q->searched || !q->l_out || !q->learned_done) return GX_ERR_NOT_READY;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    GX_HIP(hipStreamSynchronize(c->stream));
    GX_HIP(hipMemcpy(out, q->l_out, sizeof(float) * 4 * q->n_spectra, hipMemcpyDeviceToHost));
    return GX_SUCCESS;
}

namespace gx {

// the learned e-value stage of a batched search: one fused kernel (captured
// into graphs like the search)
hipError_t enqueue_learned_evalue(gx_ctx *c, gx_dqueries *q, const EvalArgs &ea, float conf, hipStream_t s) {
    gx_mlp *M = static_cast<gx_mlp *>(c->surrogate);
    LearnArgs la{};
    la.n = q->n_spectra;
    la.buckets = ea.buckets;
    la.min_cpsm = ea.min_cpsm;
    la.expect_max = ea.expect_max;
    la.best_key = ea.best_key;
    la.best_eq = ea.best_eq;
    la.best_shp = ea.best_shp;
    la.cpsms = ea.cpsms;
    la.hist = ea.hist;
    la.frag = M->frag;
    la.params = M->w;
    la.norm = M->norm;
    // sigmoid(logit) >= conf  <=>  logit >= log(conf / (1 - conf))
    la.conf_logit = conf <= 0.0f ? -INFINITY : conf >= 1.0f ? INFINITY : std::log(conf / (1.0f - conf));
    la.out = ea.out;
    la.raw = q->l_out;
    la.stats = ea.stats;
    return launch_learned_evalue(la, s);
}

int ensure_learned(gx_ctx *c, gx_dqueries *q) {
    if (!c->surrogate) return GX_ERR_NOT_READY;
    if (!q->l_out) {
        q->buf_gen++;
        if (dalloc(&q->l_out, (size_t)q->n_spectra * 4)) return GX_ERR_BAD_MEM_ALLOC;
    }
    return GX_SUCCESS;
}

}  // namespace gx
