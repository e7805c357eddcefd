// pa_filter.hip -- host side of the learned prefilter entry points
// (pa_filter_* in include/pastis_hip.h).
//
// The model lives on the device as fp32 master weights plus a copy whose
// weight matrices are rounded to bf16 (biases stay fp32); inference and the
// forward pass of training read the rounded copy.  Adam updates the master
// weights and refreshes the copy in the same kernel.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "k_filter.h"
#include "pa_ctx.h"

using namespace pa;

namespace pa {

struct FilterState {
    Buf w, wb, mean, istd, m, v, grad, scratch, loss, order, labels, feats,
        logits, bad;
    int t = 0;                        // Adam step count
    float h_mean[kNF], h_istd[kNF];
};

void filter_free(FilterState *f)
{
    if (!f)
        return;
    for (Buf *b : {&f->w, &f->wb, &f->mean, &f->istd, &f->m, &f->v, &f->grad,
                   &f->scratch, &f->loss, &f->order, &f->labels, &f->feats,
                   &f->logits, &f->bad})
        release(*b);
    delete f;
}

}  // namespace pa

namespace {

constexpr uint32_t kMaxBatch = 65536;

#define FCHK(ctx, x)                                                         \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess)                                                \
            return fail((ctx), PA_ERR_HIP,                                   \
                        std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int need_model(pa_ctx *c)
{
    if (!c)
        return PA_ERR_ARG;
    if (!c->flt)
        return fail(c, PA_ERR_ARG, "no filter model installed");
    return PA_OK;
}

int need_sketches(pa_ctx *c)
{
    if (!c)
        return PA_ERR_ARG;
    if (!c->have_rows || !c->have_cols)
        return fail(c, PA_ERR_NOSEQ, "load row and column sequences first");
    return PA_OK;
}

int upload_stdz(pa_ctx *c)
{
    FilterState *f = c->flt;
    FCHK(c, hipMemcpyAsync(f->mean.p, f->h_mean, sizeof(f->h_mean),
                           hipMemcpyHostToDevice, c->st));
    FCHK(c, hipMemcpyAsync(f->istd.p, f->h_istd, sizeof(f->h_istd),
                           hipMemcpyHostToDevice, c->st));
    return PA_OK;
}

// mean and 1/std of every feature over the n samples (double sums in sample
// order; constant features get istd 1)
void standardization(const float *feats, uint64_t n, float *mean, float *istd)
{
    for (int k = 0; k < kNF; ++k) {
        double s = 0.0;
        for (uint64_t i = 0; i < n; ++i)
            s += feats[i * kNF + k];
        const double mu = n ? s / (double)n : 0.0;
        double q = 0.0;
        for (uint64_t i = 0; i < n; ++i) {
            const double d = feats[i * kNF + k] - mu;
            q += d * d;
        }
        const double sd = n ? std::sqrt(q / (double)n) : 0.0;
        mean[k] = (float)mu;
        istd[k] = sd > 1e-6 ? (float)(1.0 / sd) : 1.0f;
    }
}

int stage_samples(pa_ctx *c, const float *feats, const uint8_t *labels,
                  uint64_t n)
{
    FilterState *f = c->flt;
    int rc;
    if ((rc = grow(c, f->feats, n * kNF * sizeof(float))) ||
        (rc = grow(c, f->labels, n)) ||
        (rc = grow(c, f->order, n * sizeof(uint32_t))))
        return rc;
    FCHK(c, hipMemcpyAsync(f->feats.p, feats, n * kNF * sizeof(float),
                           hipMemcpyHostToDevice, c->st));
    FCHK(c, hipMemcpyAsync(f->labels.p, labels, n, hipMemcpyHostToDevice,
                           c->st));
    return PA_OK;
}

int ensure_train(pa_ctx *c, uint32_t B)
{
    FilterState *f = c->flt;
    int rc;
    if ((rc = grow(c, f->scratch, (size_t)B * kScratchRow * sizeof(float))) ||
        (rc = grow(c, f->loss, sizeof(float))))
        return rc;
    return PA_OK;
}

}  // namespace

extern "C" {

int pa_filter_set_model(pa_ctx *c, const float *params, const float *mean,
                        const float *istd)
{
    if (!c || !params)
        return c ? fail(c, PA_ERR_ARG, "null parameters") : PA_ERR_ARG;
    for (int i = 0; i < kNParam; ++i)
        if (!std::isfinite(params[i]))
            return fail(c, PA_ERR_ARG, "non-finite parameter");
    hipSetDevice(c->device);
    int rc;
    if (!c->flt) {
        c->flt = new FilterState();
        for (int k = 0; k < kNF; ++k) {
            c->flt->h_mean[k] = 0.f;
            c->flt->h_istd[k] = 1.f;
        }
    }
    FilterState *f = c->flt;
    const size_t pb = kNParam * sizeof(float);
    if ((rc = grow(c, f->w, pb)) || (rc = grow(c, f->wb, pb)) ||
        (rc = grow(c, f->m, pb)) || (rc = grow(c, f->v, pb)) ||
        (rc = grow(c, f->grad, pb)) ||
        (rc = grow(c, f->mean, kNF * sizeof(float))) ||
        (rc = grow(c, f->istd, kNF * sizeof(float))) ||
        (rc = grow(c, f->bad, sizeof(uint32_t))))
        return rc;
    if (mean)
        std::memcpy(f->h_mean, mean, sizeof(f->h_mean));
    if (istd)
        std::memcpy(f->h_istd, istd, sizeof(f->h_istd));
    FCHK(c, hipMemcpyAsync(f->w.p, params, pb, hipMemcpyHostToDevice, c->st));
    FCHK(c, hipMemsetAsync(f->m.p, 0, pb, c->st));
    FCHK(c, hipMemsetAsync(f->v.p, 0, pb, c->st));
    FCHK(c, round_weights(P<float>(f->w), P<float>(f->wb), c->st));
    if ((rc = upload_stdz(c)))
        return rc;
    f->t = 0;
    FCHK(c, hipStreamSynchronize(c->st));
    return PA_OK;
}

int pa_filter_get_model(pa_ctx *c, float *params, float *mean, float *istd)
{
    int rc;
    if ((rc = need_model(c)))
        return rc;
    hipSetDevice(c->device);
    FilterState *f = c->flt;
    if (params) {
        FCHK(c, hipMemcpyAsync(params, f->w.p, kNParam * sizeof(float),
                               hipMemcpyDeviceToHost, c->st));
        FCHK(c, hipStreamSynchronize(c->st));
    }
    if (mean)
        std::memcpy(mean, f->h_mean, sizeof(f->h_mean));
    if (istd)
        std::memcpy(istd, f->h_istd, sizeof(f->h_istd));
    return PA_OK;
}

int pa_filter_features_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                           float *d_feats, void *stream)
{
    int rc;
    if ((rc = need_sketches(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_feats)
        return fail(c, PA_ERR_ARG, "null pointer");
    hipSetDevice(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->st;
    if ((rc = grow(c, c->ctr, 32 * sizeof(uint32_t))))
        return rc;
    FCHK(c, features(d_pairs, n, P<SeqSketch>(c->rsk), (uint32_t)c->rlen.size(),
                     P<SeqSketch>(c->csk), (uint32_t)c->clen.size(), d_feats,
                     P<uint32_t>(c->ctr) + 31, s));
    return PA_OK;
}

int pa_filter_features(pa_ctx *c, const pa_pair *pairs, uint64_t n,
                       float *feats)
{
    int rc;
    if ((rc = need_sketches(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!pairs || !feats)
        return fail(c, PA_ERR_ARG, "null pointer");
    for (uint64_t i = 0; i < n; ++i)
        if (pairs[i].r >= c->rlen.size() || pairs[i].c >= c->clen.size())
            return fail(c, PA_ERR_INDEX,
                        "pair " + std::to_string(i) +
                            " names a sequence outside the loaded sets");
    hipSetDevice(c->device);
    if ((rc = grow(c, c->pairs, n * sizeof(pa_pair))) ||
        (rc = grow(c, c->scores, n * kNF * sizeof(float))))
        return rc;
    FCHK(c, hipMemcpyAsync(c->pairs.p, pairs, n * sizeof(pa_pair),
                           hipMemcpyHostToDevice, c->st));
    if ((rc = pa_filter_features_dev(c, P<pa_pair>(c->pairs), n,
                                     P<float>(c->scores), nullptr)))
        return rc;
    FCHK(c, hipMemcpyAsync(feats, c->scores.p, n * kNF * sizeof(float),
                           hipMemcpyDeviceToHost, c->st));
    FCHK(c, hipStreamSynchronize(c->st));
    return PA_OK;
}

int pa_filter_grad(pa_ctx *c, const float *feats, const uint8_t *labels,
                   uint64_t n, float pos_weight, float *grad, float *loss)
{
    int rc;
    if ((rc = need_model(c)))
        return rc;
    if (!feats || !labels || !grad || n == 0)
        return fail(c, PA_ERR_ARG, "null pointer or empty batch");
    if (n > kMaxBatch)
        return fail(c, PA_ERR_LIMIT, "gradient batch larger than 65536");
    hipSetDevice(c->device);
    FilterState *f = c->flt;
    if ((rc = stage_samples(c, feats, labels, n)) ||
        (rc = ensure_train(c, (uint32_t)n)))
        return rc;
    std::vector<uint32_t> ord(n);
    for (uint64_t i = 0; i < n; ++i)
        ord[i] = (uint32_t)i;
    FCHK(c, hipMemcpyAsync(f->order.p, ord.data(), n * 4,
                           hipMemcpyHostToDevice, c->st));
    FCHK(c, hipMemsetAsync(f->loss.p, 0, sizeof(float), c->st));
    FCHK(c, train_step(P<float>(f->feats), P<float>(f->mean), P<float>(f->istd),
                       P<uint8_t>(f->labels), P<uint32_t>(f->order), 0,
                       (uint32_t)n, P<float>(f->w), P<float>(f->wb),
                       P<float>(f->m), P<float>(f->v), P<float>(f->grad),
                       P<float>(f->scratch), P<float>(f->loss), pos_weight, 0.f,
                       0.f, 0.f, 0.f, 0, false, c->st));
    float ls = 0.f;
    FCHK(c, hipMemcpyAsync(grad, f->grad.p, kNParam * sizeof(float),
                           hipMemcpyDeviceToHost, c->st));
    FCHK(c, hipMemcpyAsync(&ls, f->loss.p, sizeof(float),
                           hipMemcpyDeviceToHost, c->st));
    FCHK(c, hipStreamSynchronize(c->st));
    if (loss)
        *loss = ls / (float)n;
    return PA_OK;
}

int pa_filter_train(pa_ctx *c, const float *feats, const uint8_t *labels,
                    uint64_t n, const pa_train_config *tc, float *epoch_loss)
{
    int rc;
    if ((rc = need_model(c)))
        return rc;
    if (!feats || !labels || !tc || n == 0)
        return fail(c, PA_ERR_ARG, "null pointer or empty training set");
    if (tc->batch == 0 || tc->batch > kMaxBatch || tc->epochs == 0)
        return fail(c, PA_ERR_ARG, "batch must be in 1..65536, epochs >= 1");
    if (n > 0xffffffffull)
        return fail(c, PA_ERR_LIMIT, "training set larger than 2^32 samples");
    if (!(tc->lr > 0.f) || !(tc->beta1 >= 0.f && tc->beta1 < 1.f) ||
        !(tc->beta2 >= 0.f && tc->beta2 < 1.f) || !(tc->eps > 0.f) ||
        !(tc->pos_weight > 0.f))
        return fail(c, PA_ERR_ARG, "bad optimizer settings");
    hipSetDevice(c->device);
    FilterState *f = c->flt;
    if (tc->standardize) {
        standardization(feats, n, f->h_mean, f->h_istd);
        if ((rc = upload_stdz(c)))
            return rc;
    }
    const uint32_t B = (uint32_t)std::min<uint64_t>(tc->batch, n);
    if ((rc = stage_samples(c, feats, labels, n)) || (rc = ensure_train(c, B)))
        return rc;
    std::vector<uint32_t> ord(n);
    uint64_t rng = tc->seed;
    for (uint32_t e = 0; e < tc->epochs; ++e) {
        for (uint64_t i = 0; i < n; ++i)
            ord[i] = (uint32_t)i;
        if (tc->seed)   // Fisher-Yates with splitmix64
            for (uint64_t i = n - 1; i > 0; --i) {
                const uint64_t j = splitmix64(rng) % (i + 1);
                std::swap(ord[i], ord[j]);
            }
        // the order buffer is consumed by this epoch's kernels before the
        // next copy is queued on the same stream
        FCHK(c, hipMemcpyAsync(f->order.p, ord.data(), n * 4,
                               hipMemcpyHostToDevice, c->st));
        FCHK(c, hipMemsetAsync(f->loss.p, 0, sizeof(float), c->st));
        for (uint64_t base = 0; base < n; base += B) {
            const uint32_t b = (uint32_t)std::min<uint64_t>(B, n - base);
            ++f->t;
            FCHK(c, train_step(P<float>(f->feats), P<float>(f->mean),
                               P<float>(f->istd), P<uint8_t>(f->labels),
                               P<uint32_t>(f->order), base, b, P<float>(f->w),
                               P<float>(f->wb), P<float>(f->m), P<float>(f->v),
                               P<float>(f->grad), P<float>(f->scratch),
                               P<float>(f->loss), tc->pos_weight, tc->lr,
                               tc->beta1, tc->beta2, tc->eps, f->t, true,
                               c->st));
        }
        float ls = 0.f;
        FCHK(c, hipMemcpyAsync(&ls, f->loss.p, sizeof(float),
                               hipMemcpyDeviceToHost, c->st));
        FCHK(c, hipStreamSynchronize(c->st));
        if (epoch_loss)
            epoch_loss[e] = ls / (float)n;
    }
    return PA_OK;
}

int pa_filter_logits(pa_ctx *c, const float *feats, uint64_t n, float *logits)
{
    int rc;
    if ((rc = need_model(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!feats || !logits)
        return fail(c, PA_ERR_ARG, "null pointer");
    hipSetDevice(c->device);
    FilterState *f = c->flt;
    if ((rc = grow(c, f->feats, n * kNF * sizeof(float))) ||
        (rc = grow(c, f->logits, n * sizeof(float))))
        return rc;
    FCHK(c, hipMemcpyAsync(f->feats.p, feats, n * kNF * sizeof(float),
                           hipMemcpyHostToDevice, c->st));
    FCHK(c, logits_of(P<float>(f->feats), n, P<float>(f->wb), P<float>(f->mean),
                      P<float>(f->istd), P<float>(f->logits), c->st));
    FCHK(c, hipMemcpyAsync(logits, f->logits.p, n * sizeof(float),
                           hipMemcpyDeviceToHost, c->st));
    FCHK(c, hipStreamSynchronize(c->st));
    return PA_OK;
}

int pa_filter_score_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                        float *d_logits, void *stream)
{
    int rc;
    if ((rc = need_model(c)) || (rc = need_sketches(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_logits)
        return fail(c, PA_ERR_ARG, "null pointer");
    hipSetDevice(c->device);
    FilterState *f = c->flt;
    hipStream_t s = stream ? (hipStream_t)stream : c->st;
    FCHK(c, score_pairs(d_pairs, n, P<SeqSketch>(c->rsk),
                        (uint32_t)c->rlen.size(), P<SeqSketch>(c->csk),
                        (uint32_t)c->clen.size(), P<float>(f->wb),
                        P<float>(f->mean), P<float>(f->istd), d_logits,
                        P<uint32_t>(f->bad), s));
    return PA_OK;
}

int pa_filter_select_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                         float tau, pa_pair *d_kept, uint32_t *d_kept_index,
                         uint64_t *n_kept, void *stream)
{
    int rc;
    if ((rc = need_model(c)) || (rc = need_sketches(c)))
        return rc;
    if (!n_kept)
        return fail(c, PA_ERR_ARG, "null pointer");
    *n_kept = 0;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_kept)
        return fail(c, PA_ERR_ARG, "null pointer");
    if (n > 0xffffffffull)
        return fail(c, PA_ERR_LIMIT, "more than 2^32 pairs");
    hipSetDevice(c->device);
    FilterState *f = c->flt;
    hipStream_t s = stream ? (hipStream_t)stream : c->st;
    if ((rc = grow(c, f->logits, n * sizeof(float))) ||
        (rc = grow(c, f->loss, 2 * sizeof(uint32_t))))
        return rc;
    uint32_t *cnt = P<uint32_t>(f->loss);
    FCHK(c, hipMemsetAsync(cnt, 0, sizeof(uint32_t), s));
    FCHK(c, hipMemsetAsync(f->bad.p, 0, sizeof(uint32_t), s));
    FCHK(c, score_pairs(d_pairs, n, P<SeqSketch>(c->rsk),
                        (uint32_t)c->rlen.size(), P<SeqSketch>(c->csk),
                        (uint32_t)c->clen.size(), P<float>(f->wb),
                        P<float>(f->mean), P<float>(f->istd),
                        P<float>(f->logits), P<uint32_t>(f->bad), s));
    FCHK(c, select_pairs(d_pairs, P<float>(f->logits), n, tau, d_kept,
                         d_kept_index, cnt, s));
    uint32_t h[2] = {0, 0};
    FCHK(c, hipMemcpyAsync(&h[0], cnt, 4, hipMemcpyDeviceToHost, s));
    FCHK(c, hipMemcpyAsync(&h[1], f->bad.p, 4, hipMemcpyDeviceToHost, s));
    FCHK(c, hipStreamSynchronize(s));
    if (h[1])
        return fail(c, PA_ERR_INDEX,
                    std::to_string(h[1]) +
                        " pairs name sequences outside the loaded sets");
    *n_kept = h[0];
    return PA_OK;
}

}  // extern "C"
