// k_filter.hip -- learned candidate prefilter (features + MLP 16-64-64-1).
//
// The filter sits beside the reference's k-mer prefilter (count <= ckthr ||
// score <= mosthr, src/DistPWRunner.cpp:48-54): it scores each candidate
// pair from cheap per-pair features and drops pairs the exact aligner would
// reject.  Weights and activations are bf16 (round to nearest even), every
// dot product accumulates in fp32 in a fixed order with explicitly rounded
// multiplies and adds, so the forward pass matches the CPU oracle bit for bit
// and training is deterministic.
//
// Features (16, fp32) of pair (r, c) with lengths Lr, Lc, shared k-mer count,
// max k-mer self-score kscore and seeds (ro1, co1), (ro2, co2):
//   0 log2(Lr)/16          1 log2(Lc)/16          2 min(L)/max(L)
//   3 log2(1+count)/8      4 kscore/64            5 (co1-ro1)/max(L)
//   6 count>1 ? |(co2-ro2)-(co1-ro1)|/max(L) : 0
//   7 ro1/Lr  8 co1/Lc  9 count>1 ? ro2/Lr : 0  10 count>1 ? co2/Lc : 0
//   11 count/max(1, min(L)-5)
//   12 murphy10 6-mer MinHash agreement (16 hashes)
//   13 sum_i min(pr_i, pc_i), 14 sum_i pr_i*pc_i over the murphy10
//      composition fractions
//   15 (min(ro1,co1) + min(Lr-ro1, Lc-co1)) / max(L): length of the seed
//      diagonal's overlap relative to the longer sequence
#include "k_filter.h"

// Every fp32 operation below rounds on its own, as in the CPU oracle: hipcc
// contracts a * b + c into an FMA by default (also through __fmul_rn /
// __fadd_rn, whose bodies come from the HIP headers), so the Makefile builds
// this file with -ffp-contract=off.  fmaf is used where the product of two
// bf16 values is exact, which makes the FMA equal to the separately rounded
// multiply and add.

namespace pa {
namespace {

// residue code (ARNDCQEGHILKMFPSTWYVBZX*J) -> murphy10 class, as
// Murphy10ProteinAlph's init() numbers them (inc/Alphabet.hpp:89-103,
// src/Alphabet.cpp:31-51); codes 10..14 are B Z X * J
__constant__ uint8_t kMurphy[25] = {0, 1, 2, 2, 3, 2, 2, 4, 5, 6, 6, 1, 6,
                                    7, 8, 9, 9, 7, 7, 6, 10, 11, 12, 13, 14};

// xorshift-multiply 32-bit mixer ("lowbias32")
__device__ inline uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ inline float bf16r(float x)
{
    uint32_t u = __float_as_uint(x);
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    return __uint_as_float(u);
}

__global__ void k_seq_sketch(const uint64_t *__restrict__ offs, uint64_t nseq,
                             SeqSketch *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseq)
        return;
    const uint8_t *s = (const uint8_t *)(uintptr_t)offs[i];
    const int L = (int)(offs[i + 1] - offs[i]);
    uint32_t cnt[10] = {0};
    uint32_t mh[kSketch];
    for (int h = 0; h < kSketch; ++h)
        mh[h] = 0xffffffffu;
    uint32_t code = 0, tot = 0;
    const uint32_t p5 = 15u * 15u * 15u * 15u * 15u;
    for (int j = 0; j < L; ++j) {
        const uint32_t m = kMurphy[s[j]];
        if (m < 10) {
            cnt[m]++;
            tot++;
        }
        code = (code % p5) * 15u + m;
        if (j >= kK - 1) {
            for (int h = 0; h < kSketch; ++h) {
                const uint32_t v = mix32(code ^ (0x9e3779b9u * (uint32_t)(h + 1)));
                mh[h] = v < mh[h] ? v : mh[h];
            }
        }
    }
    SeqSketch o;
    const float ft = (float)(tot ? tot : 1u);
    for (int k = 0; k < 10; ++k)
        o.comp[k] = __fdiv_rn((float)cnt[k], ft);
    for (int h = 0; h < kSketch; ++h)
        o.mh[h] = mh[h];
    o.len = (uint32_t)L;
    o.pad = 0;
    out[i] = o;
}

// log2 of a non-negative integer, correctly rounded to fp32 (double log2)
__device__ inline float lg2(float x) { return (float)log2((double)x); }

__device__ inline void pair_features(const pa_pair &p, const SeqSketch &R,
                                     const SeqSketch &C, float *f)
{
    const float Lr = (float)R.len, Lc = (float)C.len;
    const float Lr1 = Lr > 1.f ? Lr : 1.f, Lc1 = Lc > 1.f ? Lc : 1.f;
    const float mx = Lr1 > Lc1 ? Lr1 : Lc1, mn = Lr < Lc ? Lr : Lc;
    const bool two = p.count > 1;
    const float ro1 = p.seed[0], co1 = p.seed[1], ro2 = p.seed[2],
                co2 = p.seed[3];
    const float d1 = __fsub_rn(co1, ro1);
    f[0] = __fdiv_rn(lg2(Lr1), 16.f);
    f[1] = __fdiv_rn(lg2(Lc1), 16.f);
    f[2] = __fdiv_rn(mn, mx);
    f[3] = __fdiv_rn(lg2(__fadd_rn(1.f, (float)p.count)), 8.f);
    f[4] = __fdiv_rn((float)p.kscore, 64.f);
    f[5] = __fdiv_rn(d1, mx);
    f[6] = two ? __fdiv_rn(fabsf(__fsub_rn(__fsub_rn(co2, ro2), d1)), mx) : 0.f;
    f[7] = __fdiv_rn(ro1, Lr1);
    f[8] = __fdiv_rn(co1, Lc1);
    f[9] = two ? __fdiv_rn(ro2, Lr1) : 0.f;
    f[10] = two ? __fdiv_rn(co2, Lc1) : 0.f;
    const float den = __fsub_rn(mn, 5.f);
    f[11] = __fdiv_rn((float)p.count, den > 1.f ? den : 1.f);
    int agree = 0;
    for (int h = 0; h < kSketch; ++h)
        agree += (R.mh[h] == C.mh[h] && R.mh[h] != 0xffffffffu);
    f[12] = __fdiv_rn((float)agree, (float)kSketch);
    float smin = 0.f, sdot = 0.f;
    for (int k = 0; k < 10; ++k) {
        smin = __fadd_rn(smin, R.comp[k] < C.comp[k] ? R.comp[k] : C.comp[k]);
        sdot = __fadd_rn(sdot, __fmul_rn(R.comp[k], C.comp[k]));
    }
    f[13] = smin;
    f[14] = sdot;
    // seeds lie inside their sequences, so both differences are >= 0
    const float a = ro1 < co1 ? ro1 : co1;
    const float b1 = __fsub_rn(Lr, ro1), b2 = __fsub_rn(Lc, co1);
    f[15] = __fdiv_rn(__fadd_rn(a, b1 < b2 ? b1 : b2), mx);
}

// Pairs naming a sequence outside the sets get NaN features and logits and
// are counted in *bad; nothing outside the sets is read.
__global__ void k_features(const pa_pair *__restrict__ pairs, uint64_t n,
                           const SeqSketch *__restrict__ rs, uint32_t nr,
                           const SeqSketch *__restrict__ cs, uint32_t nc,
                           float *__restrict__ out, uint32_t *bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const pa_pair p = pairs[i];
    float f[kNF];
    if (p.r < nr && p.c < nc) {
        pair_features(p, rs[p.r], cs[p.c], f);
    } else {
        for (int k = 0; k < kNF; ++k)
            f[k] = __builtin_nanf("");
        atomicAdd(bad, 1u);
    }
    for (int k = 0; k < kNF; ++k)
        out[i * kNF + k] = f[k];
}

// ---- MLP ------------------------------------------------------------------
// LDS model: bf16-rounded weights as fp32, fp32 biases, mean, 1/std
struct LdsModel {
    float w[kNParam];
    float mean[kNF], istd[kNF];
};

__device__ inline void load_model(LdsModel &m, const float *wb,
                                  const float *mean, const float *istd)
{
    for (int i = threadIdx.x; i < kNParam; i += blockDim.x)
        m.w[i] = wb[i];
    for (int i = threadIdx.x; i < kNF; i += blockDim.x) {
        m.mean[i] = mean[i];
        m.istd[i] = istd[i];
    }
    __syncthreads();
}

// logit of standardized, bf16-rounded features xb
__device__ inline float mlp_logit(const float *w, const float *xb)
{
    float h1[kH];
    for (int j = 0; j < kH; ++j) {
        float acc = 0.f;
        for (int k = 0; k < kNF; ++k)
            acc = fmaf(xb[k], w[kOffW1 + k * kH + j], acc);  // exact product
        acc = __fadd_rn(acc, w[kOffB1 + j]);
        h1[j] = bf16r(acc > 0.f ? acc : 0.f);
    }
    float z3 = 0.f;
    for (int j = 0; j < kH; ++j) {
        float acc = 0.f;
        for (int k = 0; k < kH; ++k)
            acc = fmaf(h1[k], w[kOffW2 + k * kH + j], acc);
        acc = __fadd_rn(acc, w[kOffB2 + j]);
        const float h2 = bf16r(acc > 0.f ? acc : 0.f);
        z3 = fmaf(h2, w[kOffW3 + j], z3);
    }
    return __fadd_rn(z3, w[kOffB3]);
}

__device__ inline void standardize(const float *f, const LdsModel &m,
                                   float *xb)
{
    for (int k = 0; k < kNF; ++k)
        xb[k] = bf16r(__fmul_rn(__fsub_rn(f[k], m.mean[k]), m.istd[k]));
}

__global__ __launch_bounds__(256) void k_score(
    const pa_pair *__restrict__ pairs, uint64_t n,
    const SeqSketch *__restrict__ rs, uint32_t nr,
    const SeqSketch *__restrict__ cs, uint32_t nc, const float *wb,
    const float *mean, const float *istd, float *__restrict__ logit_out,
    uint32_t *bad)
{
    __shared__ LdsModel m;
    load_model(m, wb, mean, istd);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const pa_pair p = pairs[i];
    if (p.r >= nr || p.c >= nc) {
        logit_out[i] = __builtin_nanf("");
        atomicAdd(bad, 1u);
        return;
    }
    float f[kNF], xb[kNF];
    pair_features(p, rs[p.r], cs[p.c], f);
    standardize(f, m, xb);
    logit_out[i] = mlp_logit(m.w, xb);
}

// logits of precomputed features (tests and training-set evaluation)
__global__ __launch_bounds__(256) void k_logits_of(
    const float *__restrict__ feats, uint64_t n, const float *wb,
    const float *mean, const float *istd, float *__restrict__ logit_out)
{
    __shared__ LdsModel m;
    load_model(m, wb, mean, istd);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float xb[kNF];
    standardize(feats + i * kNF, m, xb);
    logit_out[i] = mlp_logit(m.w, xb);
}

// keep pairs whose logit >= tau_logit; order of the kept list is arbitrary.
// The block counts its kept pairs in LDS and reserves one range with a
// single global atomic.
__global__ __launch_bounds__(256) void k_select(
    const pa_pair *__restrict__ pairs, const float *__restrict__ logits,
    uint64_t n, float tau, pa_pair *__restrict__ kept,
    uint32_t *__restrict__ kept_index, uint32_t *count)
{
    __shared__ uint32_t nk, base;
    if (threadIdx.x == 0)
        nk = 0;
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool keep = i < n && logits[i] >= tau;
    const uint32_t rank = keep ? atomicAdd(&nk, 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0)
        base = nk ? atomicAdd(count, nk) : 0u;
    __syncthreads();
    if (keep) {
        const uint32_t pos = base + rank;
        kept[pos] = pairs[i];
        if (kept_index)
            kept_index[pos] = (uint32_t)i;
    }
}

// ---- training -------------------------------------------------------------
// Per sample of the batch: standardize, forward, weighted BCE, backward to
// dz1, dz2, dz3.  Scratch rows (fp32): xb[16] h1b[64] h2b[64] dz1[64] dz2[64]
// dz3[1].
__global__ __launch_bounds__(256) void k_train_fwd_bwd(
    const float *__restrict__ feats, const float *__restrict__ mean,
    const float *__restrict__ istd, const uint8_t *__restrict__ labels,
    const uint32_t *__restrict__ order, uint64_t base, uint32_t B,
    const float *wb, float pos_weight, float *__restrict__ scratch,
    float *__restrict__ loss_sum)
{
    __shared__ float w[kNParam];
    for (int i = threadIdx.x; i < kNParam; i += blockDim.x)
        w[i] = wb[i];
    __syncthreads();
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B)
        return;
    const uint32_t idx = order[base + s];
    float *row = scratch + (size_t)s * kScratchRow;
    float *h1b = row + kNF, *h2b = h1b + kH, *dz1 = h2b + kH, *dz2 = dz1 + kH;
    float xb[kNF], z1[kH], z2[kH], h1[kH];
    for (int k = 0; k < kNF; ++k) {
        xb[k] = bf16r(__fmul_rn(__fsub_rn(feats[(size_t)idx * kNF + k], mean[k]),
                                istd[k]));
        row[k] = xb[k];
    }
    for (int j = 0; j < kH; ++j) {
        float acc = 0.f;
        for (int k = 0; k < kNF; ++k)
            acc = fmaf(xb[k], w[kOffW1 + k * kH + j], acc);  // exact product
        z1[j] = __fadd_rn(acc, w[kOffB1 + j]);
        h1[j] = bf16r(z1[j] > 0.f ? z1[j] : 0.f);
        h1b[j] = h1[j];
    }
    float z3 = 0.f;
    for (int j = 0; j < kH; ++j) {
        float acc = 0.f;
        for (int k = 0; k < kH; ++k)
            acc = fmaf(h1[k], w[kOffW2 + k * kH + j], acc);
        z2[j] = __fadd_rn(acc, w[kOffB2 + j]);
        const float h2 = bf16r(z2[j] > 0.f ? z2[j] : 0.f);
        h2b[j] = h2;
        z3 = fmaf(h2, w[kOffW3 + j], z3);
    }
    z3 = __fadd_rn(z3, w[kOffB3]);
    const float p = (float)(1.0 / (1.0 + exp(-(double)z3)));
    const float y = labels[idx] ? 1.f : 0.f;
    const float eps = 1e-7f;
    const float lp = logf(p > eps ? p : eps), lq = logf(1.f - p > eps ? 1.f - p : eps);
    atomicAdd(loss_sum, -(pos_weight * y * lp + (1.f - y) * lq));
    // dL/dz3 of the mean loss
    const float dz3 = __fdiv_rn(
        __fsub_rn(__fmul_rn(__fsub_rn(1.f, y), p),
                  __fmul_rn(__fmul_rn(pos_weight, y), __fsub_rn(1.f, p))),
        (float)B);
    row[kScratchRow - 1] = dz3;
    for (int j = 0; j < kH; ++j) {
        const float dh2 = __fmul_rn(dz3, w[kOffW3 + j]);
        dz2[j] = z2[j] > 0.f ? dh2 : 0.f;
    }
    for (int i = 0; i < kH; ++i) {
        float acc = 0.f;
        for (int k = 0; k < kH; ++k)
            acc = __fadd_rn(acc, __fmul_rn(dz2[k], w[kOffW2 + i * kH + k]));
        dz1[i] = z1[i] > 0.f ? acc : 0.f;
    }
}

// One thread per parameter: sums over the batch in sample order.
__global__ void k_train_wgrad(const float *__restrict__ scratch, uint32_t B,
                              float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam)
        return;
    // row offsets: xb 0, h1b kNF, h2b kNF+kH, dz1 kNF+2kH, dz2 kNF+3kH,
    // dz3 kScratchRow-1; ao = -1 marks a bias (single factor)
    constexpr int oH1 = kNF, oH2 = kNF + kH, oD1 = kNF + 2 * kH,
                  oD2 = kNF + 3 * kH, oD3 = kScratchRow - 1;
    int ao, bo;
    if (p < kOffB1) {          // W1[i][j] += xb[i] * dz1[j]
        ao = p / kH; bo = oD1 + p % kH;
    } else if (p < kOffW2) {   // b1[j] += dz1[j]
        ao = -1; bo = oD1 + (p - kOffB1);
    } else if (p < kOffB2) {   // W2[i][j] += h1b[i] * dz2[j]
        const int q = p - kOffW2;
        ao = oH1 + q / kH; bo = oD2 + q % kH;
    } else if (p < kOffW3) {   // b2[j] += dz2[j]
        ao = -1; bo = oD2 + (p - kOffB2);
    } else if (p < kOffB3) {   // w3[i] += h2b[i] * dz3
        ao = oH2 + (p - kOffW3); bo = oD3;
    } else {                   // b3 += dz3
        ao = -1; bo = oD3;
    }
    float acc = 0.f;
    for (uint32_t s = 0; s < B; ++s) {
        const float *row = scratch + (size_t)s * kScratchRow;
        const float v = ao >= 0 ? __fmul_rn(row[ao], row[bo]) : row[bo];
        acc = __fadd_rn(acc, v);
    }
    grad[p] = acc;
}

__global__ void k_adam(float *__restrict__ w, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ g,
                       float *__restrict__ wb, float lr, float beta1,
                       float beta2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam)
        return;
    const float gp = g[p];
    const float mp = __fadd_rn(__fmul_rn(beta1, m[p]), __fmul_rn(1.f - beta1, gp));
    const float vp = __fadd_rn(__fmul_rn(beta2, v[p]),
                               __fmul_rn(__fmul_rn(1.f - beta2, gp), gp));
    m[p] = mp;
    v[p] = vp;
    const float mh = __fdiv_rn(mp, c1), vh = __fdiv_rn(vp, c2);
    const float step = __fdiv_rn(__fmul_rn(lr, mh), __fadd_rn(__fsqrt_rn(vh), eps));
    const float nw = __fsub_rn(w[p], step);
    w[p] = nw;
    wb[p] = p >= kOffB1 && p < kOffW2 ? nw
          : p >= kOffB2 && p < kOffW3 ? nw
          : p == kOffB3 ? nw : bf16r(nw);
}

// bf16 copy of the master weights (biases stay fp32)
__global__ void k_round_weights(const float *__restrict__ w,
                                float *__restrict__ wb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam)
        return;
    const bool bias = (p >= kOffB1 && p < kOffW2) ||
                      (p >= kOffB2 && p < kOffW3) || p == kOffB3;
    wb[p] = bias ? w[p] : bf16r(w[p]);
}

inline unsigned nblk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t seq_sketch(const uint64_t *offs, uint64_t nseq, SeqSketch *out,
                      hipStream_t s)
{
    if (!nseq)
        return hipSuccess;
    hipLaunchKernelGGL(k_seq_sketch, dim3(nblk(nseq, 128)), dim3(128), 0, s,
                       offs, nseq, out);
    return hipGetLastError();
}

hipError_t features(const pa_pair *pairs, uint64_t n, const SeqSketch *rs,
                    uint32_t nr, const SeqSketch *cs, uint32_t nc, float *out,
                    uint32_t *bad, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3(nblk(n, 256)), dim3(256), 0, s, pairs,
                       n, rs, nr, cs, nc, out, bad);
    return hipGetLastError();
}

hipError_t score_pairs(const pa_pair *pairs, uint64_t n, const SeqSketch *rs,
                       uint32_t nr, const SeqSketch *cs, uint32_t nc,
                       const float *wb, const float *mean, const float *istd,
                       float *logits, uint32_t *bad, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_score, dim3(nblk(n, 256)), dim3(256), 0, s, pairs, n,
                       rs, nr, cs, nc, wb, mean, istd, logits, bad);
    return hipGetLastError();
}

hipError_t logits_of(const float *feats, uint64_t n, const float *wb,
                     const float *mean, const float *istd, float *logits,
                     hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_logits_of, dim3(nblk(n, 256)), dim3(256), 0, s, feats,
                       n, wb, mean, istd, logits);
    return hipGetLastError();
}

hipError_t select_pairs(const pa_pair *pairs, const float *logits, uint64_t n,
                        float tau, pa_pair *kept, uint32_t *kept_index,
                        uint32_t *count, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_select, dim3(nblk(n, 256)), dim3(256), 0, s, pairs,
                       logits, n, tau, kept, kept_index, count);
    return hipGetLastError();
}

hipError_t train_step(const float *feats, const float *mean,
                      const float *istd, const uint8_t *labels,
                      const uint32_t *order, uint64_t base, uint32_t B,
                      float *w, float *wb, float *m, float *v, float *grad,
                      float *scratch, float *loss_sum, float pos_weight,
                      float lr, float beta1, float beta2, float eps, int t,
                      bool update, hipStream_t s)
{
    hipLaunchKernelGGL(k_train_fwd_bwd, dim3(nblk(B, 256)), dim3(256), 0, s,
                       feats, mean, istd, labels, order, base, B, wb, pos_weight, scratch,
                       loss_sum);
    hipLaunchKernelGGL(k_train_wgrad, dim3(nblk(kNParam, 64)), dim3(64), 0, s,
                       scratch, B, grad);
    if (update) {
        const float c1 = (float)(1.0 - pow((double)beta1, (double)t));
        const float c2 = (float)(1.0 - pow((double)beta2, (double)t));
        hipLaunchKernelGGL(k_adam, dim3(nblk(kNParam, 256)), dim3(256), 0, s,
                           w, m, v, grad, wb, lr, beta1, beta2, eps, c1, c2);
    }
    return hipGetLastError();
}

hipError_t round_weights(const float *w, float *wb, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_weights, dim3(nblk(kNParam, 256)), dim3(256), 0,
                       s, w, wb);
    return hipGetLastError();
}

}  // namespace pa
