// k_filter.h -- learned candidate prefilter: per-sequence sketches, pair
// features, MLP 16-64-64-1 inference, selection and training kernels.
#pragma once

#include "pa_internal.h"

namespace pa {

constexpr int kNF = 16;           // features per pair
constexpr int kH = 64;            // hidden width
constexpr int kSketch = 16;       // MinHash values per sequence
constexpr int kK = 6;             // murphy10 k-mer length of the sketch
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kNF * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kNParam = kOffB3 + 1;
static_assert(kNParam == 5313, "parameter layout");
// training scratch row: xb[16] h1b[64] h2b[64] dz1[64] dz2[64] dz3
constexpr int kScratchRow = kNF + 4 * kH + 1;

// Sketch of one sequence: murphy10 composition fractions, MinHash of its
// murphy10 6-mers (0xffffffff when the sequence is shorter than 6).
struct SeqSketch {
    float comp[10];
    uint32_t mh[kSketch];
    uint32_t len;
    uint32_t pad;
};
static_assert(sizeof(SeqSketch) == 112, "SeqSketch layout");

hipError_t seq_sketch(const uint64_t *offs, uint64_t nseq, SeqSketch *out,
                      hipStream_t s);
hipError_t features(const pa_pair *pairs, uint64_t n, const SeqSketch *rs,
                    uint32_t nr, const SeqSketch *cs, uint32_t nc, float *out,
                    uint32_t *bad, hipStream_t s);
hipError_t score_pairs(const pa_pair *pairs, uint64_t n, const SeqSketch *rs,
                       uint32_t nr, const SeqSketch *cs, uint32_t nc,
                       const float *wb, const float *mean, const float *istd,
                       float *logits, uint32_t *bad, hipStream_t s);
hipError_t logits_of(const float *feats, uint64_t n, const float *wb,
                     const float *mean, const float *istd, float *logits,
                     hipStream_t s);
hipError_t select_pairs(const pa_pair *pairs, const float *logits, uint64_t n,
                        float tau, pa_pair *kept, uint32_t *kept_index,
                        uint32_t *count, hipStream_t s);
hipError_t train_step(const float *feats, const float *mean,
                      const float *istd, const uint8_t *labels,
                      const uint32_t *order, uint64_t base, uint32_t B,
                      float *w, float *wb, float *m, float *v, float *grad,
                      float *scratch, float *loss_sum, float pos_weight,
                      float lr, float beta1, float beta2, float eps, int t,
                      bool update, hipStream_t s);
hipError_t round_weights(const float *w, float *wb, hipStream_t s);

}  // namespace pa
