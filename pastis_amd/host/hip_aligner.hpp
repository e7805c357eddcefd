// hip_aligner.hpp -- the aligner operator on MI355X.
//
// PWAlign restates the reference's abstract operator (inc/align/PWAlign.hpp:
// 20-82) with the same virtual signatures; HipAligner implements it over the C
// ABI of include/pastis_hip.h:
//   HipAligner(gap_open, gap_ext)                      -- SeqanFullAligner
//       (inc/align/SeqanFullAligner.hpp:35-37, --sfa)
//   HipAligner(gap_open, gap_ext, seed_len, xdrop, seed_cnt)
//                                                      -- SeqanXdropAligner
//       (inc/align/SeqanXdropAligner.hpp:38-42, --sxa)
// The constructors take the reference's CLI values and apply them the way the
// reference's Score(gap_open, gap_ext) call does (see pa_config_for_cli).
//
// The sequence sets come through SequenceSource instead of DistFastaData (an
// MPI/CombBLAS object); a source over DistFastaData is a short shim (it walks
// my_nbrs exactly like construct_seqs_bl does).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pastis_hip.h"
#include "pw_types.hpp"

namespace pastis_amd {

// Row and column sequence sets of one block, as ASCII residues.
class SequenceSource {
public:
    virtual ~SequenceSource() = default;
    virtual uint64_t nrseqs() const = 0;
    virtual uint64_t ncseqs() const = 0;
    // pointer to the residues of row / column sequence i and its length
    virtual const char *rseq(uint64_t i, uint32_t &len) const = 0;
    virtual const char *cseq(uint64_t i, uint32_t &len) const = 0;
    // true on a diagonal block: the column set is the row set
    // (SeqanXdropAligner.cpp:82-87)
    virtual bool diagonal() const { return false; }
};

// In-memory source over concatenated residues and offsets.
class VectorSource : public SequenceSource {
public:
    VectorSource(std::vector<char> res, std::vector<uint64_t> offs)
        : rres_(std::move(res)), roffs_(std::move(offs)), diag_(true) {}
    VectorSource(std::vector<char> rres, std::vector<uint64_t> roffs,
                 std::vector<char> cres, std::vector<uint64_t> coffs)
        : rres_(std::move(rres)), roffs_(std::move(roffs)),
          cres_(std::move(cres)), coffs_(std::move(coffs)), diag_(false) {}
    uint64_t nrseqs() const override { return roffs_.size() - 1; }
    uint64_t ncseqs() const override
    {
        return diag_ ? nrseqs() : coffs_.size() - 1;
    }
    const char *rseq(uint64_t i, uint32_t &len) const override
    {
        len = (uint32_t)(roffs_[i + 1] - roffs_[i]);
        return rres_.data() + roffs_[i];
    }
    const char *cseq(uint64_t i, uint32_t &len) const override
    {
        if (diag_)
            return rseq(i, len);
        len = (uint32_t)(coffs_[i + 1] - coffs_[i]);
        return cres_.data() + coffs_[i];
    }
    bool diagonal() const override { return diag_; }

private:
    std::vector<char> rres_;
    std::vector<uint64_t> roffs_;
    std::vector<char> cres_;
    std::vector<uint64_t> coffs_;
    bool diag_;
};

// inc/align/PWAlign.hpp:20-82 with SequenceSource for DistFastaData.
class PWAlign {
public:
    virtual ~PWAlign() {}
    virtual void construct_seqs(std::shared_ptr<SequenceSource> dfd) = 0;
    virtual void construct_seqs_bl(std::shared_ptr<SequenceSource> dfd) = 0;
    virtual void aln_batch(std::tuple<uint64_t, uint64_t, CommonKmerLight *> *,
                           uint64_t, uint64_t, uint64_t, uint64_t,
                           const params_t &)
    {
    }
    virtual void aln_batch(std::tuple<uint64_t, uint64_t, CommonKmerLoc *> *,
                           uint64_t, uint64_t, uint64_t, uint64_t,
                           const params_t &)
    {
    }
    virtual size_t rseq_len(uint64_t) { return 0; }
    virtual size_t cseq_len(uint64_t) { return 0; }
};

// Counters of the calls made on one HipAligner (the reference keeps them in
// parops->tp->mat_stats: aln_pairs, aln_pair_lens).
struct HipAlignerStats {
    uint64_t aln_pairs = 0;       // pairs handed to aln_batch
    uint64_t filtered_out = 0;    // dropped by the learned filter
    uint64_t aligned = 0;         // pairs aligned exactly
    uint64_t passed = 0;          // pairs whose score_aln was written
    double aln_pair_lens = 0;     // summed lengths of aligned pairs
};

class HipAligner : public PWAlign {
public:
    // --sfa: full Smith-Waterman
    HipAligner(int gap_open, int gap_ext, int device = 0,
               int gap_model = PA_GAP_AS_CALLED);
    // --sxa: seed-and-extend X-drop
    HipAligner(int gap_open, int gap_ext, uint16_t seed_len, int xdrop,
               int seed_cnt, int device = 0, int gap_model = PA_GAP_AS_CALLED);
    ~HipAligner() override;
    HipAligner(const HipAligner &) = delete;
    HipAligner &operator=(const HipAligner &) = delete;

    void construct_seqs(std::shared_ptr<SequenceSource> dfd) override;
    void construct_seqs_bl(std::shared_ptr<SequenceSource> dfd) override;

    // Full SW reads Light tuples (SeqanFullAligner.cpp:173-266); X-drop needs
    // the seeds of Loc tuples (SeqanXdropAligner.cpp:190-361).  Like the
    // reference, an X-drop aligner ignores Light batches; a full aligner also
    // accepts Loc batches (the seeds are not used).
    void aln_batch(std::tuple<uint64_t, uint64_t, CommonKmerLight *> *mattuples,
                   uint64_t beg, uint64_t end, uint64_t bl_roffset,
                   uint64_t bl_coffset, const params_t &params) override;
    void aln_batch(std::tuple<uint64_t, uint64_t, CommonKmerLoc *> *mattuples,
                   uint64_t beg, uint64_t end, uint64_t bl_roffset,
                   uint64_t bl_coffset, const params_t &params) override;
    size_t rseq_len(uint64_t i) override;
    size_t cseq_len(uint64_t i) override;

    // Learned prefilter: pairs whose logit is below tau are not aligned (their
    // score_aln stays 0, so the caller's post-prune drops them).  params is
    // PA_FILTER_NPARAM floats, mean/istd PA_FILTER_NF floats each.
    void set_filter(const float *params, const float *mean, const float *istd,
                    float tau);
    void clear_filter() { filter_on_ = false; }

    // Launch-shape autotuning (pa_autotune) on the first aln_batch call, on
    // an evenly spaced sample of that call's pairs; on by default, off when
    // the environment sets PA_AUTOTUNE=0.  The library caches the choice per
    // device, so later aligners on the device install it without timing.
    void set_autotune(bool on) { autotune_ = on; }

    const HipAlignerStats &stats() const { return stats_; }
    pa_ctx *context() { return ctx_; }

private:
    template <class NT>
    void run(std::tuple<uint64_t, uint64_t, NT *> *mattuples, uint64_t beg,
             uint64_t end, uint64_t bl_roffset, uint64_t bl_coffset,
             const params_t &params, bool seeds);
    void load(const SequenceSource &src);
    void check(int rc, const char *what);
    void device_reserve(uint64_t n);
    void host_reserve(uint64_t n);
    void tune(const pa_pair *pairs, uint64_t n);

    pa_ctx *ctx_ = nullptr;
    pa_config cfg_{};
    std::vector<uint32_t> rlen_, clen_;
    // Two page-locked staging slots: while the GPU aligns the pairs of one
    // chunk, a host thread writes the previous chunk's results back into
    // the tuples and packs the next chunk into the other slot.
    struct Slot {
        pa_pair *pairs = nullptr;
        pa_result *results = nullptr;
        uint32_t *index = nullptr;   // kept pairs (learned filter)
        uint64_t n = 0, kept = 0;
    } slot_[2];
    uint64_t h_cap_ = 0;
    void *d_pairs_ = nullptr, *d_kept_ = nullptr, *d_index_ = nullptr,
         *d_results_ = nullptr;
    uint64_t d_cap_ = 0;
    bool filter_on_ = false;
    bool autotune_ = true, tuned_ = false;
    float tau_ = 0.f;
    HipAlignerStats stats_;
};

}  // namespace pastis_amd
