// hip_aligner.cpp -- HipAligner over the C ABI (include/pastis_hip.h).
#include "hip_aligner.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <future>
#include <stdexcept>

namespace pastis_amd {

namespace {

// Pairs per call into the library: bounds the staging buffers whatever the
// caller's batch size (the reference's default batch is 1e10 tuples).  The
// environment variable PA_HOST_CHUNK overrides it (tests use small chunks to
// run the double-buffered loop over many chunks).
constexpr uint64_t kChunk = 1ull << 22;

uint64_t chunk_pairs()
{
    const char *e = std::getenv("PA_HOST_CHUNK");
    if (e && *e) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0)
            return (uint64_t)v;
    }
    return kChunk;
}
// Pairs the autotuner times on the first call.
constexpr uint64_t kTuneSample = 262144;

}  // namespace

void HipAligner::check(int rc, const char *what)
{
    if (rc == PA_OK)
        return;
    std::string msg = std::string("HipAligner: ") + what + " failed: " +
                      pa_strerror(rc);
    const char *detail = ctx_ ? pa_last_error(ctx_) : "";
    if (detail && *detail)
        msg += std::string(": ") + detail;
    std::fprintf(stderr, "%s\n", msg.c_str());
    throw std::runtime_error(msg);
}

HipAligner::HipAligner(int gap_open, int gap_ext, int device, int gap_model)
{
    check(pa_config_for_cli(&cfg_, gap_open, gap_ext, 6, 0, 2, gap_model),
          "pa_config_for_cli");
    cfg_.mode = PA_MODE_SW;
    cfg_.device = device;
    check(pa_create(&cfg_, &ctx_), "pa_create");
}

HipAligner::HipAligner(int gap_open, int gap_ext, uint16_t seed_len, int xdrop,
                       int seed_cnt, int device, int gap_model)
{
    check(pa_config_for_cli(&cfg_, gap_open, gap_ext, seed_len, xdrop,
                            seed_cnt, gap_model),
          "pa_config_for_cli");
    cfg_.mode = PA_MODE_XDROP;
    cfg_.device = device;
    check(pa_create(&cfg_, &ctx_), "pa_create");
}

HipAligner::~HipAligner()
{
    if (ctx_) {
        for (void *p : {d_pairs_, d_kept_, d_index_, d_results_})
            if (p)
                pa_device_free(ctx_, p);
        for (Slot &sl : slot_)
            for (void *p : {(void *)sl.pairs, (void *)sl.results,
                            (void *)sl.index})
                if (p)
                    pa_host_free(ctx_, p);
        pa_destroy(ctx_);
    }
}

void HipAligner::load(const SequenceSource &src)
{
    auto gather = [](const SequenceSource &s, bool rows,
                     std::vector<char> &res, std::vector<uint64_t> &offs,
                     std::vector<uint32_t> &lens) {
        const uint64_t n = rows ? s.nrseqs() : s.ncseqs();
        offs.assign(n + 1, 0);
        lens.assign(n, 0);
        uint64_t tot = 0;
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t L = 0;
            rows ? s.rseq(i, L) : s.cseq(i, L);
            lens[i] = L;
            tot += L;
            offs[i + 1] = tot;
        }
        res.resize(tot);
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t L = 0;
            const char *p = rows ? s.rseq(i, L) : s.cseq(i, L);
            std::copy(p, p + L, res.begin() + (ptrdiff_t)offs[i]);
        }
    };
    std::vector<char> res;
    std::vector<uint64_t> offs;
    gather(src, true, res, offs, rlen_);
    if (src.diagonal()) {
        check(pa_load_seqs(ctx_, PA_SIDE_BOTH, (const uint8_t *)res.data(),
                           offs.data(), rlen_.size(), PA_ENC_ASCII),
              "pa_load_seqs");
        clen_ = rlen_;
        return;
    }
    check(pa_load_seqs(ctx_, PA_SIDE_ROW, (const uint8_t *)res.data(),
                       offs.data(), rlen_.size(), PA_ENC_ASCII),
          "pa_load_seqs");
    gather(src, false, res, offs, clen_);
    check(pa_load_seqs(ctx_, PA_SIDE_COL, (const uint8_t *)res.data(),
                       offs.data(), clen_.size(), PA_ENC_ASCII),
          "pa_load_seqs");
}

// SeqanFullAligner.cpp:28-88 / SeqanXdropAligner.cpp:32-105: all row and
// column sequences of the rank's neighbourhood.
void HipAligner::construct_seqs(std::shared_ptr<SequenceSource> dfd)
{
    load(*dfd);
}

// SeqanFullAligner.cpp:92-168: the sequences of the current block.
void HipAligner::construct_seqs_bl(std::shared_ptr<SequenceSource> dfd)
{
    load(*dfd);
}

size_t HipAligner::rseq_len(uint64_t i)
{
    return i < rlen_.size() ? rlen_[i] : 0;
}

size_t HipAligner::cseq_len(uint64_t i)
{
    return i < clen_.size() ? clen_[i] : 0;
}

void HipAligner::set_filter(const float *params, const float *mean,
                            const float *istd, float tau)
{
    check(pa_filter_set_model(ctx_, params, mean, istd), "pa_filter_set_model");
    tau_ = tau;
    filter_on_ = true;
}

void HipAligner::device_reserve(uint64_t n)
{
    if (n <= d_cap_)
        return;
    for (void **p : {&d_pairs_, &d_kept_, &d_index_, &d_results_})
        if (*p) {
            pa_device_free(ctx_, *p);
            *p = nullptr;
        }
    check(pa_device_alloc(ctx_, n * sizeof(pa_pair), &d_pairs_), "alloc");
    check(pa_device_alloc(ctx_, n * sizeof(pa_pair), &d_kept_), "alloc");
    check(pa_device_alloc(ctx_, n * sizeof(uint32_t), &d_index_), "alloc");
    check(pa_device_alloc(ctx_, n * sizeof(pa_result), &d_results_), "alloc");
    d_cap_ = n;
}

void HipAligner::host_reserve(uint64_t n)
{
    if (n <= h_cap_)
        return;
    for (Slot &sl : slot_) {
        for (void *p : {(void *)sl.pairs, (void *)sl.results, (void *)sl.index})
            if (p)
                pa_host_free(ctx_, p);
        sl = Slot{};
    }
    h_cap_ = 0;
    for (Slot &sl : slot_) {
        void *p = nullptr;
        check(pa_host_alloc(ctx_, n * sizeof(pa_pair), &p), "pa_host_alloc");
        sl.pairs = (pa_pair *)p;
        check(pa_host_alloc(ctx_, n * sizeof(pa_result), &p), "pa_host_alloc");
        sl.results = (pa_result *)p;
        check(pa_host_alloc(ctx_, n * sizeof(uint32_t), &p), "pa_host_alloc");
        sl.index = (uint32_t *)p;
    }
    h_cap_ = n;
}

void HipAligner::tune(const pa_pair *pairs, uint64_t n)
{
    tuned_ = true;
    const char *env = std::getenv("PA_AUTOTUNE");
    if ((env && std::string(env) == "0") || n == 0)
        return;
    const uint64_t m = std::min(n, kTuneSample);
    std::vector<pa_pair> sample(m);
    for (uint64_t i = 0; i < m; ++i)
        sample[i] = pairs[i * n / m];   // evenly spaced over the chunk
    void *d = nullptr;
    check(pa_device_alloc(ctx_, m * sizeof(pa_pair), &d), "pa_device_alloc");
    int rc = pa_memcpy_h2d(ctx_, d, sample.data(), m * sizeof(pa_pair));
    if (rc == PA_OK)
        rc = pa_autotune(ctx_, (const pa_pair *)d, m, PA_TUNE_ALIGN, m, 0,
                         nullptr);
    pa_device_free(ctx_, d);
    check(rc, "pa_autotune");
}

template <class NT> inline void seeds_of(const NT *, pa_pair &) {}
template <> inline void seeds_of(const CommonKmerLoc *v, pa_pair &p)
{
    p.seed[0] = v->first.first;    // row sequence offset
    p.seed[1] = v->first.second;   // column sequence offset
    p.seed[2] = v->second.first;
    p.seed[3] = v->second.second;
}

// The pre-stage (SeqanFullAligner.cpp:184-214, SeqanXdropAligner.cpp:201-232)
// becomes a pa_pair array; the alignment, statistics and predicate run on the
// GPU; the post-stage writes score_aln = identity / 100 for passing pairs only
// (SeqanFullAligner.cpp:253-260, SeqanXdropAligner.cpp:344-351).
//
// Chunks of chunk_pairs() tuples go through two page-locked slots: while the GPU
// aligns chunk i from one slot, a host thread writes chunk i-1's results
// back into the tuples and then packs chunk i+1 into the other slot (the
// slot chunk i-1 used).
template <class NT>
void HipAligner::run(std::tuple<uint64_t, uint64_t, NT *> *mattuples,
                     uint64_t beg, uint64_t end, uint64_t bl_roffset,
                     uint64_t bl_coffset, const params_t &params, bool seeds)
{
    if (end <= beg)
        return;
    check(pa_set_thresholds(ctx_, params.aln_cov_thr, params.aln_ani_thr),
          "pa_set_thresholds");
    stats_.aln_pairs += end - beg;
    const uint64_t total = end - beg;
    const uint64_t chunk = chunk_pairs();
    const uint64_t nch = (total + chunk - 1) / chunk;
    host_reserve(std::min(chunk, total));
    if (filter_on_)
        device_reserve(std::min(chunk, total));

    // chunk i into its slot; false when a tuple lies outside the block
    auto pack = [&](uint64_t i) -> bool {
        Slot &sl = slot_[i % 2];
        const uint64_t cb = beg + i * chunk;
        sl.n = std::min(chunk, end - cb);
        sl.kept = 0;
        for (uint64_t j = 0; j < sl.n; ++j) {
            const auto &t = mattuples[cb + j];
            const NT *v = std::get<2>(t);
            pa_pair &p = sl.pairs[j];
            p = pa_pair{};
            const uint64_t r = std::get<0>(t) + bl_roffset;
            const uint64_t c = std::get<1>(t) + bl_coffset;
            if (r >= rlen_.size() || c >= clen_.size())
                return false;
            p.r = (uint32_t)r;
            p.c = (uint32_t)c;
            p.count = v->count;
            p.kscore = v->score;
            if (seeds)
                seeds_of(v, p);
        }
        return true;
    };
    // chunk i's results back into the tuples
    auto post = [&](uint64_t i) {
        const Slot &sl = slot_[i % 2];
        const uint64_t cb = beg + i * chunk;
        stats_.aligned += sl.kept;
        stats_.filtered_out += sl.n - sl.kept;
        for (uint64_t j = 0; j < sl.kept; ++j) {
            const uint64_t k = filter_on_ ? sl.index[j] : j;
            const pa_result &res = sl.results[j];
            stats_.aln_pair_lens += (double)rlen_[sl.pairs[k].r] +
                                    (double)clen_[sl.pairs[k].c];
            if (res.flags & PA_FLAG_PASS) {
                std::get<2>(mattuples[cb + k])->score_aln = res.score_aln;
                stats_.passed++;
            }
        }
    };
    auto align = [&](Slot &sl) {
        if (!filter_on_) {
            check(pa_align(ctx_, sl.pairs, sl.n, sl.results), "pa_align");
            sl.kept = sl.n;
            return;
        }
        uint64_t k = 0;
        check(pa_memcpy_h2d(ctx_, d_pairs_, sl.pairs, sl.n * sizeof(pa_pair)),
              "pa_memcpy_h2d");
        check(pa_filter_select_dev(ctx_, (const pa_pair *)d_pairs_, sl.n, tau_,
                                   (pa_pair *)d_kept_, (uint32_t *)d_index_,
                                   &k, nullptr),
              "pa_filter_select_dev");
        if (k) {
            check(pa_align_dev(ctx_, (const pa_pair *)d_kept_, k,
                               (pa_result *)d_results_, nullptr),
                  "pa_align_dev");
            check(pa_memcpy_d2h(ctx_, sl.results, d_results_,
                                k * sizeof(pa_result)),
                  "pa_memcpy_d2h");
            check(pa_memcpy_d2h(ctx_, sl.index, d_index_,
                                k * sizeof(uint32_t)),
                  "pa_memcpy_d2h");
        }
        sl.kept = k;
    };

    if (!pack(0))
        check(PA_ERR_INDEX, "aln_batch (tuple outside the block)");
    if (autotune_ && !tuned_)
        tune(slot_[0].pairs, slot_[0].n);
    for (uint64_t i = 0; i < nch; ++i) {
        std::future<bool> side;
        if (i > 0 || i + 1 < nch)
            side = std::async(std::launch::async, [&, i]() {
                if (i > 0)
                    post(i - 1);
                return i + 1 < nch ? pack(i + 1) : true;
            });
        try {
            align(slot_[i % 2]);
        } catch (...) {
            if (side.valid())
                side.wait();
            throw;
        }
        if (side.valid() && !side.get())
            check(PA_ERR_INDEX, "aln_batch (tuple outside the block)");
    }
    post(nch - 1);
}

void HipAligner::aln_batch(
    std::tuple<uint64_t, uint64_t, CommonKmerLight *> *mattuples, uint64_t beg,
    uint64_t end, uint64_t bl_roffset, uint64_t bl_coffset,
    const params_t &params)
{
    if (cfg_.mode != PA_MODE_SW)
        return;   // as the reference's X-drop aligner: no Light overload
    run(mattuples, beg, end, bl_roffset, bl_coffset, params, false);
}

void HipAligner::aln_batch(
    std::tuple<uint64_t, uint64_t, CommonKmerLoc *> *mattuples, uint64_t beg,
    uint64_t end, uint64_t bl_roffset, uint64_t bl_coffset,
    const params_t &params)
{
    run(mattuples, beg, end, bl_roffset, bl_coffset, params,
        cfg_.mode == PA_MODE_XDROP);
}

}  // namespace pastis_amd
