// pw_capi.cpp -- C entry point that runs one block through the C++ operator
// path exactly as the reference's similarity search does: HipAligner built
// from CLI values (src/sim_search.cpp:119-124), construct_seqs_bl, then
// pw_aln_batch (src/DistPWRunner.cpp:30-160).  Used by tests and benchmarks
// to drive the C++ layer from Python.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

#include "../../include/pastis_pw.h"
#include "pw_runner.hpp"

using namespace pastis_amd;

namespace {

template <class NT>
int run_block(HipAligner &al, const params_t &prm, uint64_t bl_roffset,
              uint64_t bl_coffset, uint64_t nnz, const uint64_t *rows,
              const uint64_t *cols, const uint16_t *count,
              const int16_t *score, const uint16_t *seeds,
              float *score_aln_out, uint64_t *stats)
{
    // DCSC of the nonzeros, columns ascending, rows ascending in a column
    std::vector<uint64_t> ord(nnz);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) {
        return cols[a] != cols[b] ? cols[a] < cols[b] : rows[a] < rows[b];
    });
    Dcsc<NT> C;
    C.cp.push_back(0);
    for (uint64_t k = 0; k < nnz; ++k) {
        const uint64_t i = ord[k];
        if (C.jc.empty() || C.jc.back() != cols[i]) {
            if (!C.jc.empty())
                C.cp.push_back(C.ir.size());
            C.jc.push_back(cols[i]);
        }
        C.ir.push_back(rows[i]);
        NT v;
        v.count = count[i];
        v.score = score[i];
        if constexpr (sizeof(NT) == sizeof(CommonKmerLoc)) {
            v.first = {seeds[4 * i], seeds[4 * i + 1]};
            v.second = {seeds[4 * i + 2], seeds[4 * i + 3]};
        }
        C.numx.push_back(v);
    }
    if (!C.jc.empty())
        C.cp.push_back(C.ir.size());
    std::unordered_map<uint64_t, uint64_t> where;
    where.reserve(nnz * 2);
    for (uint64_t i = 0; i < nnz; ++i)
        where[(rows[i] << 32) | cols[i]] = i;
    if (where.size() != nnz)
        return PA_ERR_ARG;   // duplicate coordinates
    RunnerStats st = pw_aln_batch(C, al, prm, bl_roffset, bl_coffset);
    std::fill(score_aln_out, score_aln_out + nnz, 0.0f);
    for (uint64_t i = 0; i < C.nzc(); ++i)
        for (uint64_t j = C.cp[i]; j < C.cp[i + 1]; ++j)
            score_aln_out[where[(C.ir[j] << 32) | C.jc[i]]] = C.numx[j].score_aln;
    if (stats) {
        const HipAlignerStats &h = al.stats();
        const uint64_t v[8] = {st.nnz_in, st.nnz_kmer, st.nnz_aln, st.batches,
                               h.aligned, h.filtered_out, h.passed,
                               h.aln_pairs};
        std::copy(v, v + 8, stats);
    }
    return PA_OK;
}

}  // namespace

extern "C" {

// mode 0: --sfa (Light tuples), 1: --sxa (Loc tuples).  res / offs hold
// the row sequence set as ASCII residues; cres / coffs the column set, or
// NULL for a diagonal block (the column set is the row set).  Nonzero i is
// (rows[i], cols[i]) in block-local indices, aligned against row sequence
// rows[i] + bl_roffset and column sequence cols[i] + bl_coffset (the block
// offsets of DistPWRunner.cpp:127-130), with count, score and, for --sxa,
// seeds[4*i .. 4*i+3] = first.(row, col), second.(row, col).
// score_aln_out[i] receives the final value (0 when the nonzero was
// pruned).  filter_params == NULL runs without the learned filter.  stats
// (8 values, may be NULL): nnz in, after the k-mer prune, after the
// alignment prune, batches, aligned, filtered out, passed, pairs.
int pw_run_block(int mode, int gap_open, int gap_ext, int klength, int xdrop,
                 int seed_count, int gap_model, int device, const char *res,
                 const uint64_t *offs, uint64_t nseq, const char *cres,
                 const uint64_t *coffs, uint64_t ncseq, uint64_t bl_roffset,
                 uint64_t bl_coffset, uint64_t nnz,
                 const uint64_t *rows, const uint64_t *cols,
                 const uint16_t *count, const int16_t *score,
                 const uint16_t *seeds, int ckthr, float mosthr,
                 uint64_t batch_sz, double cov_thr, int ani_thr,
                 const float *filter_params, const float *filter_mean,
                 const float *filter_istd, float filter_tau,
                 float *score_aln_out, uint64_t *stats)
{
    if (!offs || (nnz && (!rows || !cols || !count || !score ||
                          !score_aln_out)) ||
        (mode == 1 && nnz && !seeds) || batch_sz == 0 ||
        (cres != nullptr) != (coffs != nullptr))
        return PA_ERR_ARG;
    try {
        params_t prm;
        prm.gap_open = gap_open;
        prm.gap_ext = gap_ext;
        prm.klength = (uint16_t)klength;
        prm.pw_aln = mode == 0 ? params_t::ALN_SEQAN_FULL
                               : params_t::ALN_SEQAN_XDROP;
        prm.aln_seqan_xdrop = xdrop;
        prm.seed_count = seed_count;
        prm.ckthr = ckthr;
        prm.mosthr = mosthr;
        prm.aln_batch_sz = batch_sz;
        prm.aln_cov_thr = cov_thr;
        prm.aln_ani_thr = ani_thr;
        std::unique_ptr<HipAligner> al;
        if (mode == 0)
            al.reset(new HipAligner(gap_open, gap_ext, device, gap_model));
        else
            al.reset(new HipAligner(gap_open, gap_ext, (uint16_t)klength,
                                    xdrop, seed_count, device, gap_model));
        std::vector<char> r(res, res + offs[nseq]);
        std::vector<uint64_t> o(offs, offs + nseq + 1);
        if (coffs) {
            std::vector<char> cr(cres, cres + coffs[ncseq]);
            std::vector<uint64_t> co(coffs, coffs + ncseq + 1);
            al->construct_seqs_bl(std::make_shared<VectorSource>(
                std::move(r), std::move(o), std::move(cr), std::move(co)));
        } else {
            al->construct_seqs_bl(
                std::make_shared<VectorSource>(std::move(r), std::move(o)));
        }
        if (filter_params)
            al->set_filter(filter_params, filter_mean, filter_istd,
                           filter_tau);
        if (mode == 0)
            return run_block<CommonKmerLight>(*al, prm, bl_roffset, bl_coffset,
                                              nnz, rows, cols, count, score,
                                              seeds, score_aln_out, stats);
        return run_block<CommonKmerLoc>(*al, prm, bl_roffset, bl_coffset, nnz,
                                        rows, cols, count, score, seeds,
                                        score_aln_out, stats);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pw_run_block: %s\n", e.what());
        return PA_ERR_HIP;
    }
}

// Times the operator path on host tuples: HipAligner::aln_batch over one
// batch of nnz tuples built once from `pairs` (block-local r, c with
// bl_roffset = bl_coffset = 0 on a diagonal block of the nseq sequences),
// `warmup` untimed and `steps` timed calls.  ms_per_call receives the mean
// wall time of a timed call (packing, staging copies, alignment, result
// write-back), passed the pairs whose score_aln the last call wrote.
// autotune = 0 turns the aligner's first-call autotuning off.
int pw_bench_aln_batch(int mode, int gap_open, int gap_ext, int klength,
                       int xdrop, int seed_count, int gap_model, int device,
                       const char *res, const uint64_t *offs, uint64_t nseq,
                       const pa_pair *pairs, uint64_t nnz, double cov_thr,
                       int ani_thr, int autotune, int warmup, int steps,
                       double *ms_per_call, uint64_t *passed)
{
    if (!res || !offs || (nnz && !pairs) || !ms_per_call || steps < 1 ||
        warmup < 0)
        return PA_ERR_ARG;
    try {
        params_t prm;
        prm.aln_cov_thr = cov_thr;
        prm.aln_ani_thr = ani_thr;
        std::unique_ptr<HipAligner> al;
        if (mode == 0)
            al.reset(new HipAligner(gap_open, gap_ext, device, gap_model));
        else
            al.reset(new HipAligner(gap_open, gap_ext, (uint16_t)klength,
                                    xdrop, seed_count, device, gap_model));
        al->set_autotune(autotune != 0);
        al->construct_seqs_bl(std::make_shared<VectorSource>(
            std::vector<char>(res, res + offs[nseq]),
            std::vector<uint64_t>(offs, offs + nseq + 1)));
        std::vector<CommonKmerLoc> vals(nnz);
        std::vector<std::tuple<uint64_t, uint64_t, CommonKmerLoc *>> tup(nnz);
        for (uint64_t i = 0; i < nnz; ++i) {
            CommonKmerLoc &v = vals[i];
            v.count = pairs[i].count;
            v.score = pairs[i].kscore;
            v.first = {pairs[i].seed[0], pairs[i].seed[1]};
            v.second = {pairs[i].seed[2], pairs[i].seed[3]};
            tup[i] = std::make_tuple((uint64_t)pairs[i].r,
                                     (uint64_t)pairs[i].c, &vals[i]);
        }
        double ms = 0;
        for (int it = 0; it < warmup + steps; ++it) {
            for (auto &v : vals)
                v.score_aln = 0.0f;
            const auto t0 = std::chrono::steady_clock::now();
            al->aln_batch(tup.data(), 0, nnz, 0, 0, prm);
            const auto t1 = std::chrono::steady_clock::now();
            if (it >= warmup)
                ms += std::chrono::duration<double, std::milli>(t1 - t0)
                          .count();
        }
        *ms_per_call = ms / steps;
        if (passed) {
            uint64_t k = 0;
            for (const auto &v : vals)
                k += v.score_aln != 0.0f;
            *passed = k;
        }
        return PA_OK;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pw_bench_aln_batch: %s\n", e.what());
        return PA_ERR_HIP;
    }
}

}  // extern "C"
