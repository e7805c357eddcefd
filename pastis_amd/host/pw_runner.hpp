// pw_runner.hpp -- the batch loop that feeds the aligner operator, restated
// from pw_aln_batch (src/DistPWRunner.cpp:30-160) over a plain DCSC block
// instead of a CombBLAS SpParMat.
//
//   1. k-mer prune: drop nonzeros with count <= ckthr || score <= mosthr
//      (DistPWRunner.cpp:48-54; score is a short compared as a float);
//   2. tuple formation: column-major walk over the nonzero columns,
//      (row index, column index, pointer to the value) (:89-105);
//   3. batch loop: batch_cnt = l_nnz / bsz + 1 calls of aln_batch on
//      [b*bsz, min((b+1)*bsz, l_nnz)), the last one possibly empty (:116-131);
//   4. post-prune: drop nonzeros whose score_aln is 0 (:149-153).
#pragma once

#include <cstdint>
#include <tuple>
#include <vector>

#include "hip_aligner.hpp"

namespace pastis_amd {

// Doubly compressed sparse columns: nzc nonzero columns jc[i] whose rows are
// ir[cp[i] .. cp[i+1]) with values numx[...] (CombBLAS Dcsc's fields).
template <class NT> struct Dcsc {
    std::vector<uint64_t> jc, cp, ir;
    std::vector<NT> numx;
    uint64_t nzc() const { return jc.size(); }
    uint64_t nnz() const { return numx.size(); }

    // Keeps the nonzeros for which drop() is false, preserving order.
    template <class F> void prune(F drop)
    {
        std::vector<uint64_t> njc, ncp{0}, nir;
        std::vector<NT> nx;
        for (uint64_t i = 0; i < nzc(); ++i) {
            for (uint64_t j = cp[i]; j < cp[i + 1]; ++j)
                if (!drop(numx[j])) {
                    nir.push_back(ir[j]);
                    nx.push_back(numx[j]);
                }
            if (nir.size() > ncp.back()) {
                njc.push_back(jc[i]);
                ncp.push_back(nir.size());
            }
        }
        jc.swap(njc);
        cp.swap(ncp);
        ir.swap(nir);
        numx.swap(nx);
    }
};

struct RunnerStats {
    uint64_t nnz_in = 0, nnz_kmer = 0, nnz_aln = 0, batches = 0;
};

// pw_aln_batch for one block: bri / bci offsets as bl_roffset / bl_coffset.
template <class NT>
RunnerStats pw_aln_batch(Dcsc<NT> &C, PWAlign &pwa, const params_t &params,
                         uint64_t bl_roffset = 0, uint64_t bl_coffset = 0)
{
    RunnerStats st;
    st.nnz_in = C.nnz();
    const int ckthr = params.ckthr;
    const float mosthr = params.mosthr;
    C.prune([&](const NT &v) {
        return v.count <= ckthr || (float)v.score <= mosthr;
    });
    st.nnz_kmer = C.nnz();
    const uint64_t l_nnz = C.nnz();
    if (l_nnz != 0) {
        const uint64_t bsz = params.aln_batch_sz;
        const uint64_t batch_cnt = l_nnz / bsz + 1;
        std::vector<std::tuple<uint64_t, uint64_t, NT *>> t(l_nnz);
        uint64_t k = 0;
        for (uint64_t i = 0; i < C.nzc(); ++i)
            for (uint64_t j = C.cp[i]; j < C.cp[i + 1]; ++j)
                t[k++] = std::make_tuple(C.ir[j], C.jc[i], &C.numx[j]);
        for (uint64_t b = 0; b < batch_cnt; ++b) {
            const uint64_t beg = b * bsz;
            const uint64_t end = (b + 1) * bsz > l_nnz ? l_nnz : (b + 1) * bsz;
            pwa.aln_batch(t.data(), beg, end, bl_roffset, bl_coffset, params);
            st.batches++;
        }
    }
    C.prune([](const NT &v) { return v.score_aln == 0.0f; });
    st.nnz_aln = C.nnz();
    return st;
}

}  // namespace pastis_amd
