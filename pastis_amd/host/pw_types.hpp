// pw_types.hpp -- the value types the reference's aligner operator works on,
// restated with the same field order and sizes so that tuple arrays built by
// the reference (or by pw_runner.hpp) can be handed over unchanged.
//
//   CommonKmerLoc / CommonKmerLight : inc/Types.hpp:163-266
//   params_t (alignment fields)     : inc/Types.hpp:36-93, defaults of
//                                     src/main.cpp:97-176
#pragma once

#include <cstdint>
#include <utility>

namespace pastis_amd {

// A nonzero of the candidate matrix with the offsets of two shared k-mers
// (first = first shared k-mer, second = last one; offsets are relative to the
// sequence starts, row sequence first).
struct CommonKmerLoc {
    uint16_t count = 1;          // shared k-mers
    int16_t score = -1;          // max k-mer self score
    float score_aln = 0.0f;      // written by the aligner for passing pairs
    std::pair<uint16_t, uint16_t> first{0, 0};
    std::pair<uint16_t, uint16_t> second{0, 0};
};
static_assert(sizeof(CommonKmerLoc) == 16, "CommonKmerLoc layout");

// A nonzero without k-mer locations.
struct CommonKmerLight {
    uint16_t count = 1;
    int16_t score = -1;
    float score_aln = 0.0f;
};
static_assert(sizeof(CommonKmerLight) == 8, "CommonKmerLight layout");

// The alignment-related run parameters.
struct params_t {
    enum PwAln { ALN_NONE = 0, ALN_SEQAN_FULL, ALN_SEQAN_XDROP };
    int gap_open = -11;
    int gap_ext = -2;
    uint16_t klength = 6;
    PwAln pw_aln = ALN_NONE;
    int aln_seqan_xdrop = 49;
    uint64_t aln_batch_sz = 10000000000ull;
    double aln_cov_thr = 0.7;
    int aln_ani_thr = 30;
    int seed_count = 2;
    int ckthr = 0;
    float mosthr = -1.0f;
};

}  // namespace pastis_amd
