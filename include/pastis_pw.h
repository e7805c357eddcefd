/*
 * pastis_pw.h -- C entry point of the C++ operator layer (libpastis_pw.so):
 * HipAligner (a pastis::PWAlign, inc/align/PWAlign.hpp:20-82) driven by the
 * batch loop of pw_aln_batch (src/DistPWRunner.cpp:30-160) over one
 * diagonal block, as src/sim_search.cpp:119-135 drives SeqanFullAligner /
 * SeqanXdropAligner.  The C++ classes themselves are in
 * pastis_amd/host/hip_aligner.hpp and pw_runner.hpp.
 */
#ifndef PASTIS_PW_H
#define PASTIS_PW_H

#include <stdint.h>

#include "pastis_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mode 0 = --sfa (CommonKmerLight tuples), 1 = --sxa (CommonKmerLoc).
 * res / offs: ASCII residues of the nseq row sequences; cres / coffs: the
 * ncseq column sequences, or NULL for a diagonal block (column set = row
 * set).  Nonzero i is (rows[i], cols[i]) in block-local indices, aligned
 * against row sequence rows[i] + bl_roffset and column sequence cols[i] +
 * bl_coffset (the block offsets pw_aln_batch hands to aln_batch,
 * src/DistPWRunner.cpp:127-130), with count[i], score[i] and, in mode 1,
 * seeds[4i..4i+3] = first.(row, col), second.(row, col); coordinates must be
 * unique.  ckthr / mosthr: the k-mer prune of DistPWRunner.cpp:48-54;
 * batch_sz: params.aln_batch_sz; cov_thr / ani_thr: params.aln_cov_thr /
 * aln_ani_thr.  filter_params (PA_FILTER_NPARAM floats, NULL = no learned
 * filter), filter_mean / filter_istd (16 floats each) and filter_tau
 * configure the learned prefilter.  score_aln_out[i] receives the nonzero's
 * final score_aln (0 when pruned).  stats (8 values, may be NULL): nnz in,
 * after the k-mer prune, after the alignment prune, batches, pairs aligned,
 * pairs dropped by the filter, pairs passed, pairs handed to aln_batch.
 * Returns a PA_* status of include/pastis_hip.h. */
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
                 float *score_aln_out, uint64_t *stats);

/* Times the operator path on host tuples: HipAligner::aln_batch
 * (inc/align/PWAlign.hpp:45-63) over one batch of nnz tuples made once from
 * pairs (r, c = block-local indices of a diagonal block of the nseq
 * sequences, count, kscore, seeds), warmup untimed and steps timed calls.
 * ms_per_call receives the mean wall time of a timed call: packing into
 * pa_pair, page-locked staging copies, the GPU alignment and the write-back
 * of score_aln.  passed (may be NULL) receives the pairs the last call
 * wrote.  autotune = 0 turns the aligner's first-call autotuning off. */
int pw_bench_aln_batch(int mode, int gap_open, int gap_ext, int klength,
                       int xdrop, int seed_count, int gap_model, int device,
                       const char *res, const uint64_t *offs, uint64_t nseq,
                       const pa_pair *pairs, uint64_t nnz, double cov_thr,
                       int ani_thr, int autotune, int warmup, int steps,
                       double *ms_per_call, uint64_t *passed);

#ifdef __cplusplus
}
#endif
#endif /* PASTIS_PW_H */
