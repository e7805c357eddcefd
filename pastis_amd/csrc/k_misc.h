// k_misc.h -- launchers of the task, binning and result kernels.
#pragma once

#include "pa_internal.h"

namespace pa {

// Bins: 2 * DP class (dp_class under the context's DpShapes) + (1 for the
// int32 kernels).  Inside a bin
// the tasks are ordered by kBuckets log-spaced buckets of their step count,
// longest first: the tasks a wave runs together are alike, and the
// persistent grid ends on short tasks.
constexpr int kClasses = 4;
constexpr int kBins = 2 * kClasses;
constexpr int kBuckets = 64;

hipError_t fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t s);
// *dst += *src (one thread; a device-side counter)
hipError_t accumulate(const uint32_t *src, unsigned long long *dst,
                      hipStream_t s);
// Pairs naming sequences outside the nr row / nc column sequences become
// empty tasks (and X-drop tasks whose seed does not fit are flagged
// kXdInvalid), so no kernel reads outside the sequence store.
hipError_t build_sw_tasks(const pa_pair *pairs, uint64_t n,
                          const uint64_t *roffs, const uint64_t *coffs,
                          uint64_t nr, uint64_t nc, DpTask *tasks,
                          unsigned long long *cells, hipStream_t s);
hipError_t build_xd_tasks(const pa_pair *pairs, uint64_t n, int seed_cnt,
                          int seed_len, const uint64_t *roffs,
                          const uint64_t *coffs, uint64_t nr, uint64_t nc,
                          XdTask *tasks, hipStream_t s);
hipError_t build_nw_tasks(const XdTask *xt, uint64_t nt, const XdOut *xo,
                          DpTask *tasks, ResBound rows, ResBound cols,
                          unsigned long long *cells,
                          unsigned long long *xcells, hipStream_t s);
// Counting sort of tasks [a, b) into per-bin lists: hist and fill are
// kBins * kBuckets counters that start at zero, counts receives the kBins
// list lengths.
// rng (optional): the task range [rng[0], rng[1]) read on the device, within
// [a, b) (chunks sized on the GPU, k_trace_chunks)
hipError_t bin_tasks(const DpTask *tasks, uint32_t a, uint32_t b,
                     const DpShapes &sh, int local, int maxs, int open,
                     int ext, uint32_t *lists,
                     uint32_t cap, uint32_t *counts, uint32_t *hist,
                     uint32_t *fill, unsigned long long *bin_cells,
                     hipStream_t s, const uint32_t *rng = nullptr);
hipError_t trace_sizes(const DpTask *tasks, uint64_t n, const DpShapes &sh,
                       uint64_t *sizes, hipStream_t s);
hipError_t empty_out(const DpTask *tasks, uint32_t a, uint32_t b, int local,
                     int open, int ext, DpOut *out, hipStream_t s,
                     const uint32_t *rng = nullptr);
// Device-side chunking of a traced batch (see k_trace_chunks).
hipError_t trace_chunks(const uint64_t *toffs, uint32_t n, uint64_t cap,
                        int nchunks, uint32_t *ranges, uint64_t *bases,
                        uint32_t *err, hipStream_t s);
hipError_t mark_flags(const uint32_t *list, const uint32_t *count,
                      uint32_t max_n, uint8_t *flags, uint8_t bit,
                      hipStream_t s);
// err (optional, device): when nonzero the batch's trace did not fit the
// captured chunks of a graph, and every result is PA_FLAG_NOTRACE
hipError_t finalize_sw(const DpTask *tasks, const DpOut *dp,
                       const uint8_t *tflags, uint64_t n, double cov_thr,
                       int ani_thr, pa_result *out, hipStream_t s,
                       const uint32_t *err = nullptr);
hipError_t finalize_xd(const XdTask *xt, const XdOut *xo, const DpOut *dp,
                       const uint8_t *tflags, uint64_t n, int seed_cnt,
                       double cov_thr, int ani_thr, pa_result *out,
                       hipStream_t s, const uint32_t *err = nullptr);
hipError_t scores_out(const DpOut *dp, uint64_t n, pa_score *out,
                      hipStream_t s);
hipError_t extents_out(const XdOut *xo, uint64_t n, int seed, pa_extent *out,
                       hipStream_t s);

}  // namespace pa
