// pa_ctx.h -- the context behind the C ABI's opaque pa_ctx handle, shared by
// the host-side sources of libpastis_hip.so.
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "k_misc.h"

namespace pa {

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Timer {
    int kind;  // 0 sw, 1 xdrop, 2 nw, 3 trace, 4 total
    hipEvent_t a, b;
};

struct FilterState;   // learned prefilter (pa_filter.hip)
void filter_free(FilterState *f);
struct KmerState;     // k-mer candidate generation (k_kmer.hip)
void kmer_free(KmerState *k);

}  // namespace pa

namespace pa {
inline pa_tuning default_tuning()
{
    pa_tuning t{};
    t.dp_wg_per_cu = 0;
    t.xd_wg_per_cu = 0;
    t.xd_ring = 16;
    t.xd_coop = 0;
    for (int c = 0; c < 3; ++c)
        t.dp_edge[c] = (int32_t)kDefaultShapes.edge[c];
    for (int c = 0; c < 4; ++c)
        t.dp_shape[c] = kDefaultShapes.shape[c];
    t.tb_stage = 1;
    t.xd_wave = 1;
    t.tb_overlap = 0;
    return t;
}
inline DpShapes shapes_of(const pa_tuning &t)
{
    DpShapes sh{};
    for (int c = 0; c < 3; ++c)
        sh.edge[c] = (uint32_t)t.dp_edge[c];
    for (int c = 0; c < 4; ++c)
        sh.shape[c] = (uint8_t)t.dp_shape[c];
    return sh;
}
}  // namespace pa

struct pa_ctx {
    using Buf = pa::Buf;
    using Timer = pa::Timer;
    pa_config cfg{};
    int8_t mat[32 * 32];
    int maxs = 0;
    int device = 0;
    hipStream_t st = nullptr;
    // Tracebacks of traced DP batches run on st2, each DP class's as soon as
    // that class's DP kernels are done (fork events), while the next class's
    // DP runs; the join event hands the results back to the batch's stream.
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork[pa::kClasses] = {};
    hipEvent_t ev_join = nullptr;
    std::string err;
    // sequence sets
    Buf rcodes, ccodes, roffs, coffs;
    // k-mer letters of the sets, one byte per residue at the residue's
    // offset in rcodes / ccodes: the letter's code 0..24, or 255 for a
    // letter that no k-mer may hold (KmerOps.cpp:66-73)
    Buf rkc, ckc;
    uint64_t seq_gen = 0;         // bumped whenever a sequence set changes
    Buf rsk, csk;                 // per-sequence filter sketches
    // byte offset of the residue-bound prefix sums inside rcodes / ccodes
    size_t rpre_off = 0, cpre_off = 0;
    bool shared_cols = false;
    std::vector<uint32_t> rlen, clen;
    uint32_t max_rlen = 0, max_clen = 0;
    bool have_rows = false, have_cols = false;
    // workspace
    Buf pairs, results, scores, tasks, dpout, tflags, sizes, toffs, lists,
        sat, ctr, trace, bnd, xdt, xdo, ovf, cub, cells, dmat, ext;
    size_t trace_cap = 0;
    Buf bincells;                 // u64 cells per (DP mode, bin)
    Buf tri_flag, tri_pos;        // similarity-triple compaction
    struct KStat {
        uint64_t launches = 0;
        double ms = 0;
    } kstat[3][pa::kBins + pa::kClasses];   // DP mode x (bins, int32 re-runs)
    std::vector<Timer> pending;
    std::vector<hipEvent_t> free_events;
    pa_stats stats{};
    pa::FilterState *flt = nullptr;
    pa::KmerState *kmer = nullptr;
    int ncu = 256;                // compute units of the device
    // hipGraphs of score-only SW batches (pa_sw_scores_graph), one per
    // (pairs, n, out, stream), least recently used dropped beyond kGraphs.
    // They share the workspace, so they replay one after another on a
    // stream (as the batches of a pw_aln_batch loop do).
    bool capturing = false;       // no event timers while a stream captures
    struct GraphEntry {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        const void *pairs = nullptr, *out = nullptr;
        uint64_t n = 0;
        hipStream_t s = nullptr;
        uint64_t gen = 0, used = 0;
        int kind = 0;             // 0 pa_sw_scores_graph, 1 pa_align_graph
    };
    // Traced batches captured by pa_align_graph: while capturing, g_chunks
    // (> 0) trace-store chunks of g_cap bytes are sized on the device
    // (gchunk: task ranges, base offsets, overflow flag) and the X-drop
    // global-ring pass always runs, on the kept scratch xdscr.
    int g_chunks = 0;
    uint64_t g_cap = 0;
    Buf gchunk, xdscr;
    int xdscr_cap = 0;           // ring slots of xdscr per lane
    static constexpr int kGraphs = 8;
    std::vector<GraphEntry> graphs;
    uint64_t graph_tick = 0;
    // Workspace generation: bumped whenever a buffer a graph may have
    // captured is reallocated (grow), the sequence sets change or the launch
    // shapes change; a graph captured at another generation is recaptured.
    uint64_t gen = 0;
    pa_tuning tune = pa::default_tuning();
    pa::DpShapes shapes = pa::kDefaultShapes;   // derived from tune
    // the scoring matrix fits 4-bit profiles (kPmatNib) with scores nib_min
    // .. nib_min + 15; if not, shapes replaces the 4-bit-profile shapes
    bool nib_ok = true;
    int nib_min = 0;
    // No event timers (pa_sw_scores_noalloc: event objects are allocated).
    bool quiet = false;
    // Cap on the task slots (boundary rows) of a DP grid (0 = none); the
    // boundary-row store is sized for the grids.
    uint64_t slot_cap = 0;
};

namespace pa {

// kernel shapes of a tuning on this context (4-bit-profile shapes replaced
// when the matrix does not fit them)
inline DpShapes shapes_for(const pa_ctx *c, const pa_tuning &t)
{
    DpShapes sh = shapes_of(t);
    if (!c->nib_ok)
        for (int k = 0; k < 4; ++k)
            sh.shape[k] = (uint8_t)shape_byte_fallback(sh.shape[k]);
    return sh;
}

inline int fail(pa_ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    return code;
}

#define HIPCHK(ctx, x)                                                       \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess)                                                \
            return fail((ctx), PA_ERR_HIP,                                   \
                        std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

// Buffer of at least `bytes`, allocated with a quarter of slack (at most up
// to `limit`) so that slightly larger requests reuse it.
inline int grow(pa_ctx *c, Buf &b, size_t bytes, size_t limit = SIZE_MAX)
{
    if (b.cap >= bytes && b.p)
        return PA_OK;
    if (c)
        ++c->gen;   // captured graphs hold the old address
    if (b.p)
        hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = std::max(bytes, std::min(bytes + bytes / 4 + 256, limit));
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        return fail(c, PA_ERR_NOMEM,
                    "device allocation of " + std::to_string(want) + " bytes");
    }
    b.cap = want;
    return PA_OK;
}

inline void release(Buf &b)
{
    if (b.p)
        hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

template <class T> inline T *P(Buf &b) { return (T *)b.p; }

// Score-only local alignment of n device pairs on the context's workspace
// (pa_sw_scores_dev after its argument checks).
int sw_scores_impl(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                   pa_score *d_out, hipStream_t s);
// Resolves this device's scratch array for pa_sw_scores_noalloc
// (pa_noalloc.hip); called once per context at creation.
void noalloc_init(int device);

}  // namespace pa
