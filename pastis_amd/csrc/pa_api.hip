// pa_api.hip -- host side of the C ABI declared in include/pastis_hip.h.
//
// A context owns one HIP stream, the uploaded row/column sequence sets and a
// workspace that grows on demand and is reused across batches.  Every batch
// runs as a fixed sequence of kernels on the context's stream:
//   SW mode   : build tasks -> bin -> traced local DP (int16, int32 re-run of
//               saturated tasks) -> traceback -> predicate
//   X-drop    : build seed tasks -> X-drop extension -> build infix tasks ->
//               bin -> traced global DP -> traceback -> best seed + predicate
// The traced DP writes 4-bit codes into a trace store; batches whose trace
// would exceed the store are split into chunks at task boundaries.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "k_filter.h"
#include "k_misc.h"
#include "pa_ctx.h"

using namespace pa;

namespace {

const char *kStatus[] = {"ok",
                         "invalid argument",
                         "HIP runtime error",
                         "out of memory",
                         "sequence longer than 65535",
                         "residue code outside 0..24",
                         "pair index outside the loaded sequences",
                         "seed outside its sequences",
                         "sequences not loaded",
                         "no usable gfx950 device",
                         "size limit exceeded"};

const char kAlphabet[] = "ARNDCQEGHILKMFPSTWYVBZX*J";

// inc/ScoreMat.hpp:104-131 (BLOSUM62, letter order of src/Alphabet.cpp:26)
const int8_t kBlosum62[25 * 25] = {
     4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0,-2,-1, 0,-4,-1,
    -1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3,-1, 0,-1,-4,-2,
    -2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3, 3, 0,-1,-4,-3,
    -2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3, 4, 1,-1,-4,-4,
     0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1,-3,-3,-2,-4,-1,
    -1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2, 0, 3,-1,-4,-2,
    -1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4,-3,
     0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3,-1,-2,-1,-4,-4,
    -2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3, 0, 0,-1,-4,-3,
    -1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3,-3,-3,-1,-4, 2,
    -1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-3,-1,-4, 4,
    -1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2, 0, 1,-1,-4,-2,
    -1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1,-3,-1,-1,-4, 2,
    -2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1,-3,-3,-1,-4, 0,
    -1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2,-2,-1,-2,-4,-3,
     1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2, 0, 0, 0,-4,-2,
     0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0,-1,-1, 0,-4,-1,
    -3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3,-4,-3,-2,-4,-2,
    -2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1,-3,-2,-1,-4,-1,
     0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4,-3,-2,-1,-4, 1,
    -2,-1, 3, 4,-3, 0, 1,-1, 0,-3,-4, 0,-3,-3,-2, 0,-1,-4,-3,-3, 4, 1,-1,-4,-4,
    -1, 0, 0, 1,-3, 3, 4,-2, 0,-3,-3, 1,-1,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4,-3,
     0,-1,-1,-1,-2,-1,-1,-1,-1,-1,-1,-1,-1,-1,-2, 0, 0,-2,-1,-1,-1,-1,-1,-4,-1,
    -4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4, 1,-4,
    -1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-3,-1,-4, 4,
};

// Work-queue and counter slots (uint32) of the per-chunk counter block.
enum {
    C_BIN = 0,          // kBins list lengths
    C_BINQ = 8,         // kBins work-queue heads
    C_SAT = 16,         // kClasses saturation-list lengths
    C_SATQ = 20,        // kClasses re-run queue heads
    C_XDQ = 24,         // X-drop queue head
    C_XDOVF = 25,       // X-drop overflow count
    C_XDOVFQ = 26,      // X-drop overflow queue head
    C_XDN = 27,         // X-drop task count
    C_XDDUMMY = 28,     // overflow sink of the wide re-run
    C_ERR = 29,         // validation errors
    C_XDLONG = 30,      // X-drop items left by the 16-lane kernel
    C_XDLONGQ = 31,     // their queue head
    C_XDWOVF = 32,      // X-drop items left by the 128-column band kernel
    C_XDWOVFQ = 33,     // their queue head
    C_XDW2OVF = 34,     // items left by the 256-column band kernel
    C_XDW2OVFQ = 35,    // their queue head
    C_HIST = 36,        // kBins x kBuckets histogram of the counting sort
    C_FILL = C_HIST + kBins * kBuckets,   // its running slots
    C_N = C_FILL + kBins * kBuckets
};

constexpr int kMaxWgPerCu = 32;     // one-wave workgroups per CU (8 per SIMD)
constexpr int kXdWideLanes = 65536;  // lanes of the wide (global scratch) re-run
constexpr int kXdGraphLanes = 4096;  // its lanes in a captured batch

}  // namespace

namespace {



hipEvent_t get_event(pa_ctx *c)
{
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Event timer of one region on a stream; inactive while the context's
// stream is being captured into a graph.
struct Scope {
    pa_ctx *c;
    Timer t;
    hipStream_t s;
    bool on;
    Scope(pa_ctx *c_, int kind, hipStream_t s_)
        : c(c_), s(s_), on(!c_->capturing && !c_->quiet)
    {
        if (!on)
            return;
        t.kind = kind;
        t.a = get_event(c);
        t.b = get_event(c);
        hipEventRecord(t.a, s);
    }
    ~Scope()
    {
        if (!on)
            return;
        hipEventRecord(t.b, s);
        c->pending.push_back(t);
    }
};

void resolve_timers(pa_ctx *c)
{
    for (auto &t : c->pending) {
        float ms = 0.f;
        hipEventSynchronize(t.b);
        hipEventElapsedTime(&ms, t.a, t.b);
        if (t.kind >= 100) {
            auto &k = c->kstat[(t.kind - 100) / 16][(t.kind - 100) % 16];
            k.launches++;
            k.ms += ms;
            c->free_events.push_back(t.a);
            c->free_events.push_back(t.b);
            continue;
        }
        switch (t.kind) {
        case 0: c->stats.ms_sw += ms; break;
        case 1: c->stats.ms_xdrop += ms; break;
        case 2: c->stats.ms_nw += ms; break;
        case 3: c->stats.ms_trace += ms; break;
        default: c->stats.ms_total += ms; break;
        }
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
}

int ensure_common(pa_ctx *c, uint64_t nt)
{
    int rc;
    if ((rc = grow(c, c->ctr, C_N * sizeof(uint32_t))))
        return rc;
    if ((rc = grow(c, c->tasks, nt * sizeof(DpTask))))
        return rc;
    if ((rc = grow(c, c->dpout, nt * sizeof(DpOut))))
        return rc;
    if ((rc = grow(c, c->tflags, nt)))
        return rc;
    if ((rc = grow(c, c->lists, (size_t)kBins * nt * sizeof(uint32_t))))
        return rc;
    if ((rc = grow(c, c->sat, (size_t)kClasses * nt * sizeof(uint32_t))))
        return rc;
    return PA_OK;
}

// Persistent DP grid: occupancy-limited workgroups per CU, capped by the
// tuning and by the number of tasks.
// persistent DP workgroups per CU of a class's kernel
int dp_wg_per_cu(pa_ctx *c, int cls, bool local, bool trace, bool i16)
{
    int per_cu = dp_occupancy(c->shapes.shape[cls], local, trace, i16);
    if (c->tune.dp_wg_per_cu > 0)
        per_cu = std::min(per_cu, (int)c->tune.dp_wg_per_cu);
    return std::min(per_cu, kMaxWgPerCu);
}

int dp_grid(pa_ctx *c, uint64_t ntasks, int cls, bool local, bool trace,
            bool i16)
{
    uint64_t g = std::min<uint64_t>(
        ntasks, (uint64_t)c->ncu * dp_wg_per_cu(c, cls, local, trace, i16));
    if (c->slot_cap > 0)
        g = std::min<uint64_t>(
            g, c->slot_cap / (uint64_t)shape_slots(c->shapes.shape[cls]));
    return (int)std::max<uint64_t>(1, g);
}

int xd_grid(pa_ctx *c, uint64_t nitems)
{
    int per_cu = xdrop_occupancy(c->tune.xd_ring, c->cfg.xdrop,
                                 c->cfg.xdrop_gap);
    if (c->tune.xd_wg_per_cu > 0)
        per_cu = std::min(per_cu, (int)c->tune.xd_wg_per_cu);
    per_cu = std::min(per_cu, kMaxWgPerCu);
    return (int)std::max<uint64_t>(
        1, std::min<uint64_t>((nitems + 63) / 64, (uint64_t)c->ncu * per_cu));
}

int ensure_bnd(pa_ctx *c, bool local, bool trace)
{
    // one boundary row per task slot of the persistent workgroups of the
    // run's DP grids: per class, its grid (dp_grid) times the task slots of
    // its shape's one-wave workgroup
    size_t slots = 1;
    for (int cls = 0; cls < kClasses; ++cls)
        for (int i16 = 0; i16 < 2; ++i16)
            slots = std::max(slots,
                             (size_t)dp_grid(c, ~0ull, cls, local, trace,
                                             i16 != 0) *
                                 shape_slots(c->shapes.shape[cls]));
    const size_t need =
        slots * 2u * (size_t)(c->max_clen + 2) * sizeof(int32_t);
    return grow(c, c->bnd, need);
}

// Overflow flag of the device-sized trace chunks (after the ranges and
// bases in gchunk).
uint32_t *gchunk_err(pa_ctx *c)
{
    return (uint32_t *)((uint64_t *)(P<uint32_t>(c->gchunk) +
                                     2 * c->g_chunks) +
                        c->g_chunks);
}

// The trace store at `bytes`, or as large as the device allows: the store
// limit (trace_cap) is lowered to leave kTraceReserve of the device memory
// free now for everything else, and halves when the allocation still fails
// (down to 64 MiB); the batch then runs in more chunks.
constexpr uint64_t kTraceReserve = (uint64_t)16 << 30;

int grow_trace(pa_ctx *c, uint64_t bytes)
{
    if (c->trace.cap < std::min<uint64_t>(bytes, c->trace_cap)) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
            // the store's own memory comes back when it is reallocated
            const uint64_t room = (uint64_t)fr + c->trace.cap;
            const uint64_t limit =
                room > kTraceReserve + ((uint64_t)64 << 20)
                    ? room - kTraceReserve
                    : (uint64_t)64 << 20;
            c->trace_cap = std::min<uint64_t>(c->trace_cap, limit);
        }
    }
    for (;;) {
        const int rc = grow(c, c->trace,
                            std::min<uint64_t>(bytes, c->trace_cap),
                            c->trace_cap);
        if (rc != PA_ERR_NOMEM || c->trace_cap <= ((uint64_t)64 << 20))
            return rc;
        c->trace_cap /= 2;
    }
}

// Runs the DP kernels (and the traceback when traced) on tasks [0, nt).
int run_dp(pa_ctx *c, uint64_t nt, bool local, bool trace, int kind,
           hipStream_t s)
{
    int rc;
    if ((rc = ensure_bnd(c, local, trace)))
        return rc;
    // a fill kernel, not a memset node, so that the sequence can be captured
    HIPCHK(c, fill_u32(P<uint32_t>(c->tflags), 0, (uint32_t)((nt + 3) / 4), s));
    std::vector<std::pair<uint64_t, uint64_t>> chunks;
    std::vector<uint64_t> chunk_base;   // trace offset of each chunk's first
    if (trace && c->g_chunks > 0) {
        // captured batch: the trace store (g_cap bytes) is split into
        // g_chunks chunks on the device; no host synchronisation
        if ((rc = grow(c, c->sizes, nt * sizeof(uint64_t) + 8)))
            return rc;
        if ((rc = grow(c, c->toffs, nt * sizeof(uint64_t) + 8)))
            return rc;
        HIPCHK(c, trace_sizes(P<DpTask>(c->tasks), nt, c->shapes,
                              P<uint64_t>(c->sizes), s));
        // sizes[nt] = 0 so that toffs[nt] is the total (a fill kernel: no
        // memset node in a captured graph)
        HIPCHK(c, fill_u32((uint32_t *)(P<uint64_t>(c->sizes) + nt), 0, 2, s));
        size_t tmp = 0;
        HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(
                      nullptr, tmp, P<uint64_t>(c->sizes),
                      P<uint64_t>(c->toffs), (int)(nt + 1), s));
        if (tmp > c->cub.cap)
            return fail(c, PA_ERR_HIP, "scan workspace grew during capture");
        HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(
                      c->cub.p, tmp, P<uint64_t>(c->sizes),
                      P<uint64_t>(c->toffs), (int)(nt + 1), s));
        HIPCHK(c, trace_chunks(P<uint64_t>(c->toffs), (uint32_t)nt, c->g_cap,
                               c->g_chunks, P<uint32_t>(c->gchunk),
                               (uint64_t *)(P<uint32_t>(c->gchunk) +
                                            2 * c->g_chunks),
                               gchunk_err(c), s));
        for (int k = 0; k < c->g_chunks; ++k)
            chunks.push_back({0, nt});
    } else if (trace) {
        if ((rc = grow(c, c->sizes, nt * sizeof(uint64_t) + 8)))
            return rc;
        if ((rc = grow(c, c->toffs, nt * sizeof(uint64_t) + 8)))
            return rc;
        HIPCHK(c, trace_sizes(P<DpTask>(c->tasks), nt, c->shapes,
                              P<uint64_t>(c->sizes), s));
        size_t tmp = 0;
        HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(
                      nullptr, tmp, P<uint64_t>(c->sizes),
                      P<uint64_t>(c->toffs), (int)(nt + 1), s));
        if ((rc = grow(c, c->cub, tmp)))
            return rc;
        // sizes[nt] = 0 so that toffs[nt] is the total
        HIPCHK(c, hipMemsetAsync(P<uint64_t>(c->sizes) + nt, 0, 8, s));
        HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(
                      c->cub.p, tmp, P<uint64_t>(c->sizes),
                      P<uint64_t>(c->toffs), (int)(nt + 1), s));
        uint64_t total = 0;
        HIPCHK(c, hipMemcpyAsync(&total, P<uint64_t>(c->toffs) + nt, 8,
                                 hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        if (total <= c->trace_cap &&
            (rc = grow_trace(c, std::max<uint64_t>(total, 64))) == PA_OK &&
            c->trace.cap >= total) {
            chunks.push_back({0, nt});
        } else {
            if (rc && rc != PA_ERR_NOMEM)
                return rc;
            if ((rc = grow_trace(c, std::min<uint64_t>(c->trace_cap, total))))
                return rc;
            // cut into chunks of at most trace_cap bytes on the device
            // (k_trace_chunks); only the ranges and bases come back
            int nch = (int)std::min<uint64_t>(total / c->trace_cap + 2,
                                              1u << 20);
            for (;;) {
                const size_t bytes = (size_t)nch * 16 + 4;
                if ((rc = grow(c, c->gchunk, bytes)))
                    return rc;
                uint32_t *rg = P<uint32_t>(c->gchunk);
                uint64_t *bs = (uint64_t *)(rg + 2 * nch);
                uint32_t *err = (uint32_t *)(bs + nch);
                HIPCHK(c, fill_u32(err, 0, 1, s));
                HIPCHK(c, trace_chunks(P<uint64_t>(c->toffs), (uint32_t)nt,
                                       c->trace_cap, nch, rg, bs, err, s));
                std::vector<uint8_t> h(bytes);
                HIPCHK(c, hipMemcpyAsync(h.data(), rg, bytes,
                                         hipMemcpyDeviceToHost, s));
                HIPCHK(c, hipStreamSynchronize(s));
                const uint32_t *hr = (const uint32_t *)h.data();
                const uint64_t *hb = (const uint64_t *)(hr + 2 * nch);
                const uint32_t herr = *(const uint32_t *)(hb + nch);
                // the chunker stops with the error flag either at a task
                // larger than the store (the ranges then end at nt) or when
                // the chunks run out before the tasks do
                const bool ran_out = hr[2 * (nch - 1) + 1] < nt;
                if (herr && !ran_out)
                    return fail(c, PA_ERR_LIMIT,
                                "one alignment needs more trace memory than "
                                "the trace store holds");
                if (herr) {   // more chunks than estimated
                    nch *= 2;
                    continue;
                }
                for (int k = 0; k < nch; ++k)
                    if (hr[2 * k] < hr[2 * k + 1]) {
                        chunks.push_back({hr[2 * k], hr[2 * k + 1]});
                        chunk_base.push_back(hb[k]);
                    }
                break;
            }
        }
    } else {
        chunks.push_back({0, nt});
    }

    uint32_t *ctr = P<uint32_t>(c->ctr);
    uint32_t *lists = P<uint32_t>(c->lists);
    uint32_t *sat = P<uint32_t>(c->sat);
    const int dmode = local ? (trace ? 1 : 0) : 2;
    // tracebacks on the second stream, overlapping later classes' DP
    // (also inside a pa_align_graph capture: the event fork and join
    // bring the second stream into the graph)
    // the second stream (also inside a pa_align_graph capture: the event
    // fork and join bring it into the graph)
    const bool two = c->st2 && (!c->capturing || c->g_chunks > 0);
    // walks of each class beside the next classes' DP (tb_overlap 1), or
    // of all classes but the last beside the last class's DP (2), or all
    // after the DP (0)
    const bool split = trace && two && c->tune.tb_overlap == 1;
    const bool beside = trace && two && c->tune.tb_overlap == 2;
    const bool dev_chunks = trace && c->g_chunks > 0;
    for (size_t ci = 0; ci < chunks.size(); ++ci) {
        const auto &ch = chunks[ci];
        const uint32_t a = (uint32_t)ch.first, b = (uint32_t)ch.second;
        uint64_t base = 0;
        if (trace && !dev_chunks && chunks.size() > 1)
            base = chunk_base[ci];
        // device-sized chunk ci: task range and trace base from the device
        const uint32_t *rng =
            dev_chunks ? P<uint32_t>(c->gchunk) + 2 * ci : nullptr;
        const uint64_t *tbase =
            dev_chunks ? (const uint64_t *)(P<uint32_t>(c->gchunk) +
                                            2 * c->g_chunks) + ci
                       : nullptr;
        HIPCHK(c, fill_u32(ctr, 0, C_XDQ, s));
        HIPCHK(c, fill_u32(ctr + C_HIST, 0, C_N - C_HIST, s));
        HIPCHK(c, bin_tasks(P<DpTask>(c->tasks), a, b, c->shapes,
                            local ? 1 : 0, c->maxs,
                            c->cfg.gap_open, c->cfg.gap_ext, lists,
                            (uint32_t)nt, ctr + C_BIN, ctr + C_HIST,
                            ctr + C_FILL,
                            (unsigned long long *)P<uint64_t>(c->bincells) +
                                dmode * kBins,
                            s, rng));
        HIPCHK(c, empty_out(P<DpTask>(c->tasks), a, b, local ? 1 : 0,
                            c->cfg.gap_open, c->cfg.gap_ext,
                            P<DpOut>(c->dpout), s, rng));
        DpArgs args{};
        args.codes = nullptr;  // task offsets hold absolute device addresses
        args.tasks = P<DpTask>(c->tasks);
        args.mat = P<int8_t>(c->dmat);
        const uint32_t *pmat_byte =
            (const uint32_t *)(P<int8_t>(c->dmat) +
                               (!local ? kPmatGlobal
                                       : trace ? kPmatLocal : kPmatUntraced));
        const uint32_t *pmat_nib =
            (const uint32_t *)(P<int8_t>(c->dmat) + kPmatNib);
        args.nib_min = c->nib_min;
        args.out = P<DpOut>(c->dpout);
        args.trace = trace ? P<uint8_t>(c->trace) - base : nullptr;
        args.trace_off = trace ? P<uint64_t>(c->toffs) : nullptr;
        args.trace_base = tbase;
        args.bnd = P<int32_t>(c->bnd);
        args.bnd_stride = 2u * (c->max_clen + 2);
        args.gap_open = c->cfg.gap_open;
        args.gap_ext = c->cfg.gap_ext;
        TbArgs tb{};
        tb.codes = nullptr;
        tb.tasks = P<DpTask>(c->tasks);
        tb.trace = trace ? P<uint8_t>(c->trace) - base : nullptr;
        tb.trace_off = P<uint64_t>(c->toffs);
        tb.trace_base = tbase;
        tb.out = P<DpOut>(c->dpout);
        tb.local = local ? 1 : 0;
        // traceback of one class's tasks: its two bin lists (bin = 2 * class
        // + arithmetic)
        // Local tasks always run in int16 (saturated ones re-run in int32
        // afterwards), so their int32 bins stay empty and are not launched.
        const int nw = local ? 1 : 2;
        auto traceback = [&](int cls, hipStream_t ts) -> int {
            Scope sc(c, 3, ts);
            for (int w = 0; w < nw; ++w) {
                const int bin = 2 * cls + w;
                tb.list = lists + (size_t)bin * nt;
                tb.count = ctr + C_BIN + bin;
                HIPCHK(c, launch_traceback(tb, c->shapes.shape[cls],
                                           c->tune.tb_stage != 0,
                                           (int)((b - a + kTbThreads - 1) /
                                                 kTbThreads),
                                           ts));
                c->stats.launches++;
            }
            return PA_OK;
        };
        {
            Scope sc(c, kind, s);
            // DP of class rcl's bin w (0: int16, 1: int32)
            auto dp_bin = [&](int rcl, int w) -> int {
                // short tasks are claimed many per atomic, long ones singly
                // (the 16-lane classes claim rounds of four tasks)
                args.chunk = rcl <= 2 ? 4u : 1u;
                args.pmat = rows_nib(shape_rows(c->shapes.shape[rcl]))
                                ? pmat_nib
                                : pmat_byte;
                {
                    const int bin = 2 * rcl + w;
                    args.list = lists + (size_t)bin * nt;
                    args.count = ctr + C_BIN + bin;
                    args.queue = ctr + C_BINQ + bin;
                    args.sat_list = sat + (size_t)rcl * nt;
                    args.sat_count = ctr + C_SAT + rcl;
                    {
                        Scope k(c, 100 + dmode * 16 + bin, s);
                        HIPCHK(c, launch_dp(c->shapes.shape[rcl], local, trace,
                                            w == 0, args,
                                            dp_grid(c, b - a, rcl, local, trace,
                                                    w == 0),
                                            s));
                    }
                    c->stats.launches++;
                }
                return PA_OK;
            };
            // The int32 bins (global tasks that may leave the int16 range,
            // usually none) of every class first, while the GPU is free: a
            // launch queued between the classes waits for the CUs that the
            // walks of the previous class hold, even when its bin is empty.
            if (nw == 2)
                for (int rcl = 0; rcl < kClasses; ++rcl)
                    if ((rc = dp_bin(rcl, 1)))
                        return rc;
            for (int rcl = 0; rcl < kClasses; ++rcl) {
                if ((rc = dp_bin(rcl, 0)))
                    return rc;
                if (local) {
                    // exact int32 re-run of the tasks whose int16 DP saturated
                    args.list = sat + (size_t)rcl * nt;
                    args.count = ctr + C_SAT + rcl;
                    args.queue = ctr + C_SATQ + rcl;
                    {
                        Scope k(c, 100 + dmode * 16 + kBins + rcl, s);
                        HIPCHK(c, launch_dp(c->shapes.shape[rcl], local, trace,
                                            false, args,
                                            dp_grid(c, b - a, rcl, local, trace,
                                                    false),
                                            s));
                    }
                    HIPCHK(c, mark_flags(args.list, args.count, b - a,
                                         P<uint8_t>(c->tflags), 1, s));
                    // tasks re-run in int32 (pa_stats.rerun_pairs)
                    HIPCHK(c, accumulate(args.count,
                                         (unsigned long long *)P<uint64_t>(
                                             c->cells) + 2,
                                         s));
                    c->stats.launches += 3;
                }
                if (split) {
                    // this class's trace is complete: it is traced back on
                    // st2, launched after the next class's DP kernels (in a
                    // captured graph, whose kernels share one queue, the DP
                    // grid is then dispatched ahead of the walks)
                    HIPCHK(c, hipEventRecord(c->ev_fork[rcl], s));
                    if (rcl > 0) {
                        HIPCHK(c, hipStreamWaitEvent(c->st2,
                                                     c->ev_fork[rcl - 1], 0));
                        if ((rc = traceback(rcl - 1, c->st2)))
                            return rc;
                    }
                }
                if (beside && rcl == kClasses - 2)
                    HIPCHK(c, hipEventRecord(c->ev_fork[rcl], s));
            }
            if (beside) {
                // launched after the last class's DP kernels (so that a
                // captured graph dispatches that grid first)
                HIPCHK(c, hipStreamWaitEvent(c->st2,
                                             c->ev_fork[kClasses - 2], 0));
                for (int cls = 0; cls < kClasses - 1; ++cls)
                    if ((rc = traceback(cls, c->st2)))
                        return rc;
            }
            if (split) {
                HIPCHK(c, hipStreamWaitEvent(c->st2,
                                             c->ev_fork[kClasses - 1], 0));
                if ((rc = traceback(kClasses - 1, c->st2)))
                    return rc;
            }
        }
        if (split) {
            // the next chunk reuses the trace store and the bin counters, and
            // the caller reads the traceback results, all on s
            HIPCHK(c, hipEventRecord(c->ev_join, c->st2));
            HIPCHK(c, hipStreamWaitEvent(s, c->ev_join, 0));
        } else if (beside) {
            // the last class's walks after its DP; the others' join
            if ((rc = traceback(kClasses - 1, s)))
                return rc;
            HIPCHK(c, hipEventRecord(c->ev_join, c->st2));
            HIPCHK(c, hipStreamWaitEvent(s, c->ev_join, 0));
        } else if (trace) {
            // after all the DP: the few long walks of the last class on the
            // second stream, beside the many short ones of the others
            if (two) {
                HIPCHK(c, hipEventRecord(c->ev_fork[kClasses - 1], s));
                HIPCHK(c, hipStreamWaitEvent(c->st2,
                                             c->ev_fork[kClasses - 1], 0));
                if ((rc = traceback(kClasses - 1, c->st2)))
                    return rc;
            }
            for (int cls = 0; cls < (two ? kClasses - 1 : kClasses); ++cls)
                if ((rc = traceback(cls, s)))
                    return rc;
            if (two) {
                HIPCHK(c, hipEventRecord(c->ev_join, c->st2));
                HIPCHK(c, hipStreamWaitEvent(s, c->ev_join, 0));
            }
        }
    }
    return PA_OK;
}

int validate_pairs_host(pa_ctx *c, const pa_pair *pairs, uint64_t n)
{
    const int k = c->cfg.seed_len;
    for (uint64_t i = 0; i < n; ++i) {
        const pa_pair &p = pairs[i];
        if (p.r >= c->rlen.size() || p.c >= c->clen.size())
            return fail(c, PA_ERR_INDEX,
                        "pair " + std::to_string(i) + " names sequence (" +
                            std::to_string(p.r) + ", " + std::to_string(p.c) +
                            ") outside the loaded sets");
        if (c->cfg.mode == PA_MODE_XDROP) {
            const uint32_t lr = c->rlen[p.r], lc = c->clen[p.c];
            for (int s = 0; s < c->cfg.seed_cnt; ++s) {
                const uint32_t ro = p.seed[2 * s], co = p.seed[2 * s + 1];
                if (ro + (uint32_t)k > lr || co + (uint32_t)k > lc)
                    return fail(c, PA_ERR_SEED,
                                "pair " + std::to_string(i) + " seed " +
                                    std::to_string(s) +
                                    " does not fit in its sequences");
            }
        }
    }
    return PA_OK;
}

int check_ready(pa_ctx *c)
{
    if (!c)
        return PA_ERR_ARG;
    if (!c->have_rows || !c->have_cols)
        return fail(c, PA_ERR_NOSEQ, "load row and column sequences first");
    return PA_OK;
}

hipStream_t pick(pa_ctx *c, void *stream)
{
    return stream ? (hipStream_t)stream : c->st;
}

int align_sw_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                 pa_result *d_out, hipStream_t s)
{
    int rc;
    if ((rc = ensure_common(c, n)))
        return rc;
    Scope tot(c, 4, s);
    HIPCHK(c, build_sw_tasks(d_pairs, n, P<uint64_t>(c->roffs),
                             P<uint64_t>(c->coffs), c->rlen.size(),
                             c->clen.size(), P<DpTask>(c->tasks),
                             (unsigned long long *)P<uint64_t>(c->cells), s));
    if ((rc = run_dp(c, n, true, true, 0, s)))
        return rc;
    HIPCHK(c, finalize_sw(P<DpTask>(c->tasks), P<DpOut>(c->dpout),
                          P<uint8_t>(c->tflags), n, c->cfg.cov_thr,
                          c->cfg.ani_thr, d_out, s,
                          c->g_chunks > 0 ? gchunk_err(c) : nullptr));
    c->stats.pairs += n;
    return PA_OK;
}

int run_xdrop(pa_ctx *c, const pa_pair *d_pairs, uint64_t n, hipStream_t s)
{
    int rc;
    const int sc = c->cfg.seed_cnt;
    const uint64_t nt = n * (uint64_t)sc;   // extension tasks
    const uint64_t ni = 2 * nt;             // work items (two directions)
    if ((rc = grow(c, c->xdt, nt * sizeof(XdTask))))
        return rc;
    if ((rc = grow(c, c->xdo, 2 * n * sizeof(XdOut))))
        return rc;
    if ((rc = grow(c, c->ovf, 2 * ni * sizeof(uint32_t))))
        return rc;
    uint32_t *ctr = P<uint32_t>(c->ctr);
    uint32_t *ovf = P<uint32_t>(c->ovf), *lng = ovf + ni;
    HIPCHK(c, fill_u32(ctr + C_XDQ, 0, C_N - C_XDQ, s));
    HIPCHK(c, fill_u32(ctr + C_XDN, (uint32_t)ni, 1, s));
    HIPCHK(c, build_xd_tasks(d_pairs, n, sc, c->cfg.seed_len,
                             P<uint64_t>(c->roffs), P<uint64_t>(c->coffs),
                             c->rlen.size(), c->clen.size(),
                             P<XdTask>(c->xdt), s));
    {
        Scope t(c, 1, s);
        const int grid = xd_grid(c, ni);
        if (c->tune.xd_coop) {
            // 16 lanes per item first; items with a wider band go to the
            // one-lane-per-item kernel through the `lng` list
            const int cg = (int)std::max<uint64_t>(
                1, std::min<uint64_t>((ni + 3) / 4,
                                      (uint64_t)c->ncu *
                                          std::min(xdrop_coop_occupancy(),
                                                   kMaxWgPerCu / 4)));
            HIPCHK(c, launch_xdrop_coop(P<XdTask>(c->xdt), nullptr,
                                        ctr + C_XDN, ctr + C_XDQ,
                                        P<int8_t>(c->dmat), c->cfg.xdrop_gap,
                                        c->cfg.xdrop, c->cfg.seed_len,
                                        P<XdOut>(c->xdo), lng, ctr + C_XDLONG,
                                        cg, s));
            HIPCHK(c, launch_xdrop(P<XdTask>(c->xdt), lng, ctr + C_XDLONG,
                                   ctr + C_XDLONGQ, P<int8_t>(c->dmat),
                                   c->cfg.xdrop_gap, c->cfg.xdrop,
                                   c->cfg.seed_len, nullptr, 0,
                                   P<XdOut>(c->xdo), ovf, ctr + C_XDOVF, grid,
                                   c->tune.xd_ring, s));
            c->stats.launches += 2;
        } else {
            HIPCHK(c, launch_xdrop(P<XdTask>(c->xdt), nullptr, ctr + C_XDN,
                                   ctr + C_XDQ, P<int8_t>(c->dmat),
                                   c->cfg.xdrop_gap, c->cfg.xdrop,
                                   c->cfg.seed_len, nullptr, 0,
                                   P<XdOut>(c->xdo), ovf, ctr + C_XDOVF, grid,
                                   c->tune.xd_ring, s));
            c->stats.launches++;
        }
    }
    // items whose band outgrew the LDS rings: 16 lanes per item with a
    // 128-column window, what outgrows it with a 256-column window, the
    // rest with rings in global memory.  Each pass's overflow list is the
    // list the pass before last consumed (`lng` is free once the per-lane
    // kernel is done).
    uint32_t *last = ovf, *last_n = ctr + C_XDOVF, *last_q = ctr + C_XDOVFQ;
    // one band pass: items of `last` in, overflow to `to`
    auto band = [&](int wide, uint32_t *to, uint32_t *to_n,
                    uint32_t *to_q) -> int {
        Scope t(c, 1, s);
        HIPCHK(c, launch_xdrop_wave(P<XdTask>(c->xdt), last, last_n, last_q,
                                    P<int8_t>(c->dmat), c->cfg.xdrop_gap,
                                    c->cfg.xdrop, c->cfg.seed_len,
                                    P<XdOut>(c->xdo), to, to_n,
                                    c->ncu * xdrop_wave_occupancy(wide), wide,
                                    s));
        c->stats.launches++;
        last = to;
        last_n = to_n;
        last_q = to_q;
        return PA_OK;
    };
    if (c->tune.xd_wave == 1 &&
        (rc = band(0, lng, ctr + C_XDWOVF, ctr + C_XDWOVFQ)))
        return rc;
    if (c->tune.xd_wave &&
        (rc = band(1, last == ovf ? lng : ovf, ctr + C_XDW2OVF,
                   ctr + C_XDW2OVFQ)))
        return rc;
    if (c->g_chunks > 0) {
        // captured batch: no host synchronisation; the global-ring pass
        // always runs, on the kept scratch (its grid exits when its list
        // is empty)
        Scope t(c, 1, s);
        HIPCHK(c, launch_xdrop(P<XdTask>(c->xdt), last, last_n, last_q,
                               P<int8_t>(c->dmat), c->cfg.xdrop_gap,
                               c->cfg.xdrop, c->cfg.seed_len,
                               P<int32_t>(c->xdscr), c->xdscr_cap,
                               P<XdOut>(c->xdo), last == ovf ? lng : ovf,
                               ctr + C_XDDUMMY, kXdGraphLanes / 256, 0, s));
        c->stats.launches++;
        return PA_OK;
    }
    uint32_t novf = 0;
    HIPCHK(c, hipMemcpyAsync(&novf, last_n, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (novf) {
        // rings in global scratch holding a whole segment (power of two)
        int cap = 64;
        while (cap < (int)std::max(c->max_rlen, c->max_clen) + 4)
            cap <<= 1;
        // every launched lane may take an item, so the scratch covers whole
        // workgroups
        const size_t blocks2 =
            (std::min<size_t>(novf, kXdWideLanes) + 255) / 256;
        const size_t lanes2 = blocks2 * 256;
        Buf wide;
        if ((rc = grow(c, wide, lanes2 * 3 * (size_t)cap * sizeof(int32_t))))
            return rc;
        {
            Scope t(c, 1, s);
            // its overflow list (never written: the rings hold a whole
            // segment) is the list the previous kernel consumed
            HIPCHK(c, launch_xdrop(P<XdTask>(c->xdt), last, last_n, last_q,
                                   P<int8_t>(c->dmat), c->cfg.xdrop_gap,
                                   c->cfg.xdrop, c->cfg.seed_len,
                                   P<int32_t>(wide), cap, P<XdOut>(c->xdo),
                                   last == ovf ? lng : ovf, ctr + C_XDDUMMY,
                                   (int)blocks2, 0, s));
            c->stats.launches++;
        }
        HIPCHK(c, hipStreamSynchronize(s));
        release(wide);
    }
    return PA_OK;
}

int align_xd_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                 pa_result *d_out, hipStream_t s)
{
    int rc;
    const uint64_t nt = n * (uint64_t)c->cfg.seed_cnt;
    if ((rc = ensure_common(c, nt)))
        return rc;
    Scope tot(c, 4, s);
    if ((rc = run_xdrop(c, d_pairs, n, s)))
        return rc;
    const ResBound rb{(uint64_t)(uintptr_t)c->rcodes.p,
                      (const uint32_t *)((char *)c->rcodes.p + c->rpre_off)};
    const ResBound cbd{(uint64_t)(uintptr_t)c->ccodes.p,
                       (const uint32_t *)((char *)c->ccodes.p + c->cpre_off)};
    HIPCHK(c, build_nw_tasks(P<XdTask>(c->xdt), nt, P<XdOut>(c->xdo),
                             P<DpTask>(c->tasks), rb, cbd,
                             (unsigned long long *)P<uint64_t>(c->cells),
                             (unsigned long long *)(P<uint64_t>(c->cells) + 1),
                             s));
    if ((rc = run_dp(c, nt, false, true, 2, s)))
        return rc;
    HIPCHK(c, finalize_xd(P<XdTask>(c->xdt), P<XdOut>(c->xdo),
                          P<DpOut>(c->dpout), P<uint8_t>(c->tflags), n,
                          c->cfg.seed_cnt, c->cfg.cov_thr, c->cfg.ani_thr,
                          d_out, s,
                          c->g_chunks > 0 ? gchunk_err(c) : nullptr));
    c->stats.pairs += n;
    return PA_OK;
}

uint8_t encode_ascii(uint8_t ch)
{
    if (ch >= 'a' && ch <= 'z')
        ch = (uint8_t)(ch - 'a' + 'A');
    for (int i = 0; i < 25; ++i)
        if ((uint8_t)kAlphabet[i] == ch)
            return (uint8_t)i;
    return 22;  // X
}

int upload_set(pa_ctx *c, bool rows, const uint8_t *res, const uint64_t *offs,
               uint64_t nseq, int enc)
{
    int rc;
    const uint64_t base = offs[0];
    const uint64_t total = offs[nseq] - base;
    std::vector<uint32_t> lens(nseq);
    uint32_t mx = 0;
    for (uint64_t i = 0; i < nseq; ++i) {
        if (offs[i + 1] < offs[i])
            return fail(c, PA_ERR_ARG, "offsets must not decrease");
        const uint64_t L = offs[i + 1] - offs[i];
        if (L > 65535)
            return fail(c, PA_ERR_SEQ_LEN,
                        "sequence " + std::to_string(i) + " has " +
                            std::to_string(L) + " residues");
        lens[i] = (uint32_t)L;
        mx = std::max(mx, (uint32_t)L);
    }
    std::vector<uint8_t> codes(total + 64, (uint8_t)kSentinel);
    // k-mer letters: a residue upper-cased with & 0xDF must be one of the 25
    // letters (add_kmers, src/kmer/KmerOps.cpp:61-73); '*' never is, since
    // '*' & 0xDF is not a letter
    std::vector<uint8_t> kc(total + 64, 255);
    for (uint64_t i = 0; i < total; ++i) {
        uint8_t x = res[base + i];
        uint8_t k = 255;
        if (enc == PA_ENC_ASCII) {
            const uint8_t up = (uint8_t)(x & 0xDF);
            k = up < 128 ? encode_ascii(up) : 22;
            if (k == 22 && up != 'X')
                k = 255;
            if (k == 23)
                k = 255;
            x = encode_ascii(x);
        } else if (x >= 25) {
            return fail(c, PA_ERR_RESIDUE,
                        "residue code " + std::to_string(x) + " at offset " +
                            std::to_string(base + i));
        } else {
            k = x == 23 ? 255 : x;
        }
        codes[i] = x;
        kc[i] = k;
    }
    // Residue-bound prefix sums (ResBound) after the codes: the best score of
    // each code against any code, in either operand position, at least 0.
    uint32_t best[32] = {};
    for (int x = 0; x < kAlpha; ++x) {
        int m = 0;
        for (int y = 0; y < kAlpha; ++y)
            m = std::max(m, (int)std::max(c->mat[x * 32 + y], c->mat[y * 32 + x]));
        best[x] = (uint32_t)m;
    }
    const size_t pre_off = (codes.size() + 3) & ~(size_t)3;
    std::vector<uint32_t> pre(total + 1);
    uint32_t acc = 0;   // wraps around; span differences stay exact
    for (uint64_t i = 0; i < total; ++i) {
        pre[i] = acc;
        acc += best[codes[i]];
    }
    pre[total] = acc;
    Buf &cb = rows ? c->rcodes : c->ccodes;
    Buf &ob = rows ? c->roffs : c->coffs;
    Buf &kb = rows ? c->rkc : c->ckc;
    if ((rc = grow(c, cb, pre_off + pre.size() * sizeof(uint32_t))) ||
        (rc = grow(c, kb, kc.size())))
        return rc;
    HIPCHK(c, hipMemcpyAsync(kb.p, kc.data(), kc.size(), hipMemcpyHostToDevice,
                             c->st));
    (rows ? c->rpre_off : c->cpre_off) = pre_off;
    if ((rc = grow(c, ob, (nseq + 1) * sizeof(uint64_t))))
        return rc;
    std::vector<uint64_t> abs(nseq + 1);
    const uint64_t dev = (uint64_t)(uintptr_t)cb.p;
    for (uint64_t i = 0; i <= nseq; ++i)
        abs[i] = dev + (offs[i] - base);
    HIPCHK(c, hipMemcpyAsync(cb.p, codes.data(), codes.size(),
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync((char *)cb.p + pre_off, pre.data(),
                             pre.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(ob.p, abs.data(), abs.size() * 8,
                             hipMemcpyHostToDevice, c->st));
    Buf &sb = rows ? c->rsk : c->csk;
    if ((rc = grow(c, sb, std::max<uint64_t>(nseq, 1) * sizeof(SeqSketch))))
        return rc;
    HIPCHK(c, seq_sketch(P<uint64_t>(ob), nseq, P<SeqSketch>(sb), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (rows) {
        c->rlen.swap(lens);
        c->max_rlen = mx;
        c->have_rows = true;
    } else {
        c->clen.swap(lens);
        c->max_clen = mx;
        c->have_cols = true;
    }
    return PA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
namespace {
void destroy_streams(pa_ctx *c)
{
    for (auto &e : c->ev_fork)
        if (e)
            hipEventDestroy(e);
    if (c->ev_join)
        hipEventDestroy(c->ev_join);
    if (c->st2)
        hipStreamDestroy(c->st2);
    if (c->st)
        hipStreamDestroy(c->st);
}
}  // namespace

extern "C" {

const char *pa_strerror(int status)
{
    if (status < 0 || status >= (int)(sizeof(kStatus) / sizeof(kStatus[0])))
        return "unknown status";
    return kStatus[status];
}

const char *pa_last_error(pa_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "no context";
}

int pa_config_for_cli(pa_config *cfg, int gap_open_cli, int gap_ext_cli,
                      int klength, int xdrop, int seed_count, int gap_model)
{
    if (!cfg)
        return PA_ERR_ARG;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->mode = xdrop > 0 ? PA_MODE_XDROP : PA_MODE_SW;
    if (gap_model == PA_GAP_AS_CALLED) {
        // Score(gapExtend = -g, gapOpen = -e)
        cfg->gap_open = gap_ext_cli;
        cfg->gap_ext = gap_open_cli;
        cfg->xdrop_gap = gap_open_cli;
    } else if (gap_model == PA_GAP_AS_DOCUMENTED) {
        cfg->gap_open = gap_open_cli;
        cfg->gap_ext = gap_ext_cli;
        cfg->xdrop_gap = gap_ext_cli;
    } else {
        return PA_ERR_ARG;
    }
    cfg->xdrop = xdrop;
    cfg->seed_len = klength;
    cfg->seed_cnt = seed_count;
    cfg->ani_thr = 30;    // src/main.cpp:110
    cfg->cov_thr = 0.7;   // src/main.cpp:109
    cfg->device = 0;
    cfg->matrix = nullptr;
    return PA_OK;
}

int pa_create(const pa_config *cfg, pa_ctx **out)
{
    if (!cfg || !out)
        return PA_ERR_ARG;
    *out = nullptr;
    if (cfg->mode != PA_MODE_SW && cfg->mode != PA_MODE_XDROP)
        return PA_ERR_ARG;
    // the traced int16 kernels keep 4*score + tag in int16 and use -128 as
    // the below-everything profile entry
    if (cfg->gap_open > -1 || cfg->gap_open < -31 || cfg->gap_ext > -1 ||
        cfg->gap_ext < -31)
        return PA_ERR_ARG;
    if (cfg->mode == PA_MODE_XDROP &&
        (cfg->xdrop <= 0 || cfg->xdrop_gap > -1 || cfg->seed_len < 1 ||
         cfg->seed_cnt < 1 || cfg->seed_cnt > 2))
        return PA_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device ||
        cfg->device < 0)
        return PA_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return PA_ERR_NODEV;
    pa_ctx *c = new pa_ctx();
    c->cfg = *cfg;
    c->device = cfg->device;
    std::memset(c->mat, 0, sizeof(c->mat));
    for (int a = 0; a < 25; ++a)
        for (int b = 0; b < 25; ++b) {
            const int8_t v = cfg->matrix ? cfg->matrix[a * 32 + b]
                                         : kBlosum62[a * 25 + b];
            if (v < -32 || v > 31) {
                delete c;
                return PA_ERR_ARG;
            }
            c->mat[a * 32 + b] = v;
            c->maxs = std::max<int>(c->maxs, v);
        }
    c->cfg.matrix = nullptr;
    c->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    bool ok = hipSetDevice(c->device) == hipSuccess &&
              hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) ==
                  hipSuccess &&
              hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) ==
                  hipSuccess &&
              hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) ==
                  hipSuccess;
    for (int k = 0; ok && k < kClasses; ++k)
        ok = hipEventCreateWithFlags(&c->ev_fork[k], hipEventDisableTiming) ==
             hipSuccess;
    if (!ok) {
        destroy_streams(c);
        delete c;
        return PA_ERR_HIP;
    }
    // the plain table, then the DP kernels' profile matrices (kPmat*)
    std::vector<int8_t> dm(kPmatBytes, (int8_t)-128);
    std::memcpy(dm.data(), c->mat, sizeof(c->mat));
    for (int q = 0; q < kAlpha; ++q)
        for (int x = 0; x < kAlpha; ++x) {
            const int sc = c->mat[q * 32 + x];
            dm[kPmatUntraced + q * 32 + x] = (int8_t)sc;
            dm[kPmatLocal + q * 32 + x] = (int8_t)(4 * sc + 2);
            dm[kPmatGlobal + q * 32 + x] = (int8_t)(4 * sc + 2 + (q == x));
        }
    // 4-bit profiles: scores over the alphabet within 16 values of a
    // negative minimum; nibble 0 (the minimum) outside the alphabet
    int lo = 127, hi = -128;
    for (int q = 0; q < kAlpha; ++q)
        for (int x = 0; x < kAlpha; ++x) {
            lo = std::min(lo, (int)c->mat[q * 32 + x]);
            hi = std::max(hi, (int)c->mat[q * 32 + x]);
        }
    c->nib_ok = lo < 0 && hi - lo <= 15;
    c->nib_min = lo;
    for (int q = 0; q < 32; ++q)
        for (int x = 0; x < 32; ++x)
            dm[kPmatNib + q * 32 + x] =
                (int8_t)(c->nib_ok && q < kAlpha && x < kAlpha
                             ? c->mat[q * 32 + x] - lo
                             : 0);
    c->shapes = shapes_for(c, c->tune);
    if (grow(c, c->dmat, kPmatBytes) ||
        hipMemcpy(c->dmat.p, dm.data(), kPmatBytes, hipMemcpyHostToDevice) !=
            hipSuccess) {
        pa_destroy(c);
        return PA_ERR_NOMEM;
    }
    if (grow(c, c->bincells, 3 * kBins * sizeof(uint64_t)) ||
        hipMemset(c->bincells.p, 0, 3 * kBins * sizeof(uint64_t)) !=
            hipSuccess ||
        grow(c, c->cells, 4 * sizeof(uint64_t)) ||
        hipMemset(c->cells.p, 0, 4 * sizeof(uint64_t)) != hipSuccess) {
        pa_destroy(c);
        return PA_ERR_NOMEM;
    }
    noalloc_init(c->device);
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    // the trace store may take up to 70 % of the HBM free now (it is
    // allocated as batches need it, and halves if an allocation fails)
    c->trace_cap = (size_t)(fr * 0.7);
    if (const char *e = std::getenv("PA_TRACE_BYTES"))
        c->trace_cap = std::max<size_t>(std::strtoull(e, nullptr, 10), 1 << 20);
    *out = c;
    return PA_OK;
}

void pa_destroy(pa_ctx *c)
{
    if (!c)
        return;
    hipSetDevice(c->device);
    if (c->st)
        hipStreamSynchronize(c->st);
    resolve_timers(c);
    for (auto &e : c->graphs) {
        hipGraphExecDestroy(e.x);
        hipGraphDestroy(e.g);
    }
    c->graphs.clear();
    if (c->shared_cols) {
        c->ccodes = Buf();
        c->coffs = Buf();
        c->csk = Buf();
        c->ckc = Buf();
    }
    release(c->csk);
    filter_free(c->flt);
    c->flt = nullptr;
    kmer_free(c->kmer);
    c->kmer = nullptr;
    for (Buf *b : {&c->rcodes, &c->ccodes, &c->roffs, &c->coffs, &c->pairs,
                   &c->results, &c->scores, &c->tasks, &c->dpout, &c->tflags,
                   &c->sizes, &c->toffs, &c->lists, &c->sat, &c->ctr,
                   &c->trace, &c->bnd, &c->xdt, &c->xdo, &c->ovf, &c->rsk,
                   &c->rkc, &c->ckc, &c->tri_flag, &c->tri_pos,
                   &c->cub, &c->cells, &c->dmat, &c->ext, &c->bincells,
                   &c->gchunk, &c->xdscr})
        release(*b);
    for (auto e : c->free_events)
        hipEventDestroy(e);
    destroy_streams(c);
    delete c;
}

int pa_get_tuning(pa_ctx *c, pa_tuning *out)
{
    if (!c || !out)
        return PA_ERR_ARG;
    *out = c->tune;
    return PA_OK;
}

int pa_set_tuning(pa_ctx *c, const pa_tuning *t)
{
    if (!c || !t)
        return PA_ERR_ARG;
    if (t->dp_wg_per_cu < 0 || t->dp_wg_per_cu > kMaxWgPerCu ||
        t->xd_wg_per_cu < 0 || t->xd_wg_per_cu > kMaxWgPerCu ||
        (t->xd_ring != 16 && t->xd_ring != 32))
        return fail(c, PA_ERR_ARG, "tuning outside its range");
    if (t->xd_coop != 0 && t->xd_coop != 1)
        return fail(c, PA_ERR_ARG, "xd_coop must be 0 or 1");
    if (t->tb_stage != 0 && t->tb_stage != 1)
        return fail(c, PA_ERR_ARG, "tb_stage must be 0 or 1");
    if (t->xd_wave < 0 || t->xd_wave > 2)
        return fail(c, PA_ERR_ARG, "xd_wave must be 0, 1 or 2");
    if (t->tb_overlap < 0 || t->tb_overlap > 2)
        return fail(c, PA_ERR_ARG, "tb_overlap must be 0, 1 or 2");
    for (int k = 0; k < 3; ++k)
        if (t->dp_edge[k] < 1 || t->dp_edge[k] > 65535 ||
            (k > 0 && t->dp_edge[k] <= t->dp_edge[k - 1]))
            return fail(c, PA_ERR_ARG,
                        "dp_edge must increase strictly within 1..65535");
    for (int k = 0; k < 4; ++k)
        if (t->dp_shape[k] < 0 || t->dp_shape[k] >= kShapes)
            return fail(c, PA_ERR_ARG, "dp_shape outside pa_dp_shape");
    c->tune = *t;
    c->shapes = shapes_for(c, *t);
    ++c->gen;   // grids and class shapes of captured graphs follow the tuning
    return PA_OK;
}

int pa_set_thresholds(pa_ctx *c, double cov_thr, int ani_thr)
{
    if (!c)
        return PA_ERR_ARG;
    if (!(cov_thr >= 0.0 && cov_thr <= 1e9))
        return fail(c, PA_ERR_ARG, "coverage threshold must be >= 0");
    c->cfg.cov_thr = cov_thr;
    c->cfg.ani_thr = ani_thr;
    return PA_OK;
}

int pa_load_seqs(pa_ctx *c, int side, const uint8_t *residues,
                 const uint64_t *offsets, uint64_t nseq, int encoding)
{
    if (!c || !offsets || (nseq && !residues))
        return PA_ERR_ARG;
    if (side < PA_SIDE_ROW || side > PA_SIDE_BOTH ||
        (encoding != PA_ENC_CODES && encoding != PA_ENC_ASCII))
        return fail(c, PA_ERR_ARG, "bad side or encoding");
    hipSetDevice(c->device);
    // captured graphs hold the set sizes and the boundary-row stride
    ++c->gen;
    ++c->seq_gen;
    int rc;
    if (side == PA_SIDE_ROW || side == PA_SIDE_BOTH) {
        if (c->shared_cols) {
            // the column set aliased the row buffers; detach it first
            c->ccodes = Buf();
            c->coffs = Buf();
            c->csk = Buf();
            c->ckc = Buf();
            c->have_cols = false;
            c->shared_cols = false;
        }
        if ((rc = upload_set(c, true, residues, offsets, nseq, encoding)))
            return rc;
    }
    if (side == PA_SIDE_COL) {
        if (c->shared_cols) {
            c->ccodes = Buf();
            c->coffs = Buf();
            c->csk = Buf();
            c->ckc = Buf();
            c->shared_cols = false;
        }
        if ((rc = upload_set(c, false, residues, offsets, nseq, encoding)))
            return rc;
    }
    if (side == PA_SIDE_BOTH) {
        if (!c->shared_cols) {
            release(c->ccodes);
            release(c->coffs);
            release(c->csk);
            release(c->ckc);
        }
        c->ccodes = c->rcodes;
        c->coffs = c->roffs;
        c->csk = c->rsk;
        c->ckc = c->rkc;
        c->cpre_off = c->rpre_off;
        c->clen = c->rlen;
        c->max_clen = c->max_rlen;
        c->have_cols = true;
        c->shared_cols = true;
    }
    return PA_OK;
}

int pa_align_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                 pa_result *d_out, void *stream)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_out)
        return fail(c, PA_ERR_ARG, "null pointer");
    if (n > 0x7fffffffull / 2)
        return fail(c, PA_ERR_LIMIT, "batch larger than 2^30 pairs");
    hipSetDevice(c->device);
    hipStream_t s = pick(c, stream);
    return c->cfg.mode == PA_MODE_SW ? align_sw_dev(c, d_pairs, n, d_out, s)
                                     : align_xd_dev(c, d_pairs, n, d_out, s);
}

int pa_align(pa_ctx *c, const pa_pair *pairs, uint64_t n, pa_result *out)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!pairs || !out)
        return fail(c, PA_ERR_ARG, "null pointer");
    if ((rc = validate_pairs_host(c, pairs, n)))
        return rc;
    hipSetDevice(c->device);
    if ((rc = grow(c, c->pairs, n * sizeof(pa_pair))))
        return rc;
    if ((rc = grow(c, c->results, n * sizeof(pa_result))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->pairs.p, pairs, n * sizeof(pa_pair),
                             hipMemcpyHostToDevice, c->st));
    if ((rc = pa_align_dev(c, P<pa_pair>(c->pairs), n, P<pa_result>(c->results),
                           nullptr)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->results.p, n * sizeof(pa_result),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    return PA_OK;
}

}  // extern "C"

int pa::sw_scores_impl(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                       pa_score *d_out, hipStream_t s)
{
    int rc;
    if ((rc = ensure_common(c, n)))
        return rc;
    Scope tot(c, 4, s);
    HIPCHK(c, build_sw_tasks(d_pairs, n, P<uint64_t>(c->roffs),
                             P<uint64_t>(c->coffs), c->rlen.size(),
                             c->clen.size(), P<DpTask>(c->tasks),
                             (unsigned long long *)P<uint64_t>(c->cells), s));
    if ((rc = run_dp(c, n, true, false, 0, s)))
        return rc;
    HIPCHK(c, scores_out(P<DpOut>(c->dpout), n, d_out, s));
    c->stats.pairs += n;
    return PA_OK;
}

extern "C" {

int pa_sw_scores_dev(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                     pa_score *d_out, void *stream)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_out)
        return fail(c, PA_ERR_ARG, "null pointer");
    hipSetDevice(c->device);
    return sw_scores_impl(c, d_pairs, n, d_out, pick(c, stream));
}

int pa_sw_scores_graph(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                       pa_score *d_out, void *stream)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_out)
        return fail(c, PA_ERR_ARG, "null pointer");
    if (n > 0xffffffffull)
        return fail(c, PA_ERR_LIMIT, "more than 2^32 pairs");
    hipSetDevice(c->device);
    hipStream_t s = pick(c, stream);
    // graphs of an older workspace generation hold freed addresses
    auto &G = c->graphs;
    for (size_t i = 0; i < G.size();) {
        if (G[i].gen != c->gen) {
            hipGraphExecDestroy(G[i].x);
            hipGraphDestroy(G[i].g);
            G.erase(G.begin() + (long)i);
        } else {
            ++i;
        }
    }
    pa_ctx::GraphEntry *hit = nullptr;
    for (auto &e : G)
        if (e.kind == 0 && e.pairs == d_pairs && e.n == n &&
            e.out == d_out && e.s == s)
            hit = &e;
    if (!hit) {
        // every buffer the batch uses exists before the capture starts; a
        // reallocation here invalidates the other graphs
        if ((rc = ensure_common(c, n)) || (rc = ensure_bnd(c, true, false)))
            return rc;
        for (size_t i = 0; i < G.size();) {
            if (G[i].gen != c->gen || (G.size() >= (size_t)pa_ctx::kGraphs &&
                                       i == 0)) {
                hipGraphExecDestroy(G[i].x);
                hipGraphDestroy(G[i].g);
                G.erase(G.begin() + (long)i);
            } else {
                ++i;
            }
        }
        HIPCHK(c, hipStreamSynchronize(s));
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        rc = build_sw_tasks(d_pairs, n, P<uint64_t>(c->roffs),
                            P<uint64_t>(c->coffs), c->rlen.size(),
                            c->clen.size(), P<DpTask>(c->tasks),
                            (unsigned long long *)P<uint64_t>(c->cells), s) !=
                     hipSuccess
                 ? PA_ERR_HIP
                 : PA_OK;
        if (!rc)
            rc = run_dp(c, n, true, false, 0, s);
        if (!rc && scores_out(P<DpOut>(c->dpout), n, d_out, s) != hipSuccess)
            rc = PA_ERR_HIP;
        c->capturing = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc || e != hipSuccess) {
            if (g)
                hipGraphDestroy(g);
            return fail(c, rc ? rc : PA_ERR_HIP, "graph capture failed");
        }
        pa_ctx::GraphEntry ent;
        ent.g = g;
        if (hipGraphInstantiate(&ent.x, g, nullptr, nullptr, 0) !=
            hipSuccess) {
            hipGraphDestroy(g);
            return fail(c, PA_ERR_HIP, "graph instantiation failed");
        }
        ent.pairs = d_pairs;
        ent.n = n;
        ent.out = d_out;
        ent.s = s;
        ent.gen = c->gen;
        G.push_back(ent);
        hit = &G.back();
        c->stats.graph_captures++;
    }
    hit->used = ++c->graph_tick;
    HIPCHK(c, hipGraphLaunch(hit->x, s));
    // least recently used first, so that eviction takes entry 0
    std::sort(G.begin(), G.end(),
              [](const pa_ctx::GraphEntry &a, const pa_ctx::GraphEntry &b) {
                  return a.used < b.used;
              });
    c->stats.pairs += n;
    return PA_OK;
}

int pa_align_graph(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                   pa_result *d_out, void *stream)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_out)
        return fail(c, PA_ERR_ARG, "null pointer");
    if (n > 0x7fffffffull / 2)
        return fail(c, PA_ERR_LIMIT, "batch larger than 2^30 pairs");
    hipSetDevice(c->device);
    hipStream_t s = pick(c, stream);
    auto &G = c->graphs;
    auto drop = [&](size_t i) {
        hipGraphExecDestroy(G[i].x);
        hipGraphDestroy(G[i].g);
        G.erase(G.begin() + (long)i);
    };
    for (size_t i = 0; i < G.size();) {
        if (G[i].gen != c->gen)
            drop(i);
        else
            ++i;
    }
    pa_ctx::GraphEntry *hit = nullptr;
    for (auto &e : G)
        if (e.kind == 1 && e.pairs == d_pairs && e.n == n && e.out == d_out &&
            e.s == s)
            hit = &e;
    if (!hit) {
        // 1. the batch eagerly: results for this call, every buffer at its
        //    size, and the batch's trace size
        if ((rc = pa_align_dev(c, d_pairs, n, d_out, stream)))
            return rc;
        const uint64_t nt = c->cfg.mode == PA_MODE_SW
                                ? n
                                : n * (uint64_t)c->cfg.seed_cnt;
        uint64_t total = 0;
        HIPCHK(c, hipMemcpyAsync(&total, P<uint64_t>(c->toffs) + nt, 8,
                                 hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        // 2. the captured layout: a store of twice this batch's trace
        //    (within the context's trace memory), split into as many chunks
        //    as this batch needs plus one
        uint64_t cap = std::max<uint64_t>(
            64, std::min<uint64_t>(c->trace_cap, 2 * total + (1u << 20)));
        if ((rc = grow_trace(c, cap)))
            return rc;
        cap = std::min<uint64_t>(cap, c->trace_cap);   // if the store shrank
        const int chunks = (int)std::min<uint64_t>(
            64, (total + cap - 1) / cap + 1);
        if ((rc = grow(c, c->gchunk, (size_t)chunks * 16 + 16)))
            return rc;
        if (c->cfg.mode == PA_MODE_XDROP) {
            int rcap = 64;
            while (rcap < (int)std::max(c->max_rlen, c->max_clen) + 4)
                rcap <<= 1;
            if ((rc = grow(c, c->xdscr, (size_t)kXdGraphLanes * 3 *
                                            (size_t)rcap * sizeof(int32_t))))
                return rc;
            c->xdscr_cap = rcap;
        }
        c->g_cap = cap;
        for (size_t i = 0; i < G.size();) {
            if (G[i].gen != c->gen ||
                (G.size() >= (size_t)pa_ctx::kGraphs && i == 0))
                drop(i);
            else
                ++i;
        }
        // 3. capture
        HIPCHK(c, hipStreamSynchronize(s));
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        c->g_chunks = chunks;
        // the chunk overflow flag starts at zero on every replay
        rc = fill_u32(gchunk_err(c), 0, 1, s) != hipSuccess ? PA_ERR_HIP
                                                            : PA_OK;
        const uint64_t gen0 = c->gen;
        if (!rc)
            rc = c->cfg.mode == PA_MODE_SW
                     ? align_sw_dev(c, d_pairs, n, d_out, s)
                     : align_xd_dev(c, d_pairs, n, d_out, s);
        c->g_chunks = 0;
        c->capturing = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (!rc && c->gen != gen0)
            rc = fail(c, PA_ERR_HIP, "workspace reallocated during capture");
        if (rc || e != hipSuccess) {
            if (g)
                hipGraphDestroy(g);
            return rc ? rc : fail(c, PA_ERR_HIP, "graph capture failed");
        }
        pa_ctx::GraphEntry ent;
        ent.g = g;
        if (hipGraphInstantiate(&ent.x, g, nullptr, nullptr, 0) !=
            hipSuccess) {
            hipGraphDestroy(g);
            return fail(c, PA_ERR_HIP, "graph instantiation failed");
        }
        ent.pairs = d_pairs;
        ent.n = n;
        ent.out = d_out;
        ent.s = s;
        ent.gen = c->gen;
        ent.kind = 1;
        G.push_back(ent);
        c->stats.graph_captures++;
        std::sort(G.begin(), G.end(),
                  [](const pa_ctx::GraphEntry &a,
                     const pa_ctx::GraphEntry &b) { return a.used < b.used; });
        // this call's results came from the eager run
        for (auto &x : G)
            if (x.kind == 1 && x.pairs == d_pairs && x.n == n &&
                x.out == d_out && x.s == s)
                x.used = ++c->graph_tick;
        return PA_OK;
    }
    hit->used = ++c->graph_tick;
    HIPCHK(c, hipGraphLaunch(hit->x, s));
    std::sort(G.begin(), G.end(),
              [](const pa_ctx::GraphEntry &a, const pa_ctx::GraphEntry &b) {
                  return a.used < b.used;
              });
    c->stats.pairs += n;
    return PA_OK;
}

int pa_sw_scores(pa_ctx *c, const pa_pair *pairs, uint64_t n, pa_score *out)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (n == 0)
        return PA_OK;
    if (!pairs || !out)
        return fail(c, PA_ERR_ARG, "null pointer");
    for (uint64_t i = 0; i < n; ++i)
        if (pairs[i].r >= c->rlen.size() || pairs[i].c >= c->clen.size())
            return fail(c, PA_ERR_INDEX,
                        "pair " + std::to_string(i) + " outside the loaded sets");
    hipSetDevice(c->device);
    if ((rc = grow(c, c->pairs, n * sizeof(pa_pair))))
        return rc;
    if ((rc = grow(c, c->scores, n * sizeof(pa_score))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->pairs.p, pairs, n * sizeof(pa_pair),
                             hipMemcpyHostToDevice, c->st));
    if ((rc = pa_sw_scores_dev(c, P<pa_pair>(c->pairs), n,
                               P<pa_score>(c->scores), nullptr)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->scores.p, n * sizeof(pa_score),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    return PA_OK;
}

int pa_xdrop_extents(pa_ctx *c, const pa_pair *pairs, uint64_t n, int seed,
                     pa_extent *out)
{
    int rc;
    if ((rc = check_ready(c)))
        return rc;
    if (c->cfg.mode != PA_MODE_XDROP)
        return fail(c, PA_ERR_ARG, "context is not in X-drop mode");
    if (seed < 0 || seed >= c->cfg.seed_cnt)
        return fail(c, PA_ERR_ARG, "seed index");
    if (n == 0)
        return PA_OK;
    if (!pairs || !out)
        return fail(c, PA_ERR_ARG, "null pointer");
    if ((rc = validate_pairs_host(c, pairs, n)))
        return rc;
    hipSetDevice(c->device);
    if ((rc = ensure_common(c, n * 2)))
        return rc;
    if ((rc = grow(c, c->pairs, n * sizeof(pa_pair))))
        return rc;
    if ((rc = grow(c, c->ext, n * sizeof(pa_extent))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->pairs.p, pairs, n * sizeof(pa_pair),
                             hipMemcpyHostToDevice, c->st));
    if ((rc = run_xdrop(c, P<pa_pair>(c->pairs), n, c->st)))
        return rc;
    HIPCHK(c, extents_out(P<XdOut>(c->xdo), n, seed, P<pa_extent>(c->ext),
                          c->st));
    HIPCHK(c, hipMemcpyAsync(out, c->ext.p, n * sizeof(pa_extent),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    return PA_OK;
}

int pa_kmer_prune(const pa_pair *pairs, uint64_t n, int ckthr, float mosthr,
                  uint8_t *keep, uint64_t *kept)
{
    if ((n && (!pairs || !keep)))
        return PA_ERR_ARG;
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; ++i) {
        // DistPWRunner.cpp:48-54: prune if count <= ckthr || score <= mosthr
        const bool prune = (int)pairs[i].count <= ckthr ||
                           (float)pairs[i].kscore <= mosthr;
        keep[i] = prune ? 0 : 1;
        k += !prune;
    }
    if (kept)
        *kept = k;
    return PA_OK;
}

int pa_device_alloc(pa_ctx *c, size_t bytes, void **ptr)
{
    if (!c || !ptr)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess)
        return fail(c, PA_ERR_NOMEM, "pa_device_alloc");
    return PA_OK;
}

int pa_device_free(pa_ctx *c, void *ptr)
{
    if (!c)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipFree(ptr));
    return PA_OK;
}

int pa_host_alloc(pa_ctx *c, size_t bytes, void **ptr)
{
    if (!c || !ptr)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) !=
        hipSuccess) {
        *ptr = nullptr;
        return fail(c, PA_ERR_NOMEM, "pa_host_alloc");
    }
    return PA_OK;
}

int pa_host_free(pa_ctx *c, void *ptr)
{
    if (!c)
        return PA_ERR_ARG;
    HIPCHK(c, hipHostFree(ptr));
    return PA_OK;
}

int pa_memcpy_h2d(pa_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PA_OK;
}

int pa_memcpy_d2h(pa_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PA_OK;
}

int pa_synchronize(pa_ctx *c)
{
    if (!c)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    return PA_OK;
}

int pa_get_stats(pa_ctx *c, pa_stats *out)
{
    if (!c || !out)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    if (c->cells.p) {
        uint64_t cells[3] = {0, 0, 0};
        HIPCHK(c, hipMemcpy(cells, c->cells.p, 24, hipMemcpyDeviceToHost));
        c->stats.dp_cells = cells[0];
        c->stats.xdrop_cells = cells[1];
        c->stats.rerun_pairs = cells[2];
    }
    *out = c->stats;
    return PA_OK;
}

int pa_reset_stats(pa_ctx *c)
{
    if (!c)
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    c->stats = pa_stats{};
    for (auto &m : c->kstat)
        for (auto &k : m)
            k = pa_ctx::KStat{};
    if (c->cells.p)
        HIPCHK(c, hipMemset(c->cells.p, 0, 32));
    HIPCHK(c, hipMemset(c->bincells.p, 0, 3 * kBins * sizeof(uint64_t)));
    return PA_OK;
}

int pa_get_kernel_stats(pa_ctx *c, pa_kernel_stat *out, int cap, int *n)
{
    if (!c || !n || (cap > 0 && !out))
        return PA_ERR_ARG;
    hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->st));
    resolve_timers(c);
    uint64_t cells[3 * kBins];
    HIPCHK(c, hipMemcpy(cells, c->bincells.p, sizeof(cells),
                        hipMemcpyDeviceToHost));
    static const char *mname[3] = {"sw", "swt", "nwt"};
    int k = 0;
    for (int m = 0; m < 3; ++m)
        for (int b = 0; b < kBins + kClasses; ++b) {
            const auto &ks = c->kstat[m][b];
            if (!ks.launches)
                continue;
            if (k < cap) {
                pa_kernel_stat &o = out[k];
                std::memset(&o, 0, sizeof(o));
                // mode, row class, the class's current kernel shape
                // (lanes x rows per lane), arithmetic
                static const char *shname[kShapes] = {
                    "l16r4", "l16r8", "l64r4", "l64r8", "l32r8", "l8r8"};
                const int cls = b < kBins ? b / 2 : b - kBins;
                const char *sh = shname[c->shapes.shape[cls]];
                if (b < kBins)
                    std::snprintf(o.name, sizeof(o.name), "%s.c%d.%s.%s",
                                  mname[m], cls, sh, b % 2 ? "i32" : "i16");
                else
                    std::snprintf(o.name, sizeof(o.name), "%s.c%d.%s.rerun",
                                  mname[m], cls, sh);
                o.launches = ks.launches;
                o.ms = ks.ms;
                o.cells = b < kBins ? cells[m * kBins + b] : 0;
            }
            ++k;
        }
    *n = k;
    return PA_OK;
}

}  // extern "C"
