// k_xdrop_wave.hip -- gapped X-drop seed extension with one 16-lane row of a
// wave per work item (four items per wave), for the items whose band
// outgrows the per-lane kernel's LDS rings (k_xdrop.hip).  Same algorithm,
// step for step, as k_xdrop_pl and the oracle's or_xdrop_one (SeqAn's
// one-direction GappedXDrop as called at
// src/align/SeqanXdropAligner.cpp:281-283), so extents and evaluated-cell
// counts are bit-identical.
//
// The cells of one anti-diagonal depend only on the two before it, so the 16
// lanes of an item evaluate 16 consecutive columns at once.  The item's three
// live anti-diagonals stay in registers: lane j holds, for each of kChunks
// chunks c, column base + 16c + j of each (a window of kWin columns starting
// at `base`).  The value of column col - 1 comes from the neighbouring lane
// through DPP row_shr:1, whose first lane keeps the DPP "old" operand, lane 15
// of the chunk below rotated in with row_ror:1.  The window moves up a chunk
// when the oldest anti-diagonal's range has left its first chunk; an item
// whose band outgrows the window is handed on to the global-ring kernel
// through the overflow list.  The anti-diagonal's best score is a row
// maximum (row_ror 1, 2, 4, 8) and the band update two ballots (the first
// column from the left and the last from the right that stay live).  Query
// residues sit in registers per column; database residues are loaded two
// anti-diagonals ahead and turned into substitution scores one ahead, so no
// step waits on memory.  A row that finishes its item takes the next one
// from a range its wave claims with one queue atomic.
#include <limits.h>

#include "pa_internal.h"

namespace pa {
namespace {

constexpr int32_t kUndef = INT_MIN / 4;
constexpr int kW = 16;                // lanes per item
constexpr uint32_t kClaim = 4;        // items claimed per queue atomic
// window of the main pass (128 columns) and of the pass that takes its
// overflow (256 columns; few items, so its register use does not matter)
constexpr int kBandChunks = 8;
constexpr int kWideChunks = 16;

struct WaveArgs {
    const XdTask *tasks;
    const uint32_t *list;   // work items (k = 2 * task + dir)
    const uint32_t *count;
    uint32_t *queue;
    const int8_t *mat;
    int gap, X, seed_len;
    XdOut *out;
    uint32_t *ovf_list, *ovf_count;
};

// lane j of each 16-lane row reads lane j - K (mod 16) of its row (every
// lane has a source, so no "old" operand: the compiler can fold the move
// into the instruction that uses it)
template <int K> __device__ inline int32_t ror(int32_t v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x120 + K, 0xf, 0xf, false);
}

// Column col - 1 of the lanes' columns of a chunk: lanes 1..15 take lane
// j - 1 of `cur`, lane 0 keeps the DPP "old" operand, lane 15 of `below`.
__device__ inline int32_t left_of(int32_t cur, int32_t below)
{
    return __builtin_amdgcn_update_dpp(ror<1>(below), cur, 0x111, 0xf, 0xf,
                                       false);
}

__device__ inline int32_t row_max(int32_t v)
{
    int32_t o = ror<1>(v);
    v = o > v ? o : v;
    o = ror<2>(v);
    v = o > v ? o : v;
    o = ror<4>(v);
    v = o > v ? o : v;
    o = ror<8>(v);
    return o > v ? o : v;
}

__device__ inline int32_t row_min(int32_t v)
{
    int32_t o = ror<1>(v);
    v = o < v ? o : v;
    o = ror<2>(v);
    v = o < v ? o : v;
    o = ror<4>(v);
    v = o < v ? o : v;
    o = ror<8>(v);
    return o < v ? o : v;
}

// maximum over the wave of a non-negative value
__device__ inline int wave_max(int v)
{
    const int r = row_max(v);
    const int a = __builtin_amdgcn_readlane(r, 0);
    const int b = __builtin_amdgcn_readlane(r, 16);
    const int c = __builtin_amdgcn_readlane(r, 32);
    const int d = __builtin_amdgcn_readlane(r, 48);
    const int ab = a > b ? a : b, cd = c > d ? c : d;
    return ab > cd ? ab : cd;
}

// Writes the half of XdOut that item k (direction k & 1) owns.
__device__ inline void wave_write(const WaveArgs &a, uint32_t k, int32_t score,
                                  int col, int row, bool upd, uint32_t nc,
                                  uint16_t flags)
{
    const XdTask T = a.tasks[k >> 1];
    XdOut PA_GLOBAL &o = glb(a.out)[T.id];
    const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
    if ((k & 1) == 0) {
        o.beg_h = upd ? beg_h - row : beg_h;
        o.beg_v = upd ? beg_v - col : beg_v;
        o.score_left = score;
        o.cells_left = nc;
        o.flags_left = flags;
    } else {
        const int end_h = beg_h + a.seed_len, end_v = beg_v + a.seed_len;
        o.end_h = upd ? end_h + row : end_h;
        o.end_v = upd ? end_v + col : end_v;
        o.score_right = score;
        o.cells_right = nc;
        o.flags_right = flags;
    }
}

enum : int { kIdle = 0, kRun = 1, kDone = 2 };

// NCH chunks of 16 columns per lane: a window of 16 * NCH columns.
template <int NCH>
__global__ __launch_bounds__(64) void k_xdrop_band(WaveArgs a)
{
    constexpr int kChunks = NCH;
    constexpr int kWin = kW * kChunks;
    __shared__ int8_t mat[32 * 32];
    const int lane = threadIdx.x, j = lane & (kW - 1), g = lane / kW;
    for (int i = lane; i < 32 * 32; i += 64)
        mat[i] = a.mat[i];
    __syncthreads();
    const uint32_t n = *a.count;
    const int gap = a.gap, X = a.X;

    // Item of this lane's row; the same in all 16 lanes of the row.
    int state = kIdle;
    uint32_t k = 0;
    const uint8_t PA_GLOBAL *qb = nullptr;
    const uint8_t PA_GLOBAL *db = nullptr;
    int qstep = 1, dstep = 1, cols = 2, rows = 2;
    int d = 0, minCol = 0, maxCol = 0, best = 0, base = 0;
    int off1 = 0, off2 = 0, off3 = 0, l1 = 0, l2 = 0, l3 = 0;
    uint32_t nc = 0;
    // this lane's column base + kW*c + j of anti-diagonals d-2, d-1, d
    int32_t A1[kChunks], A2[kChunks], A3[kChunks];
    int q[kChunks];             // query residue of the column
    int32_t sc[kChunks];        // substitution score of the next step's cell
    int B0[kChunks], B1[kChunks];   // database residues, loaded ahead
    // column - 1 of A2 in this lane's columns, saved by each step: the next
    // step's column - 1 of A1 (its diagonal input); kept by the 128-column
    // pass only (the 256-column pass is at its register limit)
    constexpr bool kKeepDL = NCH <= 8;
    int32_t DL[kKeepDL ? kChunks : 1];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        A1[c] = A2[c] = A3[c] = kUndef;
        q[c] = 0;
        sc[c] = 0;
        B0[c] = B1[c] = 0;
        if (kKeepDL)
            DL[c] = kUndef;
    }
    // DL does not hold this row's values (new item, or the window moved)
    bool fresh = true;
    int cw_prev = -1;   // chunks the previous step computed
    uint32_t cbase = 0, cend = 0;   // claimed, not yet handed out
    // highest chunk any running row reaches in the coming step (set at the
    // end of each step; a new item starts in chunk 0)
    int cw_next = 0;

    auto qres = [&](int col) {   // query residue of column col (clamped)
        col = col < 1 ? 1 : (col > cols - 1 ? cols - 1 : col);
        return (int)qb[qstep > 0 ? col : -col];
    };
    auto dres = [&](int row) {   // database residue of row (clamped)
        row = row < 1 ? 1 : (row > rows - 1 ? rows - 1 : row);
        return (int)db[dstep > 0 ? row : -row];
    };
    // group value of column x of anti-diagonal array A (x in the window)
    auto at = [&](const int32_t *A, int x) {
        int32_t v = INT_MIN;
#pragma unroll
        for (int c = 0; c < kChunks; ++c)
            v = base + kW * c + j == x ? A[c] : v;
        return row_max(v);
    };

    // One half of the loop: claims for idle rows, then one anti-diagonal of
    // every running item.  LOAD receives the database residues of the step
    // after next, LOOK holds those of the next step (loaded a half before).
    auto half = [&](int *LOAD, int *LOOK) {
        // ---- claims (wave-uniform block; every lane runs it) -----------
        uint64_t need = __ballot(state == kIdle && j == 0);
        while (need) {
            if (cbase >= cend) {
                uint32_t b = atomicAdd(a.queue, lane == 0 ? kClaim : 0u);
                b = __builtin_amdgcn_readfirstlane(b);
                cbase = b;
                cend = b + kClaim;
            }
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t take = min(cnt, cend - cbase);
            // rank of this row among the rows asking
            const uint64_t mine = 1ull << (g * kW);
            const uint32_t rank = (uint32_t)__popcll(need & (mine - 1ull));
            const bool asks = (need & mine) != 0;
            if (asks && rank < take) {
                const uint32_t idx = cbase + rank;
                if (idx >= n) {
                    state = kDone;
                } else {
                    k = glb(a.list)[idx];
                    const XdTask T = a.tasks[k >> 1];
                    const uint8_t PA_GLOBAL *h = glb_addr<const uint8_t>(T.h_off);
                    const uint8_t PA_GLOBAL *v = glb_addr<const uint8_t>(T.v_off);
                    const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
                    if ((k & 1) == 0) {
                        cols = beg_v + 1;
                        rows = beg_h + 1;
                        qb = v + (cols - 1);
                        qstep = -1;
                        db = h + (rows - 1);
                        dstep = -1;
                    } else {
                        cols = (int)T.nv - beg_v - a.seed_len + 1;
                        rows = (int)T.nh - beg_h - a.seed_len + 1;
                        qb = v + beg_v + a.seed_len - 1;
                        qstep = 1;
                        db = h + beg_h + a.seed_len - 1;
                        dstep = 1;
                    }
                    // item_init: anti-diagonals 0 (A2) and 1 (A3); the
                    // step rotates them into A1 and A2
                    d = 1;
                    minCol = 1;
                    maxCol = 2;
                    best = 0;
                    base = 0;
                    off1 = off2 = off3 = 0;
                    l1 = 0;
                    l2 = 1;
                    l3 = 2;
                    nc = 0;
                    const int32_t g1 = -gap > X ? kUndef : gap;
#pragma unroll
                    for (int c = 0; c < kChunks; ++c) {
                        const int col = kW * c + j;
                        A2[c] = col == 0 ? 0 : kUndef;
                        A3[c] = col <= 1 ? g1 : kUndef;
                        q[c] = qres(col);
                        // scores of step 2 (rows 2 - col) and residues of
                        // step 3 (rows 3 - col)
                        sc[c] = mat[q[c] * 32 + dres(2 - col)];
                        LOOK[c] = dres(3 - col);
                    }
                    fresh = true;
                    state = kRun;
                }
            }
            cbase += take;
            for (uint32_t t = 0; t < take; ++t)
                need &= need - 1;
        }

        // ---- one anti-diagonal of each running item ---------------------
        // Every lane runs this part (the DPP moves, ballots and row maxima
        // need whole rows, and the emulation of the tests needs the whole
        // wave): rows without an item compute on stale registers, which the
        // next claim resets, and only memory accesses and state changes are
        // guarded by `run`.
        bool run = state == kRun;
        ++d;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            A1[c] = A2[c];
            A2[c] = A3[c];
        }
        l1 = l2;
        l2 = l3;
        off1 = off2;
        off2 = off3;
        off3 = minCol - 1;
        l3 = maxCol + 1 - off3;
        const int maxCol0 = maxCol;
        if (run && maxCol0 - base >= kWin) {
            // the band left the window: restart in the global-ring kernel
            if (j == 0) {
                const uint32_t pos = atomicAdd(a.ovf_count, 1u);
                a.ovf_list[pos] = k;
                wave_write(a, k, 0, 0, 0, false, 0, 1);
            }
            state = kIdle;
            run = false;
        }
        const bool cond = (long)d * gap > (long)best - X;
        const int32_t lim = best - X;
        const int32_t dgap = d * gap;
        int32_t adbest = dgap;
        // the chunks any running row reads or writes this step: columns up
        // to its maxCol0 (the others hold columns no step reads)
        const int cw = cw_next;
        // The diagonal input (column - 1 of A1) is the left input the
        // previous step computed (column - 1 of its A2), unless a running
        // row started an item or moved its window since, or the chunk is
        // new this step.
        const bool stale = __ballot(run && fresh) != 0;
        // The new minCol is the first column c in [minCol, maxCol0] that
        // does not satisfy (A3[c] undefined, c - 1 in A2's range, A2[c-1]
        // undefined), or maxCol0 + 1 when all do; the new maxCol is one past
        // the last column x in [off3, maxCol0) with A3[x] or A2[x] defined
        // (off3 when none), plus one.  Each lane keeps its own candidates;
        // one row minimum and maximum per step combine them.  x = col -
        // minCol turns the ranges into single unsigned comparisons.
        const int w = maxCol0 - minCol;
        const int ecol0 = cond && off3 == 0 ? 0 : -1;   // edge columns
        const int ecol1 = cond && d == maxCol0 ? maxCol0 : -1;
        const int dlim = off2 + 1 + l2;   // c - 1 in A2's range: c < dlim
        int32_t lo = INT_MAX, hi = INT_MIN;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            if (c > cw)
                break;
            const int col = base + kW * c + j;
            const int x = col - minCol;
            const int32_t u = A2[c];
            const int32_t l = left_of(u, c ? A2[c - 1] : kUndef);
            int32_t dl;
            if (!kKeepDL || stale || c > cw_prev)
                dl = left_of(A1[c], c ? A1[c - 1] : kUndef);
            else
                dl = DL[c];
            if constexpr (kKeepDL)
                DL[c] = l;
            int32_t tmp = (l > u ? l : u) + gap;
            const int32_t dg = dl + sc[c];
            tmp = dg > tmp ? dg : tmp;
            const bool inb = (uint32_t)x < (uint32_t)w;
            const int32_t val = tmp < lim ? kUndef : tmp;
            const bool edge = col == ecol0 || col == ecol1;
            const int32_t a3 = inb ? val : (edge ? dgap : kUndef);
            A3[c] = a3;
            adbest = inb && val > adbest ? val : adbest;
            const bool u3 = a3 == kUndef;
            if ((uint32_t)x <= (uint32_t)w && !(u3 && col < dlim && l == kUndef))
                lo = col < lo ? col : lo;
            if ((uint32_t)(x + 1) < (uint32_t)(w + 1) && !(u3 && u == kUndef))
                hi = col + 1;
        }
        fresh = false;
        cw_prev = cw;
        nc += (uint32_t)(maxCol0 - minCol);
        adbest = row_max(adbest);
        best = adbest > best ? adbest : best;
        lo = row_min(lo);
        hi = row_max(hi);
        int nmin = lo == INT_MAX ? maxCol0 + 1 : lo;
        int nmax = hi == INT_MIN ? off3 : hi;
        ++nmax;
        if (nmin < d + 2 - rows)
            nmin = d + 2 - rows;
        if (nmax > cols)
            nmax = cols;
        minCol = nmin;
        maxCol = nmax;
        const bool ended = run && minCol >= maxCol;
        if (__ballot(ended)) {
            // item_end: the "longest extension" end point
            int lcol = l3 + off3 - 2;
            int lrow = d - lcol;
            int32_t ls = at(A3, lcol);
            const int32_t v2 = at(A2, l2 + off2 - 2);
            const int32_t v3 = at(A2, l2 + off2 - 3);
            if (ls == kUndef) {
                if (v2 != kUndef) {
                    lcol = l2 + off2 - 2;
                    lrow = d - 1 - lcol;
                    ls = v2;
                } else if (l2 > 2 && v3 != kUndef) {
                    lcol = l2 + off2 - 3;
                    lrow = d - 1 - lcol;
                    ls = v3;
                }
            }
            // the first maximum above kUndef of anti-diagonal d - 2: the
            // largest value, and among equal values the smallest column
            int32_t bv = kUndef;
            int bc = INT_MAX;
#pragma unroll
            for (int c = 0; c < kChunks; ++c) {
                const int col = base + kW * c + j;
                if (col >= off1 && col < off1 + l1 && A1[c] > bv) {
                    bv = A1[c];
                    bc = col;
                }
            }
            // row reduction of (value, column)
            int32_t ov = ror<1>(bv), oc = ror<1>(bc);
            if (ov > bv || (ov == bv && oc < bc)) {
                bv = ov;
                bc = oc;
            }
            ov = ror<2>(bv);
            oc = ror<2>(bc);
            if (ov > bv || (ov == bv && oc < bc)) {
                bv = ov;
                bc = oc;
            }
            ov = ror<4>(bv);
            oc = ror<4>(bc);
            if (ov > bv || (ov == bv && oc < bc)) {
                bv = ov;
                bc = oc;
            }
            ov = ror<8>(bv);
            oc = ror<8>(bc);
            if (ov > bv || (ov == bv && oc < bc)) {
                bv = ov;
                bc = oc;
            }
            if (ls == kUndef && bv > kUndef) {
                ls = bv;
                lcol = bc;
                lrow = d - 2 - lcol;
            }
            if (ended && j == 0)
                wave_write(a, k, ls, lcol, lrow, ls != kUndef, nc, 0);
        }
        if (ended) {
            state = kIdle;
            run = false;
        }
        // move the window up while its first chunk is below the oldest
        // range the next steps read (A1 of the next step starts at off2)
        while (run && base + kW <= off2) {
#pragma unroll
            for (int c = 0; c < kChunks - 1; ++c) {
                A1[c] = A1[c + 1];
                A2[c] = A2[c + 1];
                A3[c] = A3[c + 1];
                q[c] = q[c + 1];
                LOOK[c] = LOOK[c + 1];
            }
            base += kW;
            fresh = true;
            const int col = base + kW * (kChunks - 1) + j;
            A1[kChunks - 1] = A2[kChunks - 1] = A3[kChunks - 1] = kUndef;
            q[kChunks - 1] = qres(col);
            LOOK[kChunks - 1] = dres(d + 1 - col);
        }
        // scores of the next step from the residues loaded a half ago, and
        // the residues of the step after, for the chunks the next steps can
        // reach (the band's top moves up at most one column per step)
        // (the next step computes chunks up to the reach, the one after at
        // most one chunk more: its band's top is at most two columns higher)
        const int reach = wave_max(run ? (maxCol - base) / kW : 0);
        cw_next = min(kChunks - 1, reach);
        const int pw = min(kChunks - 1, reach);
        if (run) {
#pragma unroll
            for (int c = 0; c < kChunks; ++c) {
                if (c > pw + 1)
                    break;
                const int col = base + kW * c + j;
                if (c <= pw)
                    sc[c] = mat[q[c] * 32 + LOOK[c]];
                LOAD[c] = dres(d + 2 - col);
            }
        }
    };

    for (;;) {
        half(B0, B1);
        if (!__ballot(state != kDone))
            break;
        half(B1, B0);
        if (!__ballot(state != kDone))
            break;
    }
}

}  // namespace

hipError_t launch_xdrop_wave(const XdTask *tasks, const uint32_t *list,
                             const uint32_t *count, uint32_t *queue,
                             const int8_t *mat, int gap, int xdrop,
                             int seed_len, XdOut *out, uint32_t *ovf_list,
                             uint32_t *ovf_count, int grid, int wide,
                             hipStream_t s)
{
    WaveArgs a{tasks, list, count, queue, mat, gap, xdrop, seed_len,
               out, ovf_list, ovf_count};
    if (wide)
        hipLaunchKernelGGL(k_xdrop_band<kWideChunks>, dim3(grid), dim3(64), 0,
                           s, a);
    else
        hipLaunchKernelGGL(k_xdrop_band<kBandChunks>, dim3(grid), dim3(64), 0,
                           s, a);
    return hipGetLastError();
}

int xdrop_wave_occupancy(int wide)
{
    static int cache[2] = {0, 0};
    if (!cache[wide ? 1 : 0]) {
        int n = 0;
        if (wide)
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, k_xdrop_band<kWideChunks>, 64, 0);
        else
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, k_xdrop_band<kBandChunks>, 64, 0);
        cache[wide ? 1 : 0] = n > 0 ? n : 1;
    }
    return cache[wide ? 1 : 0];
}

}  // namespace pa
