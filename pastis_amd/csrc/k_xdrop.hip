// k_xdrop.hip -- gapped X-drop seed extension (SeqAn's GappedXDrop on a
// simple seed, as called at src/align/SeqanXdropAligner.cpp:281-283), one
// work item per lane.
//
// A work item is one direction (left or right) of one seed of one pair; the
// two directions of a seed are independent and write disjoint fields of its
// XdOut.  The band update, the clipping at the segment ends and the "longest
// extension" end point follow the oracle's or_xdrop_one step for step, so
// extents and evaluated-cell counts are bit-identical.
//
// Storage of the three live anti-diagonals: one array of `slots` cells per
// anti-diagonal, column c of an anti-diagonal whose range starts at column
// off in slot c - off.  The algorithm only reads columns inside the range
// written on that anti-diagonal ([off, maxCol]), so the arrays are exact
// while that range fits; a wider range reports overflow.
//   - k_xdrop_pl<RING, T>: RING-slot arrays in LDS ([3][RING][64 lanes] of
//     int8 or int16 T, RING = 16 or 32).  A cell of anti-diagonal d is kept
//     only if it is at least lim_d = best - X (best before d), and no cell
//     exceeds best plus the largest matrix score (<= 31, as pa_create
//     requires), so an array holds v - lim_d - bias, in [-bias, X + 31 -
//     bias], and the type's minimum for an undefined cell; bias is 64 for
//     int8 (X <= 160, gap >= -32) and 0 for int16 (X <= 30000, gap >=
//     -1000).  Each anti-diagonal keeps its own lim_d + bias (its "frame");
//     the cell arithmetic runs in the frame of anti-diagonal d-1.  Items of
//     calls outside the type's X and gap bounds overflow at once to the
//     int32 kernels;
//   - k_xdrop_pl<0, int32_t>: int32 arrays in global scratch sized to the
//     longest sequence (the overflow path), frames all zero.
// Scheduling: each lane advances its own item one anti-diagonal per loop
// trip and takes the next item from the queue as soon as its item ends (one
// atomic per wave for all lanes that need an item), so a long extension
// occupies one lane, not a whole wave.
#include <limits.h>

#include <type_traits>

#include "pa_internal.h"

namespace pa {
namespace {

constexpr int32_t kUndef = INT_MIN / 4;

template <class T> struct Rings {
    T *base;
    int ring_stride;   // elements between the three anti-diagonals
    int slot_stride;   // elements between consecutive slots
    int slots;
    // stored value of an undefined cell
    static constexpr int32_t kSent =
        sizeof(T) == 1 ? INT8_MIN : sizeof(T) == 2 ? INT16_MIN : kUndef;
    // values are kept relative to their anti-diagonal's frame, lim_d + kBias
    static constexpr bool kRel = sizeof(T) < 4;
    static constexpr int32_t kBias = sizeof(T) == 1 ? 64 : 0;
    // the X-drop and gap values the encoding covers: defined cells in
    // [-kBias, X + 31 - kBias], and an undefined cell, after the shifts of
    // the cell loop, below every defined candidate and below the limit
    static constexpr int kXMax = sizeof(T) == 1 ? 160 : 30000;
    static constexpr int kGapMin = sizeof(T) == 1 ? -32 : -1000;
    static __host__ __device__ inline bool covers(int X, int gap)
    {
        return !kRel || (X >= 0 && X <= kXMax && gap >= kGapMin && gap <= 0);
    }
    // column col of anti-diagonal array r, whose range starts at off
    __device__ inline T &at(int r, int col, int off) const
    {
        return base[r * ring_stride + (col - off) * slot_stride];
    }
    // the value of a cell (kUndef if undefined) from its stored form
    __device__ inline int32_t get(int r, int col, int off,
                                  int32_t frame) const
    {
        const int32_t v = at(r, col, off);
        return v == kSent ? kUndef : v + frame;
    }
};

// Resumable state of one direction's extension.
struct Item {
    const uint8_t PA_GLOBAL *qs, *ds;   // query (V, columns), database (H, rows)
    int cols, rows;           // segment lengths + 1
    bool right;
    int d, minCol, maxCol, best;
    int dgap;                 // d * gap
    int off1, off2, off3, l1, l2, l3;
    int r1, r2, r3;           // arrays of anti-diagonals d-2, d-1, d
    int32_t f1, f2, f3;       // their frames (stored value = value - frame)
    uint32_t nc;              // evaluated cells
    // residue staging (LDS kernels): the request issued at the start of the
    // previous anti-diagonal is stored at the start of this one
    const uint8_t PA_GLOBAL *qb, *db;   // column c: qb[qstep * c], row r:
    int qstep, dstep;                   // db[dstep * r]
    int qhi, rhi;             // highest column / row requested
    int rq, rcol, rd, rrow;   // request in flight: residues, column, row
};

// Residues of column col (query) and row row (database) of the item.
__device__ inline int q_at(const Item &s, int col)
{
    return s.qb[s.qstep > 0 ? col : -col];   // qstep is +1 or -1
}
__device__ inline int d_at(const Item &s, int row)
{
    return s.db[s.dstep > 0 ? row : -row];
}

// Per-lane LDS rings of residues, as many slots as the cell arrays (RING),
// indexed by column and row.  An anti-diagonal that runs has maxCol - minCol
// <= RING - 2, minCol never decreases, maxCol grows by at most one per step
// and d - maxCol never decreases.  So the column stored at the start of step
// d (maxCol of step d-1, <= minCol(d) + RING - 2) never lands on the slot of
// a column >= minCol(d) still to be read, and the row stored then (d -
// minCol(d-1)) is less than RING above the lowest row read from then on
// (d - maxCol(d) + 1, with maxCol(d) - minCol(d-1) <= RING - 1).
struct ResRings {
    uint8_t *q, *d;   // slot stride 64 (one byte per lane)
    int mask;         // slots - 1
    __device__ inline uint8_t &qs(int col) const
    {
        return q[(col & mask) * 64];
    }
    __device__ inline uint8_t &ds(int row) const
    {
        return d[(row & mask) * 64];
    }
};

// Stores the request in flight into the rings, then requests column `col`
// and row `row` (one already requested, or outside the segments, is skipped:
// the load then reads column / row 1 again and is not stored).
__device__ inline void stage(Item &s, const ResRings &Rr, int col, int row)
{
    if (s.rcol)
        Rr.qs(s.rcol) = (uint8_t)s.rq;
    if (s.rrow)
        Rr.ds(s.rrow) = (uint8_t)s.rd;
    const bool qv = col > s.qhi && col < s.cols;
    const bool dv = row > s.rhi && row < s.rows;
    s.rcol = qv ? col : 0;
    s.rrow = dv ? row : 0;
    s.rq = q_at(s, qv ? col : 1);
    s.rd = d_at(s, dv ? row : 1);
    s.qhi = qv ? col : s.qhi;
    s.rhi = dv ? row : s.rhi;
}

template <bool STAGE, class T>
__device__ inline void item_init(Item &s, const Rings<T> &A,
                                 const ResRings &Rr, int gap, int X)
{
    s.r1 = 0;
    s.r2 = 1;
    s.r3 = 2;
    s.l1 = 0;
    s.l2 = 1;
    s.l3 = 2;
    // anti-diagonals 0 and 1 in the frame of the first limit (best 0):
    // frames never decrease
    const int32_t f0 = A.kRel ? -X + A.kBias : 0;
    s.f1 = s.f2 = s.f3 = f0;
    A.at(s.r2, 0, 0) = 0 - f0;
    A.at(s.r3, 0, 0) = -gap > X ? A.kSent : gap - f0;
    A.at(s.r3, 1, 0) = -gap > X ? A.kSent : gap - f0;
    s.minCol = 1;
    s.maxCol = 2;
    s.off1 = s.off2 = s.off3 = 0;
    s.d = 1;
    s.dgap = gap;
    s.best = 0;
    s.nc = 0;
    if (STAGE) {
        // anti-diagonal 2 reads column 1 and row 1
        Rr.qs(1) = (uint8_t)q_at(s, 1);
        Rr.ds(1) = (uint8_t)d_at(s, 1);
        s.qhi = 1;
        s.rhi = 1;
        s.rcol = s.rrow = 0;
    }
}

// One anti-diagonal.  Returns 0 while the band is not empty, 1 when the
// extension ended, 2 when the written range outgrew the arrays.
template <bool STAGE, class T>
__device__ inline int item_step(Item &s, const Rings<T> &A,
                                const ResRings &Rr, const int8_t *mat,
                                int gap, int X)
{
    // anti-diagonal d+1 (the next one) reads columns < maxCol(d+1) <=
    // maxCol + 1 and rows <= d + 1 - minCol(d+1) <= d + 1 - minCol
    if (STAGE)
        stage(s, Rr, s.maxCol, s.d + 2 - s.minCol);
    const int d = ++s.d;
    const int t = s.r1;
    s.r1 = s.r2;
    s.r2 = s.r3;
    s.r3 = t;
    s.l1 = s.l2;
    s.l2 = s.l3;
    s.off1 = s.off2;
    s.off2 = s.off3;
    s.off3 = s.minCol - 1;
    s.l3 = s.maxCol + 1 - s.off3;
    const int32_t lim = s.best - X;
    s.f1 = s.f2;
    s.f2 = s.f3;
    s.f3 = A.kRel ? lim + A.kBias : 0;
    if (s.l3 > A.slots)
        return 2;
    const int r1 = s.r1, r2 = s.r2, r3 = s.r3;
    const int off3 = s.off3, maxCol0 = s.maxCol;
    // _initAntiDiag3: edge entries, first column / first row cells
    A.at(r3, off3, off3) = A.kSent;
    A.at(r3, maxCol0, off3) = A.kSent;
    // d * gap as a running sum; with narrow cells (|gap| <= 1000, X <=
    // 30000) it cannot overflow int32, else the comparison is made in 64 bits
    const int32_t dgap = (s.dgap += gap);
    if (A.kRel ? dgap > s.best - X : (long)d * gap > (long)s.best - X) {
        if (off3 == 0)
            A.at(r3, off3, off3) = dgap - s.f3;
        if (d - maxCol0 == 0)
            A.at(r3, maxCol0, off3) = dgap - s.f3;
    }
    const int cols = s.cols, rows = s.rows;
    // the cells are computed in the frame of anti-diagonal d-1 (r2): c12
    // moves an r1 value into it, lim2 is the limit and the frame of r3
    // there.  An undefined stored value (kSent) stays below every defined
    // candidate and below the limit after these shifts.
    const int32_t c12 = s.f1 - s.f2;
    const int32_t lim2 = lim - s.f2;
    const int32_t st2 = s.f3 - s.f2;
    int32_t adbest = dgap - s.f2;
    // the up value of one cell is the left value of the next (r2 is not
    // written in this loop)
    const int off1 = s.off1, off2 = s.off2;
    int32_t l = A.at(r2, s.minCol - 1, off2);
    // cell pointers of column minCol (r1: minCol - 1), stepped per column
    const T *p1 = &A.at(r1, s.minCol - 1, off1);
    const T *p2 = &A.at(r2, s.minCol, off2);
    T *p3 = &A.at(r3, s.minCol, off3);
    const int ss = A.slot_stride;
    for (int col = s.minCol; col < maxCol0; ++col) {
        int qp, dp;
        if (s.right) {
            qp = col - 1;
            dp = d - col - 1;
        } else {
            qp = cols - 1 - col;
            dp = rows - 1 + col - d;
        }
        const int32_t u = *p2;
        int32_t tmp = (l > u ? l : u) + gap;
        const int qr = STAGE ? (int)Rr.qs(col) : (int)s.qs[qp];
        const int dr = STAGE ? (int)Rr.ds(d - col) : (int)s.ds[dp];
        const int32_t dg = *p1 + c12 + mat[qr * 32 + dr];
        tmp = dg > tmp ? dg : tmp;
        // kept (>= the limit) or undefined, as selects: no branch per cell
        const bool keep = tmp >= lim2;
        *p3 = keep ? tmp - st2 : A.kSent;
        p1 += ss;
        p2 += ss;
        p3 += ss;
        const int32_t kv = keep ? tmp : adbest;
        adbest = kv > adbest ? kv : adbest;
        l = u;
    }
    adbest += s.f2;
    s.nc += (uint32_t)(maxCol0 - s.minCol);
    s.best = adbest > s.best ? adbest : s.best;
    int minCol = s.minCol, maxCol = maxCol0;
    while (minCol - off3 < s.l3 && A.at(r3, minCol, off3) == A.kSent &&
           minCol - off2 - 1 < s.l2 &&
           A.at(r2, minCol - 1, off2) == A.kSent)
        ++minCol;
    while (maxCol - off3 > 0 && A.at(r3, maxCol - 1, off3) == A.kSent &&
           A.at(r2, maxCol - 1, off2) == A.kSent)
        --maxCol;
    ++maxCol;
    if (minCol < d + 2 - rows)
        minCol = d + 2 - rows;
    if (maxCol > cols)
        maxCol = cols;
    s.minCol = minCol;
    s.maxCol = maxCol;
    return minCol < maxCol ? 0 : 1;
}

// "Longest extension" end point: both ends, end of query, end of database,
// else the best cell of the last live anti-diagonal.
template <class T>
__device__ inline void item_end(const Item &s, const Rings<T> &A,
                                int32_t &score, int &col, int &row, bool &upd)
{
    const int d = s.d;
    int lcol = s.l3 + s.off3 - 2;
    int lrow = d - lcol;
    int32_t ls = A.get(s.r3, lcol, s.off3, s.f3);
    if (ls == kUndef) {
        if (A.get(s.r2, s.l2 + s.off2 - 2, s.off2, s.f2) != kUndef) {
            lcol = s.l2 + s.off2 - 2;
            lrow = d - 1 - lcol;
            ls = A.get(s.r2, lcol, s.off2, s.f2);
        } else if (s.l2 > 2 &&
                   A.get(s.r2, s.l2 + s.off2 - 3, s.off2, s.f2) != kUndef) {
            lcol = s.l2 + s.off2 - 3;
            lrow = d - 1 - lcol;
            ls = A.get(s.r2, lcol, s.off2, s.f2);
        }
    }
    if (ls == kUndef) {
        for (int i = 0; i < s.l1; ++i) {
            const int32_t v = A.get(s.r1, s.off1 + i, s.off1, s.f1);
            if (v > ls) {
                ls = v;
                lcol = i + s.off1;
                lrow = d - 2 - lcol;
            }
        }
    }
    score = ls;
    upd = ls != kUndef;
    col = lcol;
    row = lrow;
}

struct XdArgs {
    const XdTask *tasks;
    const uint32_t *list;   // work items (k = 2 * task + dir), NULL = 0..count
    const uint32_t *count;
    uint32_t *queue;
    const int8_t *mat;
    int gap, X, seed_len;
    int32_t *scratch;       // global-scratch arrays (RING == 0)
    int cap;                // their slots (power of two)
    XdOut *out;
    uint32_t *ovf_list, *ovf_count;
};

// Writes the half of XdOut that item k (direction k & 1) owns.
__device__ inline void item_write(const XdArgs &a, uint32_t k, int32_t score,
                                  int col, int row, bool upd, uint32_t nc,
                                  uint16_t flags)
{
    const XdTask T = a.tasks[k >> 1];
    XdOut PA_GLOBAL &o = glb(a.out)[T.id];
    if (T.pad & kXdInvalid) {   // empty extents [0, 0) on both sequences
        if ((k & 1) == 0) {
            o.beg_h = o.beg_v = 0;
            o.score_left = 0;
            o.cells_left = 0;
            o.flags_left = 0;
        } else {
            o.end_h = o.end_v = 0;
            o.score_right = 0;
            o.cells_right = 0;
            o.flags_right = 0;
        }
        return;
    }
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

enum : int { kIdle = 0, kRun = 1, kExhausted = 2, kEnd = 3 };
constexpr uint32_t kChunk = 64;   // items claimed per queue atomic
// Lanes whose item ended wait (kEnd) until this many have, or until no lane
// of the wave still runs: the end-point search, the result write and the
// claims then run once for all of them instead of in most steps for one.
// With 24 waves per CU, 8 measured best (4: as fast, 16: 1.5 % and 32: 8 %
// slower on the X-drop workload).
constexpr int kEndBatch = 8;

// the LDS variants: at most 80 VGPRs, six waves per SIMD (their LDS, 6 KB
// per wave with int8 cells, allows more)
template <int RING, class Cell>
__global__ __launch_bounds__(RING ? 64 : 256)
PA_WAVES_PER_EU(RING ? 6 : 1) void k_xdrop_pl(XdArgs a)
{
    static_assert((RING > 0) == (sizeof(Cell) < 4), "LDS cells are narrow");
    __shared__ int8_t mat[32 * 32];
    __shared__ Cell ring[RING ? 3 * RING * 64 : 1];
    __shared__ uint8_t res[RING ? 2 * RING * 64 : 1];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x)
        mat[i] = a.mat[i];
    __syncthreads();
    const uint32_t n = *a.count;
    const int lane = threadIdx.x & 63;
    Rings<Cell> A;
    if (RING) {
        A = Rings<Cell>{ring + lane, RING * 64, 64, RING};
    } else {
        // (Cell is int32_t here)
        Cell *buf = (Cell *)a.scratch + (size_t)(blockIdx.x * blockDim.x +
                                           threadIdx.x) *
                                      (size_t)(3 * a.cap);
        A = Rings<Cell>{buf, a.cap, 1, a.cap};
    }
    const ResRings Rr{res + lane, res + RING * 64 + lane, RING - 1};
    constexpr bool ST = RING > 0;
    const int gap = a.gap, X = a.X;
    // parameters the narrow cell encoding does not cover (the host picks
    // the variant that does): every item overflows to the int32 kernels
    const bool fits = Rings<Cell>::covers(X, gap);
    int state = kIdle;
    uint32_t k = 0;
    Item s{};
    uint32_t cbase = 0, cend = 0;   // claimed, not yet handed out
    for (;;) {
        // lanes without an item take the next ones from the wave's claimed
        // range, which is refilled kChunk items per queue atomic; every lane
        // runs this block (wave-uniform control flow)
        uint64_t need = __ballot(state == kIdle);
        while (need) {
            if (cbase >= cend) {
                uint32_t b = atomicAdd(a.queue, lane == 0 ? kChunk : 0u);
                b = __builtin_amdgcn_readfirstlane(b);
                cbase = b;
                cend = b + kChunk;
            }
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t take = min(cnt, cend - cbase);
            const uint32_t rank =
                (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
            if (((need >> lane) & 1ull) && rank < take) {
                const uint32_t idx = cbase + rank;
                if (idx >= n) {
                    state = kExhausted;
                } else {
                    k = a.list ? glb(a.list)[idx] : idx;
                    const XdTask T = a.tasks[k >> 1];
                    const uint8_t PA_GLOBAL *h = glb_addr<const uint8_t>(T.h_off);
                    const uint8_t PA_GLOBAL *v = glb_addr<const uint8_t>(T.v_off);
                    const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
                    s.right = (k & 1) != 0;
                    if (!s.right) {   // query V[0, beg_v), database H[0, beg_h)
                        s.qs = v;
                        s.cols = beg_v + 1;
                        s.ds = h;
                        s.rows = beg_h + 1;
                        // reversed: column c is qs[cols-1-c], row r ds[rows-1-r]
                        s.qb = s.qs + (s.cols - 1);
                        s.qstep = -1;
                        s.db = s.ds + (s.rows - 1);
                        s.dstep = -1;
                    } else {          // V[end_v, nv), H[end_h, nh)
                        s.qs = v + beg_v + a.seed_len;
                        s.cols = (int)T.nv - beg_v - a.seed_len + 1;
                        s.ds = h + beg_h + a.seed_len;
                        s.rows = (int)T.nh - beg_h - a.seed_len + 1;
                        s.qb = s.qs - 1;
                        s.qstep = 1;
                        s.db = s.ds - 1;
                        s.dstep = 1;
                    }
                    if ((T.pad & kXdInvalid) || s.rows <= 1 || s.cols <= 1) {
                        item_write(a, k, 0, 0, 0, false, 0, 0);
                    } else {
                        item_init<ST>(s, A, Rr, gap, X);
                        state = kRun;
                    }
                }
            }
            cbase += take;
            // the lanes served leave `need`; a lane whose item was trivial
            // asks again on the next trip
            for (uint32_t t = 0; t < take; ++t)
                need &= need - 1;
        }
        if (!__ballot(state != kExhausted))
            break;
        if (state == kRun) {
            const int st = fits ? item_step<ST>(s, A, Rr, mat, gap, X) : 2;
            if (st == 1) {
                state = kEnd;
            } else if (st == 2) {
                const uint32_t pos = atomicAdd(a.ovf_count, 1u);
                a.ovf_list[pos] = k;
                item_write(a, k, 0, 0, 0, false, 0, 1);
                state = kIdle;
            }
        }
        const uint64_t ending = __ballot(state == kEnd);
        if (ending && (__popcll(ending) >= kEndBatch ||
                       !__ballot(state == kRun))) {
            if (state == kEnd) {
                // the arrays still hold the item's last anti-diagonals
                int32_t score;
                int col, row;
                bool upd;
                item_end(s, A, score, col, row, upd);
                item_write(a, k, score, col, row, upd, s.nc, 0);
                state = kIdle;
            }
        }
    }
}

}  // namespace

// The LDS-ring variant for `ring` slots: int8 cells where the call's X-drop
// value and gap allow them, else int16.
template <class F> static auto pick_pl(int ring, int xdrop, int gap, F &&f)
{
    const bool narrow = Rings<int8_t>::covers(xdrop, gap);
    if (ring == 16)
        return narrow ? f(k_xdrop_pl<16, int8_t>) : f(k_xdrop_pl<16, int16_t>);
    return narrow ? f(k_xdrop_pl<32, int8_t>) : f(k_xdrop_pl<32, int16_t>);
}

hipError_t launch_xdrop(const XdTask *tasks, const uint32_t *list,
                        const uint32_t *count, uint32_t *queue,
                        const int8_t *mat, int gap, int xdrop, int seed_len,
                        int32_t *scratch, int cap, XdOut *out,
                        uint32_t *ovf_list, uint32_t *ovf_count, int grid,
                        int ring, hipStream_t s)
{
    XdArgs a{tasks, list, count, queue, mat, gap, xdrop, seed_len,
             scratch, cap, out, ovf_list, ovf_count};
    if (ring == 0) {
        hipLaunchKernelGGL((k_xdrop_pl<0, int32_t>), dim3(grid), dim3(256), 0,
                           s, a);
    } else if (ring == 16 || ring == 32) {
        pick_pl(ring, xdrop, gap, [&](auto k) {
            hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, a);
            return 0;
        });
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int xdrop_occupancy(int ring, int xdrop, int gap)
{
    // per variant: ring 16 / 32, int8 / int16
    static int cache[4] = {0, 0, 0, 0};
    const int i = (ring == 16 ? 0 : 2) +
                  (Rings<int8_t>::covers(xdrop, gap) ? 0 : 1);
    if (!cache[i]) {
        const int n = pick_pl(ring, xdrop, gap, [](auto k) {
            int m = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&m, k, 64, 0) !=
                hipSuccess)
                m = 0;
            return m;
        });
        cache[i] = n > 0 ? n : 1;
    }
    return cache[i];
}

}  // namespace pa
