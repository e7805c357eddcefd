// k_xdrop.hip -- gapped X-drop seed extension (SeqAn's GappedXDrop on a
// simple seed, as called at src/align/SeqanXdropAligner.cpp:281-283).
//
// A work item is one direction (left or right) of one seed of one pair; the
// two directions of a seed are independent and write disjoint fields of its
// XdOut.  One work item per lane.  The band update, the clipping at the
// segment ends and the "longest extension" end point follow the oracle's
// or_xdrop_one step for step, so extents and evaluated-cell counts are
// bit-identical.
//
// Storage of the three live anti-diagonals: a ring per anti-diagonal indexed
// by absolute column (slot = column mod ring size).  The algorithm only reads
// columns inside the range written on that anti-diagonal ([off, maxCol]), so
// the ring is exact while that range fits; a wider range reports overflow.
//   - k_xdrop_lds<RING>: RING-slot rings in LDS ([3][RING][64 lanes] int32,
//     RING = 16 or 32: fewer slots leave room for more waves per CU);
//   - k_xdrop_wide: rings in global scratch sized to the longest sequence
//     (the overflow path).
// Scheduling: a first pass runs with an anti-diagonal budget; items still
// alive when it runs out are recomputed from scratch by a second pass, so
// each pass keeps lanes of a wave on similar amounts of work.
#include <limits.h>

#include "pa_internal.h"

namespace pa {
namespace {

constexpr int32_t kUndef = INT_MIN / 4;

enum : uint16_t { kXdOvf = 1, kXdLong = 2 };

struct Rings {
    int32_t *base;
    int ring_stride;   // elements between the three rings
    int slot_stride;   // elements between consecutive slots
    int mask;          // slots - 1
    int slots;
    __device__ inline int32_t &at(int r, int col) const
    {
        return base[r * ring_stride + (col & mask) * slot_stride];
    }
};

struct Dir {
    int32_t score;
    int32_t col, row;
    bool upd;
    uint16_t status;   // 0 done, kXdOvf, kXdLong
};

__device__ Dir xd_dir(const uint8_t *qs, int qlen, const uint8_t *ds, int dlen,
                      bool right, const int8_t *mat, int gap, int X,
                      const Rings &A, int budget, uint32_t *cells)
{
    Dir res = {0, 0, 0, false, 0};
    const int cols = qlen + 1, rows = dlen + 1;
    if (rows == 1 || cols == 1)
        return res;
    int r1 = 0, r2 = 1, r3 = 2;   // rings of anti-diagonals d-2, d-1, d
    int l1 = 0, l2 = 1, l3 = 2;
    A.at(r2, 0) = 0;
    A.at(r3, 0) = -gap > X ? kUndef : gap;
    A.at(r3, 1) = -gap > X ? kUndef : gap;
    int minCol = 1, maxCol = 2;
    int off1 = 0, off2 = 0, off3 = 0;
    int d = 1;
    int32_t best = 0;
    uint32_t nc = 0;
    while (minCol < maxCol) {
        if (budget > 0 && d > budget) {
            res.status = kXdLong;
            return res;
        }
        ++d;
        const int t = r1;
        r1 = r2;
        r2 = r3;
        r3 = t;
        l1 = l2;
        l2 = l3;
        off1 = off2;
        off2 = off3;
        off3 = minCol - 1;
        l3 = maxCol + 1 - off3;
        if (l3 > A.slots) {
            res.status = kXdOvf;
            return res;
        }
        // _initAntiDiag3: edge entries, first column / first row cells
        A.at(r3, off3) = kUndef;
        A.at(r3, maxCol) = kUndef;
        if ((long)d * gap > (long)best - X) {
            if (off3 == 0)
                A.at(r3, off3) = d * gap;
            if (d - maxCol == 0)
                A.at(r3, maxCol) = d * gap;
        }
        int32_t adbest = d * gap;
        const int32_t lim = best - X;
        for (int col = minCol; col < maxCol; ++col) {
            int qp, dp;
            if (right) {
                qp = col - 1;
                dp = d - col - 1;
            } else {
                qp = cols - 1 - col;
                dp = rows - 1 + col - d;
            }
            const int32_t l = A.at(r2, col - 1), u = A.at(r2, col);
            int32_t tmp = (l > u ? l : u) + gap;
            const int32_t dg =
                A.at(r1, col - 1) + mat[(int)qs[qp] * 32 + ds[dp]];
            tmp = dg > tmp ? dg : tmp;
            if (tmp < lim) {
                A.at(r3, col) = kUndef;
            } else {
                A.at(r3, col) = tmp;
                adbest = tmp > adbest ? tmp : adbest;
            }
        }
        nc += (uint32_t)(maxCol - minCol);
        best = adbest > best ? adbest : best;
        while (minCol - off3 < l3 && A.at(r3, minCol) == kUndef &&
               minCol - off2 - 1 < l2 && A.at(r2, minCol - 1) == kUndef)
            ++minCol;
        while (maxCol - off3 > 0 && A.at(r3, maxCol - 1) == kUndef &&
               A.at(r2, maxCol - 1) == kUndef)
            --maxCol;
        ++maxCol;
        if (minCol < d + 2 - rows)
            minCol = d + 2 - rows;
        if (maxCol > cols)
            maxCol = cols;
    }
    *cells += nc;
    // longest extension: both ends, end of query, end of database, else
    // the best cell of the last live anti-diagonal
    int lcol = l3 + off3 - 2;
    int lrow = d - lcol;
    int32_t ls = A.at(r3, lcol);
    if (ls == kUndef) {
        if (A.at(r2, l2 + off2 - 2) != kUndef) {
            lcol = l2 + off2 - 2;
            lrow = d - 1 - lcol;
            ls = A.at(r2, lcol);
        } else if (l2 > 2 && A.at(r2, l2 + off2 - 3) != kUndef) {
            lcol = l2 + off2 - 3;
            lrow = d - 1 - lcol;
            ls = A.at(r2, lcol);
        }
    }
    if (ls == kUndef) {
        for (int i = 0; i < l1; ++i) {
            const int32_t v = A.at(r1, off1 + i);
            if (v > ls) {
                ls = v;
                lcol = i + off1;
                lrow = d - 2 - lcol;
            }
        }
    }
    res.score = ls;
    if (ls != kUndef) {
        res.upd = true;
        res.col = lcol;
        res.row = lrow;
    }
    return res;
}

// One work item: direction `dir` of extension task T.
__device__ inline void run_item(const XdTask &T, int dir, const int8_t *mat,
                                int gap, int X, int seed_len, const Rings &A,
                                int budget, uint32_t k, XdOut *out,
                                uint32_t *ovf_list, uint32_t *ovf_count,
                                uint32_t *long_list, uint32_t *long_count)
{
    const uint8_t *h = (const uint8_t *)(uintptr_t)T.h_off;
    const uint8_t *v = (const uint8_t *)(uintptr_t)T.v_off;
    const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
    const int end_h = beg_h + seed_len, end_v = beg_v + seed_len;
    uint32_t cells = 0;
    Dir r;
    if (dir == 0)  // left: query = V[0, beg_v), database = H[0, beg_h)
        r = xd_dir(v, beg_v, h, beg_h, false, mat, gap, X, A, budget, &cells);
    else           // right: query = V[end_v, nv), database = H[end_h, nh)
        r = xd_dir(v + end_v, (int)T.nv - end_v, h + end_h, (int)T.nh - end_h,
                   true, mat, gap, X, A, budget, &cells);
    XdOut &o = out[T.id];
    if (r.status) {
        uint32_t *lst = r.status == kXdOvf ? ovf_list : long_list;
        uint32_t *cnt = r.status == kXdOvf ? ovf_count : long_count;
        const uint32_t pos = atomicAdd(cnt, 1u);  // no wave op follows
        lst[pos] = k;
    }
    if (dir == 0) {
        o.beg_h = r.upd ? beg_h - r.row : beg_h;
        o.beg_v = r.upd ? beg_v - r.col : beg_v;
        o.score_left = r.score;
        o.cells_left = cells;
        o.flags_left = r.status;
    } else {
        o.end_h = r.upd ? end_h + r.row : end_h;
        o.end_v = r.upd ? end_v + r.col : end_v;
        o.score_right = r.score;
        o.cells_right = cells;
        o.flags_right = r.status;
    }
}

struct XdArgs {
    const XdTask *tasks;
    const uint32_t *list;   // work items (k = 2 * task + dir), NULL = 0..count
    const uint32_t *count;
    uint32_t *queue;
    const int8_t *mat;
    int gap, X, seed_len, budget;
    int32_t *scratch;       // wide kernel only
    int cap;                // wide kernel: ring slots (power of two)
    XdOut *out;
    uint32_t *ovf_list, *ovf_count, *long_list, *long_count;
};

template <int RING>
__global__ __launch_bounds__(64) void k_xdrop_lds(XdArgs a)
{
    __shared__ int8_t mat[32 * 32];
    __shared__ int32_t ring[3 * RING * 64];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x)
        mat[i] = a.mat[i];
    __syncthreads();
    const uint32_t n = *a.count;
    const int lane = threadIdx.x;
    const Rings A{ring + lane, RING * 64, 64, RING - 1, RING};
    for (;;) {
        // all lanes join the atomic (lane 0 adds 64) and no divergent
        // continue follows, so the wave re-enters the loop together
        uint32_t base = atomicAdd(a.queue, lane == 0 ? 64u : 0u);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= n)
            break;
        const uint32_t j = base + (uint32_t)lane;
        if (j < n) {
            const uint32_t k = a.list ? a.list[j] : j;
            const XdTask T = a.tasks[k >> 1];
            run_item(T, (int)(k & 1), mat, a.gap, a.X, a.seed_len, A, a.budget,
                     k, a.out, a.ovf_list, a.ovf_count, a.long_list,
                     a.long_count);
        }
    }
}

__global__ __launch_bounds__(256) void k_xdrop_wide(XdArgs a)
{
    __shared__ int8_t mat[32 * 32];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x)
        mat[i] = a.mat[i];
    __syncthreads();
    const uint32_t n = *a.count;
    const int lane = threadIdx.x & 63;
    int32_t *buf = a.scratch + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) *
                                   (size_t)(3 * a.cap);
    const Rings A{buf, a.cap, 1, a.cap - 1, a.cap};
    for (;;) {
        uint32_t base = atomicAdd(a.queue, lane == 0 ? 64u : 0u);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= n)
            break;
        const uint32_t j = base + (uint32_t)lane;
        if (j < n) {
            const uint32_t k = a.list ? a.list[j] : j;
            const XdTask T = a.tasks[k >> 1];
            run_item(T, (int)(k & 1), mat, a.gap, a.X, a.seed_len, A, 0, k,
                     a.out, a.ovf_list, a.ovf_count, a.long_list,
                     a.long_count);
        }
    }
}

}  // namespace

hipError_t launch_xdrop(const XdTask *tasks, const uint32_t *list,
                        const uint32_t *count, uint32_t *queue,
                        const int8_t *mat, int gap, int xdrop, int seed_len,
                        int budget, int32_t *scratch, int cap, XdOut *out,
                        uint32_t *ovf_list, uint32_t *ovf_count,
                        uint32_t *long_list, uint32_t *long_count, int grid,
                        bool wide, int ring, hipStream_t s)
{
    XdArgs a{tasks, list, count, queue, mat, gap, xdrop, seed_len, budget,
             scratch, cap, out, ovf_list, ovf_count, long_list, long_count};
    if (wide)
        hipLaunchKernelGGL(k_xdrop_wide, dim3(grid), dim3(256), 0, s, a);
    else if (ring == 16)
        hipLaunchKernelGGL(k_xdrop_lds<16>, dim3(grid), dim3(64), 0, s, a);
    else if (ring == 32)
        hipLaunchKernelGGL(k_xdrop_lds<32>, dim3(grid), dim3(64), 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

int xdrop_occupancy(int ring)
{
    static int cache[2] = {0, 0};
    const int i = ring == 16 ? 0 : 1;
    if (!cache[i]) {
        int n = 0;
        if (ring == 16)
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_xdrop_lds<16>,
                                                         64, 0);
        else
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_xdrop_lds<32>,
                                                         64, 0);
        cache[i] = n > 0 ? n : 1;
    }
    return cache[i];
}

}  // namespace pa
